/* CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * Sequential restatement of the library's geometric multigrid solve (the
 * added `gmg` solver; the reference has only a V-cycle skeleton whose coarse
 * correction never reaches p, src/solver-mg.c:299-354). Same algorithm, same
 * operand order, so the GPU path can be checked pointwise:
 *   - operator: 7-point Laplacian with Neumann ghost copy
 *     (reference src/solver-rb.c:66-69, 129-148), h_l = 2^l h_0;
 *   - smoother: red-black SOR, odd cells (i+j+k odd) first, face ghosts after
 *     the second colour;
 *   - restriction: (1/8) sum of the children's residuals;
 *   - prolongation: cell-centred trilinear with clamped coarse indices;
 *   - coarsest level: exact solve in the DCT-II eigenbasis of the Neumann
 *     Laplacian, constant mode dropped;
 *   - FP64 stop test ||rhs - A p|| / ||rhs|| <= tol after each cycle;
 *   - mixed: FP32 V-cycle on the FP64 residual, p += e (defect correction);
 *   - grouped (orcGmgSolveGrouped): the first test of a solve comes after
 *     `firstGroup` cycles (mixed: that many FP32 cycles on one defect, the
 *     later ones continuing from e), later tests after every cycle.
 * Pins (tests/test_oracle_gmg.py): convergence to the manufactured
 * eigenfunction solution, V-cycle contraction factor bounds, exactness of
 * the restriction/prolongation on constants and linear fields, the coarse
 * DCT solve against a dense least-squares solve, the FP32 path agreeing
 * with the FP64 path to the tolerance, and the grouped loop: a group of one
 * is the plain loop bit for bit, an FP64 group of m cycles gives the plain
 * loop's m-th iterate bit for bit (the stop test does not touch p), and a
 * mixed group converges at the plain loop's rate.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXLEV 12
#define COARSE_MAX_CELLS 2048

typedef struct {
  int ni, nj, nk;
  size_t sj, sk, n;
} G3;

static G3 g3(int ni, int nj, int nk)
{
  G3 g;
  g.ni = ni;
  g.nj = nj;
  g.nk = nk;
  g.sj = (size_t)ni + 2;
  g.sk = g.sj * ((size_t)nj + 2);
  g.n  = g.sk * ((size_t)nk + 2);
  return g;
}

static inline size_t IX(const G3 *g, int i, int j, int k)
{
  return (size_t)k * g->sk + (size_t)j * g->sj + (size_t)i;
}

static void coefs(double dx, double dy, double dz, double omega, double out[4])
{
  double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  out[0] = 1.0 / dx2;
  out[1] = 1.0 / dy2;
  out[2] = 1.0 / dz2;
  out[3] = omega * 0.5 * (dx2 * dy2 * dz2) / (dy2 * dz2 + dx2 * dz2 + dx2 * dy2);
}

/* Level operations for one precision T (instantiated for double and float). */
#define DEFINE_OPS(T, S)                                                                     \
  typedef struct {                                                                           \
    G3 g;                                                                                    \
    T *x, *b;                                                                                \
    T cf[4];                                                                                 \
  } Lev##S;                                                                                  \
  static inline T lap##S(const T *x, size_t c, const G3 *g, const T *cf)                     \
  {                                                                                          \
    T xc = x[c];                                                                             \
    return (x[c + 1] - (T)2 * xc + x[c - 1]) * cf[0] +                                       \
           (x[c + g->sj] - (T)2 * xc + x[c - g->sj]) * cf[1] +                               \
           (x[c + g->sk] - (T)2 * xc + x[c - g->sk]) * cf[2];                                \
  }                                                                                          \
  static void ghosts##S(T *x, const G3 *g)                                                   \
  {                                                                                          \
    for (int j = 1; j <= g->nj; j++)                                                         \
      for (int i = 1; i <= g->ni; i++) {                                                     \
        x[IX(g, i, j, 0)]         = x[IX(g, i, j, 1)];                                       \
        x[IX(g, i, j, g->nk + 1)] = x[IX(g, i, j, g->nk)];                                   \
      }                                                                                      \
    for (int k = 1; k <= g->nk; k++)                                                         \
      for (int i = 1; i <= g->ni; i++) {                                                     \
        x[IX(g, i, 0, k)]         = x[IX(g, i, 1, k)];                                       \
        x[IX(g, i, g->nj + 1, k)] = x[IX(g, i, g->nj, k)];                                   \
      }                                                                                      \
    for (int k = 1; k <= g->nk; k++)                                                         \
      for (int j = 1; j <= g->nj; j++) {                                                     \
        x[IX(g, 0, j, k)]         = x[IX(g, 1, j, k)];                                       \
        x[IX(g, g->ni + 1, j, k)] = x[IX(g, g->ni, j, k)];                                   \
      }                                                                                      \
  }                                                                                          \
  static void sweep##S(T *x, const T *b, const G3 *g, const T *cf)                           \
  {                                                                                          \
    for (int parity = 1; parity >= 0; parity--)                                              \
      for (int k = 1; k <= g->nk; k++)                                                       \
        for (int j = 1; j <= g->nj; j++)                                                     \
          for (int i = ((1 + j + k) % 2 == parity) ? 1 : 2; i <= g->ni; i += 2) {            \
            size_t c = IX(g, i, j, k);                                                       \
            T r      = b[c] - lap##S(x, c, g, cf);                                           \
            x[c]     = x[c] - cf[3] * r;                                                     \
          }                                                                                  \
    ghosts##S(x, g);                                                                         \
  }                                                                                          \
  static void restrict##S(const Lev##S *f, Lev##S *c)                                        \
  {                                                                                          \
    memset(c->x, 0, c->g.n * sizeof(T));                                                     \
    for (int K = 1; K <= c->g.nk; K++)                                                       \
      for (int J = 1; J <= c->g.nj; J++)                                                     \
        for (int I = 1; I <= c->g.ni; I++) {                                                 \
          T s = 0;                                                                           \
          for (int dk = 0; dk < 2; dk++)                                                     \
            for (int dj = 0; dj < 2; dj++)                                                   \
              for (int di = 0; di < 2; di++) {                                               \
                size_t q = IX(&f->g, 2 * I - 1 + di, 2 * J - 1 + dj, 2 * K - 1 + dk);        \
                s += f->b[q] - lap##S(f->x, q, &f->g, f->cf);                                \
              }                                                                              \
          c->b[IX(&c->g, I, J, K)] = (T)0.125 * s;                                           \
        }                                                                                    \
  }                                                                                          \
  static int clampi##S(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }    \
  static void prolong##S(Lev##S *f, const Lev##S *c)                                         \
  {                                                                                          \
    const G3 *gc = &c->g;                                                                    \
    const T a = (T)0.75, q = (T)0.25;                                                        \
    for (int k = 1; k <= f->g.nk; k++)                                                       \
      for (int j = 1; j <= f->g.nj; j++)                                                     \
        for (int i = 1; i <= f->g.ni; i++) {                                                 \
          int I = (i + 1) / 2, J = (j + 1) / 2, K = (k + 1) / 2;                             \
          int I2 = clampi##S(I + ((i & 1) ? -1 : 1), 1, gc->ni);                             \
          int J2 = clampi##S(J + ((j & 1) ? -1 : 1), 1, gc->nj);                             \
          int K2 = clampi##S(K + ((k & 1) ? -1 : 1), 1, gc->nk);                             \
          const T *xc = c->x;                                                                \
          T c11 = a * xc[IX(gc, I, J, K)] + q * xc[IX(gc, I2, J, K)];                        \
          T c21 = a * xc[IX(gc, I, J2, K)] + q * xc[IX(gc, I2, J2, K)];                      \
          T c12 = a * xc[IX(gc, I, J, K2)] + q * xc[IX(gc, I2, J, K2)];                      \
          T c22 = a * xc[IX(gc, I, J2, K2)] + q * xc[IX(gc, I2, J2, K2)];                    \
          T e   = a * (a * c11 + q * c21) + q * (a * c12 + q * c22);                         \
          size_t x = IX(&f->g, i, j, k);                                                     \
          f->x[x] = f->x[x] + e;                                                             \
        }                                                                                    \
    ghosts##S(f->x, &f->g);                                                                  \
  }                                                                                          \
  static void coarse##S(Lev##S *L, const T *tab, const int *o)                             \
  {                                                                                          \
    const G3 *g = &L->g;                                                                     \
    int ni = g->ni, nj = g->nj, nk = g->nk, N = ni * nj * nk;                                \
    T *W1 = malloc((size_t)N * sizeof(T)), *W2 = malloc((size_t)N * sizeof(T));              \
    const T *Ci = tab + o[0], *Cj = tab + o[1], *Ck = tab + o[2];                            \
    const T *Di = tab + o[3], *Dj = tab + o[4], *Dk = tab + o[5];                            \
    const T *Li = tab + o[6], *Lj = tab + o[7], *Lk = tab + o[8];                            \
    for (int q = 0; q < N; q++) {                                                            \
      int i = q % ni, j = (q / ni) % nj, k = q / (ni * nj);                                  \
      W1[q] = L->b[IX(g, i + 1, j + 1, k + 1)];                                              \
    }                                                                                        \
    for (int q = 0; q < N; q++) {                                                            \
      int m = q % ni, row = q - m;                                                           \
      T s = 0;                                                                               \
      for (int i = 0; i < ni; i++) s += Ci[m * ni + i] * W1[row + i];                        \
      W2[q] = s;                                                                             \
    }                                                                                        \
    for (int q = 0; q < N; q++) {                                                            \
      int mi = q % ni, m = (q / ni) % nj, k = q / (ni * nj);                                 \
      T s = 0;                                                                               \
      for (int j = 0; j < nj; j++) s += Cj[m * nj + j] * W2[(k * nj + j) * ni + mi];         \
      W1[q] = s;                                                                             \
    }                                                                                        \
    for (int q = 0; q < N; q++) {                                                            \
      int mi = q % ni, mj = (q / ni) % nj, m = q / (ni * nj);                                \
      T s = 0;                                                                               \
      for (int k = 0; k < nk; k++) s += Ck[m * nk + k] * W1[(k * nj + mj) * ni + mi];        \
      T lam = Li[mi] + Lj[mj] + Lk[m];                                                       \
      W2[q] = (q == 0) ? (T)0 : s / lam;                                                     \
    }                                                                                        \
    for (int q = 0; q < N; q++) {                                                            \
      int mi = q % ni, mj = (q / ni) % nj, k = q / (ni * nj);                                \
      T s = 0;                                                                               \
      for (int m = 0; m < nk; m++) s += Dk[m * nk + k] * W2[(m * nj + mj) * ni + mi];        \
      W1[q] = s;                                                                             \
    }                                                                                        \
    for (int q = 0; q < N; q++) {                                                            \
      int mi = q % ni, j = (q / ni) % nj, k = q / (ni * nj);                                 \
      T s = 0;                                                                               \
      for (int m = 0; m < nj; m++) s += Dj[m * nj + j] * W1[(k * nj + m) * ni + mi];         \
      W2[q] = s;                                                                             \
    }                                                                                        \
    for (int q = 0; q < N; q++) {                                                            \
      int i = q % ni, row = q - i, j = (q / ni) % nj, k = q / (ni * nj);                     \
      T s = 0;                                                                               \
      for (int m = 0; m < ni; m++) s += Di[m * ni + i] * W2[row + m];                        \
      L->x[IX(g, i + 1, j + 1, k + 1)] = s;                                                  \
    }                                                                                        \
    ghosts##S(L->x, g);                                                                      \
    free(W1);                                                                                \
    free(W2);                                                                                \
  }                                                                                          \
  static void vcycle##S(Lev##S *L, int l, int nl, int nu1, int nu2, const T *tab, const int *o) \
  {                                                                                          \
    if (l == nl - 1) {                                                                       \
      coarse##S(&L[l], tab, o);                                                              \
      return;                                                                                \
    }                                                                                        \
    for (int s = 0; s < nu1; s++) sweep##S(L[l].x, L[l].b, &L[l].g, L[l].cf);                \
    restrict##S(&L[l], &L[l + 1]);                                                           \
    vcycle##S(L, l + 1, nl, nu1, nu2, tab, o);                                               \
    prolong##S(&L[l], &L[l + 1]);                                                            \
    for (int s = 0; s < nu2; s++) sweep##S(L[l].x, L[l].b, &L[l].g, L[l].cf);                \
  }

DEFINE_OPS(double, D)
DEFINE_OPS(float, F)

/* level count: halve while all extents are even and >= 4, then keep halving
 * until the coarsest fits the one-workgroup solve (as the GPU library) */
int orcGmgLevels(int ni, int nj, int nk, int want)
{
  int nl = 1;
  if (want < 1) want = 1;
  if (want > MAXLEV) want = MAXLEV;
  while (nl < want && ni % 2 == 0 && nj % 2 == 0 && nk % 2 == 0 && ni >= 4 && nj >= 4 && nk >= 4) {
    ni /= 2; nj /= 2; nk /= 2; nl++;
  }
  while ((long long)ni * nj * nk > COARSE_MAX_CELLS) {
    if (ni % 2 || nj % 2 || nk % 2 || ni < 4 || nj < 4 || nk < 4 || nl >= MAXLEV) return -1;
    ni /= 2; nj /= 2; nk /= 2; nl++;
  }
  return nl;
}

/* DCT-II tables of a level (same layout and closed forms as the GPU library):
 * C[m][i] = cos(pi m (i+1/2)/n), D = s_m C, s_0 = 1/n, s_m = 2/n,
 * L[m] = -4/h^2 sin^2(pi m/(2n)). o[9] receives the offsets. */
static double *dctTables(const G3 *g, double hx, double hy, double hz, int o[9])
{
  int n[3] = {g->ni, g->nj, g->nk};
  double h[3] = {hx, hy, hz};
  int tot = 0;
  for (int d = 0; d < 3; d++) { o[d] = tot; tot += n[d] * n[d]; }
  for (int d = 0; d < 3; d++) { o[3 + d] = tot; tot += n[d] * n[d]; }
  for (int d = 0; d < 3; d++) { o[6 + d] = tot; tot += n[d]; }
  double *t = malloc((size_t)tot * sizeof(double));
  for (int d = 0; d < 3; d++) {
    int nn = n[d];
    for (int m = 0; m < nn; m++) {
      double sm = (m == 0 ? 1.0 : 2.0) / nn;
      for (int i = 0; i < nn; i++) {
        double cv                 = cos(M_PI * m * (i + 0.5) / nn);
        t[o[d] + m * nn + i]      = cv;
        t[o[3 + d] + m * nn + i]  = sm * cv;
      }
      double sn      = sin(M_PI * m / (2.0 * nn));
      t[o[6 + d] + m] = -4.0 / (h[d] * h[d]) * sn * sn;
    }
  }
  return t;
}

/* Test hook: the coarse solve alone on a padded FP64 array pair. */
void orcGmgCoarseSolve(int ni, int nj, int nk, double hx, double hy, double hz, const double *b,
    double *x)
{
  LevD L;
  memset(&L, 0, sizeof L);
  L.g = g3(ni, nj, nk);
  L.b = (double *)b;
  L.x = x;
  int o[9];
  double *t = dctTables(&L.g, hx, hy, hz, o);
  coarseD(&L, t, o);
  free(t);
}

/* cycles before the next stop test: the first group, then one */
static int groupSize(int done, int firstGroup, int maxCycles)
{
  int m = done == 0 && firstGroup > 1 ? firstGroup : 1;
  return m < maxCycles - done ? m : maxCycles - done;
}

static double resNorm(const G3 *g, const double *p, const double *rhs, const double *cf, float *r32,
    double *nb)
{
  double a0 = 0.0, a1 = 0.0;
  for (int k = 1; k <= g->nk; k++)
    for (int j = 1; j <= g->nj; j++)
      for (int i = 1; i <= g->ni; i++) {
        size_t c = IX(g, i, j, k);
        double r = rhs[c] - lapD(p, c, g, cf);
        if (r32) r32[c] = (float)r;
        a0 += r * r;
        a1 += rhs[c] * rhs[c];
      }
  *nb = sqrt(a1);
  return sqrt(a0);
}

/* Solves A p = rhs on the padded arrays p, rhs (rhs loses its mean when
 * `project`). The first stop test comes after min(firstGroup, maxCycles)
 * cycles, the later ones after each cycle. Returns the number of cycles, -1
 * if the grid cannot be coarsened; *relOut receives the final relative
 * residual. */
int orcGmgSolveGrouped(int ni, int nj, int nk, double dx, double dy, double dz, double *p, double *rhs,
    int levels, int nu1, int nu2, double omega, double tol, int maxCycles, int project,
    int mixed, int firstGroup, double *relOut)
{
  int nl = orcGmgLevels(ni, nj, nk, levels);
  if (nl < 0) return -1;
  G3 g0 = g3(ni, nj, nk);
  if (project) {
    double s = 0.0;
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++) s += rhs[IX(&g0, i, j, k)];
    double m = s / ((double)ni * nj * nk);
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++) rhs[IX(&g0, i, j, k)] = rhs[IX(&g0, i, j, k)] - m;
  }
  ghostsD(p, &g0);
  LevD LD[MAXLEV];
  LevF LF[MAXLEV];
  memset(LD, 0, sizeof(LD));
  memset(LF, 0, sizeof(LF));
  double cf0[4];
  coefs(dx, dy, dz, omega, cf0);
  double hx = dx, hy = dy, hz = dz;
  int a = ni, b = nj, c = nk;
  int o[9];
  double *tab64 = NULL;
  float *tab32  = NULL;
  for (int l = 0; l < nl; l++) {
    double cf[4];
    coefs(hx, hy, hz, omega, cf);
    G3 g = g3(a, b, c);
    if (l == nl - 1) {
      int tot = 0;
      tab64   = dctTables(&g, hx, hy, hz, o);
      tot     = o[8] + g.nk;
      tab32   = malloc((size_t)tot * sizeof(float));
      for (int q = 0; q < tot; q++) tab32[q] = (float)tab64[q];
    }
    LD[l].g = g;
    LF[l].g = g;
    for (int q = 0; q < 4; q++) {
      LD[l].cf[q] = cf[q];
      LF[l].cf[q] = (float)cf[q];
    }
    if (mixed) {
      LF[l].x = calloc(g.n, sizeof(float));
      LF[l].b = calloc(g.n, sizeof(float));
    } else if (l > 0) {
      LD[l].x = calloc(g.n, sizeof(double));
      LD[l].b = calloc(g.n, sizeof(double));
    }
    a /= 2; b /= 2; c /= 2;
    hx *= 2; hy *= 2; hz *= 2;
  }
  double nb = 0.0, rel = 0.0;
  int cycles = 0;
  if (!mixed) {
    LD[0].x = p;
    LD[0].b = rhs;
    double rn = resNorm(&g0, p, rhs, cf0, NULL, &nb);
    rel       = nb > 0.0 ? rn / nb : 0.0;
    while (nb > 0.0 && rel > tol && cycles < maxCycles) {
      int m = groupSize(cycles, firstGroup, maxCycles);
      for (int q = 0; q < m; q++) vcycleD(LD, 0, nl, nu1, nu2, tab64, o);
      cycles += m;
      rn  = resNorm(&g0, p, rhs, cf0, NULL, &nb);
      rel = rn / nb;
    }
  } else {
    double rn = resNorm(&g0, p, rhs, cf0, LF[0].b, &nb);
    rel       = nb > 0.0 ? rn / nb : 0.0;
    while (nb > 0.0 && rel > tol && cycles < maxCycles) {
      int m = groupSize(cycles, firstGroup, maxCycles);
      memset(LF[0].x, 0, g0.n * sizeof(float));
      for (int q = 0; q < m; q++) vcycleF(LF, 0, nl, nu1, nu2, tab32, o);
      for (int k = 1; k <= nk; k++)
        for (int j = 1; j <= nj; j++)
          for (int i = 1; i <= ni; i++) {
            size_t q = IX(&g0, i, j, k);
            p[q]     = p[q] + (double)LF[0].x[q];
          }
      ghostsD(p, &g0);
      cycles += m;
      rn  = resNorm(&g0, p, rhs, cf0, LF[0].b, &nb);
      rel = rn / nb;
    }
  }
  for (int l = 0; l < nl; l++) {
    free(LF[l].x);
    free(LF[l].b);
    if (l > 0) {
      free(LD[l].x);
      free(LD[l].b);
    }
  }
  free(tab64);
  free(tab32);
  if (relOut) *relOut = rel;
  return cycles;
}

/* the plain loop: a stop test after every cycle */
int orcGmgSolve(int ni, int nj, int nk, double dx, double dy, double dz, double *p, double *rhs,
    int levels, int nu1, int nu2, double omega, double tol, int maxCycles, int project,
    int mixed, double *relOut)
{
  return orcGmgSolveGrouped(ni, nj, nk, dx, dy, dz, p, rhs, levels, nu1, nu2, omega, tol, maxCycles,
      project, mixed, 1, relOut);
}

/* Test hooks: one prolongation (x_f += P x_c) and one residual restriction
 * on FP64 arrays, fine extents (ni, nj, nk), coarse = fine / 2. */
void orcGmgProlong(int ni, int nj, int nk, const double *xc, double *xf)
{
  LevD f, c;
  memset(&f, 0, sizeof f);
  memset(&c, 0, sizeof c);
  f.g = g3(ni, nj, nk);
  c.g = g3(ni / 2, nj / 2, nk / 2);
  f.x = xf;
  c.x = (double *)xc;
  prolongD(&f, &c);
}

void orcGmgRestrict(int ni, int nj, int nk, double dx, double dy, double dz, const double *x,
    const double *b, double *bc)
{
  LevD f, c;
  memset(&f, 0, sizeof f);
  memset(&c, 0, sizeof c);
  f.g = g3(ni, nj, nk);
  c.g = g3(ni / 2, nj / 2, nk / 2);
  double cf[4];
  coefs(dx, dy, dz, 1.0, cf);
  for (int q = 0; q < 4; q++) f.cf[q] = cf[q];
  f.x = (double *)x;
  f.b = (double *)b;
  double *xc = calloc(c.g.n, sizeof(double));
  c.x = xc;
  c.b = bc;
  restrictD(&f, &c);
  free(xc);
}
