/* CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch C restatement of NuSiF's Chorin-projection time step
 * (reference: src/discretization.c, src/solver-rb.c, src/solver-mg.c, sequential
 * semantics of src/comm.c). It exports the same C API as the reference
 * (discretization.h / solver.h) so tests can drive the compiled reference, this
 * oracle and the GPU library through one harness.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this code, and only as the checker or the timed CPU baseline. The GPU library
 * never links or calls it.
 *
 * Built twice by oracle/Makefile:
 *   -DORC_SOLVER_MG=0  -> solve() is the red-black SOR of src/solver-rb.c
 *   -DORC_SOLVER_MG=1  -> solve() is the observable behaviour of src/solver-mg.c
 * with strict IEEE flags (-O2 -fno-fast-math -ffp-contract=off) for parity, and
 * with the reference's own flags (-O3 -ffast-math) for the CPU baseline.
 *
 * Pins (tests/test_oracle_*.py): every function is checked against vectors
 * produced by the compiled reference (oracle/_ref, fixtures under
 * tests/golden/), plus closed forms (divergence of polynomial fields, constant
 * pressure gradients, the CFL formula, the discrete Neumann eigenfunction).
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "comm.h"
#include "discretization.h"
#include "parameter.h"
#include "solver.h"

#ifndef ORC_SOLVER_MG
#define ORC_SOLVER_MG 0
#endif

/* Padded, i-fastest, k-slowest layout with one ghost layer
 * (reference src/util.h:19-34). */
typedef struct {
  int ni, nj, nk; /* interior cells */
  size_t sj, sk;  /* strides of j and k */
} Box;

static Box boxOf(const CommType *c)
{
  Box b;
  b.ni = c->imaxLocal;
  b.nj = c->jmaxLocal;
  b.nk = c->kmaxLocal;
  b.sj = (size_t)(b.ni + 2);
  b.sk = b.sj * (size_t)(b.nj + 2);
  return b;
}

static inline size_t at(const Box *b, int i, int j, int k)
{
  return (size_t)k * b->sk + (size_t)j * b->sj + (size_t)i;
}

static size_t paddedCount(const Box *b) { return b->sk * (size_t)(b->nk + 2); }

/* Problem extras that the reference hard-codes: lid speed (the reference's
 * `2.0 - U` is lid speed 1.0, src/discretization.c:376) and the canal inflow
 * value (2.0, src/discretization.c:384). */
static double gLid    = 1.0;
static double gInflow = 2.0;
static int gLastIterations = 0;

void orcSetProblemExtras(double lid, double inflow)
{
  gLid    = lid;
  gInflow = inflow;
}

int orcLastIterations(void) { return gLastIterations; }

static double *newField(size_t n, double value)
{
  double *a = NULL;
  if (posix_memalign((void **)&a, 64, n * sizeof(double)) != 0 || a == NULL) {
    fprintf(stderr, "oracle: out of memory\n");
    exit(EXIT_FAILURE);
  }
  for (size_t q = 0; q < n; q++) a[q] = value;
  return a;
}

/* src/discretization.c:49-117 */
void initDiscretization(Discretization *s, Parameter *prm)
{
  s->problem  = prm->name;
  s->bcLeft   = prm->bcLeft;
  s->bcRight  = prm->bcRight;
  s->bcBottom = prm->bcBottom;
  s->bcTop    = prm->bcTop;
  s->bcFront  = prm->bcFront;
  s->bcBack   = prm->bcBack;

  s->grid.imax    = prm->imax;
  s->grid.jmax    = prm->jmax;
  s->grid.kmax    = prm->kmax;
  s->grid.xlength = prm->xlength;
  s->grid.ylength = prm->ylength;
  s->grid.zlength = prm->zlength;
  s->grid.dx      = prm->xlength / prm->imax;
  s->grid.dy      = prm->ylength / prm->jmax;
  s->grid.dz      = prm->zlength / prm->kmax;

  s->eps     = prm->eps;
  s->omega   = prm->omg;
  s->itermax = prm->itermax;
  s->re      = prm->re;
  s->gx      = prm->gx;
  s->gy      = prm->gy;
  s->gz      = prm->gz;
  s->dt      = prm->dt;
  s->te      = prm->te;
  s->tau     = prm->tau;
  s->gamma   = prm->gamma;

  Box b    = boxOf(&s->comm);
  size_t n = paddedCount(&b);
  /* every padded cell, ghosts included, starts at the init value */
  s->u   = newField(n, prm->u_init);
  s->v   = newField(n, prm->v_init);
  s->w   = newField(n, prm->w_init);
  s->p   = newField(n, prm->p_init);
  s->rhs = newField(n, 0.0);
  s->f   = newField(n, 0.0);
  s->g   = newField(n, 0.0);
  s->h   = newField(n, 0.0);

  double dx = s->grid.dx, dy = s->grid.dy, dz = s->grid.dz;
  double inv2 = 1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz);
  s->dtBound  = 0.5 * s->re * 1.0 / inv2;
}

/* Velocity ghost fill of one face. `n` is the face-normal component, `t1`, `t2`
 * the tangential ones; the index maps below place them. The per-type rules
 * are those of src/discretization.c:119-332:
 *   NOSLIP : normal wall value 0, tangential ghost = -inside
 *   SLIP   : normal wall value 0, tangential ghost = +inside
 *   OUTFLOW: normal wall value = next inner normal value, tangential ghost = inside
 *   PERIODIC: nothing */
typedef enum { FACE_TOP, FACE_BOTTOM, FACE_LEFT, FACE_RIGHT, FACE_FRONT, FACE_BACK } Face;

static void fillFace(Discretization *s, Face face, int type)
{
  Box b     = boxOf(&s->comm);
  double *u = s->u, *v = s->v, *w = s->w;
  int ni = b.ni, nj = b.nj, nk = b.nk;
  if (type != NOSLIP && type != SLIP && type != OUTFLOW) return; /* PERIODIC: no-op */
  double tsign = (type == NOSLIP) ? -1.0 : 1.0;

  switch (face) {
  case FACE_TOP: /* normal V; wall at j=nj, ghost row j=nj+1 */
  case FACE_BOTTOM: {
    int top = (face == FACE_TOP);
    int jg = top ? nj + 1 : 0, jin = top ? nj : 1, jw = top ? nj : 0;
    int jn = top ? nj - 1 : 1; /* OUTFLOW source of the normal component */
    for (int k = 1; k <= nk; k++)
      for (int i = 1; i <= ni; i++) {
        if (type == OUTFLOW) {
          u[at(&b, i, jg, k)] = u[at(&b, i, jin, k)];
          v[at(&b, i, jw, k)] = v[at(&b, i, jn, k)];
          w[at(&b, i, jg, k)] = w[at(&b, i, jin, k)];
        } else {
          u[at(&b, i, jg, k)] = tsign == -1.0 ? -u[at(&b, i, jin, k)] : u[at(&b, i, jin, k)];
          v[at(&b, i, jw, k)] = 0.0;
          w[at(&b, i, jg, k)] = tsign == -1.0 ? -w[at(&b, i, jin, k)] : w[at(&b, i, jin, k)];
        }
      }
  } break;
  case FACE_LEFT: /* normal U */
  case FACE_RIGHT: {
    int right = (face == FACE_RIGHT);
    int ig = right ? ni + 1 : 0, iin = right ? ni : 1, iw = right ? ni : 0;
    int in = right ? ni - 1 : 1;
    for (int k = 1; k <= nk; k++)
      for (int j = 1; j <= nj; j++) {
        if (type == OUTFLOW) {
          u[at(&b, iw, j, k)] = u[at(&b, in, j, k)];
          v[at(&b, ig, j, k)] = v[at(&b, iin, j, k)];
          w[at(&b, ig, j, k)] = w[at(&b, iin, j, k)];
        } else {
          u[at(&b, iw, j, k)] = 0.0;
          v[at(&b, ig, j, k)] = tsign == -1.0 ? -v[at(&b, iin, j, k)] : v[at(&b, iin, j, k)];
          w[at(&b, ig, j, k)] = tsign == -1.0 ? -w[at(&b, iin, j, k)] : w[at(&b, iin, j, k)];
        }
      }
  } break;
  case FACE_FRONT: /* normal W */
  case FACE_BACK: {
    int back = (face == FACE_BACK);
    int kg = back ? nk + 1 : 0, kin = back ? nk : 1, kw = back ? nk : 0;
    int kn = back ? nk - 1 : 1;
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++) {
        if (type == OUTFLOW) {
          u[at(&b, i, j, kg)] = u[at(&b, i, j, kin)];
          v[at(&b, i, j, kg)] = v[at(&b, i, j, kin)];
          w[at(&b, i, j, kw)] = w[at(&b, i, j, kn)];
        } else {
          u[at(&b, i, j, kg)] = tsign == -1.0 ? -u[at(&b, i, j, kin)] : u[at(&b, i, j, kin)];
          v[at(&b, i, j, kg)] = tsign == -1.0 ? -v[at(&b, i, j, kin)] : v[at(&b, i, j, kin)];
          w[at(&b, i, j, kw)] = 0.0;
        }
      }
  } break;
  }
}

/* src/discretization.c:119-332. Faces are applied strictly in the order
 * TOP, BOTTOM, LEFT, RIGHT, FRONT, BACK: later faces read values written by
 * earlier ones at shared edges. */
void setBoundaryConditions(Discretization *s)
{
  fillFace(s, FACE_TOP, s->bcTop);
  fillFace(s, FACE_BOTTOM, s->bcBottom);
  fillFace(s, FACE_LEFT, s->bcLeft);
  fillFace(s, FACE_RIGHT, s->bcRight);
  fillFace(s, FACE_FRONT, s->bcFront);
  fillFace(s, FACE_BACK, s->bcBack);
}

/* src/discretization.c:364-389. The lid covers 1 <= i < imax, 1 <= k < kmax only. */
void setSpecialBoundaryCondition(Discretization *s)
{
  Box b = boxOf(&s->comm);
  double *u = s->u;
  if (strcmp(s->problem, "dcavity") == 0) {
    double twoLid = 2.0 * gLid;
    for (int k = 1; k < b.nk; k++)
      for (int i = 1; i < b.ni; i++)
        u[at(&b, i, b.nj + 1, k)] = twoLid - u[at(&b, i, b.nj, k)];
  } else if (strcmp(s->problem, "canal") == 0) {
    for (int k = 1; k <= b.nk; k++)
      for (int j = 1; j <= b.nj; j++) u[at(&b, 0, j, k)] = gInflow;
  }
}

/* src/discretization.c:391-401: max |a| over the whole padded array (ghost
 * edges and corners included), seeded with DBL_MIN. */
static double absMax(const Box *b, const double *a)
{
  double m   = DBL_MIN;
  size_t n   = paddedCount(b);
  for (size_t q = 0; q < n; q++) {
    double x = fabs(a[q]);
    m        = (m > x) ? m : x;
  }
  return m;
}

/* src/discretization.c:431-453 */
void computeTimestep(Discretization *s)
{
  Box b     = boxOf(&s->comm);
  double dt = s->dtBound;
  double um = absMax(&b, s->u), vm = absMax(&b, s->v), wm = absMax(&b, s->w);
  if (um > 0) dt = (dt > s->grid.dx / um) ? s->grid.dx / um : dt;
  if (vm > 0) dt = (dt > s->grid.dy / vm) ? s->grid.dy / vm : dt;
  if (wm > 0) dt = (dt > s->grid.dz / wm) ? s->grid.dz / wm : dt;
  s->dt = dt * s->tau;
}

/* One donor-cell convective flux difference, in the operand order of the
 * reference's expressions (src/discretization.c:491-594):
 *   c*((a1*b1) - (a0*b0)) + gamma*c*(|a1|*d1 + |a0|*d0),  c = inv*0.25 */
static inline double donor(double inv, double gamma, double a1, double b1, double d1,
    double a0, double b0, double d0)
{
  return inv * 0.25 * (a1 * b1 - a0 * b0) + gamma * inv * 0.25 * (fabs(a1) * d1 + fabs(a0) * d0);
}

static inline double d2(double inv, double xp, double x, double xm)
{
  return inv * inv * (xp - 2.0 * x + xm);
}

/* src/discretization.c:455-652. Note dvwdz (src/discretization.c:543-550) uses
 * V(i,j,k+1) in both halves of its k-1/2 flux; reproduced as written. */
void computeFG(Discretization *s)
{
  Box b = boxOf(&s->comm);
  const double *u = s->u, *v = s->v, *w = s->w;
  double *f = s->f, *g = s->g, *h = s->h;
  double gam = s->gamma, dt = s->dt;
  double iRe = 1.0 / s->re;
  double iDx = 1.0 / s->grid.dx, iDy = 1.0 / s->grid.dy, iDz = 1.0 / s->grid.dz;
  const size_t sj = b.sj, sk = b.sk;

  for (int k = 1; k <= b.nk; k++)
    for (int j = 1; j <= b.nj; j++)
      for (int i = 1; i <= b.ni; i++) {
        size_t c = at(&b, i, j, k);
        double U = u[c], V = v[c], W = w[c];

        /* F at the x-face (i+1/2, j, k) */
        double du2dx = donor(iDx, gam, U + u[c + 1], U + u[c + 1], U - u[c + 1],
            U + u[c - 1], U + u[c - 1], U - u[c - 1]);
        double duvdy = donor(iDy, gam, V + v[c + 1], U + u[c + sj], U - u[c + sj],
            v[c - sj] + v[c - sj + 1], U + u[c - sj], U - u[c - sj]);
        double duwdz = donor(iDz, gam, W + w[c + 1], U + u[c + sk], U - u[c + sk],
            w[c - sk] + w[c - sk + 1], U + u[c - sk], U - u[c - sk]);
        double lapU = d2(iDx, u[c + 1], U, u[c - 1]) + d2(iDy, u[c + sj], U, u[c - sj]) +
                      d2(iDz, u[c + sk], U, u[c - sk]);
        f[c] = U + dt * (iRe * lapU - du2dx - duvdy - duwdz + s->gx);

        /* G at the y-face (i, j+1/2, k) */
        double duvdx = donor(iDx, gam, U + u[c + sj], V + v[c + 1], V - v[c + 1],
            u[c - 1] + u[c - 1 + sj], V + v[c - 1], V - v[c - 1]);
        double dv2dy = donor(iDy, gam, V + v[c + sj], V + v[c + sj], V - v[c + sj],
            V + v[c - sj], V + v[c - sj], V - v[c - sj]);
        double dvwdz = donor(iDz, gam, W + w[c + sj], V + v[c + sk], V - v[c + sk],
            w[c - sk] + w[c - sk + sj], V + v[c + sk], V - v[c + sk]);
        double lapV = d2(iDx, v[c + 1], V, v[c - 1]) + d2(iDy, v[c + sj], V, v[c - sj]) +
                      d2(iDz, v[c + sk], V, v[c - sk]);
        g[c] = V + dt * (iRe * lapV - duvdx - dv2dy - dvwdz + s->gy);

        /* H at the z-face (i, j, k+1/2) */
        double duwdx = donor(iDx, gam, U + u[c + sk], W + w[c + 1], W - w[c + 1],
            u[c - 1] + u[c - 1 + sk], W + w[c - 1], W - w[c - 1]);
        double dvwdy = donor(iDy, gam, V + v[c + sk], W + w[c + sj], W - w[c + sj],
            v[c - sj + sk] + v[c - sj], W + w[c - sj], W - w[c - sj]);
        double dw2dz = donor(iDz, gam, W + w[c + sk], W + w[c + sk], W - w[c + sk],
            W + w[c - sk], W + w[c - sk], W - w[c - sk]);
        double lapW = d2(iDx, w[c + 1], W, w[c - 1]) + d2(iDy, w[c + sj], W, w[c - sj]) +
                      d2(iDz, w[c + sk], W, w[c - sk]);
        h[c] = W + dt * (iRe * lapW - duwdx - dvwdy - dw2dz + s->gz);
      }

  /* wall faces of F, G, H take the wall velocity (src/discretization.c:599-651) */
  for (int k = 1; k <= b.nk; k++)
    for (int j = 1; j <= b.nj; j++) {
      f[at(&b, 0, j, k)]    = u[at(&b, 0, j, k)];
      f[at(&b, b.ni, j, k)] = u[at(&b, b.ni, j, k)];
    }
  for (int k = 1; k <= b.nk; k++)
    for (int i = 1; i <= b.ni; i++) {
      g[at(&b, i, 0, k)]    = v[at(&b, i, 0, k)];
      g[at(&b, i, b.nj, k)] = v[at(&b, i, b.nj, k)];
    }
  for (int j = 1; j <= b.nj; j++)
    for (int i = 1; i <= b.ni; i++) {
      h[at(&b, i, j, 0)]    = w[at(&b, i, j, 0)];
      h[at(&b, i, j, b.nk)] = w[at(&b, i, j, b.nk)];
    }
}

/* src/discretization.c:334-362 */
void computeRHS(Discretization *s)
{
  Box b = boxOf(&s->comm);
  double idx = 1.0 / s->grid.dx, idy = 1.0 / s->grid.dy, idz = 1.0 / s->grid.dz;
  double idt = 1.0 / s->dt;
  const double *f = s->f, *g = s->g, *h = s->h;
  for (int k = 1; k <= b.nk; k++)
    for (int j = 1; j <= b.nj; j++)
      for (int i = 1; i <= b.ni; i++) {
        size_t c    = at(&b, i, j, k);
        s->rhs[c] = ((f[c] - f[c - 1]) * idx + (g[c] - g[c - b.sj]) * idy +
                        (h[c] - h[c - b.sk]) * idz) *
                    idt;
      }
}

/* src/discretization.c:403-429: interior mean removal, ghosts untouched. */
void normalizePressure(Discretization *s)
{
  Box b = boxOf(&s->comm);
  double sum = 0.0;
  for (int k = 1; k <= b.nk; k++)
    for (int j = 1; j <= b.nj; j++)
      for (int i = 1; i <= b.ni; i++) sum += s->p[at(&b, i, j, k)];
  sum /= (s->grid.imax * s->grid.jmax * s->grid.kmax);
  for (int k = 1; k <= b.nk; k++)
    for (int j = 1; j <= b.nj; j++)
      for (int i = 1; i <= b.ni; i++) s->p[at(&b, i, j, k)] = s->p[at(&b, i, j, k)] - sum;
}

/* src/discretization.c:654-681 */
void adaptUV(Discretization *s)
{
  Box b = boxOf(&s->comm);
  double fx = s->dt / s->grid.dx, fy = s->dt / s->grid.dy, fz = s->dt / s->grid.dz;
  const double *p = s->p;
  for (int k = 1; k <= b.nk; k++)
    for (int j = 1; j <= b.nj; j++)
      for (int i = 1; i <= b.ni; i++) {
        size_t c = at(&b, i, j, k);
        s->u[c]  = s->f[c] - (p[c + 1] - p[c]) * fx;
        s->v[c]  = s->g[c] - (p[c + b.sj] - p[c]) * fy;
        s->w[c]  = s->h[c] - (p[c + b.sk] - p[c]) * fz;
      }
}

/* ------------------------------------------------------------------------- */
/* Pressure solvers                                                          */
/* ------------------------------------------------------------------------- */

/* src/solver-rb.c:17-25, src/solver-mg.c:356-389 (the MG buffers are not
 * needed: the coarse-grid work of src/solver-mg.c never reaches p). */
void initSolver(Solver *s, Discretization *d, Parameter *prm)
{
  s->eps        = prm->eps;
  s->omega      = prm->omg;
  s->itermax    = prm->itermax;
  s->levels     = prm->levels;
  s->presmooth  = prm->presmooth;
  s->postsmooth = prm->postsmooth;
  s->grid       = &d->grid;
  s->comm       = &d->comm;
  s->problem    = prm->name;
  s->r = s->e = NULL;
}

typedef struct {
  Box b;
  double idx2, idy2, idz2, factor;
} Sweep;

static Sweep sweepOf(const Solver *s)
{
  Sweep w;
  w.b = boxOf(s->comm);
  double dx2 = s->grid->dx * s->grid->dx;
  double dy2 = s->grid->dy * s->grid->dy;
  double dz2 = s->grid->dz * s->grid->dz;
  w.idx2 = 1.0 / dx2;
  w.idy2 = 1.0 / dy2;
  w.idz2 = 1.0 / dz2;
  /* src/solver-rb.c:46-47 */
  w.factor = s->omega * 0.5 * (dx2 * dy2 * dz2) / (dy2 * dz2 + dx2 * dz2 + dx2 * dy2);
  return w;
}

/* r = rhs - Lap7(p) at cell c (src/solver-rb.c:66-69) */
static inline double defect(const Sweep *w, const double *p, const double *rhs, size_t c)
{
  size_t sj = w->b.sj, sk = w->b.sk;
  return rhs[c] - ((p[c + 1] - 2.0 * p[c] + p[c - 1]) * w->idx2 +
                      (p[c + sj] - 2.0 * p[c] + p[c - sj]) * w->idy2 +
                      (p[c + sk] - 2.0 * p[c] + p[c - sk]) * w->idz2);
}

/* One colour of a red-black sweep: cells with (i+j+k) % 2 == parity, in
 * k, j, i order. Pass 0 of the reference is parity 1 (i+j+k odd), pass 1 is
 * parity 0 (src/solver-rb.c:55-79). With `accumulate`, r^2 is added onto `acc`
 * one cell at a time (the reference adds onto the carried residual,
 * src/solver-rb.c:72) and the new sum is returned. */
static double colourPass(
    const Sweep *w, double *p, const double *rhs, int parity, double acc, int accumulate)
{
  for (int k = 1; k <= w->b.nk; k++)
    for (int j = 1; j <= w->b.nj; j++) {
      int i0 = ((1 + j + k) % 2 == parity) ? 1 : 2;
      for (int i = i0; i <= w->b.ni; i += 2) {
        size_t c = at(&w->b, i, j, k);
        double r = defect(w, p, rhs, c);
        p[c] -= w->factor * r;
        if (accumulate) acc += r * r;
      }
    }
  return acc;
}

/* Neumann ghost copy over the interior range of each face
 * (src/solver-rb.c:129-148, src/solver-mg.c:153-174). */
static void pressureGhosts(const Box *b, double *p)
{
  for (int j = 1; j <= b->nj; j++)
    for (int i = 1; i <= b->ni; i++) {
      p[at(b, i, j, 0)]         = p[at(b, i, j, 1)];
      p[at(b, i, j, b->nk + 1)] = p[at(b, i, j, b->nk)];
    }
  for (int k = 1; k <= b->nk; k++)
    for (int i = 1; i <= b->ni; i++) {
      p[at(b, i, 0, k)]         = p[at(b, i, 1, k)];
      p[at(b, i, b->nj + 1, k)] = p[at(b, i, b->nj, k)];
    }
  for (int k = 1; k <= b->nk; k++)
    for (int j = 1; j <= b->nj; j++) {
      p[at(b, 0, j, k)]         = p[at(b, 1, j, k)];
      p[at(b, b->ni + 1, j, k)] = p[at(b, b->ni, j, k)];
    }
}

/* src/solver-rb.c:27-172. The residual starts at 1.0 once per call and is
 * carried: res <- (res + r_1^2 + r_2^2 + ...) / N after each full iteration. */
static double solveRB(Solver *s, double *p, const double *rhs)
{
  Sweep w      = sweepOf(s);
  double epssq = s->eps * s->eps;
  double n     = (double)(s->grid->imax * s->grid->jmax * s->grid->kmax);
  double res   = 1.0;
  int it       = 0;
  while (res >= epssq && it < s->itermax) {
    double acc = colourPass(&w, p, rhs, 1, res, 1);
    acc        = colourPass(&w, p, rhs, 0, acc, 1);
    pressureGhosts(&w.b, p);
    res = acc / n;
    it++;
  }
  gLastIterations = it;
  return res;
}

/* Observable behaviour of src/solver-mg.c:299-354 at the finest level: the
 * coarse-grid correction adds e[0], which is never written (it stays 0), so
 * only the fine-grid smoothing and the residual (seeded with 1.0,
 * src/solver-mg.c:252) reach p and the return value. */
static double residualMS(const Sweep *w, const Solver *s, const double *p, const double *rhs)
{
  double acc = 1.0;
  /* same colour order as the reference's calculateResidual */
  for (int parity = 1; parity >= 0; parity--)
    for (int k = 1; k <= w->b.nk; k++)
      for (int j = 1; j <= w->b.nj; j++) {
        int i0 = ((1 + j + k) % 2 == parity) ? 1 : 2;
        for (int i = i0; i <= w->b.ni; i += 2) {
          double r = defect(w, p, rhs, at(&w->b, i, j, k));
          acc += r * r;
        }
      }
  return acc / (double)(s->grid->imax * s->grid->jmax * s->grid->kmax);
}

static double solveMG(Solver *s, double *p, const double *rhs)
{
  Sweep w = sweepOf(s);
  int coarsest = (s->levels - 1 == 0) || w.b.ni < 2 || w.b.nj < 2 || w.b.nk < 2;
  if (coarsest) {
    for (int q = 0; q < s->presmooth; q++) {
      colourPass(&w, p, rhs, 1, 0.0, 0);
      colourPass(&w, p, rhs, 0, 0.0, 0);
    }
    gLastIterations = s->presmooth;
    return residualMS(&w, s, p, rhs);
  }
  for (int q = 0; q < s->presmooth; q++) {
    colourPass(&w, p, rhs, 1, 0.0, 0);
    colourPass(&w, p, rhs, 0, 0.0, 0);
    pressureGhosts(&w.b, p);
  }
  double res = residualMS(&w, s, p, rhs);
  pressureGhosts(&w.b, p); /* after the (null) correction */
  for (int q = 0; q < s->postsmooth; q++) {
    colourPass(&w, p, rhs, 1, 0.0, 0);
    colourPass(&w, p, rhs, 0, 0.0, 0);
    pressureGhosts(&w.b, p);
  }
  gLastIterations = s->presmooth + s->postsmooth;
  return res;
}

double solve(Solver *s, double *p, const double *rhs)
{
#if ORC_SOLVER_MG
  return solveMG(s, p, rhs);
#else
  return solveRB(s, p, rhs);
#endif
}

/* Free the eight fields (the reference never frees them). */
void orcFreeDiscretization(Discretization *s)
{
  free(s->u); free(s->v); free(s->w); free(s->p);
  free(s->rhs); free(s->f); free(s->g); free(s->h);
  s->u = s->v = s->w = s->p = s->rhs = s->f = s->g = s->h = NULL;
}

/* Sequential commCollectResult (src/comm.c:436-476): interior p and
 * face velocities averaged to cell centres. */
void commCollectResult(CommType *c, double *ug, double *vg, double *wg, double *pg,
    double *u, double *v, double *w, double *p, int kmax, int jmax, int imax)
{
  (void)kmax; (void)jmax; (void)imax;
  Box b = boxOf(c);
  size_t q = 0;
  for (int k = 1; k <= b.nk; k++)
    for (int j = 1; j <= b.nj; j++)
      for (int i = 1; i <= b.ni; i++, q++) {
        size_t x = at(&b, i, j, k);
        pg[q] = p[x];
        ug[q] = (u[x] + u[x - 1]) / 2.0;
        vg[q] = (v[x] + v[x - b.sj]) / 2.0;
        wg[q] = (w[x] + w[x - b.sk]) / 2.0;
      }
}
