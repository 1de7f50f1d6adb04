/* The eight step functions of discretization.h on the GPU.
 *
 * Reference: src/discretization.c. Fields keep the reference layout
 * (i fastest, one ghost layer, src/util.h:19-34) so host staging copies are
 * plain memcpy. Kernels are compiled with -ffp-contract=off and evaluate
 * every expression in the reference's operand order, so each cell value is
 * bit-identical to a strict IEEE build of the reference; only reductions
 * (sums) differ in order.
 *
 * Launch shapes: 64 x 4 workgroups over (i, j); a 64-wide row keeps one
 * wavefront on one contiguous 512 B stretch of a row (coalesced HBM
 * streaming). */
#include <float.h>
#include <math.h>
#include <string.h>

#include "nusif_internal.h"
#include "reduce.h"

void nusifRegisterCtx(NusifCtx *c);
void nusifUnregisterCtx(NusifCtx *c);

#define TX 64
#define TY 4

/* ------------------------------------------------------------------------ */
/* computeTimestep: max |u|, |v|, |w| over the padded arrays                */
/* ------------------------------------------------------------------------ */
/* src/discretization.c:391-401: every padded cell counts (ghost edges and
 * corners included), seed DBL_MIN. */
__global__ __launch_bounds__(256) void kAbsMax3(const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, long long n,
    double *__restrict__ partials, DevScalars *sc)
{
  __shared__ double lds[4];
  double mu = DBL_MIN, mv = DBL_MIN, mw = DBL_MIN;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
    double a = fabs(u[q]), b = fabs(v[q]), c = fabs(w[q]);
    mu       = mu > a ? mu : a;
    mv       = mv > b ? mv : b;
    mw       = mw > c ? mw : c;
  }
  mu = blockMax(mu, lds);
  mv = blockMax(mv, lds);
  mw = blockMax(mw, lds);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x + 0] = mu;
    partials[3 * blockIdx.x + 1] = mv;
    partials[3 * blockIdx.x + 2] = mw;
  }
}

/* dt from the three maxima, with the operations of computeTimestep on the
 * host (src/discretization.c:431-453): on = 0 leaves dt alone */
struct DtArgs {
  int on;
  double dtBound, dx, dy, dz, tau;
};

/* scHost: device address of the pinned, host-coherent mirror (also written) */
__global__ __launch_bounds__(1024) void kFoldMax3(const double *partials, int nb, DevScalars *sc, DevScalars *scHost,
    DtArgs da)
{
  __shared__ double lds[16]; /* one per wavefront */
  /* the three maxima in one pass, eight partials in flight per thread; a
   * maximum does not depend on the order it is taken in */
  double m[3] = {DBL_MIN, DBL_MIN, DBL_MIN};
  const int st = blockDim.x;
  int q        = threadIdx.x;
  for (; q + 7 * st < nb; q += 8 * st) {
    double x[8][3];
#pragma unroll
    for (int e = 0; e < 8; e++)
#pragma unroll
      for (int c = 0; c < 3; c++) x[e][c] = partials[3LL * (q + e * st) + c];
#pragma unroll
    for (int e = 0; e < 8; e++)
#pragma unroll
      for (int c = 0; c < 3; c++) m[c] = m[c] > x[e][c] ? m[c] : x[e][c];
  }
  for (; q < nb; q += st)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double x = partials[3LL * q + c];
      m[c]     = m[c] > x ? m[c] : x;
    }
  double r[3];
  for (int c = 0; c < 3; c++) {
    r[c] = blockMax(m[c], lds);
    if (threadIdx.x == 0) {
      sc->absmax[c]     = (unsigned long long)__double_as_longlong(r[c]);
      scHost->absmax[c] = (unsigned long long)__double_as_longlong(r[c]);
    }
  }
  if (threadIdx.x == 0 && da.on) {
    double dt = da.dtBound;
    if (r[0] > 0) dt = (dt > da.dx / r[0]) ? da.dx / r[0] : dt;
    if (r[1] > 0) dt = (dt > da.dy / r[1]) ? da.dy / r[1] : dt;
    if (r[2] > 0) dt = (dt > da.dz / r[2]) ? da.dz / r[2] : dt;
    dt         = dt * da.tau;
    sc->dt     = dt;
    scHost->dt = dt;
  }
  if (threadIdx.x == 0) __threadfence_system();
}

static int reductionBlocks(long long n)
{
  long long b = (n + 256 * 8 - 1) / (256 * 8);
  if (b > 2048) b = 2048;
  if (b < 1) b = 1;
  return (int)b;
}

/* Maxima over the ghost shell of the padded box: the planes k = 0, nk+1,
 * the rows j = 0, nj+1 and the cells i = 0, ni+1 of the planes between.
 * Each workgroup also folds its share of the nInterior per-workgroup maxima
 * adaptUV left (interior), so the final fold reads one partial per shell
 * workgroup instead of one per adaptUV workgroup (a maximum does not depend
 * on the order it is taken in). */
#define SHELL_BLOCKS 1024
/* 32-bit index arithmetic (the shell of any grid this library allocates
 * has far fewer than 2^31 cells; checked at launch): 64-bit divisions were
 * most of this kernel's time */
__global__ __launch_bounds__(256) void kShellMax3(const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, Geo g, double *__restrict__ partials,
    const double *__restrict__ interior, int nInterior)
{
  __shared__ double lds[4];
  const int plane = (int)g.sk, sj = g.sj, rowsPerPlane = 2 * g.sj + 2 * g.nj;
  const int total = 2 * plane + g.nk * rowsPerPlane;
  double mu = DBL_MIN, mv = DBL_MIN, mw = DBL_MIN;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    long long c;
    if (q < 2 * plane) {
      c = q < plane ? q : (long long)(g.nk + 1) * g.sk + (q - plane);
    } else {
      int r = q - 2 * plane;
      int k = r / rowsPerPlane;
      int e = r - k * rowsPerPlane;
      k += 1;
      if (e < 2 * sj) {
        int j = e < sj ? 0 : g.nj + 1;
        c = g.at(e < sj ? e : e - sj, j, k);
      } else {
        int f = e - 2 * sj; /* cells i = 0 and ni+1 of rows 1..nj */
        c = g.at((f & 1) ? g.ni + 1 : 0, (f >> 1) + 1, k);
      }
    }
    double a = fabs(u[c]), b = fabs(v[c]), d = fabs(w[c]);
    mu = mu > a ? mu : a;
    mv = mv > b ? mv : b;
    mw = mw > d ? mw : d;
  }
  const int per = (nInterior + (int)gridDim.x - 1) / (int)gridDim.x;
  const int qEnd = min(nInterior, ((int)blockIdx.x + 1) * per);
  for (int q = (int)blockIdx.x * per + (int)threadIdx.x; q < qEnd; q += (int)blockDim.x) {
    const double a = interior[3 * q], b = interior[3 * q + 1], d = interior[3 * q + 2];
    mu = mu > a ? mu : a;
    mv = mv > b ? mv : b;
    mw = mw > d ? mw : d;
  }
  mu = blockMax(mu, lds);
  mv = blockMax(mv, lds);
  mw = blockMax(mw, lds);
  if (threadIdx.x == 0) {
    partials[3 * blockIdx.x + 0] = mu;
    partials[3 * blockIdx.x + 1] = mv;
    partials[3 * blockIdx.x + 2] = mw;
  }
}

/* max |u|, |v|, |w| over the padded arrays into scal->absmax. After an
 * adaptUV that nothing has overwritten since, only the ghost shell is read
 * (the interior-range maxima are adaptUV's); otherwise the whole arrays. */
void launchAbsMax3(NusifCtx *c, const Discretization *d)
{
  DtArgs da = {0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (d) da = {1, d->dtBound, d->grid.dx, d->grid.dy, d->grid.dz, d->tau};
  if (c->cflValid) {
    double *shell = c->cflPartials + 3LL * c->cflInteriorBlocks;
    kShellMax3<<<SHELL_BLOCKS, 256, 0, nusifStream()>>>(c->u, c->v, c->w, c->geo, shell, c->cflPartials,
        c->cflInteriorBlocks);
    kFoldMax3<<<1, 1024, 0, nusifStream()>>>(shell, SHELL_BLOCKS, c->scal, c->scalHostDev, da);
    return;
  }
  int nb = reductionBlocks(c->geo.n);
  NUSIF_TIMED("absmax", (kAbsMax3<<<nb, 256, 0, nusifStream()>>>(c->u, c->v, c->w, c->geo.n,
                            c->partials, c->scal)));
  kFoldMax3<<<1, 1024, 0, nusifStream()>>>(c->partials, nb, c->scal, c->scalHostDev, da);
}

/* device address of dt while the loop keeps it on the device, else null
 * (the kernels then take d->dt from their arguments) */
static const double *dtDevice(NusifCtx *c) { return c->dtOnDevice ? &c->scal->dt : nullptr; }

/* ------------------------------------------------------------------------ */
/* Velocity boundary conditions (src/discretization.c:119-332)              */
/* ------------------------------------------------------------------------ */
enum { F_TOP = 0, F_BOTTOM, F_LEFT, F_RIGHT, F_FRONT, F_BACK };

/* Applies one face's rule at face point (a, b) (1-based tangential indices).
 * TOP/BOTTOM: (a, b) = (i, k); LEFT/RIGHT: (j, k); FRONT/BACK: (i, j). */
__device__ inline void velocityFace(
    double *u, double *v, double *w, const Geo &g, int face, int type, int a, int b)
{
  if (type != NOSLIP && type != SLIP && type != OUTFLOW) return; /* PERIODIC: no-op */
  bool outflow = (type == OUTFLOW);
  double s     = (type == NOSLIP) ? -1.0 : 1.0;
  switch (face) {
  case F_TOP:
  case F_BOTTOM: {
    if (a > g.ni || b > g.nk) return;
    bool top = face == F_TOP;
    int jg = top ? g.nj + 1 : 0, jin = top ? g.nj : 1, jw = top ? g.nj : 0;
    int jn = top ? g.nj - 1 : 1;
    double uu = u[g.at(a, jin, b)], ww = w[g.at(a, jin, b)];
    u[g.at(a, jg, b)] = outflow ? uu : (s < 0.0 ? -uu : uu);
    v[g.at(a, jw, b)] = outflow ? v[g.at(a, jn, b)] : 0.0;
    w[g.at(a, jg, b)] = outflow ? ww : (s < 0.0 ? -ww : ww);
  } break;
  case F_LEFT:
  case F_RIGHT: {
    if (a > g.nj || b > g.nk) return;
    bool right = face == F_RIGHT;
    int ig = right ? g.ni + 1 : 0, iin = right ? g.ni : 1, iw = right ? g.ni : 0;
    int in = right ? g.ni - 1 : 1;
    double vv = v[g.at(iin, a, b)], ww = w[g.at(iin, a, b)];
    double un = outflow ? u[g.at(in, a, b)] : 0.0;
    u[g.at(iw, a, b)] = un;
    v[g.at(ig, a, b)] = outflow ? vv : (s < 0.0 ? -vv : vv);
    w[g.at(ig, a, b)] = outflow ? ww : (s < 0.0 ? -ww : ww);
  } break;
  default: { /* FRONT / BACK */
    if (a > g.ni || b > g.nj) return;
    bool back = face == F_BACK;
    int kg = back ? g.nk + 1 : 0, kin = back ? g.nk : 1, kw = back ? g.nk : 0;
    int kn = back ? g.nk - 1 : 1;
    double uu = u[g.at(a, b, kin)], vv = v[g.at(a, b, kin)];
    double wn = outflow ? w[g.at(a, b, kn)] : 0.0;
    u[g.at(a, b, kg)] = outflow ? uu : (s < 0.0 ? -uu : uu);
    v[g.at(a, b, kg)] = outflow ? vv : (s < 0.0 ? -vv : vv);
    w[g.at(a, b, kw)] = wn;
  } break;
  }
}

/* Two opposite faces in one launch. They touch disjoint cells and read
 * nothing the other writes as long as the normal extent is >= 2, so the pair
 * is order-free; pairs themselves run in the reference's order. */
__global__ __launch_bounds__(256) void kVelocityFaces(double *u, double *v, double *w, Geo g,
    int faceA, int typeA, int faceB, int typeB)
{
  int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int b = blockIdx.y * blockDim.y + threadIdx.y + 1;
  velocityFace(u, v, w, g, faceA, typeA, a, b);
  if (faceB >= 0) velocityFace(u, v, w, g, faceB, typeB, a, b);
}

static void faceExtent(const Geo &g, int face, int *ea, int *eb)
{
  if (face == F_TOP || face == F_BOTTOM) {
    *ea = g.ni;
    *eb = g.nk;
  } else if (face == F_LEFT || face == F_RIGHT) {
    *ea = g.nj;
    *eb = g.nk;
  } else {
    *ea = g.ni;
    *eb = g.nj;
  }
}

static void launchFaces(NusifCtx *c, int fa, int ta, int fb, int tb)
{
  int ea, eb;
  faceExtent(c->geo, fa, &ea, &eb);
  dim3 blk(64, 4);
  dim3 grd((ea + 63) / 64, (eb + 3) / 4);
  kVelocityFaces<<<grd, blk, 0, nusifStream()>>>(c->u, c->v, c->w, c->geo, fa, ta, fb, tb);
  c->cflValid = false; /* wall faces of the interior range rewritten */
}

void launchVelocityBC(NusifCtx *c, const Discretization *d)
{
  const Geo &g = c->geo;
  /* reference order: TOP, BOTTOM, LEFT, RIGHT, FRONT, BACK */
  if (g.nj >= 2) launchFaces(c, F_TOP, d->bcTop, F_BOTTOM, d->bcBottom);
  else {
    launchFaces(c, F_TOP, d->bcTop, -1, 0);
    launchFaces(c, F_BOTTOM, d->bcBottom, -1, 0);
  }
  if (g.ni >= 2) launchFaces(c, F_LEFT, d->bcLeft, F_RIGHT, d->bcRight);
  else {
    launchFaces(c, F_LEFT, d->bcLeft, -1, 0);
    launchFaces(c, F_RIGHT, d->bcRight, -1, 0);
  }
  if (g.nk >= 2) launchFaces(c, F_FRONT, d->bcFront, F_BACK, d->bcBack);
  else {
    launchFaces(c, F_FRONT, d->bcFront, -1, 0);
    launchFaces(c, F_BACK, d->bcBack, -1, 0);
  }
}

/* src/discretization.c:364-389. Lid: 1 <= i < imax, 1 <= k < kmax only. */
__global__ __launch_bounds__(256) void kLid(double *u, Geo g, double twoLid)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i >= g.ni || k >= g.nk) return;
  u[g.at(i, g.nj + 1, k)] = twoLid - u[g.at(i, g.nj, k)];
}

__global__ __launch_bounds__(256) void kInflow(double *u, Geo g, double value, const double *profile)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int k = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (j > g.nj || k > g.nk) return;
  u[g.at(0, j, k)] = profile ? profile[(long long)k * (g.nj + 2) + j] : value;
}

void launchSpecialBC(NusifCtx *c, const Discretization *d)
{
  const Geo &g = c->geo;
  dim3 blk(64, 4);
  if (strcmp(d->problem, "dcavity") == 0) {
    dim3 grd((g.ni + 63) / 64, (g.nk + 3) / 4);
    kLid<<<grd, blk, 0, nusifStream()>>>(c->u, g, 2.0 * c->opt.lid);
  } else if (strcmp(d->problem, "canal") == 0) {
    dim3 grd((g.nj + 63) / 64, (g.nk + 3) / 4);
    kInflow<<<grd, blk, 0, nusifStream()>>>(c->u, g, c->opt.uInflow, c->inflow);
  }
}

/* ------------------------------------------------------------------------ */
/* The boundary work of a time step in one launch                           */
/* ------------------------------------------------------------------------ */
/* computeTimestep's maxima over the ghost shell, then setBoundaryConditions
 * and setSpecialBoundaryCondition (src/main.c:61-63), after an adaptUV that
 * nothing has overwritten since: one launch and the fold instead of the
 * shell pass, the fold and four boundary launches.
 *
 * The reference applies the faces in the order TOP, BOTTOM, LEFT, RIGHT,
 * FRONT, BACK (src/discretization.c:119-332), then the lid or the inflow.
 * Every written cell here gets the value that order gives it:
 * - A face reads a cell an earlier face writes only where a high face writes
 *   its wall velocity inside the interior range (TOP v(i,nj,k), RIGHT
 *   u(ni,j,k)): LEFT/RIGHT read v(iin,nj,k), FRONT/BACK read v(i,nj,kin)
 *   and u(ni,j,kin). Those reads apply the earlier face's rule to the
 *   original neighbour it copies (v(.,nj-1,.), u(ni-1,.,.)), which no face
 *   writes.
 * - A face reads the original of a cell a later face writes only at
 *   u(ni,jin,k) (TOP/BOTTOM read it, RIGHT writes it) and w(.,.,nk)
 *   (TOP/BOTTOM and LEFT/RIGHT read it, BACK writes it). Those ghost
 *   updates move into the task that writes the cell, before its write:
 *   TOP/BOTTOM's u ghost at i = ni into the LEFT/RIGHT task of row jin,
 *   TOP/BOTTOM's and LEFT/RIGHT's w ghosts of plane nk into the FRONT/BACK
 *   task of that column.
 * - The lid writes u(i,nj+1,k) for i < ni, k < nk over TOP's ghost from
 *   u(i,nj,k), which no face writes; the inflow writes u(0,j,k) over LEFT's
 *   wall value. Neither cell is read by a face.
 * - The maxima read each shell cell's old value in the task that writes it
 *   (before the write); shell cells no task writes (the shell's edges and
 *   corners, and the high ghosts u(ni+1,.,.), v(.,nj+1,.), w(.,.,nk+1)) are
 *   read by the tasks of their faces or by edge tasks. A maximum does not
 *   depend on the order it is taken in, so dt is the same bits.
 * Tasks: TOP/BOTTOM per (i, k), LEFT/RIGHT per junction of two rows,
 * FRONT/BACK per (i, j), one per edge or corner cell; each needs ni, nj,
 * nk >= 3. */
struct BoundaryStep {
  int tTop, tBottom, tLeft, tRight, tFront, tBack;
  int special; /* 0 none, 1 lid, 2 inflow */
  double twoLid, inflow;
  const double *profile;
  int maxOn;                /* maxima (adaptive dt) */
  const double *interior;   /* adaptUV's per-workgroup maxima */
  int nInterior;
  double *partials;         /* per workgroup */
};

__device__ inline bool bcActive(int t) { return t == NOSLIP || t == SLIP || t == OUTFLOW; }
/* a ghost from its interior neighbour, as velocityFace forms it (a wall
 * velocity is the cell next to it for OUTFLOW, else 0, the cell read only
 * for OUTFLOW: a row end's extra cache line less on LEFT / RIGHT) */
__device__ inline double bcGhost(int t, double x) { return t == OUTFLOW ? x : (t == NOSLIP ? -x : x); }

__global__ __launch_bounds__(256) void kBoundaryStep(double *__restrict__ u, double *__restrict__ v,
    double *__restrict__ w, Geo g, BoundaryStep a)
{
  __shared__ double lds[4];
  const int ni = g.ni, nj = g.nj, nk = g.nk;
  const int nTB = ni * nk, nLR = (nj + 1) * nk, nFB = ni * nj;
  const int nEi = 4 * ni, nEj = 4 * nj, nEk = 4 * nk;
  const int total = nTB + nLR + nFB + nEi + nEj + nEk + 8;
  double mu = DBL_MIN, mv = DBL_MIN, mw = DBL_MIN;
  auto mx   = [](double m, double x) { return m > x ? m : x; };
  auto take = [&](long long c, bool tu, bool tv, bool tw) {
    if (!a.maxOn) return;
    if (tu) mu = mx(mu, fabs(u[c]));
    if (tv) mv = mx(mv, fabs(v[c]));
    if (tw) mw = mx(mw, fabs(w[c]));
  };
  /* Each task issues all its loads before its first store: a store to u, v
   * or w may alias a later load of the same array as far as the compiler
   * knows, so loads after a store wait for it (one memory latency per
   * store-load turn); no cell a task stores is loaded after its store. */
  const bool mOn = a.maxOn != 0;
  auto amx = [&](double &m, double x) { m = mx(m, fabs(x)); };
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    /* LEFT / RIGHT first: their row ends are the slow accesses, so their
     * workgroups start first. A task takes the junction of rows J and J + 1
     * of plane k: RIGHT's cells of row J and LEFT's of row J + 1, which lie
     * next to each other in memory (the end of a row, then the start of the
     * next). */
    int t = q;
    if (t < nLR) {
      const int J = t % (nj + 1), k = 1 + t / (nj + 1);
      const bool hasR = J >= 1, hasL = J + 1 <= nj, wk = k < nk;
      const int jR = J, jL = J + 1, tR = a.tRight, tL = a.tLeft;
      /* RIGHT's row: TOP / BOTTOM's u ghost at i = ni moves here (RIGHT
       * writes u(ni, jin, k)); v(ni, nj, k) as TOP left it */
      const bool gMove = hasR && (jR == nj || jR == 1);
      const int tG = jR == nj ? a.tTop : a.tBottom;
      const long long cR = g.at(ni + 1, jR, k), cG = g.at(ni, jR == nj ? nj + 1 : 0, k), cL = g.at(0, jL, k);
      double mRu = 0.0, mRv = 0.0, mRw = 0.0, mGu = 0.0, gSrc = 0.0, vR = 0.0, wR = 0.0, uR = 0.0;
      double mLu = 0.0, mLv = 0.0, mLw = 0.0, vL = 0.0, wL = 0.0, uL = 0.0;
      if (hasR) {
        if (mOn) {
          mRu = u[cR];
          mRv = v[cR];
          if (wk) mRw = w[cR];
          if (gMove) mGu = u[cG];
        }
        if (gMove) gSrc = u[g.at(ni, jR, k)];
        if (bcActive(tR)) {
          vR = bcActive(a.tTop) && jR == nj ? (a.tTop == OUTFLOW ? v[g.at(ni, nj - 1, k)] : 0.0) : v[g.at(ni, jR, k)];
          if (wk) wR = w[g.at(ni, jR, k)];
          if (tR == OUTFLOW) uR = u[g.at(ni - 1, jR, k)];
        }
      }
      if (hasL) {
        if (mOn) {
          mLu = u[cL];
          mLv = v[cL];
          if (wk) mLw = w[cL];
        }
        if (bcActive(tL)) {
          vL = bcActive(a.tTop) && jL == nj ? (a.tTop == OUTFLOW ? v[g.at(1, nj - 1, k)] : 0.0) : v[g.at(1, jL, k)];
          if (wk) wL = w[g.at(1, jL, k)];
          if (tL == OUTFLOW) uL = u[g.at(1, jL, k)];
        }
      }
      if (mOn) {
        if (hasR) {
          amx(mu, mRu);
          amx(mv, mRv);
          if (wk) amx(mw, mRw);
          if (gMove) amx(mu, mGu);
        }
        if (hasL) {
          amx(mu, mLu);
          amx(mv, mLv);
          if (wk) amx(mw, mLw);
        }
      }
      if (hasR) {
        if (gMove && bcActive(tG)) u[cG] = bcGhost(tG, gSrc);
        if (bcActive(tR)) {
          v[cR] = bcGhost(tR, vR);
          if (wk) w[cR] = bcGhost(tR, wR);
          u[g.at(ni, jR, k)] = tR == OUTFLOW ? uR : 0.0;
        }
      }
      if (hasL) {
        if (bcActive(tL)) {
          v[cL] = bcGhost(tL, vL);
          if (wk) w[cL] = bcGhost(tL, wL);
          u[cL] = tL == OUTFLOW ? uL : 0.0;
        }
        if (a.special == 2) u[cL] = a.profile ? a.profile[(long long)k * (nj + 2) + jL] : a.inflow;
      }
      continue;
    }
    t -= nLR;
    if (t < nTB) { /* TOP and BOTTOM at (i, k) */
      const int i = 1 + t % ni, k = 1 + t / ni;
      const bool iu = i < ni, wk = k < nk, lid = a.special == 1 && iu && wk;
      const int tT = a.tTop, tB = a.tBottom;
      const long long c0 = g.at(i, 0, k), c1 = g.at(i, nj + 1, k);
      double m0u = 0.0, m0v = 0.0, m0w = 0.0, m1u = 0.0, m1v = 0.0, m1w = 0.0;
      if (mOn) {
        if (iu) {
          m0u = u[c0];
          m1u = u[c1];
        }
        m0v = v[c0];
        m1v = v[c1];
        if (wk) {
          m0w = w[c0];
          m1w = w[c1];
        }
      }
      /* TOP reads row nj (v: nj - 1), BOTTOM row 1 */
      const double uN = iu && (bcActive(tT) || lid) ? u[g.at(i, nj, k)] : 0.0;
      const double uS = iu && bcActive(tB) ? u[g.at(i, 1, k)] : 0.0;
      const double vN = tT == OUTFLOW ? v[g.at(i, nj - 1, k)] : 0.0;
      const double vS = tB == OUTFLOW ? v[g.at(i, 1, k)] : 0.0;
      const double wN = wk && bcActive(tT) ? w[g.at(i, nj, k)] : 0.0;
      const double wS = wk && bcActive(tB) ? w[g.at(i, 1, k)] : 0.0;
      if (mOn) {
        if (iu) {
          amx(mu, m0u);
          amx(mu, m1u);
        }
        amx(mv, m0v);
        amx(mv, m1v);
        if (wk) {
          amx(mw, m0w);
          amx(mw, m1w);
        }
      }
      if (bcActive(tT)) {
        if (iu) u[c1] = bcGhost(tT, uN);
        v[g.at(i, nj, k)] = vN;
        if (wk) w[c1] = bcGhost(tT, wN);
      }
      if (bcActive(tB)) {
        if (iu) u[c0] = bcGhost(tB, uS);
        v[c0] = vS;
        if (wk) w[c0] = bcGhost(tB, wS);
      }
      if (lid) u[c1] = a.twoLid - uN;
      continue;
    }
    t -= nTB;
    if (t < nFB) { /* FRONT and BACK at (i, j) */
      const int i = 1 + t % ni, j = 1 + t / ni;
      const int tF = a.tFront, tK = a.tBack;
      const long long c0 = g.at(i, j, 0), c1 = g.at(i, j, nk + 1);
      /* TOP / BOTTOM's and LEFT / RIGHT's w ghosts of plane nk move here
       * (BACK writes w(i, j, nk)); u(ni, j, kin) as RIGHT left it, v(i, nj,
       * kin) as TOP left it */
      const bool gTB = j == nj || j == 1, gLR = i == ni || i == 1;
      const int tGTB = j == nj ? a.tTop : a.tBottom, tGLR = i == ni ? a.tRight : a.tLeft;
      const long long cGTB = g.at(i, j == nj ? nj + 1 : 0, nk), cGLR = g.at(i == ni ? ni + 1 : 0, j, nk);
      const bool rightW = bcActive(a.tRight) && i == ni, topW = bcActive(a.tTop) && j == nj;
      double m0u = 0.0, m0v = 0.0, m0w = 0.0, m1u = 0.0, m1v = 0.0, m1w = 0.0, mGt = 0.0, mGl = 0.0;
      if (mOn) {
        m0u = u[c0];
        m0v = v[c0];
        m0w = w[c0];
        m1u = u[c1];
        m1v = v[c1];
        m1w = w[c1];
        if (gTB) mGt = w[cGTB];
        if (gLR) mGl = w[cGLR];
      }
      const double wNk = (gTB && bcActive(tGTB)) || (gLR && bcActive(tGLR)) ? w[g.at(i, j, nk)] : 0.0;
      auto srcU = [&](int kin) {
        return rightW ? (a.tRight == OUTFLOW ? u[g.at(ni - 1, j, kin)] : 0.0) : u[g.at(i, j, kin)];
      };
      auto srcV = [&](int kin) {
        return topW ? (a.tTop == OUTFLOW ? v[g.at(i, nj - 1, kin)] : 0.0) : v[g.at(i, j, kin)];
      };
      const double uF = bcActive(tF) ? srcU(1) : 0.0, vF = bcActive(tF) ? srcV(1) : 0.0;
      const double wF = tF == OUTFLOW ? w[g.at(i, j, 1)] : 0.0;
      const double uK = bcActive(tK) ? srcU(nk) : 0.0, vK = bcActive(tK) ? srcV(nk) : 0.0;
      const double wK = tK == OUTFLOW ? w[g.at(i, j, nk - 1)] : 0.0;
      if (mOn) {
        amx(mu, m0u);
        amx(mv, m0v);
        amx(mw, m0w);
        amx(mu, m1u);
        amx(mv, m1v);
        amx(mw, m1w);
        if (gTB) amx(mw, mGt);
        if (gLR) amx(mw, mGl);
      }
      if (gTB && bcActive(tGTB)) w[cGTB] = bcGhost(tGTB, wNk);
      if (gLR && bcActive(tGLR)) w[cGLR] = bcGhost(tGLR, wNk);
      if (bcActive(tF)) {
        u[c0] = bcGhost(tF, uF);
        v[c0] = bcGhost(tF, vF);
        w[c0] = wF;
      }
      if (bcActive(tK)) {
        u[c1] = bcGhost(tK, uK);
        v[c1] = bcGhost(tK, vK);
        w[g.at(i, j, nk)] = wK;
      }
      continue;
    }
    t -= nFB;
    /* edge and corner cells: maxima only */
    long long c;
    if (t < nEi) {
      const int i = 1 + t % ni, e = t / ni;
      c = g.at(i, (e & 1) ? nj + 1 : 0, (e & 2) ? nk + 1 : 0);
    } else if ((t -= nEi) < nEj) {
      const int j = 1 + t % nj, e = t / nj;
      c = g.at((e & 1) ? ni + 1 : 0, j, (e & 2) ? nk + 1 : 0);
    } else if ((t -= nEj) < nEk) {
      const int k = 1 + t % nk, e = t / nk;
      c = g.at((e & 1) ? ni + 1 : 0, (e & 2) ? nj + 1 : 0, k);
    } else {
      t -= nEk;
      c = g.at((t & 1) ? ni + 1 : 0, (t & 2) ? nj + 1 : 0, (t & 4) ? nk + 1 : 0);
    }
    take(c, true, true, true);
  }
  if (!a.maxOn) return;
  /* this workgroup's share of adaptUV's maxima */
  const int per  = (a.nInterior + (int)gridDim.x - 1) / (int)gridDim.x;
  const int qEnd = min(a.nInterior, ((int)blockIdx.x + 1) * per);
  for (int q = (int)blockIdx.x * per + (int)threadIdx.x; q < qEnd; q += (int)blockDim.x) {
    mu = mx(mu, a.interior[3 * q]);
    mv = mx(mv, a.interior[3 * q + 1]);
    mw = mx(mw, a.interior[3 * q + 2]);
  }
  mu = blockMax(mu, lds);
  mv = blockMax(mv, lds);
  mw = blockMax(mw, lds);
  if (threadIdx.x == 0) {
    a.partials[3 * blockIdx.x + 0] = mu;
    a.partials[3 * blockIdx.x + 1] = mv;
    a.partials[3 * blockIdx.x + 2] = mw;
  }
}

bool boundaryStepFits(const NusifCtx *c) { return c->geo.ni >= 3 && c->geo.nj >= 3 && c->geo.nk >= 3; }

/* computeTimestep (when d is non-null and adaptive, from the shell and
 * adaptUV's maxima: c->cflValid must hold), setBoundaryConditions and
 * setSpecialBoundaryCondition as one launch and the fold */
void launchBoundaryStep(NusifCtx *c, const Discretization *d, bool withMax)
{
  const Geo &g = c->geo;
  BoundaryStep a;
  a.tTop = d->bcTop;
  a.tBottom = d->bcBottom;
  a.tLeft = d->bcLeft;
  a.tRight = d->bcRight;
  a.tFront = d->bcFront;
  a.tBack = d->bcBack;
  a.special = strcmp(d->problem, "dcavity") == 0 ? 1 : (strcmp(d->problem, "canal") == 0 ? 2 : 0);
  a.twoLid = 2.0 * c->opt.lid;
  a.inflow = c->opt.uInflow;
  a.profile = c->inflow;
  a.maxOn = withMax ? 1 : 0;
  a.interior = c->cflPartials;
  a.nInterior = c->cflInteriorBlocks;
  a.partials = c->cflPartials + 3LL * c->cflInteriorBlocks;
  if (withMax && !c->cflValid) NUSIF_FATAL("boundary step: the maxima need the adaptUV partials");
  const long long tasks = (long long)g.ni * g.nk + (long long)(g.nj + 1) * g.nk + (long long)g.ni * g.nj +
                          4LL * (g.ni + g.nj + g.nk) + 8;
  long long nb = (tasks + 255) / 256;
  if (nb > SHELL_BLOCKS) nb = SHELL_BLOCKS;
  kBoundaryStep<<<(unsigned)nb, 256, 0, nusifStream()>>>(c->u, c->v, c->w, g, a);
  if (withMax) {
    DtArgs da = {1, d->dtBound, d->grid.dx, d->grid.dy, d->grid.dz, d->tau};
    kFoldMax3<<<1, 1024, 0, nusifStream()>>>(a.partials, (int)nb, c->scal, c->scalHostDev, da);
  }
  c->cflValid = false; /* wall faces of the interior range rewritten */
}

/* ------------------------------------------------------------------------ */
/* computeFG (src/discretization.c:455-652)                                 */
/* ------------------------------------------------------------------------ */
/* c*((a1*b1) - (a0*b0)) + gamma*c*(|a1|*d1 + |a0|*d0), c = inv*0.25, in the
 * reference's operand order */
__device__ inline double donor(double inv, double gamma, double a1, double b1, double d1,
    double a0, double b0, double d0)
{
  return inv * 0.25 * (a1 * b1 - a0 * b0) + gamma * inv * 0.25 * (fabs(a1) * d1 + fabs(a0) * d0);
}

__device__ inline double d2(double inv, double xp, double x, double xm)
{
  return inv * inv * (xp - 2.0 * x + xm);
}

struct FGCoef {
  double gamma, dt, iRe, iDx, iDy, iDz, gx, gy, gz;
};

/* F, G and H of one cell from its operands, in the reference's operand
 * order (src/discretization.c:491-594). Names: X? = value at an offset of
 * the cell: E/W = i+-1, N/S = j+-1, T/B = k+-1, two letters for a diagonal
 * (uNW = U(i-1,j+1,k)). */
__device__ inline double fCell(const FGCoef &cf, double Uc, double uE, double uW, double uN, double uS,
    double uT, double uB, double Vc, double vE, double vS, double vSE, double Wc, double wE, double wB,
    double wBE)
{
  double du2dx = donor(cf.iDx, cf.gamma, Uc + uE, Uc + uE, Uc - uE, Uc + uW, Uc + uW, Uc - uW);
  double duvdy = donor(cf.iDy, cf.gamma, Vc + vE, Uc + uN, Uc - uN, vS + vSE, Uc + uS, Uc - uS);
  double duwdz = donor(cf.iDz, cf.gamma, Wc + wE, Uc + uT, Uc - uT, wB + wBE, Uc + uB, Uc - uB);
  double lapU  = d2(cf.iDx, uE, Uc, uW) + d2(cf.iDy, uN, Uc, uS) + d2(cf.iDz, uT, Uc, uB);
  return Uc + cf.dt * (cf.iRe * lapU - du2dx - duvdy - duwdz + cf.gx);
}

/* dvwdz keeps the reference's V(i,j,k+1) in both halves of its k-1/2 flux
 * (src/discretization.c:543-550) */
__device__ inline double gCell(const FGCoef &cf, double Uc, double uN, double uW, double uNW, double Vc,
    double vE, double vW, double vN, double vS, double vT, double vB, double Wc, double wN, double wB,
    double wBN)
{
  double duvdx = donor(cf.iDx, cf.gamma, Uc + uN, Vc + vE, Vc - vE, uW + uNW, Vc + vW, Vc - vW);
  double dv2dy = donor(cf.iDy, cf.gamma, Vc + vN, Vc + vN, Vc - vN, Vc + vS, Vc + vS, Vc - vS);
  double dvwdz = donor(cf.iDz, cf.gamma, Wc + wN, Vc + vT, Vc - vT, wB + wBN, Vc + vT, Vc - vT);
  double lapV  = d2(cf.iDx, vE, Vc, vW) + d2(cf.iDy, vN, Vc, vS) + d2(cf.iDz, vT, Vc, vB);
  return Vc + cf.dt * (cf.iRe * lapV - duvdx - dv2dy - dvwdz + cf.gy);
}

__device__ inline double hCell(const FGCoef &cf, double Uc, double uT, double uW, double uTW, double Vc,
    double vT, double vS, double vTS, double Wc, double wE, double wW, double wN, double wS, double wT,
    double wB)
{
  double duwdx = donor(cf.iDx, cf.gamma, Uc + uT, Wc + wE, Wc - wE, uW + uTW, Wc + wW, Wc - wW);
  double dvwdy = donor(cf.iDy, cf.gamma, Vc + vT, Wc + wN, Wc - wN, vTS + vS, Wc + wS, Wc - wS);
  double dw2dz = donor(cf.iDz, cf.gamma, Wc + wT, Wc + wT, Wc - wT, Wc + wB, Wc + wB, Wc - wB);
  double lapW  = d2(cf.iDx, wE, Wc, wW) + d2(cf.iDy, wN, Wc, wS) + d2(cf.iDz, wT, Wc, wB);
  return Wc + cf.dt * (cf.iRe * lapW - duwdx - dvwdy - dw2dz + cf.gz);
}

/* A workgroup owns a 64 x 4 (i, j) tile and marches FG_KC planes in k. Per
 * plane, the tile's u, v, w plus a one-cell ring in i and j are staged in
 * LDS (three planes rotate), and every stencil operand of F, G and H is read
 * from there: ~35 operands per cell from LDS instead of ~35 global loads
 * (one thread per cell measured 433 us at 256^3, this layout 253 us, same
 * bits). Expressions keep the reference's operand order, with
 * X? = value at (i, j, k) offset: E/W = i+-1, N/S = j+-1, T/B = k+-1.
 * The wall faces of F, G, H (src/discretization.c:599-651) are written by
 * the threads next to them: F(imax) = U(imax) replaces the interior value
 * there, F(0) = U(0) is written by i == 1, and likewise for G in j and H in
 * k. dvwdz keeps the reference's V(i,j,k+1) in both halves of its k-1/2
 * flux (src/discretization.c:543-550). */
#define FG_KC 16
__global__ __launch_bounds__(256) void kComputeFG(const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, double *__restrict__ f,
    double *__restrict__ g, double *__restrict__ h, Geo geo, FGCoef cf, int gridGx, int gridGy, int gridGz,
    const double *__restrict__ dtDev)
{
  if (dtDev) cf.dt = *dtDev;
  constexpr int HW = TX + 2, HH = TY + 2, HN = HW * HH;
  __shared__ double U[3][HH][HW], V[3][HH][HW], W[3][HH][HW];
  int bx, by, bz;
  if (!xcdBlock(gridGx, gridGy, gridGz, bx, by, bz)) return;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
  const int i = bx * TX + tx + 1, j = by * TY + ty + 1;
  const int k0 = bz * FG_KC + 1, k1 = min(k0 + FG_KC, geo.nk + 1);
  /* this thread's ring elements (two at most: HN <= 512); positions past the
   * padded box load a clamped cell, used only by cells outside the interior */
  const int e1 = tid + 256;
  const bool has1 = e1 < HN;
  const int hi0 = tid % HW, hj0 = tid / HW, hi1 = e1 % HW, hj1 = has1 ? e1 / HW : 0;
  const long long o0 = (long long)min(by * TY + hj0, geo.nj + 1) * geo.sj + min(bx * TX + hi0, geo.ni + 1);
  const long long o1 = (long long)min(by * TY + hj1, geo.nj + 1) * geo.sj + min(bx * TX + hi1, geo.ni + 1);
  /* plane k + 2 is loaded into registers while plane k is computed and
   * stored to LDS at the next step, so no step waits on global memory */
  double ru0 = 0.0, rv0 = 0.0, rw0 = 0.0, ru1 = 0.0, rv1 = 0.0, rw1 = 0.0;
  auto fetch = [&](int k) {
    long long base = (long long)min(k, geo.nk + 1) * geo.sk;
    ru0 = u[base + o0];
    rv0 = v[base + o0];
    rw0 = w[base + o0];
    if (has1) {
      ru1 = u[base + o1];
      rv1 = v[base + o1];
      rw1 = w[base + o1];
    }
  };
  auto store = [&](int s) {
    U[s][hj0][hi0] = ru0;
    V[s][hj0][hi0] = rv0;
    W[s][hj0][hi0] = rw0;
    if (has1) {
      U[s][hj1][hi1] = ru1;
      V[s][hj1][hi1] = rv1;
      W[s][hj1][hi1] = rw1;
    }
  };
  fetch(k0 - 1);
  store((k0 - 1) % 3);
  fetch(k0);
  store(k0 % 3);
  fetch(k0 + 1);
  const bool in = i <= geo.ni && j <= geo.nj;
  const int X = tx + 1, Y = ty + 1;
  for (int k = k0; k < k1; k++) {
    store((k + 1) % 3);
    if (k + 1 < k1) fetch(k + 2);
    __syncthreads();
    if (in) {
      const int m = (k + 2) % 3, c0 = k % 3, pz = (k + 1) % 3;
      const long long c = geo.at(i, j, k);
      double Uc = U[c0][Y][X], Vc = V[c0][Y][X], Wc = W[c0][Y][X];
      double uE = U[c0][Y][X + 1], uW = U[c0][Y][X - 1], uN = U[c0][Y + 1][X], uS = U[c0][Y - 1][X];
      double uT = U[pz][Y][X], uB = U[m][Y][X];
      double vE = V[c0][Y][X + 1], vW = V[c0][Y][X - 1], vN = V[c0][Y + 1][X], vS = V[c0][Y - 1][X];
      double vT = V[pz][Y][X], vB = V[m][Y][X];
      double wE = W[c0][Y][X + 1], wW = W[c0][Y][X - 1], wN = W[c0][Y + 1][X], wS = W[c0][Y - 1][X];
      double wT = W[pz][Y][X], wB = W[m][Y][X];
      double fval = i == geo.ni ? Uc
                                : fCell(cf, Uc, uE, uW, uN, uS, uT, uB, Vc, vE, vS, V[c0][Y - 1][X + 1], Wc, wE,
                                      wB, W[m][Y][X + 1]);
      double gval = j == geo.nj ? Vc
                                : gCell(cf, Uc, uN, uW, U[c0][Y + 1][X - 1], Vc, vE, vW, vN, vS, vT, vB, Wc, wN,
                                      wB, W[m][Y + 1][X]);
      double hval = k == geo.nk ? Wc
                                : hCell(cf, Uc, uT, uW, U[pz][Y][X - 1], Vc, vT, vS, V[pz][Y - 1][X], Wc, wE, wW,
                                      wN, wS, wT, wB);
      f[c] = fval;
      g[c] = gval;
      h[c] = hval;
      if (i == 1) f[c - 1] = uW;
      if (j == 1) g[c - geo.sj] = vS;
      if (k == 1) h[c - geo.sk] = wB;
    }
    __syncthreads();
  }
}

/* computeFG and computeRHS in one pass (src/discretization.c:334-362 on the
 * F, G, H this kernel has just formed): rhs(i,j,k) needs F(i-1), G(j-1) and
 * H(k-1). Inside the tile they come from the thread to the left (a wave
 * shuffle), the row below (LDS) and the previous plane (a register). At the
 * tile's low edges they are computed again here, with the same cell
 * functions, hence the same bits as the neighbouring tile's: wave 0 forms
 * G of the row below the tile, FR_TY lanes of wave 1 F of the column left
 * of it, and every thread H of the plane below its k-chunk once. Each such halo
 * cell needs one operand outside the staged ring (V(i,j0-2), U(i0-2,j),
 * W(i,j,k0-2)), loaded from global memory. At the walls the halo value is
 * the reference's face value (F(0) = U(0), G(i,0) = V(i,0), H(i,j,0) =
 * W(i,j,0)). 64 x 4 tiles (256 threads, 148 VGPRs: three workgroups per
 * CU): 251 us at 256^3 against 261 us for 64 x 8 tiles (one workgroup per
 * CU at this register count) and 201 + 125 us for the two kernels. */
#define FR_TY 4
__global__ __launch_bounds__(TX * FR_TY) void kComputeFGRHS(const double *__restrict__ u,
    const double *__restrict__ v, const double *__restrict__ w, double *__restrict__ f,
    double *__restrict__ g, double *__restrict__ h, double *__restrict__ rhs, Geo geo, FGCoef cf, double idt,
    int gridGx, int gridGy, int gridGz, const double *__restrict__ dtDev)
{
  if (dtDev) {
    cf.dt = *dtDev;
    idt   = 1.0 / cf.dt;
  }
  constexpr int HW = TX + 2, HH = FR_TY + 2, HN = HW * HH, NT = TX * FR_TY;
  __shared__ double U[3][HH][HW], V[3][HH][HW], W[3][HH][HW];
  __shared__ double Gs[FR_TY + 1][TX]; /* G of rows j0-1 .. j0+FR_TY-1 of this plane */
  __shared__ double Fh[FR_TY];         /* F(i0-1, j) of this plane */
  int bx, by, bz;
  if (!xcdBlock(gridGx, gridGy, gridGz, bx, by, bz)) return;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
  const int i0 = bx * TX + 1, j0 = by * FR_TY + 1;
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = bz * FG_KC + 1, k1 = min(k0 + FG_KC, geo.nk + 1);
  /* ring elements: two at most (HN <= 2 NT) */
  const int e1 = tid + NT;
  const bool has1 = e1 < HN;
  const int hi0 = tid % HW, hj0 = tid / HW, hi1 = e1 % HW, hj1 = has1 ? e1 / HW : 0;
  const long long o0 = (long long)min(by * FR_TY + hj0, geo.nj + 1) * geo.sj + min(bx * TX + hi0, geo.ni + 1);
  const long long o1 = (long long)min(by * FR_TY + hj1, geo.nj + 1) * geo.sj + min(bx * TX + hi1, geo.ni + 1);
  /* the halo operands: V(i, j0-2) for wave 0, U(i0-2, j0+lane) for lanes
   * 0..FR_TY-1 of wave 1 (clamped into the box; at a wall they are not used) */
  const bool gHalo = ty == 0, fHalo = ty == 1 && tx < FR_TY;
  const long long oHalo = gHalo ? (long long)max(j0 - 2, 0) * geo.sj + min(i, geo.ni + 1)
                                : (long long)min(j0 + tx, geo.nj + 1) * geo.sj + max(i0 - 2, 0);
  double ru0 = 0.0, rv0 = 0.0, rw0 = 0.0, ru1 = 0.0, rv1 = 0.0, rw1 = 0.0;
  auto fetch = [&](int k) {
    long long base = (long long)k * geo.sk;
    ru0 = u[base + o0];
    rv0 = v[base + o0];
    rw0 = w[base + o0];
    if (has1) {
      ru1 = u[base + o1];
      rv1 = v[base + o1];
      rw1 = w[base + o1];
    }
  };
  auto store = [&](int s) {
    U[s][hj0][hi0] = ru0;
    V[s][hj0][hi0] = rv0;
    W[s][hj0][hi0] = rw0;
    if (has1) {
      U[s][hj1][hi1] = ru1;
      V[s][hj1][hi1] = rv1;
      W[s][hj1][hi1] = rw1;
    }
  };
  auto fetchHalo = [&](int k) -> double {
    if (gHalo) return v[(long long)k * geo.sk + oHalo];
    if (fHalo) return u[(long long)k * geo.sk + oHalo];
    return 0.0;
  };
  const bool in = i <= geo.ni && j <= geo.nj;
  const int X = tx + 1, Y = ty + 1;
  /* W(i, j, k0-2): the operand of the halo H outside the ring */
  const double wBB = in ? w[geo.at(i, j, max(k0 - 2, 0))] : 0.0;
  fetch(k0 - 1);
  store((k0 - 1) % 3);
  fetch(k0);
  store(k0 % 3);
  double haloCur = fetchHalo(k0);
  fetch(k0 + 1);
  __syncthreads();
  /* H of the plane below the chunk (k0 - 1), as the thread that owns it
   * computes it (or W there at the bottom wall) */
  double hPrev = 0.0;
  if (in) {
    const int c0 = (k0 - 1) % 3, pz = k0 % 3;
    const double Wc = W[c0][Y][X];
    hPrev = k0 == 1 ? Wc
                    : hCell(cf, U[c0][Y][X], U[pz][Y][X], U[c0][Y][X - 1], U[pz][Y][X - 1], V[c0][Y][X],
                          V[pz][Y][X], V[c0][Y - 1][X], V[pz][Y - 1][X], Wc, W[c0][Y][X + 1], W[c0][Y][X - 1],
                          W[c0][Y + 1][X], W[c0][Y - 1][X], W[pz][Y][X], wBB);
  }
  for (int k = k0; k < k1; k++) {
    store((k + 1) % 3);
    double haloNext = 0.0;
    if (k + 1 < k1) {
      fetch(k + 2);
      haloNext = fetchHalo(k + 1);
    }
    __syncthreads();
    const int m = (k + 2) % 3, c0 = k % 3, pz = (k + 1) % 3;
    double fval = 0.0, gval = 0.0, hval = 0.0;
    if (in) {
      double Uc = U[c0][Y][X], Vc = V[c0][Y][X], Wc = W[c0][Y][X];
      double uE = U[c0][Y][X + 1], uW = U[c0][Y][X - 1], uN = U[c0][Y + 1][X], uS = U[c0][Y - 1][X];
      double uT = U[pz][Y][X], uB = U[m][Y][X];
      double vE = V[c0][Y][X + 1], vW = V[c0][Y][X - 1], vN = V[c0][Y + 1][X], vS = V[c0][Y - 1][X];
      double vT = V[pz][Y][X], vB = V[m][Y][X];
      double wE = W[c0][Y][X + 1], wW = W[c0][Y][X - 1], wN = W[c0][Y + 1][X], wS = W[c0][Y - 1][X];
      double wT = W[pz][Y][X], wB = W[m][Y][X];
      fval = i == geo.ni ? Uc
                         : fCell(cf, Uc, uE, uW, uN, uS, uT, uB, Vc, vE, vS, V[c0][Y - 1][X + 1], Wc, wE, wB,
                               W[m][Y][X + 1]);
      gval = j == geo.nj ? Vc
                         : gCell(cf, Uc, uN, uW, U[c0][Y + 1][X - 1], Vc, vE, vW, vN, vS, vT, vB, Wc, wN, wB,
                               W[m][Y + 1][X]);
      hval = k == geo.nk ? Wc
                         : hCell(cf, Uc, uT, uW, U[pz][Y][X - 1], Vc, vT, vS, V[pz][Y - 1][X], Wc, wE, wW, wN,
                               wS, wT, wB);
      const long long c = geo.at(i, j, k);
      f[c] = fval;
      g[c] = gval;
      h[c] = hval;
      if (i == 1) f[c - 1] = uW;
      if (j == 1) g[c - geo.sj] = vS;
      if (k == 1) h[c - geo.sk] = wB;
    }
    Gs[ty + 1][tx] = gval;
    if (gHalo && i <= geo.ni) { /* G(i, j0-1): row 0 of the ring */
      const int Yh = 0;
      Gs[0][tx] = j0 == 1 ? V[c0][Yh][X]
                          : gCell(cf, U[c0][Yh][X], U[c0][Yh + 1][X], U[c0][Yh][X - 1], U[c0][Yh + 1][X - 1],
                                V[c0][Yh][X], V[c0][Yh][X + 1], V[c0][Yh][X - 1], V[c0][Yh + 1][X], haloCur,
                                V[pz][Yh][X], V[m][Yh][X], W[c0][Yh][X], W[c0][Yh + 1][X], W[m][Yh][X],
                                W[m][Yh + 1][X]);
    }
    if (fHalo && j0 + tx <= geo.nj) { /* F(i0-1, j0+lane): column 0 of the ring */
      const int Yh = tx + 1;
      Fh[tx] = i0 == 1 ? U[c0][Yh][0]
                       : fCell(cf, U[c0][Yh][0], U[c0][Yh][1], haloCur, U[c0][Yh + 1][0], U[c0][Yh - 1][0],
                             U[pz][Yh][0], U[m][Yh][0], V[c0][Yh][0], V[c0][Yh][1], V[c0][Yh - 1][0],
                             V[c0][Yh - 1][1], W[c0][Yh][0], W[c0][Yh][1], W[m][Yh][0], W[m][Yh][1]);
    }
    const double fLeft = __shfl_up(fval, 1, 64);
    __syncthreads();
    if (in) {
      const double fPrev = tx > 0 ? fLeft : Fh[ty];
      const double gPrev = Gs[ty][tx];
      rhs[geo.at(i, j, k)] = ((fval - fPrev) * cf.iDx + (gval - gPrev) * cf.iDy + (hval - hPrev) * cf.iDz) * idt;
    }
    hPrev   = hval;
    haloCur = haloNext;
  }
}

void launchComputeFG(NusifCtx *c, const Discretization *d)
{
  FGCoef cf;
  cf.gamma = d->gamma;
  cf.dt    = d->dt;
  cf.iRe   = 1.0 / d->re;
  cf.iDx   = 1.0 / d->grid.dx;
  cf.iDy   = 1.0 / d->grid.dy;
  cf.iDz   = 1.0 / d->grid.dz;
  cf.gx    = d->gx;
  cf.gy    = d->gy;
  cf.gz    = d->gz;
  const Geo &g = c->geo;
  if (c->opt.fgRhs) {
    /* rhs from the new F, G, H in the same pass, into the spare rhs
     * buffer: computeRHS makes it the rhs (a pointer swap) while f, g, h
     * and dt stay as they are; until then rhs keeps its old values, as the
     * reference's computeFG leaves them */
    dim3 blk(TX, FR_TY);
    dim3 grd((g.ni + TX - 1) / TX, (g.nj + FR_TY - 1) / FR_TY, (g.nk + FG_KC - 1) / FG_KC);
    NUSIF_TIMED("computeFG_rhs", (kComputeFGRHS<<<xcdGrid(grd), blk, 0, nusifStream()>>>(c->u, c->v, c->w,
                                     c->f, c->g, c->h, c->rhsNext, g, cf, 1.0 / d->dt, grd.x, grd.y, grd.z,
                                     dtDevice(c))));
    c->rhsReady = true;
    c->rhsDt    = d->dt;
    return;
  }
  dim3 blk(TX, TY);
  dim3 grd((g.ni + TX - 1) / TX, (g.nj + TY - 1) / TY, (g.nk + FG_KC - 1) / FG_KC);
  NUSIF_TIMED("computeFG", (kComputeFG<<<xcdGrid(grd), blk, 0, nusifStream()>>>(
                                c->u, c->v, c->w, c->f, c->g, c->h, g, cf, grd.x, grd.y, grd.z, dtDevice(c))));
  c->rhsReady = false;
}

/* ------------------------------------------------------------------------ */
/* computeRHS (src/discretization.c:334-362)                                */
/* ------------------------------------------------------------------------ */
__global__ __launch_bounds__(256) void kComputeRHS(const double *__restrict__ f,
    const double *__restrict__ g, const double *__restrict__ h, double *__restrict__ rhs,
    Geo geo, double idx, double idy, double idz, double idt, int gridGx, int gridGy,
    const double *__restrict__ dtDev)
{
  if (dtDev) idt = 1.0 / *dtDev;
  int bx, by, bz;
  if (!xcdBlock(gridGx, gridGy, geo.nk, bx, by, bz)) return;
  int i = bx * blockDim.x + threadIdx.x + 1;
  int j = by * blockDim.y + threadIdx.y + 1;
  int k = bz + 1;
  if (i > geo.ni || j > geo.nj) return;
  long long c = geo.at(i, j, k);
  rhs[c] = ((f[c] - f[c - 1]) * idx + (g[c] - g[c - geo.sj]) * idy + (h[c] - h[c - geo.sk]) * idz) *
           idt;
}

void launchComputeRHS(NusifCtx *c, const Discretization *d)
{
  /* computeFG formed it from these f, g, h (with the same dt: the device's,
   * or the host's if it did not change) */
  if (c->rhsReady && (c->dtOnDevice || c->rhsDt == d->dt)) {
    double *t   = c->rhs;
    c->rhs      = c->rhsNext;
    c->rhsNext  = t;
    c->rhsReady = false;
    return;
  }
  c->rhsReady  = false;
  const Geo &g = c->geo;
  dim3 blk(TX, TY);
  dim3 grd((g.ni + TX - 1) / TX, (g.nj + TY - 1) / TY, g.nk);
  NUSIF_TIMED("computeRHS",
      (kComputeRHS<<<xcdGrid(grd), blk, 0, nusifStream()>>>(c->f, c->g, c->h, c->rhs, g,
          1.0 / d->grid.dx, 1.0 / d->grid.dy, 1.0 / d->grid.dz, 1.0 / d->dt, grd.x, grd.y, dtDevice(c))));
}

/* ------------------------------------------------------------------------ */
/* normalizePressure (src/discretization.c:403-429)                         */
/* ------------------------------------------------------------------------ */
/* Workgroup (64, 4) covers a 64 x 4 patch of (i, j) and marches all k. */
__global__ __launch_bounds__(256) void kInteriorSum(const double *__restrict__ p, Geo geo, double *partials)
{
  __shared__ double lds[4];
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  double acc = 0.0;
  if (i <= geo.ni && j <= geo.nj)
    for (int k = 1; k <= geo.nk; k++) acc += p[geo.at(i, j, k)];
  acc = blockSum(acc, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[blockIdx.x + gridDim.x * blockIdx.y] = acc;
}

/* avg = sum / (imax * jmax * kmax) (src/discretization.c:420) */
__global__ __launch_bounds__(256) void kFoldMean(const double *partials, int nb, double count, DevScalars *sc)
{
  __shared__ double lds[4];
  double s = foldSum(partials, nb, 1, 0, lds);
  if (threadIdx.x == 0) sc->sum = s / count;
}

__global__ __launch_bounds__(256) void kSubtractMean(double *__restrict__ p, Geo geo, const DevScalars *sc, int gridGx, int gridGy)
{
  int bx, by, bz;
  if (!xcdBlock(gridGx, gridGy, geo.nk, bx, by, bz)) return;
  int i = bx * blockDim.x + threadIdx.x + 1;
  int j = by * blockDim.y + threadIdx.y + 1;
  int k = bz + 1;
  if (i > geo.ni || j > geo.nj) return;
  double avg = sc->sum;
  long long c = geo.at(i, j, k);
  p[c]        = p[c] - avg;
}

void launchNormalize(NusifCtx *c)
{
  const Geo &g = c->geo;
  dim3 blk(TX, TY);
  dim3 g2((g.ni + TX - 1) / TX, (g.nj + TY - 1) / TY);
  if ((size_t)g2.x * g2.y > c->nPartials) NUSIF_FATAL("partials buffer too small");
  double count = (double)(c->owner->grid.imax * c->owner->grid.jmax * c->owner->grid.kmax);
  kInteriorSum<<<g2, blk, 0, nusifStream()>>>(c->p, g, c->partials);
  kFoldMean<<<1, 256, 0, nusifStream()>>>(c->partials, (int)(g2.x * g2.y), count, c->scal);
  dim3 g3(g2.x, g2.y, g.nk);
  kSubtractMean<<<xcdGrid(g3), blk, 0, nusifStream()>>>(c->p, g, c->scal, g3.x, g3.y);
  c->pGhostsOk = false; /* the interior moved, its ghosts did not */
}

/* ------------------------------------------------------------------------ */
/* adaptUV (src/discretization.c:654-681)                                    */
/* ------------------------------------------------------------------------ */
/* A thread updates one (i, j) column over ADAPT_KC planes (p at k+1 is kept
 * for the next plane) and keeps the maxima of |u|, |v|, |w| it wrote: the
 * interior-range part of the next computeTimestep's CFL maxima, one partial
 * per workgroup (logical block order). */
#define ADAPT_KC 4
__global__ __launch_bounds__(256) void kAdaptUV(const double *__restrict__ f,
    const double *__restrict__ g, const double *__restrict__ h, const double *__restrict__ p,
    double *__restrict__ u, double *__restrict__ v, double *__restrict__ w, Geo geo, double fx,
    double fy, double fz, double *__restrict__ partials, int gridGx, int gridGy, int gridGz,
    const double *__restrict__ dtDev, double dx, double dy, double dz)
{
  if (dtDev) { /* dt / dx etc., as the host forms them */
    const double dt = *dtDev;
    fx              = dt / dx;
    fy              = dt / dy;
    fz              = dt / dz;
  }
  __shared__ double lds[4];
  int bx, by, bz;
  if (!xcdBlock(gridGx, gridGy, gridGz, bx, by, bz)) return;
  int i = bx * blockDim.x + threadIdx.x + 1;
  int j = by * blockDim.y + threadIdx.y + 1;
  int k0 = bz * ADAPT_KC + 1, k1 = min(k0 + ADAPT_KC, geo.nk + 1);
  double mu = DBL_MIN, mv = DBL_MIN, mw = DBL_MIN;
  if (i <= geo.ni && j <= geo.nj) {
    long long c = geo.at(i, j, k0);
    double pc   = p[c];
    for (int k = k0; k < k1; k++, c += geo.sk) {
      double pz = p[c + geo.sk];
      double a  = f[c] - (p[c + 1] - pc) * fx;
      double b  = g[c] - (p[c + geo.sj] - pc) * fy;
      double d  = h[c] - (pz - pc) * fz;
      u[c]      = a;
      v[c]      = b;
      w[c]      = d;
      a         = fabs(a);
      b         = fabs(b);
      d         = fabs(d);
      mu        = mu > a ? mu : a;
      mv        = mv > b ? mv : b;
      mw        = mw > d ? mw : d;
      pc        = pz;
    }
  }
  mu = blockMax(mu, lds);
  mv = blockMax(mv, lds);
  mw = blockMax(mw, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    unsigned bid          = bx + gridGx * (by + gridGy * bz);
    partials[3 * bid + 0] = mu;
    partials[3 * bid + 1] = mv;
    partials[3 * bid + 2] = mw;
  }
}

static dim3 adaptGrid(const Geo &g) { return dim3((g.ni + TX - 1) / TX, (g.nj + TY - 1) / TY, (g.nk + ADAPT_KC - 1) / ADAPT_KC); }

void launchAdaptUV(NusifCtx *c, const Discretization *d)
{
  const Geo &g = c->geo;
  dim3 blk(TX, TY);
  dim3 grd = adaptGrid(g);
  NUSIF_TIMED("adaptUV", (kAdaptUV<<<xcdGrid(grd), blk, 0, nusifStream()>>>(c->f, c->g, c->h,
                             c->p, c->u, c->v, c->w, g, d->dt / d->grid.dx, d->dt / d->grid.dy,
                             d->dt / d->grid.dz, c->cflPartials, grd.x, grd.y, grd.z, dtDevice(c), d->grid.dx,
                             d->grid.dy, d->grid.dz)));
  c->cflValid = true;
}

/* ------------------------------------------------------------------------ */
/* Public API                                                                */
/* ------------------------------------------------------------------------ */
static double *hostField(size_t n, double value)
{
  double *a = nullptr;
  HIP_CHECK(hipHostMalloc((void **)&a, n * sizeof(double), hipHostMallocDefault));
  for (size_t q = 0; q < n; q++) a[q] = value;
  return a;
}

static double *deviceField(size_t n)
{
  double *a = nullptr;
  HIP_CHECK(hipMalloc((void **)&a, n * sizeof(double)));
  return a;
}

/* Deterministic inflow perturbation: xorshift64* stream, uniform in [-1, 1]. */
static void buildInflowProfile(NusifCtx *c)
{
  const Geo &g = c->geo;
  size_t n     = (size_t)(g.nj + 2) * (g.nk + 2);
  double *h    = (double *)malloc(n * sizeof(double));
  unsigned long long x = 0x9E3779B97F4A7C15ull ^ (unsigned long long)(c->opt.seed + 1);
  for (size_t q = 0; q < n; q++) {
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    unsigned long long r = x * 0x2545F4914F6CDD1Dull;
    double xi            = (double)(r >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    h[q]                 = c->opt.uInflow * (1.0 + c->opt.inflowNoise * xi);
  }
  HIP_CHECK(hipMalloc((void **)&c->inflow, n * sizeof(double)));
  HIP_CHECK(hipMemcpy(c->inflow, h, n * sizeof(double), hipMemcpyHostToDevice));
  free(h);
}

/* src/discretization.c:49-117 */
static void printConfiguration(const Discretization *s);

extern "C" void initDiscretization(Discretization *s, Parameter *params)
{
  s->problem  = params->name;
  s->bcLeft   = params->bcLeft;
  s->bcRight  = params->bcRight;
  s->bcBottom = params->bcBottom;
  s->bcTop    = params->bcTop;
  s->bcFront  = params->bcFront;
  s->bcBack   = params->bcBack;

  s->grid.imax    = params->imax;
  s->grid.jmax    = params->jmax;
  s->grid.kmax    = params->kmax;
  s->grid.xlength = params->xlength;
  s->grid.ylength = params->ylength;
  s->grid.zlength = params->zlength;
  s->grid.dx      = params->xlength / params->imax;
  s->grid.dy      = params->ylength / params->jmax;
  s->grid.dz      = params->zlength / params->kmax;

  s->eps     = params->eps;
  s->omega   = params->omg;
  s->itermax = params->itermax;
  s->re      = params->re;
  s->gx      = params->gx;
  s->gy      = params->gy;
  s->gz      = params->gz;
  s->dt      = params->dt;
  s->te      = params->te;
  s->tau     = params->tau;
  s->gamma   = params->gamma;

  if (s->comm.imaxLocal <= 0 || s->comm.jmaxLocal <= 0 || s->comm.kmaxLocal <= 0)
    NUSIF_FATAL("call commPartition before initDiscretization (local sizes %d %d %d)",
        s->comm.imaxLocal, s->comm.jmaxLocal, s->comm.kmaxLocal);

  NusifCtx *c = new NusifCtx();
  memset(c, 0, sizeof(*c));
  c->owner = s;
  c->geo   = makeGeo(s->comm.imaxLocal, s->comm.jmaxLocal, s->comm.kmaxLocal);
  nusifConsumeOptions(params, &c->opt);
  size_t n = (size_t)c->geo.n;

  s->u   = hostField(n, params->u_init);
  s->v   = hostField(n, params->v_init);
  s->w   = hostField(n, params->w_init);
  s->p   = hostField(n, params->p_init);
  s->rhs = hostField(n, 0.0);
  s->f   = hostField(n, 0.0);
  s->g   = hostField(n, 0.0);
  s->h   = hostField(n, 0.0);
  double *hp[8] = {s->p, s->rhs, s->f, s->g, s->h, s->u, s->v, s->w};
  memcpy(c->host, hp, sizeof(hp));

  c->p    = deviceField(n);
  c->rhs  = deviceField(n);
  c->f    = deviceField(n);
  c->g    = deviceField(n);
  c->h    = deviceField(n);
  c->u    = deviceField(n);
  c->v    = deviceField(n);
  c->w    = deviceField(n);
  c->pAlt = deviceField(n);
  c->rhsNext = deviceField(n);
  c->nPartials = 65536;
  HIP_CHECK(hipMalloc((void **)&c->partials, c->nPartials * sizeof(double)));
  {
    dim3 ag              = adaptGrid(c->geo);
    c->cflInteriorBlocks = (int)(ag.x * ag.y * ag.z);
    HIP_CHECK(hipMalloc((void **)&c->cflPartials, 3 * ((size_t)c->cflInteriorBlocks + SHELL_BLOCKS) * sizeof(double)));
    c->cflValid = false;
    const long long shell = 2LL * c->geo.sk + (long long)c->geo.nk * (2LL * c->geo.sj + 2LL * c->geo.nj);
    if (shell >= (1LL << 31)) NUSIF_FATAL("grid too large: %lld ghost-shell cells", shell);
  }
  c->rhsReady = false;
  c->rhsDt    = 0.0;
  c->saveP    = nullptr;
  c->saveRhs  = nullptr;
  c->dtOnDevice = false;
  HIP_CHECK(hipEventCreateWithFlags(&c->dtEvent, hipEventDisableTiming));
  c->solveTiming  = 0;
  c->solveEvUsed  = 0;
  c->solveEvReady = false;
  c->solveMs      = 0.0;
  c->solveCount   = 0;
  HIP_CHECK(hipMalloc((void **)&c->scal, sizeof(DevScalars)));
  HIP_CHECK(hipHostMalloc((void **)&c->scalHost, sizeof(DevScalars), hipHostMallocCoherent));
  HIP_CHECK(hipHostGetDevicePointer((void **)&c->scalHostDev, c->scalHost, 0));
  memset(c->scalHost, 0, sizeof(DevScalars));
  HIP_CHECK(hipMemcpy(c->scal, c->scalHost, sizeof(DevScalars), hipMemcpyHostToDevice));
  if (c->opt.inflowNoise != 0.0) buildInflowProfile(c);
  nusifRegisterCtx(c);
  nusifSyncDevice(s);
  HIP_CHECK(hipMemcpyAsync(c->pAlt, c->p, n * sizeof(double), hipMemcpyDeviceToDevice, nusifStream()));

  double dx = s->grid.dx, dy = s->grid.dy, dz = s->grid.dz;
  double invSqrSum = 1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz);
  s->dtBound       = 0.5 * s->re * 1.0 / invSqrSum;
  if (c->opt.verbose >= 1) printConfiguration(s);
}

/* What the reference's initDiscretization prints when built with -DVERBOSE
 * (src/discretization.c:16-47, 114-116; the sequential communicator prints
 * nothing more): the same lines, in the same formats. */
static void printConfiguration(const Discretization *s)
{
  if (!commIsMaster((CommType *)&s->comm)) return;
  const Grid &g = s->grid;
  printf("Parameters for #%s#\n", s->problem);
  printf("BC Left:%d Right:%d Bottom:%d Top:%d Front:%d Back:%d\n", s->bcLeft, s->bcRight, s->bcBottom,
      s->bcTop, s->bcFront, s->bcBack);
  printf("\tReynolds number: %.2f\n", s->re);
  printf("\tGx Gy: %.2f %.2f %.2f\n", s->gx, s->gy, s->gz);
  printf("Geometry data:\n");
  printf("\tDomain box size (x, y, z): %.2f, %.2f, %.2f\n", g.xlength, g.ylength, g.zlength);
  printf("\tCells (x, y, z): %d, %d, %d\n", g.imax, g.jmax, g.kmax);
  printf("\tCell size (dx, dy, dz): %f, %f, %f\n", g.dx, g.dy, g.dz);
  printf("Timestep parameters:\n");
  printf("\tDefault stepsize: %.2f, Final time %.2f\n", s->dt, s->te);
  printf("\tdt bound: %.6f\n", s->dtBound);
  printf("\tTau factor: %.2f\n", s->tau);
  printf("Iterative parameters:\n");
  printf("\tMax iterations: %d\n", s->itermax);
  printf("\tepsilon (stopping tolerance) : %f\n", s->eps);
  printf("\tgamma factor: %f\n", s->gamma);
  printf("\tomega (SOR relaxation): %f\n", s->omega);
}

extern "C" void nusifSyncHost(Discretization *d)
{
  NusifCtx *c   = nusifCtxOf(d);
  size_t bytes  = (size_t)c->geo.n * sizeof(double);
  double *dev[8] = {c->p, c->rhs, c->f, c->g, c->h, c->u, c->v, c->w};
  for (int q = 0; q < 8; q++)
    HIP_CHECK(hipMemcpyAsync(c->host[q], dev[q], bytes, hipMemcpyDeviceToHost, nusifStream()));
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
}

/* The spare rhs buffer mirrors rhs after every host upload: computeFG
 * writes only its interior, so the ghost cells a computeRHS swap brings in
 * are the ones rhs had (the reference's computeRHS leaves them alone). */
void nusifRhsSpareSync(NusifCtx *c)
{
  HIP_CHECK(hipMemcpyAsync(c->rhsNext, c->rhs, (size_t)c->geo.n * sizeof(double), hipMemcpyDeviceToDevice,
      nusifStream()));
  c->rhsReady = false;
}

extern "C" void nusifSyncDevice(Discretization *d)
{
  NusifCtx *c   = nusifCtxOf(d);
  size_t bytes  = (size_t)c->geo.n * sizeof(double);
  double *dev[8] = {c->p, c->rhs, c->f, c->g, c->h, c->u, c->v, c->w};
  for (int q = 0; q < 8; q++)
    HIP_CHECK(hipMemcpyAsync(dev[q], c->host[q], bytes, hipMemcpyHostToDevice, nusifStream()));
  nusifRhsSpareSync(c);
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  c->cflValid  = false;
  c->pGhostsOk = false;
}

extern "C" void nusifFreeDiscretization(Discretization *d)
{
  NusifCtx *c = nusifCtxOf(d);
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  double *dev[10] = {c->p, c->rhs, c->f, c->g, c->h, c->u, c->v, c->w, c->pAlt, c->rhsNext};
  for (int q = 0; q < 10; q++) HIP_CHECK(hipFree(dev[q]));
  HIP_CHECK(hipEventDestroy(c->dtEvent));
  if (c->solveEvReady)
    for (int q = 0; q < 2 * SOLVE_EV_POOL; q++) HIP_CHECK(hipEventDestroy(c->solveEv[q]));
  if (c->saveP) HIP_CHECK(hipFree(c->saveP));
  if (c->saveRhs) HIP_CHECK(hipFree(c->saveRhs));
  HIP_CHECK(hipFree(c->partials));
  HIP_CHECK(hipFree(c->cflPartials));
  HIP_CHECK(hipFree(c->scal));
  if (c->inflow) HIP_CHECK(hipFree(c->inflow));
  HIP_CHECK(hipHostFree(c->scalHost));
  for (int q = 0; q < 8; q++) HIP_CHECK(hipHostFree(c->host[q]));
  d->p = d->rhs = d->f = d->g = d->h = d->u = d->v = d->w = nullptr;
  nusifUnregisterCtx(c);
  delete c;
}

/* src/discretization.c:431-453. The maxima are reduced on the GPU; the dt
 * arithmetic is the reference's, on the host, so d->dt is exact. */
extern "C" void computeTimestep(Discretization *s)
{
  NusifCtx *c = nusifCtxOf(s);
  launchAbsMax3(c);
  HIP_CHECK(hipStreamSynchronize(nusifStream())); /* kFoldMax3 wrote the pinned mirror */
  double m[3];
  for (int q = 0; q < 3; q++) memcpy(&m[q], &c->scalHost->absmax[q], sizeof(double));
  double dt = s->dtBound;
  double dx = s->grid.dx, dy = s->grid.dy, dz = s->grid.dz;
  if (m[0] > 0) dt = (dt > dx / m[0]) ? dx / m[0] : dt;
  if (m[1] > 0) dt = (dt > dy / m[1]) ? dy / m[1] : dt;
  if (m[2] > 0) dt = (dt > dz / m[2]) ? dz / m[2] : dt;
  s->dt = dt * s->tau;
}

extern "C" void setBoundaryConditions(Discretization *s) { launchVelocityBC(nusifCtxOf(s), s); }
extern "C" void setSpecialBoundaryCondition(Discretization *s) { launchSpecialBC(nusifCtxOf(s), s); }
extern "C" void computeFG(Discretization *s) { launchComputeFG(nusifCtxOf(s), s); }
extern "C" void computeRHS(Discretization *s) { launchComputeRHS(nusifCtxOf(s), s); }
extern "C" void normalizePressure(Discretization *s) { launchNormalize(nusifCtxOf(s)); }
extern "C" void adaptUV(Discretization *s) { launchAdaptUV(nusifCtxOf(s), s); }
