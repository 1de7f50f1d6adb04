/* Red-black SOR pressure solves on the GPU: the reference's RB solver
 * (src/solver-rb.c:27-172) and the observable fine-grid behaviour of its MG
 * solver (src/solver-mg.c:299-402).
 *
 * Colour passes: pass 0 updates cells with i+j+k odd, pass 1 even
 * (src/solver-rb.c:55-79). A cell of one colour only reads cells of the other
 * colour, so a pass is order-free and every thread can update its cell
 * independently: p values are bit-identical to the sequential sweep; only the
 * sum of r^2 is accumulated in a different order.
 *
 * Stop test on the device: the iteration loop runs as batches of launches;
 * each kernel of iteration q returns at once when q >= stopAt, and the last
 * workgroup of iteration q's final kernel folds the residual and lowers
 * stopAt when `res >= eps^2 && it < itermax` fails (src/solver-rb.c:54). The
 * host reads the state once per batch. */
#include <limits.h>
#include <math.h>

#include <utility>

#include "nusif_internal.h"
#include "reduce.h"

#define RB_TX 64
#define RB_TY 4
#define RB_KCH 8

struct RBState {
  double res;
  int it;
  int stopAt;
};

SweepCoef sweepCoef(const Grid *g, double omega)
{
  SweepCoef c;
  double dx2 = g->dx * g->dx, dy2 = g->dy * g->dy, dz2 = g->dz * g->dz;
  c.idx2   = 1.0 / dx2;
  c.idy2   = 1.0 / dy2;
  c.idz2   = 1.0 / dz2;
  c.factor = omega * 0.5 * (dx2 * dy2 * dz2) / (dy2 * dz2 + dx2 * dz2 + dx2 * dy2);
  return c;
}

/* r = rhs - Lap7(p), operand order of src/solver-rb.c:66-69 */
__device__ inline double defect7(const double *p, double rhsc, long long c, long long sj,
    long long sk, const SweepCoef &cf)
{
  double pc = p[c];
  return rhsc - ((p[c + 1] - 2.0 * pc + p[c - 1]) * cf.idx2 +
                    (p[c + sj] - 2.0 * pc + p[c - sj]) * cf.idy2 +
                    (p[c + sk] - 2.0 * pc + p[c - sk]) * cf.idz2);
}

/* One colour pass. Thread x covers the x-th cell of its colour along a row,
 * the workgroup covers RB_TY rows and RB_KCH planes. */
__global__ __launch_bounds__(256) void kColourPass(double *__restrict__ p,
    const double *__restrict__ rhs, Geo g, SweepCoef cf, int parity, int q,
    const RBState *st, double *__restrict__ partials, int accumulate)
{
  __shared__ double lds[4];
  if (q >= st->stopAt) return;
  int x  = blockIdx.x * blockDim.x + threadIdx.x;
  int j  = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k0 = blockIdx.z * RB_KCH + 1;
  double acc = 0.0;
  if (j <= g.nj) {
    int k1 = min(k0 + RB_KCH, g.nk + 1);
    for (int k = k0; k < k1; k++) {
      int i = 2 * x + (((1 + j + k) & 1) == parity ? 1 : 2);
      if (i <= g.ni) {
        long long c = g.at(i, j, k);
        double r    = defect7(p, rhs[c], c, g.sj, g.sk, cf);
        p[c] -= cf.factor * r;
        acc += r * r;
      }
    }
  }
  if (accumulate) {
    acc = blockSum(acc, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0)
      partials[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = acc;
  }
}

/* Neumann ghost copy on the six faces over their interior ranges
 * (src/solver-rb.c:129-148): thread (a, b) serves face points of all three
 * face pairs that contain it. */
__device__ inline void ghostsAt(double *p, const Geo &g, int a, int b)
{
  if (a <= g.ni && b <= g.nj) { /* FRONT / BACK: (i, j) */
    p[g.at(a, b, 0)]        = p[g.at(a, b, 1)];
    p[g.at(a, b, g.nk + 1)] = p[g.at(a, b, g.nk)];
  }
  if (a <= g.ni && b <= g.nk) { /* BOTTOM / TOP: (i, k) */
    p[g.at(a, 0, b)]        = p[g.at(a, 1, b)];
    p[g.at(a, g.nj + 1, b)] = p[g.at(a, g.nj, b)];
  }
  if (a <= g.nj && b <= g.nk) { /* LEFT / RIGHT: (j, k) */
    p[g.at(0, a, b)]        = p[g.at(1, a, b)];
    p[g.at(g.ni + 1, a, b)] = p[g.at(g.ni, a, b)];
  }
}

__global__ __launch_bounds__(256) void kPressureGhosts(double *p, Geo g)
{
  int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int b = blockIdx.y * blockDim.y + threadIdx.y + 1;
  ghostsAt(p, g, a, b);
}

static dim3 ghostGrid(const Geo &g)
{
  int ea = g.ni > g.nj ? g.ni : g.nj;
  int eb = g.nj > g.nk ? g.nj : g.nk;
  return dim3((ea + 63) / 64, (eb + 3) / 4);
}

void launchPressureGhosts(const Geo &g, double *p, hipStream_t st)
{
  kPressureGhosts<<<ghostGrid(g), dim3(64, 4), 0, st>>>(p, g);
}

/* Iteration q's ghost fill (src/solver-rb.c:129-148) */
__global__ __launch_bounds__(256) void kRBGhosts(double *p, Geo g, int q, const RBState *st)
{
  if (q >= st->stopAt) return;
  int a = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int b = blockIdx.y * blockDim.y + threadIdx.y + 1;
  ghostsAt(p, g, a, b);
}

/* Iteration q's fold: the partials of both passes (pass 0 first; one set
 * when one launch ran the whole iteration) into the carried residual
 * (src/solver-rb.c:150-151), then the stop test (src/solver-rb.c:54). */
__global__ __launch_bounds__(256) void kRBFold(int q, RBState *st, const double *__restrict__ partials,
    int nPart, int sets, double n, double epssq, int itermax)
{
  __shared__ double lds[4];
  if (q >= st->stopAt) return;
  double s0 = foldSum(partials, nPart, 1, 0, lds);
  double s1 = sets == 2 ? foldSum(partials + nPart, nPart, 1, 0, lds) : 0.0;
  if (threadIdx.x == 0) {
    double res = (st->res + (s0 + s1)) / n;
    int it     = q + 1;
    st->res    = res;
    st->it     = it;
    if (!(res >= epssq && it < itermax)) st->stopAt = it;
  }
}

/* MG residual (src/solver-mg.c:226-289): mean square of rhs - Lap(p) over the
 * interior, seeded with 1.0 (src/solver-mg.c:252). */
__global__ __launch_bounds__(256) void kResidualMS(const double *__restrict__ p,
    const double *__restrict__ rhs, Geo g, SweepCoef cf, double *partials)
{
  __shared__ double lds[4];
  int i  = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j  = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k0 = blockIdx.z * RB_KCH + 1;
  double acc = 0.0;
  if (i <= g.ni && j <= g.nj) {
    int k1 = min(k0 + RB_KCH, g.nk + 1);
    for (int k = k0; k < k1; k++) {
      long long c = g.at(i, j, k);
      double r    = defect7(p, rhs[c], c, g.sj, g.sk, cf);
      acc += r * r;
    }
  }
  acc = blockSum(acc, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0)
    partials[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = acc;
}

__global__ __launch_bounds__(256) void kFoldMS(const double *partials, int nb, double n, RBState *st)
{
  __shared__ double lds[4];
  double s = foldSum(partials, nb, 1, 0, lds);
  if (threadIdx.x == 0) st->res = (1.0 + s) / n;
}

/* ---------------------------------------------------------------------- */
struct SorLaunch {
  dim3 blk, grd;
  int nPart;
};

static SorLaunch sorShape(const Geo &g)
{
  SorLaunch L;
  L.blk   = dim3(RB_TX, RB_TY);
  int half = (g.ni + 1) / 2;
  L.grd   = dim3((half + RB_TX - 1) / RB_TX, (g.nj + RB_TY - 1) / RB_TY,
      (g.nk + RB_KCH - 1) / RB_KCH);
  L.nPart = (int)(L.grd.x * L.grd.y * L.grd.z);
  return L;
}

static RBState *stateBuf(NusifCtx *c) { return (RBState *)&c->scal->res; }
static RBState *stateHost(NusifCtx *c) { return (RBState *)&c->scalHost->res; }

static void putState(NusifCtx *c, double res, int it, int stopAt)
{
  RBState *h = stateHost(c);
  h->res     = res;
  h->it      = it;
  h->stopAt  = stopAt;
  HIP_CHECK(hipMemcpyAsync(stateBuf(c), h, sizeof(RBState), hipMemcpyHostToDevice, nusifStream()));
}

static void getState(NusifCtx *c)
{
  HIP_CHECK(hipMemcpyAsync(stateHost(c), stateBuf(c), sizeof(RBState), hipMemcpyDeviceToHost,
      nusifStream()));
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
}

static void launchPass(NusifCtx *c, const Geo &g, const SorLaunch &L, double *p, const double *rhs,
    const SweepCoef &cf, int parity, int q, double *partials, int accumulate)
{
  NUSIF_TIMED("rb_pass", (kColourPass<<<L.grd, L.blk, 0, nusifStream()>>>(
                             p, rhs, g, cf, parity, q, stateBuf(c), partials, accumulate)));
}

/* src/solver-rb.c:27-172 */
double rbSolve(NusifSolverCtx *sc, double *p, const double *rhs, double **pOut)
{
  NusifCtx *c       = sc->dctx;
  const Solver *s   = sc->owner;
  const Geo &g      = c->geo;
  SweepCoef cf      = sweepCoef(s->grid, s->omega);
  SorLaunch L       = sorShape(g);
  double epssq      = s->eps * s->eps;
  double n          = (double)(s->grid->imax * s->grid->jmax * s->grid->kmax);
  int itermax       = s->itermax;
  /* fused: iteration q reads buf[q & 1] and writes buf[(q + 1) & 1]; the
   * partner starts as a copy of p (its edge and corner ghosts, which no
   * iteration writes, stay equal to p's) */
  const bool fused = c->opt.rbFused && rbFusedFits(g);
  const int nPart  = fused ? rbFusedTiles(g) : L.nPart;
  if ((size_t)(fused ? 1 : 2) * nPart > c->nPartials) NUSIF_FATAL("partials buffer too small");
  double *buf[2] = {p, c->pAlt};
  if (fused)
    HIP_CHECK(hipMemcpyAsync(c->pAlt, p, (size_t)g.n * sizeof(double), hipMemcpyDeviceToDevice, nusifStream()));
  *pOut = p;

  int stopAt = (1.0 >= epssq && 0 < itermax) ? itermax : 0;
  putState(c, 1.0, 0, stopAt);
  dim3 gg   = ghostGrid(g);
  int launched = 0;
  int batch    = sc->predicted > 0 ? sc->predicted : 8;
  while (launched < stopAt) {
    int end = launched + batch;
    if (end > itermax) end = itermax;
    for (int q = launched; q < end; q++) {
      if (fused) {
        launchRBFused(buf[q & 1], buf[(q + 1) & 1], rhs, g, cf, c->partials, &stateBuf(c)->stopAt, q);
        kRBFold<<<1, 256, 0, nusifStream()>>>(q, stateBuf(c), c->partials, nPart, 1, n, epssq, itermax);
        continue;
      }
      launchPass(c, g, L, p, rhs, cf, 1, q, c->partials, 1);
      launchPass(c, g, L, p, rhs, cf, 0, q, c->partials + L.nPart, 1);
      kRBGhosts<<<gg, dim3(64, 4), 0, nusifStream()>>>(p, g, q, stateBuf(c));
      kRBFold<<<1, 256, 0, nusifStream()>>>(q, stateBuf(c), c->partials, L.nPart, 2, n, epssq, itermax);
    }
    launched = end;
    getState(c);
    stopAt = stateHost(c)->stopAt;
    batch  = 8;
  }
  getState(c);
  sc->lastIterations = stateHost(c)->it;
  sc->predicted      = sc->lastIterations;
  if (fused) *pOut = buf[sc->lastIterations & 1];
  return stateHost(c)->res;
}

/* Observable behaviour of src/solver-mg.c:299-402 on p. The coarse levels'
 * work never reaches p or the returned residual (`correct` adds s->e[0],
 * which nothing writes; src/solver-mg.c:87,96,334), so it is not run. */
double mgRefSolve(NusifSolverCtx *sc, double *p, const double *rhs, double **pOut)
{
  NusifCtx *c     = sc->dctx;
  const Solver *s = sc->owner;
  const Geo &g    = c->geo;
  SweepCoef cf    = sweepCoef(s->grid, s->omega);
  SorLaunch L     = sorShape(g);
  double n        = (double)(s->grid->imax * s->grid->jmax * s->grid->kmax);
  putState(c, 1.0, 0, INT_MAX);
  bool coarsest = (s->levels - 1 == 0) || g.ni < 2 || g.nj < 2 || g.nk < 2;
  /* a sweep followed by the ghost copy is one fused launch into the partner
   * buffer (which starts as a copy of p, so the edge and corner ghosts no
   * sweep writes agree); without the ghost copy (the coarsest branch) it
   * stays two colour passes in place */
  const bool fused = c->opt.rbFused && rbFusedFits(g) && !coarsest;
  if (fused)
    HIP_CHECK(hipMemcpyAsync(c->pAlt, p, (size_t)g.n * sizeof(double), hipMemcpyDeviceToDevice, nusifStream()));
  double *other = c->pAlt;
  *pOut         = p;
  auto smooth = [&](bool ghosts) {
    if (fused) {
      launchRBFused(p, other, rhs, g, cf, nullptr, nullptr, 0);
      std::swap(p, other);
      *pOut = p;
      return;
    }
    launchPass(c, g, L, p, rhs, cf, 1, 0, c->partials, 0);
    launchPass(c, g, L, p, rhs, cf, 0, 0, c->partials, 0);
    if (ghosts) launchPressureGhosts(g, p, nusifStream());
  };
  dim3 rgrd((g.ni + 63) / 64, (g.nj + 3) / 4, (g.nk + RB_KCH - 1) / RB_KCH);
  if ((size_t)rgrd.x * rgrd.y * rgrd.z > c->nPartials) NUSIF_FATAL("partials buffer too small");
  auto residual = [&]() {
    kResidualMS<<<rgrd, dim3(64, 4), 0, nusifStream()>>>(p, rhs, g, cf, c->partials);
    kFoldMS<<<1, 256, 0, nusifStream()>>>(c->partials, (int)(rgrd.x * rgrd.y * rgrd.z), n, stateBuf(c));
  };
  if (coarsest) {
    for (int q = 0; q < s->presmooth; q++) smooth(false);
    residual();
    sc->lastIterations = s->presmooth;
  } else {
    for (int q = 0; q < s->presmooth; q++) smooth(true);
    residual();
    launchPressureGhosts(g, p, nusifStream());
    for (int q = 0; q < s->postsmooth; q++) smooth(true);
    sc->lastIterations = s->presmooth + s->postsmooth;
  }
  getState(c);
  return stateHost(c)->res;
}

/* ---- allocation-free red-black solve ---------------------------------------
 * nusifSolveRBNoAlloc: the reference's red-black SOR solve
 * (src/solver-rb.c:27-172) on device arrays the caller owns, with every
 * buffer beyond p and rhs in `scratch` below (no hipMalloc, no hipFree).
 * The kernels are those of solve() with rb_fused 0 (one launch per colour
 * pass, the ghost copy, the fold), so p, the residual and the iteration
 * count are those of that solve bit for bit.
 *
 * scratch layout (doubles): [0, 32) the RBState (res, it, stopAt; 256 B),
 * then NOALLOC_PARTIALS partial sums of r^2, nPart per colour pass: 514 KiB
 * in all, enough for every grid with imax, jmax, kmax <= 512 (nPart = 32768
 * at 512^3). Each GPU holds one copy for the life of the process; every
 * call on a GPU uses it, and the RBState is rewritten at the start of each
 * call (the partials are written before they are read), so nothing carries
 * over between calls. Calls that overlap on one GPU (two host threads) race
 * on it. */
#define NOALLOC_STATE 32
#define NOALLOC_PARTIALS 65536
#define NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[NOALLOC_STATE + NOALLOC_PARTIALS];
static double *gScratchDev[NOALLOC_MAX_DEVICES]; /* scratch's address on each device (set up on first use) */

__global__ void kRBStateInit(RBState *st, double res, int it, int stopAt)
{
  if (threadIdx.x == 0) {
    st->res    = res;
    st->it     = it;
    st->stopAt = stopAt;
  }
}

/* the device address of scratch on the current device, looked up once per
 * device with that device current; null (and a message) on failure */
static double *noAllocScratch()
{
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    fprintf(stderr, "nusif: nusifSolveRBNoAlloc: no usable current device (%d)\n", dev);
    return nullptr;
  }
  if (!gScratchDev[dev]) {
    void *ptr      = nullptr;
    hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || !ptr) {
      fprintf(stderr, "nusif: nusifSolveRBNoAlloc: hipGetSymbolAddress failed on device %d: %s\n", dev,
          hipGetErrorString(err));
      return nullptr;
    }
    gScratchDev[dev] = (double *)ptr;
  }
  return gScratchDev[dev];
}

extern "C" int nusifSolveRBNoAlloc(double *p, const double *rhs, int imax, int jmax, int kmax, double dx,
    double dy, double dz, double omega, double eps, int itermax, double *res, int *iterations)
{
  if (!p || !rhs || imax < 1 || jmax < 1 || kmax < 1 || !(dx > 0.0) || !(dy > 0.0) || !(dz > 0.0)) {
    fprintf(stderr, "nusif: nusifSolveRBNoAlloc: bad arguments (grid %d x %d x %d)\n", imax, jmax, kmax);
    return -1;
  }
  const Geo g       = makeGeo(imax, jmax, kmax);
  const SorLaunch L = sorShape(g);
  if (2LL * L.nPart > NOALLOC_PARTIALS) {
    fprintf(stderr,
        "nusif: nusifSolveRBNoAlloc: grid %d x %d x %d needs %lld partial sums, more than the %d its fixed "
        "scratch holds (every grid with imax, jmax, kmax <= 512 fits)\n",
        imax, jmax, kmax, 2LL * L.nPart, NOALLOC_PARTIALS);
    return -2;
  }
  double *base = noAllocScratch();
  if (!base) return -3;
  RBState *st      = (RBState *)base;
  double *partials = base + NOALLOC_STATE;
  Grid grid;
  grid.dx = dx;
  grid.dy = dy;
  grid.dz = dz;
  const SweepCoef cf = sweepCoef(&grid, omega);
  const double epssq = eps * eps;
  const double n     = (double)imax * jmax * kmax;
  hipStream_t s      = nusifStream();
  int stopAt         = (1.0 >= epssq && 0 < itermax) ? itermax : 0;
  kRBStateInit<<<1, 64, 0, s>>>(st, 1.0, 0, stopAt);
  const dim3 gg = ghostGrid(g);
  RBState h     = {1.0, 0, stopAt};
  int launched = 0;
  while (launched < stopAt) {
    int end = launched + 8 < itermax ? launched + 8 : itermax;
    for (int q = launched; q < end; q++) {
      kColourPass<<<L.grd, L.blk, 0, s>>>(p, rhs, g, cf, 1, q, st, partials, 1);
      kColourPass<<<L.grd, L.blk, 0, s>>>(p, rhs, g, cf, 0, q, st, partials + L.nPart, 1);
      kRBGhosts<<<gg, dim3(64, 4), 0, s>>>(p, g, q, st);
      kRBFold<<<1, 256, 0, s>>>(q, st, partials, L.nPart, 2, n, epssq, itermax);
    }
    launched = end;
    HIP_CHECK(hipMemcpyAsync(&h, st, sizeof(RBState), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    stopAt = h.stopAt;
  }
  HIP_CHECK(hipMemcpyAsync(&h, st, sizeof(RBState), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  if (res) *res = h.res;
  if (iterations) *iterations = h.it;
  return 0;
}
