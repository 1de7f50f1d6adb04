/* Internal interfaces of the MI355X NuSiF library (host side + kernel
 * launchers). Not installed; callers use include/nusif.h and the reference
 * headers. */
#ifndef NUSIF_INTERNAL_H
#define NUSIF_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "nusif.h"

#define HIP_CHECK(call)                                                                \
  do {                                                                                 \
    hipError_t err_ = (call);                                                          \
    if (err_ != hipSuccess) {                                                          \
      fprintf(stderr, "nusif: HIP error %s at %s:%d: %s\n", hipGetErrorName(err_),     \
          __FILE__, __LINE__, hipGetErrorString(err_));                                \
      exit(EXIT_FAILURE);                                                              \
    }                                                                                  \
  } while (0)

#define NUSIF_FATAL(...)                                                               \
  do {                                                                                 \
    fprintf(stderr, "nusif: " __VA_ARGS__);                                            \
    fputc('\n', stderr);                                                               \
    exit(EXIT_FAILURE);                                                                \
  } while (0)

/* Padded i-fastest grid with one ghost layer (reference src/util.h:19-34). */
struct Geo {
  int ni, nj, nk;   /* interior cells */
  int sj;           /* ni + 2 */
  long long sk;     /* (ni + 2) * (nj + 2) */
  long long n;      /* padded cell count */
  __host__ __device__ long long at(int i, int j, int k) const
  {
    return (long long)k * sk + (long long)j * sj + i;
  }
};

static inline Geo makeGeo(int ni, int nj, int nk)
{
  Geo g;
  g.ni = ni;
  g.nj = nj;
  g.nk = nk;
  g.sj = ni + 2;
  g.sk = (long long)(ni + 2) * (nj + 2);
  g.n  = g.sk * (nk + 2);
  return g;
}

/* XCD-aware workgroup order. MI355X deals workgroup ids round-robin over its
 * 8 XCDs (each with its own L2). Kernels over a (gx, gy, gz) block grid are
 * launched 1-D and give XCD x the contiguous block range
 * [x * ceil(n/8), (x+1) * ceil(n/8)), x fastest: blocks that share stencil
 * halos then share an L2. */
#define NUM_XCDS 8
static inline dim3 xcdGrid(unsigned gx, unsigned gy, unsigned gz)
{
  unsigned n = gx * gy * gz;
  return dim3(((n + NUM_XCDS - 1) / NUM_XCDS) * NUM_XCDS);
}
static inline dim3 xcdGrid(dim3 g) { return xcdGrid(g.x, g.y, g.z); }
/* Face ghosts of cell (i, j, k) from value v: the Neumann copy of the
 * pressure solvers (reference src/solver-rb.c:129-148), each ghost plane over
 * the interior of the other two directions. */
template <typename T> __device__ inline void faceGhosts(T *x, const Geo &g, int i, int j, int k, long long c, T v)
{
  if (i == 1) x[c - 1] = v;
  if (i == g.ni) x[c + 1] = v;
  if (j == 1) x[c - g.sj] = v;
  if (j == g.nj) x[c + g.sj] = v;
  if (k == 1) x[c - g.sk] = v;
  if (k == g.nk) x[c + g.sk] = v;
}

/* Logical block (bx, by, bz) of this workgroup; false for surplus ids. */
__device__ inline bool xcdBlock(int gx, int gy, int gz, int &bx, int &by, int &bz)
{
  int n   = gx * gy * gz;
  int per = (n + NUM_XCDS - 1) / NUM_XCDS;
  int t   = (int)(blockIdx.x % NUM_XCDS) * per + (int)(blockIdx.x / NUM_XCDS);
  if (t >= n) return false;
  bx = t % gx;
  by = (t / gx) % gy;
  bz = t / (gx * gy);
  return true;
}

/* Default of the `verbose` option: the reference builds with -DVERBOSE
 * (config.mk:13), so a program linked against this library prints what the
 * reference's solve() and initDiscretization print unless told otherwise. */
#ifndef NUSIF_VERBOSE_DEFAULT
#define NUSIF_VERBOSE_DEFAULT 1
#endif

/* Options carried next to a Parameter (nusifSetOption / extra .par keys). */
enum SolverKind { SOLVER_DEFAULT = 0, SOLVER_RB, SOLVER_MG, SOLVER_GMG };
struct NusifOptions {
  int nsteps;
  int solver;
  int mixed;
  double mgTol;
  int mgMaxCycles;
  double mgOmega;
  int mgProject;
  int mgTail;      /* gmg: coarse levels in one workgroup (1) or per-level kernels (0) */
  int mgGraph;     /* gmg: replay each cycle from a captured HIP graph (1) or launch eagerly (0) */
  int mgGraphCache;/* gmg: captured cycle graphs kept at most (oldest evicted first) */
  int mgTileRows;  /* gmg: fused-sweep tile height, 0 automatic, 4, 8 or 16 */
  int mgFuseProlong; /* gmg: prolongation inside the first post-smoothing sweep (1) or its own pass (0) */
  int mgSpeculate; /* gmg: enqueue cycles ahead of the convergence test (1) or one at a time (0) */
  int mgInner;     /* gmg: cycles before a solve's first convergence test (0: sized from the previous solve) */
  double lid;
  double uInflow;
  double inflowNoise;
  int seed;
  int warmstart;
  char unetWeights[512];
  int unetExtent;  /* warm start: the network runs on the first level of extent <= this (64) */
  int unetGraph;       /* warm start: replay the network's forward from a captured HIP graph */
  int unetAcceptLevel; /* warm start: the acceptance test compares residuals restricted this
                        * many levels below the network grid (2; 0: on it) */
  int warmstartSkip;   /* warm start: no network for a solve whose predecessor took at most
                        * this many cycles (0: never skipped) */
  int fgRhs;       /* computeFG also forms rhs (1, default), computeRHS then skips its pass */
  int rbFused;     /* rb, mg: one launch per sweep where the grid allows (1, default) */
  int bcFused;     /* nusifRunSteps: the step's maxima and boundary passes in one launch (1, default) */
  int verbose;     /* 0 quiet; 1 the reference's -DVERBOSE output (configuration at
                    * initDiscretization, one line per solve); 2 also -DPROFILING's
                    * solver-time line after each red-black solve */
};
void nusifDefaultOptions(NusifOptions *o);
/* Options registered for `p` (defaults when none were set). */
const NusifOptions *nusifOptionsFor(const Parameter *p);
void nusifConsumeOptions(const Parameter *p, NusifOptions *out);

/* Scalar state shared by the reduction kernels (device memory). */
struct DevScalars {
  unsigned long long absmax[3]; /* bit patterns of max |u|, |v|, |w| */
  double sum;                   /* last sum reduction */
  double res;                   /* solver residual (carried)      } RBState, */
  int it;                       /* solver iterations completed    } same    */
  int stopAt;                   /* iterations to run (stop test)  } layout  */
  unsigned int counter[4];      /* last-block counters (self-resetting) */
  double norm2[4];              /* gmg norms */
  double dt;                    /* the step's dt, formed by kFoldMax3 (device dt) */
};

struct GmgState;  /* gmg.hip */
struct UnetModel; /* unet.hip */
struct WarmStart; /* warmstart.hip */

/* One Discretization's device state. */
#define SOLVE_EV_POOL 64
struct NusifCtx {
  Discretization *owner;
  double *host[8];     /* p rhs f g h u v w host arrays, identity for lookups */
  Geo geo;
  double *p, *rhs, *f, *g, *h, *u, *v, *w; /* device fields */
  double *pAlt;        /* ping-pong partner of p for the fused RB kernel */
  double *rhsNext;     /* the rhs computeFG forms (fg_rhs); computeRHS swaps it in */
  double *saveP, *saveRhs; /* p and rhs kept aside while solve() runs on foreign arrays
                            * (allocated on first use) */
  double *partials;    /* reduction partials */
  size_t nPartials;
  DevScalars *scal;    /* device */
  DevScalars *scalHost;/* pinned, host-coherent mirror */
  DevScalars *scalHostDev; /* its device address: reduction kernels write results there directly */
  double *inflow;      /* optional per-cell inflow profile (jmax+2)*(kmax+2) */
  /* CFL maxima: adaptUV leaves per-workgroup maxima of the interior range it
   * wrote in cflPartials; while no other kernel has written u, v, w since
   * (cflValid), computeTimestep adds only the ghost shell to them */
  double *cflPartials;
  int cflInteriorBlocks;
  bool cflValid;
  /* computeFG with fg_rhs also wrote rhsNext (from the f, g, h it wrote,
   * with dt rhsDt); cleared by anything else that writes f, g, h or rhs */
  bool rhsReady;
  double rhsDt;
  /* p's face ghosts hold the Neumann copies of its interior: true after a
   * gmg solve (its last pass writes them), cleared by anything else that
   * writes p (normalizePressure, uploads, other solvers, staged arrays), so
   * the next gmg solve skips its ghost pass */
  bool pGhostsOk;
  /* while nusifRunSteps drives the loop: computeFG, computeRHS and adaptUV
   * read dt from scal->dt (formed on the device by the CFL fold), so no
   * step waits for a host round trip; dtEvent follows that fold */
  bool dtOnDevice;
  hipEvent_t dtEvent;
  /* nusifSolveTiming: an event pair around each solve of nusifRunSteps,
   * from a pool folded into solveMs / solveCount when it fills and at the
   * end of the call (no per-step wait) */
  int solveTiming;
  int solveEvUsed;
  bool solveEvReady;
  hipEvent_t solveEv[2 * SOLVE_EV_POOL];
  double solveMs;
  long long solveCount;
  NusifOptions opt;
  double t;            /* nusifRunSteps clock */
  int nt;
};

/* One Solver's device state. */
struct NusifSolverCtx {
  Solver *owner;
  NusifCtx *dctx;
  int kind;
  int lastIterations;
  double lastRel;
  int warmUsed;
  int predicted; /* iterations of the previous solve, sizes the launch batches */
  GmgState *gmg;
  WarmStart *ws; /* learned warm start / training-pair capture (gmg only) */
  double wsRes[4]; /* last warm start: fine ||r0||, fine ||r_guess||, restricted both */
  long long wsTried, wsTaken; /* network guesses formed / accepted since init */
  /* set by nusifRunSteps for one solve: the gmg cycle loop calls
   * afterFirst(afterFirstArg) once the first group of cycles is enqueued,
   * before the host waits for its norms, so the work queued there overlaps
   * that round trip. afterFirstState: 0 not called, 1 called with p final,
   * 2 called but later cycles changed p */
  void (*afterFirst)(void *) = nullptr;
  void *afterFirstArg        = nullptr;
  int afterFirstState        = 0;
};

hipStream_t nusifStream(void);
NusifCtx *nusifCtxOf(const Discretization *d);
NusifCtx *nusifCtxOfHost(const double *hostPtr, int *fieldIndex);
NusifSolverCtx *nusifSolverCtxOf(const Solver *s);

/* Kernel timing hooks (context.hip). */
struct KernelTimer;
KernelTimer *nusifTimerFor(const char *name);
void nusifTimerBegin(KernelTimer *t);
void nusifTimerEnd(KernelTimer *t);

#define NUSIF_TIMED(name, launch)                                                     \
  do {                                                                                \
    static KernelTimer *tm_ = nullptr;                                                \
    static int tmGen_       = -1;                                                     \
    extern int nusifTimerGeneration;                                                  \
    if (tmGen_ != nusifTimerGeneration) {                                             \
      tm_    = nusifTimerFor(name);                                                   \
      tmGen_ = nusifTimerGeneration;                                                  \
    }                                                                                 \
    if (tm_) nusifTimerBegin(tm_);                                                    \
    launch;                                                                           \
    if (tm_) nusifTimerEnd(tm_);                                                      \
  } while (0)

/* Launchers (discretization.hip) */
/* CFL maxima into scal (and its pinned mirror); with d, also dt (the
 * reference's arithmetic, src/discretization.c:431-453) into scal->dt */
void launchAbsMax3(NusifCtx *c, const Discretization *d = nullptr);
void launchVelocityBC(NusifCtx *c, const Discretization *d);
void launchSpecialBC(NusifCtx *c, const Discretization *d);
/* launchAbsMax3(c, d) (withMax: needs c->cflValid), launchVelocityBC and
 * launchSpecialBC as one launch and the maxima's fold; the same values.
 * Needs ni, nj, nk >= 3 (boundaryStepFits). */
bool boundaryStepFits(const NusifCtx *c);
void launchBoundaryStep(NusifCtx *c, const Discretization *d, bool withMax);
void launchNormalize(NusifCtx *c);
void launchAdaptUV(NusifCtx *c, const Discretization *d);
void launchComputeFG(NusifCtx *c, const Discretization *d);
/* gmg.hip kernels the warm start also runs: r0 = rhs - A p restricted to
 * the next level with rhs (returns the partial count; partials hold r0^2
 * and rhs^2 per workgroup; optionally r0 as FP32 on the fine grid, r32, and
 * an FP32 array zeroed, eZero), its partial count for a grid, the fold of
 * such partials into a norm pair (and its host mirror), and the FP32
 * trilinear prolongation x_f = P x_c */
int launchInputRestrict(const double *p, const double *rhs, const Geo &g, double idx2, double idy2, double idz2,
    float *rc, float *bc, const Geo &gc, double *partials, float *r32 = nullptr, float *eZero = nullptr);
void launchFoldNorms(const double *partials, int nb, double *out, double *outHost, double *extra0 = nullptr,
    double *extra0Host = nullptr);
int inputRestrictPartials(const Geo &g);
void launchProlongSet32(float *xf, const Geo &gf, const float *xc, const Geo &gc);
/* copy rhs into the spare rhs buffer (after a host upload of rhs) */
void nusifRhsSpareSync(NusifCtx *c);
void launchComputeRHS(NusifCtx *c, const Discretization *d);
void launchNormalize(NusifCtx *c);
void launchAdaptUV(NusifCtx *c, const Discretization *d);

/* Solvers */
double rbSolve(NusifSolverCtx *sc, double *p, const double *rhs, double **pOut);
double mgRefSolve(NusifSolverCtx *sc, double *p, const double *rhs, double **pOut);
double gmgSolve(NusifSolverCtx *sc, double *p, const double *rhs);
void gmgFree(GmgState *g);

/* Shared by rb / mg / gmg: colour sweep, pressure ghosts */
struct SweepCoef {
  double idx2, idy2, idz2, factor;
};
SweepCoef sweepCoef(const Grid *g, double omega);
void launchPressureGhosts(const Geo &g, double *p, hipStream_t st);
/* One red-black SOR iteration (both colours and the ghost copy) in one
 * launch, xo -> xn, for grids the fused sweep covers (gmg.hip); with
 * partials null, without the in-sweep residual and the stop test */
bool rbFusedFits(const Geo &g);
int rbFusedTiles(const Geo &g);
void launchRBFused(const double *xo, double *xn, const double *rhs, const Geo &g, const SweepCoef &sc,
    double *partials, const int *stopAt, int q);

#endif
