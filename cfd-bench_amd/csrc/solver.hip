/* Solver plugin (reference src/solver.h:35-36) and the in-library step loop.
 *
 * The reference picks its solver at link time (Makefile:19-21, SOLVER=rb|mg).
 * This library is built twice from the same sources, libnusif_rb.so and
 * libnusif_mg.so, whose default solver is NUSIF_DEFAULT_SOLVER; the added
 * option `solver` (rb | mg | gmg) overrides it per problem. */
#include <math.h>
#include <string.h>

#include "nusif_internal.h"
#include "profiler.h"
#include "timing.h"
#include "unet.h"
#include "warmstart.h"

#ifndef NUSIF_DEFAULT_SOLVER
#define NUSIF_DEFAULT_SOLVER SOLVER_RB
#endif

void nusifRegisterSolver(NusifSolverCtx *s);
void nusifUnregisterSolver(NusifSolverCtx *s);
GmgState *gmgCreate(NusifSolverCtx *sc, const Parameter *prm);

/* src/solver-rb.c:17-25 and src/solver-mg.c:356-389 */
extern "C" void initSolver(Solver *s, Discretization *d, Parameter *p)
{
  memset(s, 0, sizeof(*s));
  s->eps        = p->eps;
  s->omega      = p->omg;
  s->itermax    = p->itermax;
  s->levels     = p->levels;
  s->presmooth  = p->presmooth;
  s->postsmooth = p->postsmooth;
  s->grid       = &d->grid;
  s->comm       = &d->comm;
  s->problem    = p->name;
  s->r = s->e = nullptr; /* device hierarchy lives in the solver context */

  NusifSolverCtx *sc = new NusifSolverCtx();
  memset(sc, 0, sizeof(*sc));
  sc->owner    = s;
  sc->dctx     = nusifCtxOf(d);
  sc->kind     = sc->dctx->opt.solver != SOLVER_DEFAULT ? sc->dctx->opt.solver : NUSIF_DEFAULT_SOLVER;
  sc->lastRel  = -1.0;
  sc->warmUsed = -1;
  sc->wsTried  = 0;
  sc->wsTaken  = 0;
  if (sc->kind == SOLVER_MG && commIsMaster(s->comm))
    printf("Using Multigrid solver with %d levels\n", s->levels);
  if (sc->kind == SOLVER_GMG) sc->gmg = gmgCreate(sc, p);
  if (sc->dctx->opt.warmstart) {
    if (sc->kind != SOLVER_GMG) NUSIF_FATAL("warmstart unet needs solver gmg");
    /* A kernel-traced run of this combination ended in a host segmentation
     * fault inside the HIP runtime whose cause is not known; graph replay is
     * no faster than launching the kernels, so the pair is refused. */
    if (sc->dctx->opt.mgGraph) NUSIF_FATAL("warmstart unet does not combine with mg_graph 1");
    const char *path = sc->dctx->opt.unetWeights;
    /* a weights file that records its training extent sets the network level */
    int fileExtent = path[0] ? unetFileExtent(path) : 0;
    if (fileExtent > 0) sc->dctx->opt.unetExtent = fileExtent;
    const int levels = path[0] ? unetFileLevels(path) : 0;
    if (levels < 1) NUSIF_FATAL("warmstart unet: cannot read weights '%s' (option unet_weights)", path);
    sc->ws = warmstartCreate(sc->dctx, nullptr, false, levels);
    int nx, ny, nz;
    warmstartExtents(sc->ws, &nx, &ny, &nz);
    UnetModel *net = unetLoad(path, nx, ny, nz);
    if (!net) NUSIF_FATAL("warmstart unet: cannot read weights '%s' (option unet_weights)", path);
    warmstartSetNet(sc->ws, net);
  }
  nusifRegisterSolver(sc);
}

extern "C" void nusifFreeSolver(Solver *s)
{
  NusifSolverCtx *sc = nusifSolverCtxOf(s);
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  if (sc->gmg) gmgFree(sc->gmg);
  if (sc->ws) {
    unetFree(warmstartNet(sc->ws));
    warmstartFree(sc->ws);
  }
  nusifUnregisterSolver(sc);
  delete sc;
}

/* The per-solve output of the reference's -DVERBOSE build: red-black SOR
 * prints its iteration count and the root of its mean-square residual, then
 * printProfile (src/solver-rb.c:163-168), which with -DPROFILING (verbose 2)
 * reports the solver time accumulated in T[SOLVER] (src/profiler.c:49-74;
 * here the wall time of solve(), which returns only after the device has
 * finished the solve); the reference MG prints its returned residual
 * (src/solver-mg.c:395-397). The gmg solver reports its V-cycles and
 * relative residual in the red-black form. */
static void printSolve(NusifSolverCtx *sc, Solver *s, double res, double tStart)
{
  const int verbose = sc->dctx->opt.verbose;
  if (verbose >= 2) {
    HIP_CHECK(hipStreamSynchronize(nusifStream()));
    T[SOLVER] += getTimeStamp() - tStart;
    C[SOLVER]++;
  }
  if (sc->kind == SOLVER_MG) {
    if (commIsMaster(s->comm)) printf("Residuum: %.6f\n", res);
    return;
  }
  if (commIsMaster(s->comm))
    printf("Solver took %d iterations to reach %f\n", sc->lastIterations, sc->kind == SOLVER_RB ? sqrt(res) : res);
  printProfile(s->comm, sc->lastIterations);
}

/* p and rhs are host pointers (src/main.c:71). When they are the arrays of a
 * Discretization of this library, its device copies are solved in place;
 * other host arrays are staged through the owning context's device buffers. */
extern "C" double solve(Solver *s, double *p, const double *rhs)
{
  NusifSolverCtx *sc = nusifSolverCtxOf(s);
  NusifCtx *c        = sc->dctx;
  c->rhsReady        = false; /* a solve may rewrite rhs (staging, mean removal) */
  int ip = -1, ir = -1;
  NusifCtx *cp = nusifCtxOfHost(p, &ip);
  NusifCtx *cr = nusifCtxOfHost(rhs, &ir);
  bool staged  = !(cp == c && ip == 0 && cr == c && ir == 1);
  size_t bytes = (size_t)c->geo.n * sizeof(double);
  /* only a gmg solve of the context's own p leaves its ghosts known */
  if (staged || sc->kind != SOLVER_GMG) c->pGhostsOk = false;
  if (staged) {
    /* foreign arrays: the context's own p and rhs are set aside, the foreign
     * ones solved in their place and copied out, then p and rhs restored
     * (the reference's solve touches only the arrays it is given) */
    if (!c->saveP) {
      HIP_CHECK(hipMalloc((void **)&c->saveP, bytes));
      HIP_CHECK(hipMalloc((void **)&c->saveRhs, bytes));
    }
    HIP_CHECK(hipMemcpyAsync(c->saveP, c->p, bytes, hipMemcpyDeviceToDevice, nusifStream()));
    HIP_CHECK(hipMemcpyAsync(c->saveRhs, c->rhs, bytes, hipMemcpyDeviceToDevice, nusifStream()));
    HIP_CHECK(hipMemcpyAsync(c->p, p, bytes, hipMemcpyHostToDevice, nusifStream()));
    HIP_CHECK(hipMemcpyAsync(c->rhs, rhs, bytes, hipMemcpyHostToDevice, nusifStream()));
    nusifRhsSpareSync(c);
  }
  double res = 0.0;
  const double tStart = c->opt.verbose >= 2 ? getTimeStamp() : 0.0;
  switch (sc->kind) {
  case SOLVER_RB: {
    double *pOut = c->p;
    res          = rbSolve(sc, c->p, c->rhs, &pOut);
    if (pOut != c->p) { /* ping-pong ended on the partner buffer */
      c->pAlt = c->p;
      c->p    = pOut;
    }
  } break;
  case SOLVER_MG: {
    double *pOut = c->p;
    res          = mgRefSolve(sc, c->p, c->rhs, &pOut);
    if (pOut != c->p) { /* ping-pong ended on the partner buffer */
      c->pAlt = c->p;
      c->p    = pOut;
    }
  } break;
  case SOLVER_GMG: res = gmgSolve(sc, c->p, c->rhs); break;
  default: NUSIF_FATAL("unknown solver kind %d", sc->kind);
  }
  if (staged) {
    HIP_CHECK(hipMemcpyAsync(p, c->p, bytes, hipMemcpyDeviceToHost, nusifStream()));
    HIP_CHECK(hipMemcpyAsync(c->p, c->saveP, bytes, hipMemcpyDeviceToDevice, nusifStream()));
    HIP_CHECK(hipMemcpyAsync(c->rhs, c->saveRhs, bytes, hipMemcpyDeviceToDevice, nusifStream()));
    nusifRhsSpareSync(c);
    HIP_CHECK(hipStreamSynchronize(nusifStream()));
    c->cflValid  = false;
    c->pGhostsOk = false; /* p restored from its saved copy */
  }
  if (c->opt.verbose >= 1) printSolve(sc, s, res, tStart);
  return res;
}

extern "C" int nusifSolverIterations(Solver *s) { return nusifSolverCtxOf(s)->lastIterations; }
extern "C" double nusifSolverRelResidual(Solver *s) { return nusifSolverCtxOf(s)->lastRel; }
extern "C" int nusifWarmstartUsed(Solver *s) { return nusifSolverCtxOf(s)->warmUsed; }

extern "C" void nusifWarmstartCounts(Solver *s, long long *tried, long long *taken)
{
  NusifSolverCtx *sc = nusifSolverCtxOf(s);
  *tried             = sc->wsTried;
  *taken             = sc->wsTaken;
}

extern "C" void nusifWarmstartResiduals(Solver *s, double *out4)
{
  NusifSolverCtx *sc = nusifSolverCtxOf(s);
  for (int q = 0; q < 4; q++) out4[q] = sc->wsRes[q];
}

/* The loop body of src/main.c:60-88, driven from inside the library. */
/* elapsed times of the recorded solve event pairs into solveMs / solveCount */
static void foldSolveTimes(NusifCtx *c)
{
  for (int q = 0; q < c->solveEvUsed; q++) {
    float ms = 0.0f;
    HIP_CHECK(hipEventSynchronize(c->solveEv[2 * q + 1]));
    HIP_CHECK(hipEventElapsedTime(&ms, c->solveEv[2 * q], c->solveEv[2 * q + 1]));
    c->solveMs += ms;
    c->solveCount++;
  }
  c->solveEvUsed = 0;
}

extern "C" void nusifSolveTiming(Discretization *d, int on)
{
  NusifCtx *c = nusifCtxOf(d);
  if (on && !c->solveEvReady) {
    for (int q = 0; q < 2 * SOLVE_EV_POOL; q++) HIP_CHECK(hipEventCreate(&c->solveEv[q]));
    c->solveEvReady = true;
  }
  c->solveTiming = on ? 1 : 0;
}

extern "C" void nusifSolveTimes(Discretization *d, double *ms, long long *solves)
{
  NusifCtx *c = nusifCtxOf(d);
  foldSolveTimes(c);
  if (ms) *ms = c->solveMs;
  if (solves) *solves = c->solveCount;
  c->solveMs    = 0.0;
  c->solveCount = 0;
}

extern "C" int nusifRunSteps(Discretization *d, Solver *s, int nsteps, double *trace)
{
  NusifCtx *c        = nusifCtxOf(d);
  NusifSolverCtx *sc = nusifSolverCtxOf(s);
  /* with an adaptive dt, the CFL fold forms dt on the device and computeFG,
   * computeRHS and adaptUV read it there: the step is enqueued without
   * waiting for the maxima. The host learns dt (same bits as
   * computeTimestep's own arithmetic) once the fold has run, which the
   * solve's first read-back has already waited for. */
  const bool devDt = d->tau > 0.0;
  for (int q = 0; q < nsteps; q++) {
    /* the shell's maxima, the faces and the lid or inflow in one launch
     * where adaptUV's maxima are there (every step after the first) */
    const bool fused = c->opt.bcFused && boundaryStepFits(c) && (!devDt || c->cflValid);
    if (fused) launchBoundaryStep(c, d, devDt);
    else if (devDt) launchAbsMax3(c, d);
    if (devDt) {
      HIP_CHECK(hipEventRecord(c->dtEvent, nusifStream()));
      c->dtOnDevice = true;
    }
    if (!fused) {
      setBoundaryConditions(d);
      setSpecialBoundaryCondition(d);
    }
    computeFG(d);
    computeRHS(d);
    if (c->nt % 100 == 0) normalizePressure(d);
    const bool timeSolve = c->solveTiming && c->solveEvReady;
    if (timeSolve) {
      if (c->solveEvUsed == SOLVE_EV_POOL) foldSolveTimes(c);
      HIP_CHECK(hipEventRecord(c->solveEv[2 * c->solveEvUsed], nusifStream()));
    }
    /* gmg: adaptUV is queued as soon as the first cycle group is, so it runs
     * while the host reads that group's norms; it runs again when the solve
     * needed more cycles after it (it reads only f, g, h and p). Not with
     * solve timing, whose events bracket the solve alone. */
    sc->afterFirst      = timeSolve ? nullptr : [](void *a) { adaptUV((Discretization *)a); };
    sc->afterFirstArg   = d;
    sc->afterFirstState = 0;
    double res          = solve(s, d->p, d->rhs);
    const int early     = sc->afterFirstState;
    sc->afterFirst      = nullptr;
    if (timeSolve) HIP_CHECK(hipEventRecord(c->solveEv[2 * c->solveEvUsed++ + 1], nusifStream()));
    if (early != 1) adaptUV(d);
    if (devDt) {
      c->dtOnDevice = false;
      HIP_CHECK(hipEventSynchronize(c->dtEvent));
      d->dt = c->scalHost->dt;
    }
    if (trace) {
      trace[4 * q + 0] = c->t;
      trace[4 * q + 1] = d->dt;
      trace[4 * q + 2] = res;
      trace[4 * q + 3] = (double)nusifSolverIterations(s);
    }
    c->t += d->dt;
    c->nt++;
  }
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  foldSolveTimes(c);
  return nsteps;
}
