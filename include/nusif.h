/* nusif.h — what the MI355X library adds to the NuSiF C API.
 *
 * The reference API (discretization.h, solver.h, parameter.h, comm.h, ...) is
 * exported unchanged. Its eight fields live in HBM of the one GPU; the host
 * arrays in Discretization are staging copies:
 *   - initDiscretization fills both copies;
 *   - every step function works on the device copy only;
 *   - nusifSyncHost() downloads the device copy (commCollectResult does this
 *     by itself for the arrays it is given);
 *   - nusifSyncDevice() uploads host arrays a caller has modified.
 * Errors follow the reference convention: a message on stderr and
 * exit(EXIT_FAILURE) (reference src/allocate.c:19-33).
 *
 * None of these functions has a counterpart in the reference; each comment
 * names what it serves.
 */
#ifndef NUSIF_H
#define NUSIF_H

#include "discretization.h"
#include "parameter.h"
#include "solver.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Download all eight fields of `d` from the GPU into its host arrays. */
void nusifSyncHost(Discretization *d);
/* Upload all eight host arrays of `d` to the GPU. */
void nusifSyncDevice(Discretization *d);
/* Release the device and host memory of `d` / of a solver (the reference
 * never frees; long-running callers and tests need it). */
void nusifFreeDiscretization(Discretization *d);
void nusifFreeSolver(Solver *s);

/* Options beyond the reference's .par keys. readParameter() also accepts them
 * as .par lines. Keys are chosen so that the reference's prefix matching
 * (src/parameter.c:56-62) never mistakes them for one of its own keys.
 *   nsteps        int     stop the driver loop after this many steps (0: off)
 *   solver        rb | mg | gmg     pressure solver (default: library variant)
 *   precision     fp64 | mixed      gmg only: mixed = FP32 V-cycles inside an
 *                                   FP64 defect-correction loop
 *   mg_tol        double  gmg: relative residual ||rhs-Ap||/||rhs|| (1e-6)
 *   mg_maxcycles  int     gmg: cycle cap per solve (100)
 *   mg_omega      double  gmg: smoother relaxation (1.0)
 *   mg_project    0 | 1   gmg: remove the rhs mean (all-Neumann compatibility)
 *   mg_tail       0 | 1   gmg: run the levels of <= 4096 cells in one workgroup
 *                         (1, default) or with one kernel per pass (0); same values
 *   mg_fuse_prolong 0 | 1 gmg: each level's prolongation inside its first
 *                         post-smoothing sweep, the warm start's last prolongation
 *                         step inside the first sweep (1, default), or passes of
 *                         their own (0); same values
 *   mg_graph      0 | 1   gmg: replay each cycle from a captured HIP graph (1) or
 *                         launch its kernels one by one (0, default); same values.
 *                         Not with warmstart unet (initSolver stops with an error)
 *   mg_graph_cache int    gmg with mg_graph 1: captured cycle graphs kept, the
 *                         oldest destroyed first once the stream has drained (64)
 *   mg_tile_rows  0|4|8|16 gmg: fused-sweep tile height (0: automatic, 16 on
 *                         levels with >= 128 rows, 8 with >= 64); same values
 *   mg_inner      int     gmg: cycles before a solve's first convergence test
 *                         (1, default: test after every cycle; k > 1: the
 *                         first k cycles run back to back, in mixed mode as
 *                         FP32 cycles on one FP64 defect; 0: sized from
 *                         the previous solve, its exact count when its
 *                         group did not converge, one less when its own
 *                         contraction says that would have sufficed); the
 *                         tests after the first come after every cycle.
 *                         Ignored with mg_speculate 1.
 *   mg_speculate  0 | 1   gmg: enqueue cycles before the previous one's
 *                         convergence test is read back (1; the device skips
 *                         the ones past convergence) or one at a time (0,
 *                         default); same values
 *   lid           double  lid speed of "dcavity" (reference: 1.0)
 *   u_inflow      double  inflow velocity of "canal" (reference: 2.0)
 *   inflow_noise  double  relative uniform perturbation of the inflow (0)
 *   seed          int     seed of inflow_noise (1)
 *   warmstart     none | unet | unet_force   initial guess of the gmg solve
 *                 (unet_force: always take the network's guess, diagnostic)
 *   unet_weights  path    trained U-Net weights (nusifUnetSave format)
 *   unet_extent   int     the network runs on the first multigrid level whose
 *                         extents are all <= this (64), or the last that its
 *                         pooling levels still divide; a weights file that
 *                         records its training extent overrides it
 *   unet_graph    0 | 1   warm start: the network's forward (~40 launches) is
 *                         replayed from a HIP graph captured at its first
 *                         use (1) or launched kernel by kernel (0); same values
 *   unet_accept_level int the guess is taken when its residual, restricted
 *                         this many levels below the network grid (2; 0: on
 *                         it), is smaller than the previous pressure's: the
 *                         trilinear correction's interpolation error leaves
 *                         residual at the network grid's own scale that the
 *                         first sweeps remove, which the coarser comparison
 *                         no longer counts
 *   warmstart_skip int    warm start: a solve whose predecessor took at most
 *                         this many V-cycles starts from the previous-step
 *                         pressure without running the network (0, default:
 *                         the network runs for every solve)
 *   fg_rhs        0 | 1   computeFG also forms the right-hand side in the same
 *                         pass, into a spare buffer that computeRHS then swaps
 *                         in (1, default; computeRHS recomputes when f, g, h,
 *                         rhs or dt changed in between); 0: computeRHS runs
 *                         its own pass. Same values either way
 *   bc_fused      0 | 1   nusifRunSteps: each step's CFL maxima (after the
 *                         first step), velocity boundary conditions and lid
 *                         or inflow in one launch and a fold (1, default) or
 *                         the reference's calls one by one (0); same values
 *   rb_fused      0 | 1   rb and mg solvers: each sweep (both colour passes
 *                         and the ghost copy) in one launch where the grid
 *                         allows (imax >= 32, kmax >= 4; 1, default); 0: a
 *                         launch per colour pass. Same p either way; rb's sum
 *                         of r^2 is added in another order
 *   verbose       0 | 1 | 2  standard output of the reference's default build
 *                         (config.mk:13-14): 1 = -DVERBOSE, i.e. the
 *                         configuration at initDiscretization
 *                         (discretization.c:16-47) and one line per solve:
 *                         "Solver took %d iterations to reach %f" (rb and
 *                         gmg, solver-rb.c:163-166) or "Residuum: %.6f" (mg,
 *                         solver-mg.c:395-397); 2 = -DVERBOSE -DPROFILING,
 *                         adding printProfile's solver-time and MLUPS line
 *                         after each rb and gmg solve (profiler.c:49-74,
 *                         solve() wall time); 0 = none. Default 1, as the
 *                         reference's build; the Python binding sets 0
 * Returns 0 on success, -1 for an unknown key, -2 for a bad value. */
int nusifSetOption(Parameter *p, const char *key, const char *value);

/* The `nsteps` option of `p` (0 when unset). */
int nusifStepLimit(Parameter *p);
/* The `verbose` option of `p` (1 when unset, as the reference's build). */
int nusifVerboseLevel(Parameter *p);

/* Iterations (rb), sweeps (mg) or V-cycles (gmg) of the last solve(). */
int nusifSolverIterations(Solver *s);
/* gmg: V-cycles before the last solve's first convergence test (the
 * mg_inner group; 0 when the solve ran no cycle or the solver is not gmg). */
int nusifSolverFirstGroup(Solver *s);
/* Relative residual ||rhs - A p||_2 / ||rhs||_2 after the last gmg solve
 * (-1 for rb/mg, whose stop tests use the reference's mean-square residual). */
double nusifSolverRelResidual(Solver *s);
/* Outcome of the last warm start: 1 network guess used, 0 fell back to the
 * previous-step pressure, -1 no network. */
int nusifWarmstartUsed(Solver *s);
/* Residual norms of the last warm start: fine ||r0|| (previous-step p),
 * fine ||r_guess|| (-1 in mixed precision with the network below the finest
 * grid, whose guess is judged on the coarse grids: ||R r0 - A_n d|| against
 * ||R r0||, A_n the network grid's operator), and both restricted to the
 * acceptance level
 * (unet_accept_level below the network grid; the acceptance test compares
 * the restricted pair). */
void nusifWarmstartResiduals(Solver *s, double *out4);
/* Network guesses formed and accepted by the solves of `s` since
 * initSolver (the acceptance rate of the warm start over a run). */
void nusifWarmstartCounts(Solver *s, long long *tried, long long *taken);

/* Name of the GPU the library runs on; returns the device count. */
int nusifDeviceName(char *buf, int len);
void nusifDeviceSynchronize(void);

/* Per-kernel HIP-event timers on the library's stream (for roofline
 * reporting). nusifTimerEnable(name) starts timing every launch of that
 * kernel; nusifTimerRead returns the summed milliseconds and the launch count
 * (synchronising the stream). Names: see nusifKernelBytes. */
int nusifTimerEnable(const char *kernel);
double nusifTimerRead(const char *kernel, int *launches);
/* Stops and removes every timer (read them first). */
void nusifTimerDisableAll(void);
/* gmg solver: microseconds per launch of the finest level's red-black sweep,
 * from `reps` back-to-back launches on the solver's own buffers (changes its
 * work arrays; -1 if not applicable). */
double nusifTimeFinestSweep(Solver *s, int reps);
void nusifTimerReset(void);
/* Algorithmic HBM bytes of one launch of `kernel` on the grid of `d`
 * (each array element read once and/or written once). */
double nusifKernelBytes(Discretization *d, const char *kernel);

/* Device-to-device copy of `bytes` bytes repeated `reps` times on the library
 * stream; returns ms per copy (calibration of HBM byte counters and the
 * achievable-bandwidth reference). */
double nusifCopyProbe(size_t bytes, int reps);

/* Milliseconds per launch of three hand-written streaming kernels over
 * arrays of `bytes` bytes each, `reps` launches each: ms[0] a copy (one read,
 * one write), ms[1] two reads and one write (a smoother sweep's pattern),
 * ms[2] one read. The HBM rate a plain kernel reaches on this device. */
void nusifStreamProbe(size_t bytes, int reps, double *ms);

/* Trains the warm-start U-Net on the GPU from pressure-solve pairs the
 * solver generates itself: `runs` lid-driven cavities of grid^3 cells with
 * seeded Re ~ U[10, 1000] and lid speed ~ U[0.5, 2], `skipSteps` steps of
 * spin-up then `stepsPerRun` captured steps each (gmg, mixed precision,
 * 1e-6); then `epochs` of Adam (cosine-decayed `lr`, `batch` samples per
 * update). Writes the weights to `weightsOut` (option unet_weights reads
 * them); lossPerEpoch (may be NULL) receives the mean training loss per
 * epoch. Returns the number of training pairs. */
int nusifUnetTrainCavity(const char *weightsOut, int grid, int runs, int stepsPerRun,
    int skipSteps, int epochs, int batch, double lr, int seed, double *lossPerEpoch);
/* The same with run r on a grids[r % ngrids]^3 cavity and the network on
 * the first level of extent <= netExtent (every grid must give the same
 * network extent, e.g. 128 and 256 both give 64). */
int nusifUnetTrainCavityGrids(const char *weightsOut, const int *grids, int ngrids, int netExtent,
    int runs, int stepsPerRun, int skipSteps, int epochs, int batch, double lr, int seed,
    double *lossPerEpoch);
/* Pairs from `runs` channels of imax x jmax x kmax cells (the canal problem
 * on a domain xlength x 1 x 1: inflow u = 1 with a seeded 1% uniform
 * perturbation, a new seed per run, zero-gradient outflow, no-slip walls,
 * Re ~ U[100, 1000], the rhs mean removed), the network on the first level
 * of extents <= netExtent that four pooling levels still divide; then
 * training as above, starting from the weights in weightsIn (fine-tuning;
 * NULL or "" starts from a seeded initialisation). */
int nusifUnetTrainChannel(const char *weightsIn, const char *weightsOut, int imax, int jmax, int kmax,
    double xlength, int netExtent, int runs, int stepsPerRun, int skipSteps, int epochs, int batch, double lr,
    int seed, double *lossPerEpoch);

/* One forward and backward pass of the U-Net on the GPU, for gradient
 * checks: cfg = {levels, ch0, ch1, ch2, ch3, input channels}, x is
 * nx ny nz x cin channels-last (each extent divisible by 2^(levels-1)),
 * params in the flat layout of the weights file. Returns the loss; grad and
 * pred (may be NULL) receive the results. */
double nusifUnetLossGrad(const int *cfg, int nx, int ny, int nz, const double *params, const double *x,
    const double *target, double *grad, double *pred);
/* The inference forward on the same arguments (the warm start's: decoder
 * inputs read upsampled in place, nothing kept for a backward pass); pred
 * receives nx ny nz values, bit-identical to nusifUnetLossGrad's. */
void nusifUnetPredict(const int *cfg, int nx, int ny, int nz, const double *params, const double *x, double *pred);

/* Checks of the warm start's grid transfers on host arrays (padded layouts,
 * for tests): r0 = rhs - A p (FP64, the solvers' 7-point Neumann Laplacian
 * with spacings dx, dy, dz) and rhs averaged over the 8 children of each cell
 * of the grid of halved extents (FP32 rc, bc; extents even), and sumsq =
 * sum of r0^2; and the FP32 trilinear prolongation xf = P xc from a grid of
 * nic x njc x nkc cells to the doubled one (face ghosts written). */
void nusifRestrictResidualCheck(int ni, int nj, int nk, double dx, double dy, double dz, const double *p,
    const double *rhs, float *rc, float *bc, double *sumsq);
void nusifProlongSetCheck(int nic, int njc, int nkc, const float *xc, float *xf);

/* No counterpart in the reference: the red-black SOR solve of solve() with
 * rb_fused 0 (src/solver-rb.c:27-172) for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc). p (in/out) and rhs are DEVICE arrays of the padded
 * layout, (imax+2)(jmax+2)(kmax+2) doubles, i fastest; dx, dy, dz, omega,
 * eps and itermax are the Solver's. It calls no hipMalloc or hipFree: its
 * partial sums and stop state live in a fixed 514 KiB device array the
 * library holds per GPU for the life of the process (every grid with imax,
 * jmax, kmax <= 512 fits; larger ones return -2). The first call on a device
 * looks the array up there (and creates the library stream if nothing did).
 * Returns 0 and sets *res (the carried mean-square residual, as solve()
 * returns it) and *iterations; -1 bad arguments, -2 grid too large, -3 no
 * device. Calls that overlap on one GPU share the array and race. */
int nusifSolveRBNoAlloc(double *p, const double *rhs, int imax, int jmax, int kmax, double dx, double dy,
    double dz, double omega, double eps, int itermax, double *res, int *iterations);

/* Runs the reference time loop (src/main.c:60-88) for `nsteps` steps entirely
 * inside the library: the same calls in the same order, without the Python
 * or host-driver overhead between them. trace (may be NULL) receives
 * (t, dt, res, iterations) per step. Returns the steps run. */
int nusifRunSteps(Discretization *d, Solver *s, int nsteps, double *trace);

/* Solve time inside nusifRunSteps: with timing on, each step's solve is
 * bracketed by HIP events on the library stream (no host wait per step;
 * the reference times the solve as its SOLVER region, src/solver-rb.c:161,
 * and rates it in printProfile, src/profiler.c:49-74). nusifSolveTimes
 * returns the summed milliseconds and the number of timed solves since its
 * last call, and resets them. */
void nusifSolveTiming(Discretization *d, int on);
void nusifSolveTimes(Discretization *d, double *ms, long long *solves);

#ifdef __cplusplus
}
#endif
#endif
