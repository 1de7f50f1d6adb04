/* nusif — command-line driver: runs a .par case through the GPU library with
 * the reference driver's semantics (reference src/main.c:18-154): the same
 * call order per step, residual.dat trace, cell-centred VTK output.
 * Added: the optional .par key `nsteps` stops the loop after that many steps
 * (the reference loop only ends at t > te), and a closing line with steps/s,
 * MLUP/s and the solver's iteration total. With `verbose` 1 or 2 (1 unless
 * set) the loop prints the reference's -DVERBOSE step lines
 * (main.c:80-83), otherwise its progress bar (main.c:49-51, 84-86). */
#include <stdio.h>
#include <stdlib.h>

#include "allocate.h"
#include "discretization.h"
#include "nusif.h"
#include "parameter.h"
#include "profiler.h"
#include "progress.h"
#include "solver.h"
#include "timing.h"
#include "vtkWriter.h"

int main(int argc, char **argv)
{
  Parameter p;
  Solver s;
  Discretization d;

  commInit(&d.comm, argc, argv);
  initParameter(&p);
  if (argc != 2) {
    printf("Usage: %s <configFile>\n", argv[0]);
    exit(EXIT_SUCCESS);
  }
  FILE *fp = initResidualWriter();
  readParameter(&p, argv[1]);
  commPartition(&d.comm, p.kmax, p.jmax, p.imax);
  printParameter(&p);

  initDiscretization(&d, &p);
  initSolver(&s, &d, &p);
  initProfiler(&d.comm);
  int nsteps = nusifStepLimit(&p);

  char dev[256];
  nusifDeviceName(dev, sizeof dev);
  printf("Device: %s\n", dev);
  const int verbose = nusifVerboseLevel(&p);
  if (!verbose) initProgress(d.te);

  double t = 0.0, res = 0.0;
  int nt   = 0;
  long long sweeps = 0;
  double t0 = getTimeStamp();
  while (t <= d.te && (nsteps <= 0 || nt < nsteps)) {
    if (d.tau > 0.0) computeTimestep(&d);
    setBoundaryConditions(&d);
    setSpecialBoundaryCondition(&d);
    computeFG(&d);
    computeRHS(&d);
    if (nt % 100 == 0) normalizePressure(&d);
    res = solve(&s, d.p, d.rhs);
    adaptUV(&d);
    sweeps += nusifSolverIterations(&s);
    writeResidual(fp, t, res);
    t += d.dt;
    nt++;
    if (verbose) printf("TIME %f , TIMESTEP %f\n", t, d.dt);
    else printProgress(t);
  }
  nusifDeviceSynchronize();
  double t1 = getTimeStamp();
  if (!verbose) stopProgress();
  double cells = (double)d.grid.imax * d.grid.jmax * d.grid.kmax;
  printf("Solution took %.2fs\n", t1 - t0);
  printf("Steps: %d  steps/s: %.3f  MLUP/s (step): %.2f  solver iterations: %lld\n", nt,
      nt / (t1 - t0), cells * nt / (t1 - t0) * 1e-6, sweeps);
  fclose(fp);

  double t2     = getTimeStamp();
  size_t bytes  = (size_t)d.grid.imax * d.grid.jmax * d.grid.kmax * sizeof(double);
  double *pg    = allocate(ARRAY_ALIGNMENT, bytes);
  double *ug    = allocate(ARRAY_ALIGNMENT, bytes);
  double *vg    = allocate(ARRAY_ALIGNMENT, bytes);
  double *wg    = allocate(ARRAY_ALIGNMENT, bytes);
  commCollectResult(s.comm, ug, vg, wg, pg, d.u, d.v, d.w, d.p, d.grid.kmax, d.grid.jmax,
      d.grid.imax);
  VtkOptions opts = {.grid = s.grid, .fmt = ASCII};
  vtkOpen(&opts, s.problem);
  vtkScalar(&opts, "pressure", pg);
  vtkVector(&opts, "velocity", (VtkVector){ug, vg, wg});
  vtkClose(&opts);
  printf("Result output took %.2fs\n", getTimeStamp() - t2);

  finalizeProfiler();
  commFinalize(s.comm);
  return EXIT_SUCCESS;
}
