/* Pressure-solver plugin: struct and two-function API of the reference
 * (src/solver.h:12-36). Which solver runs is chosen by the library variant
 * (libnusif_rb.so / libnusif_mg.so mirror SOLVER=rb|mg, reference Makefile:19-21)
 * or by the added .par key `solver` (nusif.h). */
#ifndef NUSIF_SOLVER_H
#define NUSIF_SOLVER_H
#include "comm.h"
#include "discretization.h"
#include "grid.h"
#include "parameter.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  /* geometry and grid information */
  Grid *grid;
  /* arrays */
  double *p, *rhs;
  double *f, *g, *h;
  double *u, *v, *w;
  /* parameters */
  double eps, omega;
  double re, tau, gamma;
  double gx, gy, gz;
  /* time stepping */
  int itermax;
  double dt, te;
  double dtBound;
  char *problem;
  int bcLeft, bcRight, bcBottom, bcTop, bcFront, bcBack;
  /* communication */
  double **r, **e;
  int levels, presmooth, postsmooth;
  CommType *comm;
} Solver;

extern double solve(Solver *, double *, const double *);
extern void initSolver(Solver *, Discretization *, Parameter *);

#ifdef __cplusplus
}
#endif
#endif
