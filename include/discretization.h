/* Staggered-grid (MAC) discretization of the 3D incompressible Navier-Stokes
 * equations: the eight step functions of one Chorin-projection time step.
 * Struct and prototypes are those of the reference (src/discretization.h:12-41).
 *
 * In the GPU library the eight fields live in HBM; the host arrays p..w are
 * staging copies refreshed by nusifSyncHost()/commCollectResult() and pushed by
 * nusifSyncDevice() (nusif.h). */
#ifndef NUSIF_DISCRETIZATION_H
#define NUSIF_DISCRETIZATION_H

#include "comm.h"
#include "grid.h"
#include "parameter.h"

#ifdef __cplusplus
extern "C" {
#endif

enum BC { NOSLIP = 1, SLIP, OUTFLOW, PERIODIC };

typedef struct {
  /* geometry and grid information */
  Grid grid;
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
  CommType comm;
} Discretization;

/* src/discretization.h:34-41 */
extern void initDiscretization(Discretization *, Parameter *);
extern void computeRHS(Discretization *);
extern void normalizePressure(Discretization *);
extern void computeTimestep(Discretization *);
extern void setBoundaryConditions(Discretization *);
extern void setSpecialBoundaryCondition(Discretization *);
extern void computeFG(Discretization *);
extern void adaptUV(Discretization *);

#ifdef __cplusplus
}
#endif
#endif
