/* Problem parameters read from a .par file.
 * Layout-identical to the reference's Parameter (src/parameter.h:8-20);
 * the three functions keep the reference prototypes (src/parameter.h:22-24).
 * Keys this library adds to the .par format (nsteps, solver, lid, u_inflow,
 * mg_*, warmstart, ...) are kept outside the struct, see nusif.h. */
#ifndef NUSIF_PARAMETER_H
#define NUSIF_PARAMETER_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int imax, jmax, kmax;
  double xlength, ylength, zlength;
  int itermax;
  double eps, omg;
  double re, tau, gamma;
  double te, dt;
  double gx, gy, gz;
  char *name;
  int bcLeft, bcRight, bcBottom, bcTop, bcFront, bcBack;
  double u_init, v_init, w_init, p_init;
  int levels, presmooth, postsmooth;
} Parameter;

void initParameter(Parameter *);
void readParameter(Parameter *, const char *);
void printParameter(Parameter *);

#ifdef __cplusplus
}
#endif
#endif
