/* Single-domain communication interface.
 * One GPU owns the whole domain, so this is the sequential form of the
 * reference's comm layer (src/comm.h:21-69; sequential branches of src/comm.c):
 * every face is a physical boundary, exchanges/shifts/reductions are no-ops on
 * the host, and ghost layers are filled by device kernels instead. The struct is
 * layout-identical to the reference's CommType built without _MPI
 * (src/comm.h:27-38). */
#ifndef NUSIF_COMM_H
#define NUSIF_COMM_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum direction { LEFT = 0, RIGHT, BOTTOM, TOP, FRONT, BACK, NDIRS } DirectionType;
typedef enum coordinates { ICORD = 0, JCORD, KCORD, NCORDS } CoordinatesType;
typedef enum dimension { KDIM = 0, JDIM, IDIM, NDIMS } DimensionType;
enum layer { HALO = 0, BULK };
enum op { MIN = 0, MAX, SUM };

typedef struct {
  int rank;
  int size;
  int neighbours[NDIRS];
  int coords[NDIMS], dims[NDIMS];
  int imaxLocal, jmaxLocal, kmaxLocal;
} CommType;

/* src/comm.h:40-69 */
extern void commInit(CommType *c, int argc, char **argv);
extern void commPartition(CommType *c, int kmax, int jmax, int imax);
extern void commFinalize(CommType *comm);
extern void commPrintConfig(CommType *);
extern void commExchange(CommType *, double *);
extern void commShift(CommType *c, double *f, double *g, double *h);
extern void commReduceAll(double *v, int op);
extern void commReduce(double *v, double *o, int count, int op);
extern int commIsBoundary(CommType *c, DirectionType direction);
extern void commGetOffsets(CommType *c, int offsets[], int kmax, int jmax, int imax);
extern void commFreeCommunicator(CommType *comm);
extern void commUpdateDatatypes(
    CommType *oldcomm, CommType *newcomm, int imaxLocal, int jmaxLocal, int kmaxLocal);
/* Gathers cell-centred results (src/comm.c:436-476). When u, v, w, p are the
 * host arrays of a Discretization owned by the GPU library, their device copies
 * are downloaded first. */
extern void commCollectResult(CommType *c, double *ug, double *vg, double *wg,
    double *pg, double *u, double *v, double *w, double *p, int kmax, int jmax,
    int imax);

static inline bool commIsMaster(CommType *c) { return c->rank == 0; }

#ifdef __cplusplus
}
#endif
#endif
