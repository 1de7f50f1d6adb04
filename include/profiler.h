/* Region timers (reference src/profiler.h:11-37). The GPU library times its
 * kernels with HIP events instead (nusif.h); these keep main.c linking and
 * print the reference's solver MLUPS line when PROFILING is defined. */
#ifndef NUSIF_PROFILER_H
#define NUSIF_PROFILER_H
#include <stddef.h>
#include <stdio.h>
#include "comm.h"
#define MAX_STR_LENGTH 30
typedef enum { SOLVER = 0, COMM, NUMREGIONS } RegionsType;
#ifdef __cplusplus
extern "C" {
#endif
extern double T[NUMREGIONS];
extern size_t C[NUMREGIONS];
extern void printProfile(CommType *c, int it);
extern void initProfiler(CommType *c);
extern void finalizeProfiler(void);
#ifdef __cplusplus
}
#endif
#endif
