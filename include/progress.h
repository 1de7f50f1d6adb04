/* Progress bar and residual trace (reference src/progress.h:8-13). */
#ifndef NUSIF_PROGRESS_H
#define NUSIF_PROGRESS_H
#include <stdio.h>
#ifdef __cplusplus
extern "C" {
#endif
extern void initProgress(double);
extern void printProgress(double);
extern void stopProgress(void);
extern FILE *initResidualWriter(void);
extern void writeResidual(FILE *, double, double);
#ifdef __cplusplus
}
#endif
#endif
