/* Monotonic wall clock (reference src/timing.h, src/timing.c:7-19). */
#ifndef NUSIF_TIMING_H
#define NUSIF_TIMING_H
#ifdef __cplusplus
extern "C" {
#endif
extern double getTimeStamp(void);
extern double getTimeResolution(void);
#ifdef __cplusplus
}
#endif
#endif
