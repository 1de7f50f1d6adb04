/* Learned warm start of the gmg solve (warmstart.hip). */
#ifndef NUSIF_WARMSTART_H
#define NUSIF_WARMSTART_H
#include <stdint.h>

#include <vector>

#include "nusif_internal.h"

struct WarmStart;
struct UnetModel;
/* netLevels: pooling levels of the network it will run (its grid must
 * halve netLevels - 1 times) */
WarmStart *warmstartCreate(NusifCtx *c, UnetModel *net, bool capture, int netLevels);
void warmstartFree(WarmStart *w);
int warmstartLevel(const WarmStart *w);
void warmstartExtents(const WarmStart *w, int *nx, int *ny, int *nz);
void warmstartSetNet(WarmStart *w, UnetModel *net);
UnetModel *warmstartNet(WarmStart *w);
bool warmstartCapturing(const WarmStart *w);
/* FP32 level-0 buffer the caller fills with r0 = rhs - A p before the guess */
float *warmstartResidualBuffer(WarmStart *w);
/* level l's residual buffer and geometry when 1 <= l <= the network level,
 * else null */
float *warmstartLevelResidual(WarmStart *w, int l, Geo *g);
/* sum of squares of the residual restricted to the acceptance level, from
 * the fine FP32 residual, or from level 1 when it is already there (level1) */
void warmstartCoarseNormLaunch(WarmStart *w, const float *rFine, bool level1);
const double *warmstartNormsAsync(WarmStart *w);
double warmstartReferenceNorm2(const WarmStart *w);
bool warmstartPrepare(WarmStart *w, const double *p, const double *rhs, double idx2, double idy2, double idz2,
    float *r32 = nullptr, float *eZero = nullptr, double *norm2 = nullptr, double *norm2Host = nullptr);
const float *warmstartCorrection(WarmStart *w);
void warmstartForward(WarmStart *w);
void warmstartProlongCorrection(WarmStart *w, float *fine);
/* the correction prolongated from the network grid down to level 1 only
 * (nothing to do when the network runs there); returns level 1's buffer and
 * geometry (null when the network runs on the finest grid) */
const float *warmstartProlongToLevel1(WarmStart *w, Geo *g1);
void warmstartCoarseAcceptLaunch(WarmStart *w);
/* the acceptance decision made on the device after warmstartCoarseAcceptLaunch
 * (the correction zeroed when rejected); warmstartDecision reads it once the
 * stream has run it */
void warmstartDecideLaunch(WarmStart *w, int force);
int warmstartDecision(const WarmStart *w);
void warmstartCaptureBegin(WarmStart *w, const double *p, const Geo &g0);
void warmstartCaptureEnd(WarmStart *w, const double *p);
size_t warmstartSamples(const WarmStart *w);
void warmstartTakeSamples(WarmStart *w, std::vector<uint16_t *> &xs, std::vector<float *> &ts);
#endif
