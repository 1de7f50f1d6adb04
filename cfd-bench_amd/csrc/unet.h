/* Host interface of the warm-start U-Net (unet.hip). */
#ifndef NUSIF_UNET_H
#define NUSIF_UNET_H
#include <stddef.h>
#include <stdint.h>

#define UNET_MAXL 4

struct UnetModel;
/* a network on nx x ny x nz voxels (each divisible by 2^(L-1)) */
UnetModel *unetNew(int L, const int *ch, int cin, int nx, int ny, int nz);
void unetFree(UnetModel *m);
void unetInit(UnetModel *m, unsigned seed);
size_t unetParamCount(const UnetModel *m);
void unetSetParams(UnetModel *m, const float *hostParams);
void unetGetParams(UnetModel *m, float *hostParams);
void unetGetGrads(UnetModel *m, float *hostGrads);
void unetZeroGrads(UnetModel *m);
void unetSyncWeights(UnetModel *m);
/* x: nx ny nz x cin bf16 (channels-last); the prediction stays on the device */
/* forTraining: also keep the upsampled decoder inputs unetBackward reads.
 * corr (optional): the output also written, scaled by
 * (float)(sqrt(corrScal[0] / voxels) * corrH2), into a padded
 * (nx+2)(ny+2)(nz+2) FP32 array with its face ghosts. packR/packB
 * (optional, x then unused): the 8-channel input formed inside the first
 * layer from two padded FP32 arrays, scaled by 1 / sqrt(corrScal[0..1] /
 * voxels), channels 2..7 zero. */
void unetForward(UnetModel *m, const uint16_t *x, bool forTraining = false, float *corr = nullptr,
    const double *corrScal = nullptr, double corrH2 = 0.0, const float *packR = nullptr,
    const float *packB = nullptr);
const float *unetOutput(UnetModel *m);
/* after unetForward: loss against t (nx ny nz FP32), gradients added into G */
double unetBackward(UnetModel *m, const uint16_t *x, const float *t, bool wantLoss);
double unetLastLoss(UnetModel *m);
void unetAdam(UnetModel *m, float lr, int batch);
int unetSave(UnetModel *m, const char *path);
UnetModel *unetLoad(const char *path, int nx, int ny, int nz);
int unetFileExtent(const char *path);
int unetFileLevels(const char *path);
int unetLevels(const UnetModel *m);
void unetExtents(const UnetModel *m, int *nx, int *ny, int *nz);
#endif
