/* The warm-start network: a 3-D convolutional U-Net in BF16 with FP32
 * accumulation, forward and backward on MFMA (unet_kernels.h), trained on
 * the GPU with Adam. No counterpart in the reference.
 *
 * Architecture (L levels, channels ch[0..L-1], 3x3x3 convolutions, ReLU):
 *   encoder l:  conv(cin_l -> ch_l), conv(ch_l -> ch_l) = s_l; 2x2x2 average
 *               pool of s_l feeds level l+1
 *   decoder l:  nearest x2 upsample of level l+1, concatenated with s_l,
 *               conv(ch_{l+1} + ch_l -> ch_l), conv(ch_l -> ch_l)
 *   head:       1x1 convolution ch_0 -> 1 (FP32)
 * Parameters are one flat FP32 array (per conv W[cout][27][cin] then
 * b[cout]; then head W[ch_0], b) — the same layout as the CPU oracle
 * (oracle/orc_unet.c), so both can be driven with one parameter vector. */
#include <math.h>
#include <string.h>

#include <type_traits>
#include <vector>

#include "nusif_internal.h"
#include "reduce.h"
#include "unet.h"
#include "unet_kernels.h"

/* ---- elementwise kernels ------------------------------------------------- */
static inline unsigned blocksFor(size_t n) { return (unsigned)((n + 255) / 256 < 65535 * 16 ? (n + 255) / 256 : 65535 * 16); }

#define GRID_STRIDE(q, n)                                                                    \
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < (n);                    \
       q += (size_t)gridDim.x * blockDim.x)

__global__ void kFillF(float *a, size_t n, float v) { GRID_STRIDE(q, n) a[q] = v; }
__global__ void kFillH(uint16_t *a, size_t n) { GRID_STRIDE(q, n) a[q] = 0; }

/* Box extents of a tensor; the coarse partner of a fine box has half of
 * each. Voxel v = (z * ny + y) * nx + x. */
struct Ext {
  int nx, ny, nz;
  __host__ __device__ size_t voxels() const { return (size_t)nx * ny * nz; }
  __host__ __device__ size_t at(int x, int y, int z) const { return ((size_t)z * ny + y) * nx + x; }
  __host__ __device__ Ext half() const { return {nx / 2, ny / 2, nz / 2}; }
};

/* out (extents m) = average of the 8 children in `in` (extents 2 m), bf16 */
__global__ void kPoolFwd(const uint16_t *in, uint16_t *out, Ext m, int c)
{
  const Ext n  = {2 * m.nx, 2 * m.ny, 2 * m.nz};
  size_t tot   = m.voxels() * c;
  GRID_STRIDE(q, tot)
  {
    int k = (int)(q % c);
    size_t v = q / c;
    int x = (int)(v % m.nx), y = (int)((v / m.nx) % m.ny), z = (int)(v / ((size_t)m.nx * m.ny));
    float s = 0.f;
    for (int e = 0; e < 8; e++)
      s += bf2f(in[n.at(2 * x + (e & 1), 2 * y + ((e >> 1) & 1), 2 * z + (e >> 2)) * c + k]);
    out[q] = f2bf(0.125f * s);
  }
}

/* dIn (extents n) += dOut(parent) / 8 */
__global__ void kPoolBwd(const float *dOut, float *dIn, Ext n, int c)
{
  const Ext m = n.half();
  size_t tot  = n.voxels() * c;
  GRID_STRIDE(q, tot)
  {
    int k = (int)(q % c);
    size_t v = q / c;
    int x = (int)(v % n.nx), y = (int)((v / n.nx) % n.ny), z = (int)(v / ((size_t)n.nx * n.ny));
    dIn[q] += 0.125f * dOut[m.at(x / 2, y / 2, z / 2) * c + k];
  }
}

/* out (extents n) = in(parent), bf16 */
__global__ void kUpFwd(const uint16_t *in, uint16_t *out, Ext n, int c)
{
  const Ext m = n.half();
  size_t tot  = n.voxels() * c;
  GRID_STRIDE(q, tot)
  {
    int k = (int)(q % c);
    size_t v = q / c;
    int x = (int)(v % n.nx), y = (int)((v / n.nx) % n.ny), z = (int)(v / ((size_t)n.nx * n.ny));
    out[q] = in[m.at(x / 2, y / 2, z / 2) * c + k];
  }
}

/* dIn (extents m) = sum of the 8 children of dOut (extents 2 m) */
__global__ void kUpBwd(const float *dOut, float *dIn, Ext m, int c)
{
  const Ext n = {2 * m.nx, 2 * m.ny, 2 * m.nz};
  size_t tot  = m.voxels() * c;
  GRID_STRIDE(q, tot)
  {
    int k = (int)(q % c);
    size_t v = q / c;
    int x = (int)(v % m.nx), y = (int)((v / m.nx) % m.ny), z = (int)(v / ((size_t)m.nx * m.ny));
    float s = 0.f;
    for (int e = 0; e < 8; e++)
      s += dOut[n.at(2 * x + (e & 1), 2 * y + ((e >> 1) & 1), 2 * z + (e >> 2)) * c + k];
    dIn[q] = s;
  }
}

__global__ void kReluMask(float *d, const uint16_t *act, size_t n)
{
  GRID_STRIDE(q, n) if (!(bf2f(act[q]) > 0.f)) d[q] = 0.f;
}

/* y[v] = hb + sum_c hw[c] act[v][c], the channels added in order; a voxel's
 * channels are read eight at a time (one 16 B load) when c is a multiple of
 * eight */
__global__ void kHeadFwd(const uint16_t *act, int c, const float *hw, float *y, size_t nv)
{
  GRID_STRIDE(v, nv)
  {
    float s = hw[c];
    if ((c & 7) == 0) {
      for (int k0 = 0; k0 < c; k0 += 8) {
        const uint4 q      = *(const uint4 *)(act + v * c + k0);
        const uint32_t h[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int u = 0; u < 8; u++) s += hw[k0 + u] * bf2f((uint16_t)(h[u >> 1] >> (16 * (u & 1))));
      }
    } else {
      for (int k = 0; k < c; k++) s += hw[k] * bf2f(act[v * c + k]);
    }
    y[v] = s;
  }
}

/* head output scaled into a padded (nx+2)(ny+2)(nz+2) FP32 array with its
 * face ghosts: y * (float)(sqrt(scal[0] / voxels) * h2) (the warm start's
 * correction; the conv epilogue does the same where it runs the head) */
__global__ void kUnpackHead(const float *yv, Ext e, const double *scal, double h2, float *d)
{
  const size_t nv = e.voxels();
  const float sr  = (float)sqrt(scal[0] / (double)nv);
  const float f   = (float)(sr * h2);
  const long long sj = e.nx + 2, sk = (long long)(e.nx + 2) * (e.ny + 2);
  GRID_STRIDE(v, nv)
  {
    const int x = (int)(v % e.nx), y = (int)((v / e.nx) % e.ny), z = (int)(v / ((size_t)e.nx * e.ny));
    const long long c = (z + 1) * sk + (y + 1) * sj + x + 1;
    const float val   = yv[v] * f;
    d[c] = val;
    if (x == 0) d[c - 1] = val;
    if (x == e.nx - 1) d[c + 1] = val;
    if (y == 0) d[c - sj] = val;
    if (y == e.ny - 1) d[c + sj] = val;
    if (z == 0) d[c - sk] = val;
    if (z == e.nz - 1) d[c + sk] = val;
  }
}

/* dy = 2 (y - t) / nv; per-block partial sums of (y - t)^2 */
__global__ __launch_bounds__(256) void kLossGrad(const float *y, const float *t, size_t nv, float *dy, double *partials)
{
  __shared__ double lds[4];
  double acc = 0.0;
  GRID_STRIDE(v, nv)
  {
    float e = y[v] - t[v];
    dy[v]   = 2.f * e / (float)nv;
    acc += (double)e * e;
  }
  acc = blockSum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void kFoldLoss(const double *partials, int nb, double nv, double *out)
{
  __shared__ double lds[4];
  double s = foldSum(partials, nb, 1, 0, lds);
  if (threadIdx.x == 0) *out = s / nv;
}

/* dAct[v][k] = dy[v] hw[k]; grads of the head (atomics per block) */
__global__ __launch_bounds__(256) void kHeadBwd(const uint16_t *act, int c, const float *hw, const float *dy,
    float *dAct, float *ghw, size_t nv)
{
  __shared__ float part[64][17];
  float loc[17];
  for (int k = 0; k <= c; k++) loc[k] = 0.f;
  GRID_STRIDE(v, nv)
  {
    float g = dy[v];
    for (int k = 0; k < c; k++) {
      dAct[v * c + k] = g * hw[k];
      loc[k] += g * bf2f(act[v * c + k]);
    }
    loc[c] += g;
  }
  /* c <= 16: reduce over the workgroup in LDS, one atomic per parameter */
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k <= c; k++) {
    float x = loc[k];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) part[wave][k] = x;
  }
  __syncthreads();
  if (threadIdx.x <= (unsigned)c) {
    float s = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += part[w][threadIdx.x];
    atomicAdd(&ghw[threadIdx.x], s);
  }
}

/* gb[k] += sum_v d[v][k] */
__global__ __launch_bounds__(256) void kBiasGrad(const float *d, int c, size_t nv, float *gb)
{
  /* thread q handles channel q % c over a strided voxel subset */
  int k = threadIdx.x % c, lanesPerK = blockDim.x / c;
  if ((int)threadIdx.x >= lanesPerK * c) return;
  float s = 0.f;
  for (size_t v = (size_t)blockIdx.x * lanesPerK + threadIdx.x / c; v < nv; v += (size_t)gridDim.x * lanesPerK)
    s += d[v * c + k];
  atomicAdd(&gb[k], s);
}

/* bf16 weights and their transpose [ci][27][co] from the FP32 master */
__global__ void kWeightsBf16(const float *W, uint16_t *Wb, uint16_t *Wt, int co, int ci)
{
  size_t tot = (size_t)co * 27 * ci;
  GRID_STRIDE(q, tot)
  {
    int c = (int)(q % ci);
    int t = (int)((q / ci) % 27);
    int o = (int)(q / ((size_t)ci * 27));
    uint16_t h = f2bf(W[q]);
    Wb[q]      = h;
    Wt[((size_t)c * 27 + t) * co + o] = h;
  }
}

__global__ void kAdam(float *p, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2,
    float eps, float c1, float c2, float gscale)
{
  GRID_STRIDE(q, n)
  {
    float gq = g[q] * gscale;
    float mq = b1 * m[q] + (1.f - b1) * gq;
    float vq = b2 * v[q] + (1.f - b2) * gq * gq;
    m[q]     = mq;
    v[q]     = vq;
    p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
  }
}

/* ---- model ---------------------------------------------------------------- */
/* 16 x 4 x 4 output bricks of the LDS-staged forward convolution */
static int convBricks(const Ext &e)
{
  return ((e.nx + CL_BX - 1) / CL_BX) * ((e.ny + CL_BY - 1) / CL_BY) * ((e.nz + CL_BZ - 1) / CL_BZ);
}

struct ConvL {
  int cin, cout;
  size_t wOff, bOff;
};

struct LevelBuf {
  Ext e;                     /* extents */
  uint16_t *a, *s, *pl, *up, *da, *d; /* activations (bf16) */
  float *gA, *gS, *gD, *gU, *gP;      /* gradients (fp32) */
};

struct UnetModel {
  int L, ch[UNET_MAXL], cin;
  Ext e; /* extents of the finest level */
  int nconv;
  ConvL conv[4 * UNET_MAXL];
  size_t nParams, headW, headB;
  float *P, *G, *M, *V;
  uint16_t *Wb, *Wt;
  LevelBuf lv[UNET_MAXL];
  float *y, *dy;
  double *partials, *lossDev, *lossHost;
  int step;
  float *split;    /* split-K partial sums of the small-cube forward convs */
  size_t splitCap; /* floats */
};

template <typename T> static T *dalloc(size_t n)
{
  T *p = nullptr;
  HIP_CHECK(hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)));
  return p;
}

static void layout(UnetModel *m)
{
  int q = 0;
  for (int l = 0; l < m->L; l++) {
    m->conv[q++] = {l == 0 ? m->cin : m->ch[l - 1], m->ch[l], 0, 0};
    m->conv[q++] = {m->ch[l], m->ch[l], 0, 0};
  }
  for (int l = m->L - 2; l >= 0; l--) {
    m->conv[q++] = {m->ch[l + 1] + m->ch[l], m->ch[l], 0, 0};
    m->conv[q++] = {m->ch[l], m->ch[l], 0, 0};
  }
  m->nconv   = q;
  size_t off = 0;
  for (int k = 0; k < q; k++) {
    m->conv[k].wOff = off;
    off += (size_t)m->conv[k].cout * 27 * m->conv[k].cin;
    m->conv[k].bOff = off;
    off += (size_t)m->conv[k].cout;
  }
  m->headW   = off;
  m->headB   = off + m->ch[0];
  m->nParams = off + m->ch[0] + 1;
}

static void allocBuffers(UnetModel *m)
{
  Ext e = m->e;
  for (int l = 0; l < m->L; l++, e = e.half()) {
    LevelBuf &b = m->lv[l];
    b.e         = e;
    size_t v    = e.voxels();
    int c = m->ch[l], cn = l + 1 < m->L ? m->ch[l + 1] : 0, cp = l > 0 ? m->ch[l - 1] : m->cin;
    b.a  = dalloc<uint16_t>(v * c);
    b.s  = dalloc<uint16_t>(v * c);
    b.pl = l + 1 < m->L ? dalloc<uint16_t>(v / 8 * c) : nullptr;
    b.up = cn ? dalloc<uint16_t>(v * cn) : nullptr;
    b.da = cn ? dalloc<uint16_t>(v * c) : nullptr;
    b.d  = cn ? dalloc<uint16_t>(v * c) : nullptr;
    b.gA = dalloc<float>(v * c);
    b.gS = dalloc<float>(v * c);
    b.gD = dalloc<float>(v * c);
    b.gU = cn ? dalloc<float>(v * cn) : nullptr;
    b.gP = dalloc<float>(v * cp);
  }
  m->y        = dalloc<float>(m->e.voxels());
  m->dy       = dalloc<float>(m->e.voxels());
  m->partials = dalloc<double>(65536);
  m->lossDev  = dalloc<double>(1);
  HIP_CHECK(hipHostMalloc((void **)&m->lossHost, sizeof(double), hipHostMallocDefault));
}

UnetModel *unetNew(int L, const int *ch, int cin, int nx, int ny, int nz)
{
  if (L < 1 || L > UNET_MAXL) NUSIF_FATAL("unet: levels %d outside 1..%d", L, UNET_MAXL);
  if (cin % 8) NUSIF_FATAL("unet: input channels %d must be a multiple of 8", cin);
  const int div = 1 << (L - 1);
  if (nx < 1 || ny < 1 || nz < 1 || nx % div || ny % div || nz % div)
    NUSIF_FATAL("unet: extents %d x %d x %d not all divisible by 2^(L-1) = %d", nx, ny, nz, div);
  for (int l = 0; l < L; l++)
    if (ch[l] % 16 || ch[l] > 128) NUSIF_FATAL("unet: channels must be multiples of 16 up to 128");
  if (ch[0] > 16) NUSIF_FATAL("unet: the head supports at most 16 channels on level 0");
  UnetModel *m = new UnetModel();
  memset(m, 0, sizeof(*m));
  m->L   = L;
  m->cin = cin;
  m->e   = {nx, ny, nz};
  for (int l = 0; l < L; l++) m->ch[l] = ch[l];
  layout(m);
  m->P  = dalloc<float>(m->nParams);
  m->G  = dalloc<float>(m->nParams);
  m->M  = dalloc<float>(m->nParams);
  m->V  = dalloc<float>(m->nParams);
  m->Wb = dalloc<uint16_t>(m->nParams);
  m->Wt = dalloc<uint16_t>(m->nParams);
  hipStream_t st = nusifStream();
  kFillF<<<blocksFor(m->nParams), 256, 0, st>>>(m->G, m->nParams, 0.f);
  kFillF<<<blocksFor(m->nParams), 256, 0, st>>>(m->M, m->nParams, 0.f);
  kFillF<<<blocksFor(m->nParams), 256, 0, st>>>(m->V, m->nParams, 0.f);
  allocBuffers(m);
  /* split-K scratch: the largest chunks x voxels x cout over the convs whose
   * cube has fewer than 256 bricks (same order as layout()) */
  {
    size_t cap = 0;
    int q      = 0;
    auto need  = [&](int k, Ext e) {
      int bricks = convBricks(e);
      int chunks = (m->conv[k].cin + CL_CC - 1) / CL_CC;
      if (bricks < 256 && chunks > 1) cap = std::max(cap, (size_t)chunks * e.voxels() * m->conv[k].cout);
    };
    for (int l = 0; l < L; l++) {
      need(q++, m->lv[l].e);
      need(q++, m->lv[l].e);
    }
    for (int l = L - 2; l >= 0; l--) {
      need(q++, m->lv[l].e);
      need(q++, m->lv[l].e);
    }
    m->splitCap = cap;
    m->split    = cap ? dalloc<float>(cap) : nullptr;
  }
  return m;
}

void unetFree(UnetModel *m)
{
  if (!m) return;
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  void *ptrs[] = {m->P, m->G, m->M, m->V, m->Wb, m->Wt, m->y, m->dy, m->partials, m->lossDev, m->split};
  for (void *p : ptrs)
    if (p) HIP_CHECK(hipFree(p));
  for (int l = 0; l < m->L; l++) {
    LevelBuf &b = m->lv[l];
    void *q[] = {b.a, b.s, b.pl, b.up, b.da, b.d, b.gA, b.gS, b.gD, b.gU, b.gP};
    for (void *p : q)
      if (p) HIP_CHECK(hipFree(p));
  }
  HIP_CHECK(hipHostFree(m->lossHost));
  delete m;
}

size_t unetParamCount(const UnetModel *m) { return m->nParams; }

/* He-normal weights (seeded Box-Muller on the host), zero biases */
void unetInit(UnetModel *m, unsigned seed)
{
  std::vector<float> h(m->nParams, 0.f);
  unsigned long long x = 0x9E3779B97F4A7C15ull ^ (seed * 0xD1B54A32D192ED03ull + 1);
  auto uni = [&]() {
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    return ((x * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int k = 0; k < m->nconv; k++) {
    double sd = sqrt(2.0 / (27.0 * m->conv[k].cin));
    size_t nw = (size_t)m->conv[k].cout * 27 * m->conv[k].cin;
    for (size_t q = 0; q < nw; q++) {
      double u1 = uni() + 1e-12, u2 = uni();
      h[m->conv[k].wOff + q] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
    }
  }
  for (int c = 0; c < m->ch[0]; c++) h[m->headW + c] = (float)(0.1 * (uni() - 0.5));
  HIP_CHECK(hipMemcpy(m->P, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  m->step = 0;
}

/* The library stream is non-blocking: a plain hipMemcpy (null stream) does
 * not wait for the kernels queued on it, so each host copy below first
 * drains that stream (the kernels that write, or still read, the buffer). */
void unetSetParams(UnetModel *m, const float *hostParams)
{
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  HIP_CHECK(hipMemcpy(m->P, hostParams, m->nParams * sizeof(float), hipMemcpyHostToDevice));
}

void unetGetParams(UnetModel *m, float *hostParams)
{
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  HIP_CHECK(hipMemcpy(hostParams, m->P, m->nParams * sizeof(float), hipMemcpyDeviceToHost));
}

void unetGetGrads(UnetModel *m, float *hostGrads)
{
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  HIP_CHECK(hipMemcpy(hostGrads, m->G, m->nParams * sizeof(float), hipMemcpyDeviceToHost));
}

void unetZeroGrads(UnetModel *m)
{
  kFillF<<<blocksFor(m->nParams), 256, 0, nusifStream()>>>(m->G, m->nParams, 0.f);
}

/* refresh the bf16 copies (and transposes) of every conv weight block */
void unetSyncWeights(UnetModel *m)
{
  for (int k = 0; k < m->nconv; k++) {
    const ConvL &c = m->conv[k];
    size_t nw      = (size_t)c.cout * 27 * c.cin;
    kWeightsBf16<<<blocksFor(nw), 256, 0, nusifStream()>>>(m->P + c.wOff, m->Wb + c.wOff, m->Wt + c.wOff,
        c.cout, c.cin);
  }
}

/* ---- convolution launchers ---------------------------------------------- */
template <typename InT, typename OutT> static void launchConv(const ConvArgs<InT, OutT> &a)
{
  int rows = a.r0 + a.r1, nvox = (int)a.voxels();
  const int MB = 2;
  dim3 blk(256);
  auto grid = [&](int NB) { return dim3((nvox + 4 * 16 * MB - 1) / (4 * 16 * MB), (rows + 16 * NB - 1) / (16 * NB)); };
  if (rows >= 64)
    NUSIF_TIMED("unet_conv", (kConvMfma<InT, OutT, 4, MB><<<grid(4), blk, 0, nusifStream()>>>(a)));
  else if (rows >= 32)
    NUSIF_TIMED("unet_conv", (kConvMfma<InT, OutT, 2, MB><<<grid(2), blk, 0, nusifStream()>>>(a)));
  else
    NUSIF_TIMED("unet_conv", (kConvMfma<InT, OutT, 1, MB><<<grid(1), blk, 0, nusifStream()>>>(a)));
}

/* Forward convolutions on LDS-staged bricks; row tiles chosen so the grid
 * has enough workgroups. On small cubes (fewer than 256 bricks) with
 * several 32-channel chunks K is split over the grid (grid.z) into the
 * model's FP32 scratch, summed in chunk order by kConvSplitFinish, and
 * layers that still have few workgroups give each brick two K-groups of
 * four waves that split every chunk's K-steps (sums added in LDS, group
 * order): the result does not depend on scheduling. Returns false when the
 * conv ran split over the grid: its epilogue extras (pool, head) were then
 * not applied. */
static bool launchConvLds(UnetModel *m, const ConvArgs<uint16_t, uint16_t> &a)
{
  int rows = a.r0 + a.r1, cin = a.c0 + a.c1;
  unsigned bricks = (unsigned)convBricks({a.nx, a.ny, a.nz});
  int chunks      = (cin + CL_CC - 1) / CL_CC;
  int split       = 1;
  if (bricks < 256 && chunks > 1 && (size_t)chunks * a.voxels() * rows <= m->splitCap) split = chunks;
  const int nb = rows >= 64 && bricks >= 512 ? 4 : (rows >= 32 && bricks >= 512 ? 2 : 1);
  const int gb = (cin % CL_CC) / 8; /* groups of the partial chunk */
  /* two K-groups where the layer has few workgroups (measured faster below
   * ~384 workgroups with full chunks; slower for the wider grids and the
   * partial-chunk layers) */
  const unsigned wgs = bricks * (unsigned)((rows + 15) / 16) * (unsigned)split;
  const bool ks2     = bricks < 256 && nb == 1 && cin >= CL_CC && wgs < 384;
  if (a.headW && rows > 16 * nb) NUSIF_FATAL("unet: the head needs every channel of a voxel in one workgroup");
  auto grid = [&](int NB) { return dim3(bricks, (rows + 16 * NB - 1) / (16 * NB), split); };
  hipStream_t st = nusifStream();
  auto launch = [&](auto nbC, auto gbC, auto ksC) {
    constexpr int NB = decltype(nbC)::value, GB = decltype(gbC)::value, KS = decltype(ksC)::value;
    if constexpr (GB > 0) {
      if (cin < CL_CC) {
        NUSIF_TIMED("unet_conv", (kConvLds<uint16_t, NB, GB, false, KS><<<grid(NB), 256 * KS, 0, st>>>(a, m->split)));
        return;
      }
    }
    NUSIF_TIMED("unet_conv", (kConvLds<uint16_t, NB, GB, true, KS><<<grid(NB), 256 * KS, 0, st>>>(a, m->split)));
  };
  auto withGb = [&](auto nbC, auto ksC) {
    switch (gb) {
    case 0: launch(nbC, std::integral_constant<int, 0>(), ksC); break;
    case 1: launch(nbC, std::integral_constant<int, 1>(), ksC); break;
    case 2: launch(nbC, std::integral_constant<int, 2>(), ksC); break;
    default: launch(nbC, std::integral_constant<int, 3>(), ksC); break;
    }
  };
  if (ks2) withGb(std::integral_constant<int, 1>(), std::integral_constant<int, 2>());
  else if (nb == 4) withGb(std::integral_constant<int, 4>(), std::integral_constant<int, 1>());
  else if (nb == 2) withGb(std::integral_constant<int, 2>(), std::integral_constant<int, 1>());
  else withGb(std::integral_constant<int, 1>(), std::integral_constant<int, 1>());
  if (split > 1) {
    size_t total = (size_t)a.voxels() * rows;
    unsigned nbk = (unsigned)std::min<size_t>((total + 255) / 256, 4096);
    kConvSplitFinish<uint16_t><<<nbk, 256, 0, st>>>(a, m->split, split);
  }
  return split == 1;
}

/* epilogue requests of one forward conv */
struct ConvEpilogue {
  uint16_t *pool = nullptr;     /* 2x2x2 average of the result (extents / 2) */
  bool head = false;            /* the model's 1x1 head into m->y */
  bool storeAct = true;         /* keep the activation itself */
  float *corr = nullptr;        /* head output scaled into a padded correction array */
  const double *corrScal = nullptr;
  double corrH2 = 0.0;
};

/* forward conv k: out = relu(conv(in0 ++ in1) + b), with the epilogue's
 * pooling and head fused where the conv runs in one pass (otherwise as
 * separate kernels, same values) */
static void convForward(UnetModel *m, int k, const uint16_t *in0, int c0, const uint16_t *in1, int c1,
    uint16_t *out, Ext e, bool half0 = false, const ConvEpilogue &ep = ConvEpilogue(),
    const float *packR = nullptr, const float *packB = nullptr, const double *packScal = nullptr)
{
  const ConvL &c = m->conv[k];
  ConvArgs<uint16_t, uint16_t> a;
  a.in0 = in0;
  a.in1 = in1;
  a.c0 = c0;
  a.c1 = c1;
  a.A = m->Wb + c.wOff;
  a.bias = m->P + c.bOff;
  a.out0 = out;
  a.out1 = nullptr;
  a.r0 = c.cout;
  a.r1 = 0;
  a.nx = e.nx;
  a.ny = e.ny;
  a.nz = e.nz;
  a.tapSign = 1;
  a.relu = 1;
  a.accumulate = 0;
  a.half0 = half0 ? 1 : 0;
  bool fused = false;
  if ((c0 + c1) % 8 == 0) {
    ConvArgs<uint16_t, uint16_t> f = a;
    f.pool = ep.pool;
    if (packR) {
      if (c0 != 8 || c1 != 0) NUSIF_FATAL("unet: the packed input needs 8 input channels");
      f.packR    = packR;
      f.packB    = packB;
      f.packScal = packScal;
    }
    if (ep.head) {
      f.headW    = m->P + m->headW;
      f.headY    = m->y;
      f.corr     = ep.corr;
      f.corrScal = ep.corrScal;
      f.corrH2   = ep.corrH2;
      f.storeAct = ep.storeAct ? 1 : 0;
    }
    /* a grid split of K skips the epilogue (its finishing pass stores the
     * activation); the stand-alone kernels below follow */
    fused = launchConvLds(m, f);
  } else if (half0) {
    NUSIF_FATAL("unet: an upsampled input needs channels in multiples of 8");
  } else {
    launchConv(a);
  }
  if (!fused) {
    hipStream_t st = nusifStream();
    if (ep.pool) {
      size_t nout = e.half().voxels() * c.cout;
      kPoolFwd<<<blocksFor(nout), 256, 0, st>>>(out, ep.pool, e.half(), c.cout);
    }
    if (ep.head) {
      kHeadFwd<<<blocksFor(e.voxels()), 256, 0, st>>>(out, c.cout, m->P + m->headW, m->y, e.voxels());
      if (ep.corr)
        kUnpackHead<<<blocksFor(e.voxels()), 256, 0, st>>>(m->y, e, ep.corrScal, ep.corrH2, ep.corr);
    }
  }
}

/* backward of conv k given dOut (already ReLU-masked): dIn0 (= or +=),
 * dIn1 (+=) when requested, weight and bias gradients */
static void convBackward(UnetModel *m, int k, const float *dOut, const uint16_t *in0, int c0,
    const uint16_t *in1, int c1, float *dIn0, int acc0, float *dIn1, Ext e)
{
  const ConvL &c = m->conv[k];
  if (dIn0 || dIn1) {
    if (dIn1 && acc0 == 0) {
      /* split rows: dIn0 overwritten, dIn1 accumulated -> two launches */
      ConvArgs<float, float> a;
      a.in0 = dOut; a.in1 = nullptr; a.c0 = c.cout; a.c1 = 0;
      a.A = m->Wt + c.wOff; a.bias = nullptr; a.nx = e.nx; a.ny = e.ny; a.nz = e.nz; a.tapSign = -1; a.relu = 0;
      a.out0 = dIn0; a.r0 = c0; a.out1 = nullptr; a.r1 = 0; a.accumulate = 0;
      launchConv(a);
      a.A = m->Wt + c.wOff + (size_t)c0 * 27 * c.cout;
      a.out0 = dIn1; a.r0 = c1; a.accumulate = 1;
      launchConv(a);
    } else {
      ConvArgs<float, float> a;
      a.in0 = dOut; a.in1 = nullptr; a.c0 = c.cout; a.c1 = 0;
      a.A = m->Wt + c.wOff; a.bias = nullptr; a.nx = e.nx; a.ny = e.ny; a.nz = e.nz; a.tapSign = -1; a.relu = 0;
      a.out0 = dIn0; a.r0 = c0; a.out1 = dIn1; a.r1 = dIn1 ? c1 : 0; a.accumulate = acc0;
      launchConv(a);
    }
  }
  int cin = c0 + c1, nvox = (int)e.voxels();
  int chunk = 2048;
  dim3 g((cin + 31) / 32, 27 * ((c.cout + 31) / 32), (nvox + chunk - 1) / chunk);
  NUSIF_TIMED("unet_wgrad", (kConvWgrad<float><<<g, 256, 0, nusifStream()>>>(dOut, c.cout, in0, in1, c0, c1,
                                e.nx, e.ny, e.nz, m->G + c.wOff, chunk)));
  int cb = c.cout;
  kBiasGrad<<<256, 256, 0, nusifStream()>>>(dOut, cb, (size_t)nvox, m->G + c.bOff);
}

/* ---- forward / backward --------------------------------------------------- */
void unetForward(UnetModel *m, const uint16_t *x, bool forTraining, float *corr, const double *corrScal,
    double corrH2, const float *packR, const float *packB)
{
  int L = m->L;
  const uint16_t *cur = x;
  int ccur = m->cin, q = 0;
  hipStream_t st = nusifStream();
  for (int l = 0; l < L; l++) {
    LevelBuf &b = m->lv[l];
    if (l == 0 && packR) /* the first layer reads the FP32 input arrays (scaled by corrScal's norms) */
      convForward(m, q++, cur, ccur, nullptr, 0, b.a, b.e, false, ConvEpilogue(), packR, packB, corrScal);
    else
      convForward(m, q++, cur, ccur, nullptr, 0, b.a, b.e);
    ConvEpilogue ep;
    if (l + 1 < L) ep.pool = b.pl; /* pooled for the next level in the same pass */
    if (L == 1) {
      ep.head = true;
      ep.corr = corr;
      ep.corrScal = corrScal;
      ep.corrH2 = corrH2;
    }
    convForward(m, q++, b.a, m->ch[l], nullptr, 0, b.s, b.e, false, ep);
    if (l + 1 < L) {
      cur  = b.pl;
      ccur = m->ch[l];
    }
  }
  const uint16_t *prev = m->lv[L - 1].s;
  for (int l = L - 2; l >= 0; l--) {
    LevelBuf &b = m->lv[l];
    size_t nup  = b.e.voxels() * m->ch[l + 1];
    if (forTraining) { /* the backward pass reads the upsampled tensor */
      kUpFwd<<<blocksFor(nup), 256, 0, st>>>(prev, b.up, b.e, m->ch[l + 1]);
      convForward(m, q++, b.up, m->ch[l + 1], m->lv[l].s, m->ch[l], b.da, b.e);
    } else { /* the conv reads the coarse tensor through its parent index */
      convForward(m, q++, prev, m->ch[l + 1], m->lv[l].s, m->ch[l], b.da, b.e, true);
    }
    ConvEpilogue ep;
    if (l == 0) { /* the head (and the scaled correction) in the last conv's pass */
      ep.head     = true;
      ep.storeAct = forTraining; /* the backward pass reads the last activation */
      ep.corr     = corr;
      ep.corrScal = corrScal;
      ep.corrH2   = corrH2;
    }
    convForward(m, q++, b.da, m->ch[l], nullptr, 0, b.d, b.e, false, ep);
    prev = b.d;
  }
}

const float *unetOutput(UnetModel *m) { return m->y; }

/* loss of the last forward against target t; gradients accumulated into G */
double unetBackward(UnetModel *m, const uint16_t *x, const float *t, bool wantLoss)
{
  int L = m->L;
  hipStream_t st = nusifStream();
  size_t nv = m->e.voxels();
  unsigned nb = blocksFor(nv) > 1024 ? 1024 : blocksFor(nv);
  kLossGrad<<<nb, 256, 0, st>>>(m->y, t, nv, m->dy, m->partials);
  kFoldLoss<<<1, 256, 0, st>>>(m->partials, (int)nb, (double)nv, m->lossDev);
  const uint16_t *top = L > 1 ? m->lv[0].d : m->lv[0].s;
  float *dTop         = L > 1 ? m->lv[0].gD : m->lv[0].gS;
  kHeadBwd<<<256, 256, 0, st>>>(top, m->ch[0], m->P + m->headW, m->dy, dTop, m->G + m->headW, nv);
  for (int l = 0; l < L - 1; l++) {
    size_t n = m->lv[l].e.voxels() * m->ch[l];
    kFillF<<<blocksFor(n), 256, 0, st>>>(m->lv[l].gS, n, 0.f);
  }
  int q = m->nconv;
  /* decoder, top level first */
  for (int l = 0; l <= L - 2; l++) {
    LevelBuf &b = m->lv[l];
    size_t v    = b.e.voxels();
    int ib = q - 1, ia = q - 2;
    q -= 2;
    kReluMask<<<blocksFor(v * m->ch[l]), 256, 0, st>>>(b.gD, b.d, v * m->ch[l]);
    convBackward(m, ib, b.gD, b.da, m->ch[l], nullptr, 0, b.gA, 0, nullptr, b.e);
    kReluMask<<<blocksFor(v * m->ch[l]), 256, 0, st>>>(b.gA, b.da, v * m->ch[l]);
    convBackward(m, ia, b.gA, b.up, m->ch[l + 1], b.s, m->ch[l], b.gU, 0, b.gS, b.e);
    LevelBuf &nx = m->lv[l + 1];
    float *dst   = (l + 1 == L - 1) ? nx.gS : nx.gD;
    kUpBwd<<<blocksFor(nx.e.voxels() * m->ch[l + 1]), 256, 0, st>>>(b.gU, dst, nx.e, m->ch[l + 1]);
  }
  /* encoder, deepest level first */
  for (int l = L - 1; l >= 0; l--) {
    LevelBuf &b = m->lv[l];
    size_t v    = b.e.voxels();
    int ia = 2 * l, ib = 2 * l + 1;
    kReluMask<<<blocksFor(v * m->ch[l]), 256, 0, st>>>(b.gS, b.s, v * m->ch[l]);
    convBackward(m, ib, b.gS, b.a, m->ch[l], nullptr, 0, b.gA, 0, nullptr, b.e);
    kReluMask<<<blocksFor(v * m->ch[l]), 256, 0, st>>>(b.gA, b.a, v * m->ch[l]);
    const uint16_t *inl = l == 0 ? x : m->lv[l - 1].pl;
    int cinl            = l == 0 ? m->cin : m->ch[l - 1];
    convBackward(m, ia, b.gA, inl, cinl, nullptr, 0, l > 0 ? b.gP : nullptr, 0, nullptr, b.e);
    if (l > 0) {
      LevelBuf &pv = m->lv[l - 1];
      kPoolBwd<<<blocksFor(pv.e.voxels() * m->ch[l - 1]), 256, 0, st>>>(b.gP, pv.gS, pv.e, m->ch[l - 1]);
    }
  }
  if (!wantLoss) return 0.0;
  HIP_CHECK(hipMemcpyAsync(m->lossHost, m->lossDev, sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  return *m->lossHost;
}

double unetLastLoss(UnetModel *m)
{
  HIP_CHECK(hipMemcpyAsync(m->lossHost, m->lossDev, sizeof(double), hipMemcpyDeviceToHost, nusifStream()));
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  return *m->lossHost;
}

void unetAdam(UnetModel *m, float lr, int batch)
{
  m->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = 1.f - powf(b1, (float)m->step), c2 = 1.f - powf(b2, (float)m->step);
  kAdam<<<blocksFor(m->nParams), 256, 0, nusifStream()>>>(m->P, m->G, m->M, m->V, m->nParams, lr, b1, b2,
      1e-8f, c1, c2, 1.f / (float)batch);
  unetZeroGrads(m);
  unetSyncWeights(m);
}

/* ---- persistence: "NUSIFUNET1", L, ch[4], cin, nParams, FP32 params ------ */
int unetSave(UnetModel *m, const char *path)
{
  std::vector<float> h(m->nParams);
  unetGetParams(m, h.data());
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  int hdr[6] = {m->L, m->ch[0], m->ch[1], m->ch[2], m->ch[3], m->cin};
  long long np = (long long)m->nParams;
  fwrite("NUSIFUNET1", 1, 10, f);
  fwrite(hdr, sizeof(int), 6, f);
  fwrite(&np, sizeof(np), 1, f);
  fwrite(h.data(), sizeof(float), h.size(), f);
  /* trailer: the largest extent of the grid it was trained on */
  int ext = std::max(m->e.nx, std::max(m->e.ny, m->e.nz));
  fwrite("NUSIFEXT", 1, 8, f);
  fwrite(&ext, sizeof(int), 1, f);
  fclose(f);
  return 0;
}

/* the network extent recorded in a weights file, 0 if it records none */
int unetFileExtent(const char *path)
{
  FILE *f = fopen(path, "rb");
  if (!f) return 0;
  char magic[10];
  int hdr[6];
  long long np = 0;
  int ext      = 0;
  if (fread(magic, 1, 10, f) == 10 && memcmp(magic, "NUSIFUNET1", 10) == 0 &&
      fread(hdr, sizeof(int), 6, f) == 6 && fread(&np, sizeof(np), 1, f) == 1 && np > 0 &&
      fseek(f, (long)(np * (long long)sizeof(float)), SEEK_CUR) == 0) {
    char tag[8];
    if (fread(tag, 1, 8, f) != 8 || memcmp(tag, "NUSIFEXT", 8) != 0 || fread(&ext, sizeof(int), 1, f) != 1) ext = 0;
  }
  fclose(f);
  return ext;
}

/* the number of levels of the network in a weights file, 0 if unreadable */
int unetFileLevels(const char *path)
{
  FILE *f = fopen(path, "rb");
  if (!f) return 0;
  char magic[10];
  int hdr[6] = {0, 0, 0, 0, 0, 0};
  int L      = 0;
  if (fread(magic, 1, 10, f) == 10 && memcmp(magic, "NUSIFUNET1", 10) == 0 && fread(hdr, sizeof(int), 6, f) == 6)
    L = hdr[0];
  fclose(f);
  return L;
}

UnetModel *unetLoad(const char *path, int nx, int ny, int nz)
{
  FILE *f = fopen(path, "rb");
  if (!f) return nullptr;
  char magic[10];
  int hdr[6];
  long long np = 0;
  if (fread(magic, 1, 10, f) != 10 || memcmp(magic, "NUSIFUNET1", 10) != 0 ||
      fread(hdr, sizeof(int), 6, f) != 6 || fread(&np, sizeof(np), 1, f) != 1) {
    fclose(f);
    NUSIF_FATAL("unet: %s is not a weights file", path);
  }
  UnetModel *m = unetNew(hdr[0], hdr + 1, hdr[5], nx, ny, nz);
  if ((size_t)np != m->nParams) NUSIF_FATAL("unet: %s holds %lld parameters, expected %zu", path, np, m->nParams);
  std::vector<float> h(m->nParams);
  if (fread(h.data(), sizeof(float), h.size(), f) != h.size()) NUSIF_FATAL("unet: %s is truncated", path);
  fclose(f);
  unetSetParams(m, h.data());
  unetSyncWeights(m);
  return m;
}

int unetLevels(const UnetModel *m) { return m->L; }
void unetExtents(const UnetModel *m, int *nx, int *ny, int *nz)
{
  *nx = m->e.nx;
  *ny = m->e.ny;
  *nz = m->e.nz;
}
