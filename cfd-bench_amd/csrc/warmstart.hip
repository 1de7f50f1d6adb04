/* Learned warm start of the gmg pressure solve (no counterpart in the
 * reference, whose initial guess is the previous step's p: p persists
 * across steps, reference src/main.c:71).
 *
 * Inputs, on the network grid (multigrid level Ln: the first whose extents
 * are all <= unet_extent (64), or the last before that whose extents the
 * network's pooling still divides; restriction = average of the 8 children):
 *   channel 0: r0 / s_r,  r0 = rhs - A p_prev (the previous pressure against
 *              the new right-hand side), s_r = rms(r0)
 *   channel 1: rhs / s_b, s_b = rms(rhs)
 * Output y, rescaled: dp = y * s_r * (l / pi)^2, l = max(nx dx, ny dy, nz dz)
 * on the network grid (the inverse Laplacian of the smoothest mode of the
 * domain's longest side; l = n h on a cube), prolongated (trilinear) to the
 * finest grid; guess = p_prev + dp. The guess replaces p_prev only when its
 * residual rhs - A guess, restricted to the acceptance level (unet_accept_level
 * halvings below the network grid, 2 by default), has a smaller norm than r0
 * restricted there; otherwise the solve starts from p_prev. Either
 * way the multigrid iterates to the same tolerance, so the answer's accuracy
 * does not depend on the network.
 *
 * Training pairs come from the solver itself: with capture on, each solve
 * stores the input above and the target (p_solved - p_prev) restricted to
 * level Ln, divided by the same scale. */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "nusif_internal.h"
#include "reduce.h"
#include "timing.h"
#include "unet.h"
#include "unet_kernels.h"
#include "warmstart.h"

#define WTX 64
#define WTY 4

struct WsLevel {
  Geo g;
  float *r, *b, *d; /* residual, rhs, correction (padded, FP32) */
};

struct WarmStart {
  UnetModel *net;
  int Ln;                 /* network level */
  int Lc;                 /* level whose restricted residuals the acceptance compares (>= Ln) */
  int nx, ny, nz;         /* network grid extents */
  double h2;              /* output scale (l / pi)^2, l the domain's longest side */
  double cn[3];           /* 1 / h^2 per direction on the network grid */
  bool useGraph;          /* replay the forward + unpack from a captured HIP graph */
  hipGraphExec_t fwdExec; /* that graph once captured (null before) */
  WsLevel lv[12];
  uint16_t *x;            /* network input, nx ny nz voxels x 8 channels */
  double *pPrev;          /* FP64 copy of p for capture (capture only) */
  double *scal;           /* [0] sum r0^2, [1] sum b^2 on level Ln, [2] guess residual
                           * on level Lc, [3] sum r0^2 on level 0, [4] sum r0^2 on level Lc */
  double *scalHost;       /* pinned, host-coherent mirror of scal (the folds write both) */
  double *scalHostDev;    /* its device address */
  double *partials;
  /* capture */
  bool capture;
  std::vector<uint16_t *> dsX;
  std::vector<float *> dsT;
};

/* ---- kernels --------------------------------------------------------------- */
template <typename Tin>
__global__ __launch_bounds__(256) void kRestrictAvg(const Tin *__restrict__ f, Geo gf, float *__restrict__ c, Geo gc)
{
  int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int J = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int K = blockIdx.z + 1;
  if (I > gc.ni || J > gc.nj) return;
  double s = 0.0;
  for (int e = 0; e < 8; e++)
    s += (double)f[gf.at(2 * I - 1 + (e & 1), 2 * J - 1 + ((e >> 1) & 1), 2 * K - 1 + (e >> 2))];
  c[gc.at(I, J, K)] = (float)(0.125 * s);
}

/* 8-child averages of one or two level arrays (f1 -> c1, and f2 -> c2 when
 * f2 is not null), each summed as kRestrictAvg does; with partials, also each
 * workgroup's sums of c1^2 and c2^2 over the cells it wrote (stride 2, as
 * kSumSq2 lays them out), so the norms on the network grid need no pass of
 * their own */
__global__ __launch_bounds__(256) void kRestrictPair(const float *__restrict__ f1, const float *__restrict__ f2,
    Geo gf, float *__restrict__ c1, float *__restrict__ c2, Geo gc, double *partials)
{
  __shared__ double lds[4];
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int K = blockIdx.z + 1;
  double q1 = 0.0, q2 = 0.0;
  if (I <= gc.ni && J <= gc.nj) {
    const long long cc = gc.at(I, J, K);
    double s = 0.0;
    for (int e = 0; e < 8; e++)
      s += (double)f1[gf.at(2 * I - 1 + (e & 1), 2 * J - 1 + ((e >> 1) & 1), 2 * K - 1 + (e >> 2))];
    const float v1 = (float)(0.125 * s);
    c1[cc]         = v1;
    q1             = (double)v1 * v1;
    if (f2) {
      s = 0.0;
      for (int e = 0; e < 8; e++)
        s += (double)f2[gf.at(2 * I - 1 + (e & 1), 2 * J - 1 + ((e >> 1) & 1), 2 * K - 1 + (e >> 2))];
      const float v2 = (float)(0.125 * s);
      c2[cc]         = v2;
      q2             = (double)v2 * v2;
    }
  }
  if (partials) {
    q1 = blockSum(q1, lds);
    q2 = blockSum(q2, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      const unsigned bid    = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
      partials[2 * bid]     = q1;
      partials[2 * bid + 1] = q2;
    }
  }
}

/* R R f of one FP32 array in one pass: each thread takes one cell of the
 * level two below gf (gc) and its 8 children on the level between (gm),
 * each of those the 8-child average of gf; every average is summed as
 * kRestrictPair sums it, so gm and gc receive the values of two
 * kRestrictPair launches. Partials: each workgroup's sum of squares of its
 * gc cells (component 0, stride 2, as kRestrictPair lays them out). */
__global__ __launch_bounds__(256) void kRestrictChain2(const float *__restrict__ f, Geo gf, float *__restrict__ m,
    Geo gm, float *__restrict__ c, Geo gc, double *partials)
{
  __shared__ double lds[4];
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int K = blockIdx.z + 1;
  double q = 0.0;
  if (I <= gc.ni && J <= gc.nj) {
    float vm[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
      const int Im = 2 * I - 1 + (e & 1), Jm = 2 * J - 1 + ((e >> 1) & 1), Km = 2 * K - 1 + (e >> 2);
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < 8; d++)
        s += (double)f[gf.at(2 * Im - 1 + (d & 1), 2 * Jm - 1 + ((d >> 1) & 1), 2 * Km - 1 + (d >> 2))];
      vm[e]                = (float)(0.125 * s);
      m[gm.at(Im, Jm, Km)] = vm[e];
    }
    double sc = 0.0;
#pragma unroll
    for (int e = 0; e < 8; e++) sc += (double)vm[e];
    const float v    = (float)(0.125 * sc);
    c[gc.at(I, J, K)] = v;
    q = (double)v * v;
  }
  if (partials) {
    q = blockSum(q, lds);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      const unsigned bid    = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
      partials[2 * bid]     = q;
      partials[2 * bid + 1] = 0.0;
    }
  }
}

__global__ __launch_bounds__(256) void kCast(const double *__restrict__ a, float *__restrict__ b, Geo g)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > g.ni || j > g.nj) return;
  b[g.at(i, j, k)] = (float)a[g.at(i, j, k)];
}

/* per-block partial sums of r^2 and b^2 over the interior of level Ln */
/* sums of r^2 and b^2 over the interior; a workgroup covers 64 x 4 (i, j)
 * and SQ_KC planes (enough workgroups on the small network grid) */
#define SQ_KC 4
__global__ __launch_bounds__(256) void kSumSq2(const float *r, const float *b, Geo g, double *partials)
{
  __shared__ double lds[4];
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k0 = blockIdx.z * SQ_KC + 1, k1 = min(k0 + SQ_KC, g.nk + 1);
  double a0 = 0.0, a1 = 0.0;
  if (i <= g.ni && j <= g.nj)
    for (int k = k0; k < k1; k++) {
      long long c = g.at(i, j, k);
      a0 += (double)r[c] * r[c];
      a1 += (double)b[c] * b[c];
    }
  a0 = blockSum(a0, lds);
  a1 = blockSum(a1, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    unsigned bid          = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    partials[2 * bid]     = a0;
    partials[2 * bid + 1] = a1;
  }
}

static dim3 sumGrid(const Geo &g)
{
  return dim3((g.ni + 63) / 64, (g.nj + 3) / 4, (g.nk + SQ_KC - 1) / SQ_KC);
}

/* the folds below also write their sums into the pinned, host-coherent
 * mirror (outHost: its device address), which the host reads once the
 * stream has run them */
__global__ __launch_bounds__(256) void kFoldSumSq2(const double *partials, int nb, double *out, double *outHost)
{
  __shared__ double lds[4];
  double s0 = foldSum(partials, nb, 2, 0, lds), s1 = foldSum(partials, nb, 2, 1, lds);
  if (threadIdx.x == 0) {
    out[0]     = s0;
    out[1]     = s1;
    outHost[0] = s0;
    outHost[1] = s1;
    __threadfence_system();
  }
}

/* network input: channels 0, 1 as described above, 2..7 zero */
__global__ void kPack(const float *r, const float *b, Geo g, const double *scal, uint16_t *x)
{
  size_t nv = (size_t)g.ni * g.nj * g.nk;
  double cnt = (double)nv;
  float sr = (float)sqrt(scal[0] / cnt), sb = (float)sqrt(scal[1] / cnt);
  float ir = sr > 0.f ? 1.f / sr : 0.f, ib = sb > 0.f ? 1.f / sb : 0.f;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(v % g.ni) + 1, j = (int)((v / g.ni) % g.nj) + 1, k = (int)(v / ((size_t)g.ni * g.nj)) + 1;
    long long c = g.at(i, j, k);
    uint4 o;
    o.x = (uint32_t)f2bf(r[c] * ir) | ((uint32_t)f2bf(b[c] * ib) << 16);
    o.y = o.z = o.w = 0;
    *(uint4 *)(x + v * 8) = o;
  }
}

/* training target: (p - pPrev) restricted, / (s_r h2); level-0 difference */
__global__ __launch_bounds__(256) void kDiff(const double *p, const double *q, Geo g, float *d)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > g.ni || j > g.nj) return;
  long long c = g.at(i, j, k);
  d[c]        = (float)(p[c] - q[c]);
}

__global__ void kTarget(const float *d, Geo g, const double *scal, double h2, float *t)
{
  size_t nv = (size_t)g.ni * g.nj * g.nk;
  float sr  = (float)sqrt(scal[0] / (double)nv);
  float f   = sr > 0.f ? (float)(1.0 / (sr * h2)) : 0.f;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (size_t)gridDim.x * blockDim.x) {
    int i = (int)(v % g.ni) + 1, j = (int)((v / g.ni) % g.nj) + 1, k = (int)(v / ((size_t)g.ni * g.nj)) + 1;
    t[v] = d[g.at(i, j, k)] * f;
  }
}

/* 7-point Neumann Laplacian coefficients */
struct Lap {
  double idx2, idy2, idz2;
};

/* network on the finest grid (small problems): r0 and rhs as FP32 there */
__global__ __launch_bounds__(256) void kInputFine(const double *__restrict__ p, const double *__restrict__ rhs, Geo g,
    Lap cf, float *__restrict__ r, float *__restrict__ b32, double *partials)
{
  __shared__ double lds[4];
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  double rr = 0.0;
  if (i <= g.ni && j <= g.nj) {
    long long c = g.at(i, j, k);
    double xc   = p[c];
    double lap  = (p[c + 1] - 2.0 * xc + p[c - 1]) * cf.idx2 + (p[c + g.sj] - 2.0 * xc + p[c - g.sj]) * cf.idy2 +
                 (p[c + g.sk] - 2.0 * xc + p[c - g.sk]) * cf.idz2;
    double bb = rhs[c];
    double rv = bb - lap;
    r[c]      = (float)rv;
    b32[c]    = (float)bb;
    rr        = rv * rv;
  }
  rr = blockSum(rr, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0)
    partials[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = rr;
}

/* component 0 of stride-2 partials (the input pass's r0^2) */
__global__ __launch_bounds__(256) void kFoldComp0(const double *partials, int nb, double *out, double *outHost)
{
  __shared__ double lds[4];
  double s = foldSum(partials, nb, 2, 0, lds);
  if (threadIdx.x == 0) {
    *out     = s;
    *outHost = s;
    __threadfence_system();
  }
}

__global__ __launch_bounds__(256) void kFoldSum1(const double *partials, int nb, double *out, double *outHost)
{
  __shared__ double lds[4];
  double s = foldSum(partials, nb, 1, 0, lds);
  if (threadIdx.x == 0) {
    *out     = s;
    *outHost = s;
    __threadfence_system();
  }
}

/* ---- host ------------------------------------------------------------------ */
static size_t netVoxels(const WarmStart *w) { return (size_t)w->nx * w->ny * w->nz; }

static dim3 cells(const Geo &g) { return dim3((g.ni + WTX - 1) / WTX, (g.nj + WTY - 1) / WTY, g.nk); }

static float *falloc(long long n)
{
  float *p = nullptr;
  HIP_CHECK(hipMalloc((void **)&p, (size_t)n * sizeof(float)));
  HIP_CHECK(hipMemsetAsync(p, 0, (size_t)n * sizeof(float), nusifStream()));
  return p;
}

/* the first level whose extents are all <= maxExtent, halving only while
 * the halved extents stay divisible by div (the network's 2^(L-1)), so a
 * box such as 768 x 192 x 192 stops at 96 x 24 x 24 rather than at the
 * 48 x 12 x 12 that four levels of pooling cannot halve */
static int networkLevel(const Geo &g0, int maxExtent, int div)
{
  int ni = g0.ni, nj = g0.nj, nk = g0.nk, l = 0;
  while ((ni > maxExtent || nj > maxExtent || nk > maxExtent) && ni % (2 * div) == 0 && nj % (2 * div) == 0 &&
         nk % (2 * div) == 0) {
    ni /= 2;
    nj /= 2;
    nk /= 2;
    l++;
  }
  return l;
}

WarmStart *warmstartCreate(NusifCtx *c, UnetModel *net, bool capture, int netLevels)
{
  WarmStart *w = new WarmStart();
  w->net       = net;
  w->capture   = capture;
  w->useGraph  = c->opt.unetGraph != 0;
  w->fwdExec   = nullptr;
  const int div = 1 << (netLevels > 0 ? netLevels - 1 : 0);
  w->Ln        = networkLevel(c->geo, c->opt.unetExtent > 0 ? c->opt.unetExtent : 64, div);
  Geo g = c->geo;
  for (int l = 0; l <= w->Ln; l++) {
    w->lv[l].g = g;
    w->lv[l].r = falloc(g.n);
    w->lv[l].b = falloc(g.n);
    w->lv[l].d = falloc(g.n);
    if (l < w->Ln) g = makeGeo(g.ni / 2, g.nj / 2, g.nk / 2);
  }
  /* the acceptance test's level: unet_accept_level halvings below the
   * network grid, as far as the extents stay even */
  w->Lc = w->Ln;
  for (int q = 0; q < c->opt.unetAcceptLevel && w->Lc + 1 < 12; q++) {
    const Geo &gl = w->lv[w->Lc].g;
    if (gl.ni % 2 || gl.nj % 2 || gl.nk % 2 || gl.ni < 2 || gl.nj < 2 || gl.nk < 2) break;
    w->Lc++;
    w->lv[w->Lc].g = makeGeo(gl.ni / 2, gl.nj / 2, gl.nk / 2);
    w->lv[w->Lc].r = falloc(w->lv[w->Lc].g.n);
    w->lv[w->Lc].b = nullptr;
    w->lv[w->Lc].d = nullptr;
  }
  const Geo &gn = w->lv[w->Ln].g;
  if (gn.ni % div || gn.nj % div || gn.nk % div)
    NUSIF_FATAL("warm start: network grid %dx%dx%d (level %d) is not divisible by %d, which %d pooling levels need",
        gn.ni, gn.nj, gn.nk, w->Ln, div, netLevels);
  w->nx = gn.ni;
  w->ny = gn.nj;
  w->nz = gn.nk;
  {
    const Grid &gr = c->owner->grid;
    const double s = (double)(1 << w->Ln);
    w->cn[0]       = 1.0 / ((gr.dx * s) * (gr.dx * s));
    w->cn[1]       = 1.0 / ((gr.dy * s) * (gr.dy * s));
    w->cn[2]       = 1.0 / ((gr.dz * s) * (gr.dz * s));
  }
  {
    const Grid &gr = c->owner->grid; /* the domain's side lengths, on every level */
    const double l = std::max(gr.imax * gr.dx, std::max(gr.jmax * gr.dy, gr.kmax * gr.dz));
    w->h2          = (l / M_PI) * (l / M_PI);
  }
  {
    /* partials of the first input pass (two per workgroup when the network
     * is on a coarser level, one on the finest) and of the norms */
    dim3 gr = cells(w->lv[0].g);
    size_t need = w->Ln > 0 ? 2 * (size_t)inputRestrictPartials(w->lv[0].g) : (size_t)gr.x * gr.y * gr.z;
    if (need > 65536)
      NUSIF_FATAL("warm start: grid %dx%dx%d too large for the network input pass (coarsen to extents <= %d)",
          c->geo.ni, c->geo.nj, c->geo.nk, c->opt.unetExtent > 0 ? c->opt.unetExtent : 64);
  }
  HIP_CHECK(hipMalloc((void **)&w->x, netVoxels(w) * 8 * sizeof(uint16_t)));
  HIP_CHECK(hipMalloc((void **)&w->scal, 8 * sizeof(double)));
  HIP_CHECK(hipHostMalloc((void **)&w->scalHost, 8 * sizeof(double), hipHostMallocCoherent));
  HIP_CHECK(hipHostGetDevicePointer((void **)&w->scalHostDev, w->scalHost, 0));
  HIP_CHECK(hipMalloc((void **)&w->partials, 65536 * sizeof(double)));
  w->pPrev = nullptr;
  if (capture) HIP_CHECK(hipMalloc((void **)&w->pPrev, (size_t)c->geo.n * sizeof(double)));
  return w;
}

void warmstartFree(WarmStart *w)
{
  if (!w) return;
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  if (w->fwdExec) HIP_CHECK(hipGraphExecDestroy(w->fwdExec));
  for (int l = 0; l <= w->Lc; l++) {
    HIP_CHECK(hipFree(w->lv[l].r));
    if (w->lv[l].b) HIP_CHECK(hipFree(w->lv[l].b));
    if (w->lv[l].d) HIP_CHECK(hipFree(w->lv[l].d));
  }
  HIP_CHECK(hipFree(w->x));
  HIP_CHECK(hipFree(w->scal));
  HIP_CHECK(hipHostFree(w->scalHost));
  HIP_CHECK(hipFree(w->partials));
  if (w->pPrev) HIP_CHECK(hipFree(w->pPrev));
  for (auto p : w->dsX) HIP_CHECK(hipFree(p));
  for (auto p : w->dsT) HIP_CHECK(hipFree(p));
  delete w;
}

int warmstartLevel(const WarmStart *w) { return w->Ln; }
void warmstartExtents(const WarmStart *w, int *nx, int *ny, int *nz)
{
  *nx = w->nx;
  *ny = w->ny;
  *nz = w->nz;
}
void warmstartSetNet(WarmStart *w, UnetModel *net)
{
  if (w->fwdExec) { /* captured with the previous network's buffers */
    HIP_CHECK(hipStreamSynchronize(nusifStream()));
    HIP_CHECK(hipGraphExecDestroy(w->fwdExec));
    w->fwdExec = nullptr;
  }
  w->net = net;
}
UnetModel *warmstartNet(WarmStart *w) { return w->net; }
bool warmstartCapturing(const WarmStart *w) { return w->capture; }

/* src (FP32 on level l0 < Lc) restricted level by level to the acceptance
 * level Lc (lv[l].r as the levels' arrays), two levels per launch where it
 * can; its sum of squares there folded into scal[slot] and its host mirror */
static void restrictToAcceptance(WarmStart *w, const float *src, int l0, int slot)
{
  hipStream_t st = nusifStream();
  int l = l0;
  while (l < w->Lc) {
    const bool last2 = l + 2 == w->Lc, last1 = l + 1 == w->Lc;
    if (l + 2 <= w->Lc) {
      kRestrictChain2<<<cells(w->lv[l + 2].g), dim3(WTX, WTY), 0, st>>>(src, w->lv[l].g, w->lv[l + 1].r,
          w->lv[l + 1].g, w->lv[l + 2].r, w->lv[l + 2].g, last2 ? w->partials : nullptr);
      l += 2;
    } else {
      kRestrictPair<<<cells(w->lv[l + 1].g), dim3(WTX, WTY), 0, st>>>(src, nullptr, w->lv[l].g, w->lv[l + 1].r,
          nullptr, w->lv[l + 1].g, last1 ? w->partials : nullptr);
      l += 1;
    }
    src = w->lv[l].r;
  }
  const dim3 gc = cells(w->lv[w->Lc].g);
  kFoldComp0<<<1, 256, 0, st>>>(w->partials, (int)(gc.x * gc.y * gc.z), w->scal + slot, w->scalHostDev + slot);
}


/* Network input from p and rhs: r0 = rhs - A p and rhs restricted to the
 * network level (one fused pass to level 1, then 8-child averages), their
 * sums of squares there (scal[0], scal[1]) and of r0 on the finest grid
 * (scal[3]), packed into x. With r32 and the network below the finest grid,
 * the same fine-grid pass also writes r0 as FP32 into r32 (a mixed-precision
 * solve's first defect), zeroes eZero when given, and folds ||r0||^2 and
 * ||rhs||^2 into norm2[0..1] (and the host mirror norm2Host); returns
 * whether it did. */
bool warmstartPrepare(WarmStart *w, const double *p, const double *rhs, double idx2, double idy2, double idz2,
    float *r32, float *eZero, double *norm2, double *norm2Host)
{
  bool defect = false;
  hipStream_t st = nusifStream();
  Lap cf         = {idx2, idy2, idz2};
  const Geo &g0  = w->lv[0].g;
  if (w->Ln == 0) {
    dim3 gr = cells(g0);
    kInputFine<<<gr, dim3(WTX, WTY), 0, st>>>(p, rhs, g0, cf, w->lv[0].r, w->lv[0].b, w->partials);
    kFoldSum1<<<1, 256, 0, st>>>(w->partials, (int)(gr.x * gr.y * gr.z), w->scal + 3, w->scalHostDev + 3);
  } else {
    /* r0 and rhs averaged onto level 1 in the residual pass itself */
    int nb = launchInputRestrict(p, rhs, g0, idx2, idy2, idz2, w->lv[1].r, w->lv[1].b, w->lv[1].g, w->partials,
        r32, r32 ? eZero : nullptr);
    if (r32) { /* ||r0||^2 and ||rhs||^2 for the solve, ||r0||^2 also into scal[3]: one fold */
      launchFoldNorms(w->partials, nb, norm2, norm2Host, w->scal + 3, w->scalHostDev + 3);
      defect = true;
    } else {
      kFoldComp0<<<1, 256, 0, st>>>(w->partials, nb, w->scal + 3, w->scalHostDev + 3);
    }
  }
  /* down to the network level, r and rhs in one launch per level; the last
   * one also gives the sums of squares there (the partials above are folded
   * already: the stream runs in order) */
  const WsLevel &L = w->lv[w->Ln];
  if (w->Ln >= 2) {
    for (int l = 2; l <= w->Ln; l++)
      kRestrictPair<<<cells(w->lv[l].g), dim3(WTX, WTY), 0, st>>>(w->lv[l - 1].r, w->lv[l - 1].b, w->lv[l - 1].g,
          w->lv[l].r, w->lv[l].b, w->lv[l].g, l == w->Ln ? w->partials : nullptr);
    const dim3 gl = cells(L.g);
    kFoldSumSq2<<<1, 256, 0, st>>>(w->partials, (int)(gl.x * gl.y * gl.z), w->scal, w->scalHostDev);
  } else {
    const dim3 g2 = sumGrid(L.g);
    kSumSq2<<<g2, dim3(WTX, WTY), 0, st>>>(L.r, L.b, L.g, w->partials);
    kFoldSumSq2<<<1, 256, 0, st>>>(w->partials, (int)(g2.x * g2.y * g2.z), w->scal, w->scalHostDev);
  }
  /* the packed input tensor only for captured training pairs: the forward
   * forms it in its first layer */
  if (w->capture) kPack<<<1024, 256, 0, st>>>(L.r, L.b, L.g, w->scal, w->x);
  /* r0 on down to the acceptance level, its sum of squares there (scal[4]) */
  if (w->Lc > w->Ln) restrictToAcceptance(w, w->lv[w->Ln].r, w->Ln, 4);
  return defect;
}

/* The network's correction on the finest grid (FP32, interior and face
 * ghosts) from the input warmstartPrepare built: forward, unpack on the
 * network grid, prolong trilinearly to level 0. The guess is p + this
 * correction (the caller forms it together with its residual). */
const float *warmstartCorrection(WarmStart *w)
{
  warmstartForward(w);
  warmstartProlongCorrection(w, nullptr);
  return w->lv[0].d;
}

/* forward and unpack: the correction d on the network grid (lv[Ln].d).
 * With unet_graph 1 the forward's launches (convolutions with their fused
 * pooling and head, any split-K finishes) are captured once and replayed with one
 * hipGraphLaunch: they always run on the same buffers (input, weights,
 * activations, scratch, output), and the scale they read comes from device
 * memory at run time. Kernels only: no memset or copy is captured. Eager
 * while kernel timers are on (their events bracket single launches). */
extern bool nusifTimersActive();
void warmstartForward(WarmStart *w)
{
  const WsLevel &L = w->lv[w->Ln];
  hipStream_t st   = nusifStream();
  /* the last convolution's pass also runs the head and writes the scaled
   * correction (interior and face ghosts) */
  auto body = [&]() {
    if (w->capture)
      NUSIF_TIMED("unet_forward", unetForward(w->net, w->x, false, L.d, w->scal, w->h2));
    else
      NUSIF_TIMED("unet_forward", unetForward(w->net, nullptr, false, L.d, w->scal, w->h2, L.r, L.b));
  };
  if (!w->useGraph || nusifTimersActive()) {
    body();
    return;
  }
  if (!w->fwdExec) {
    hipGraph_t graph;
    HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    body();
    HIP_CHECK(hipStreamEndCapture(st, &graph));
    HIP_CHECK(hipGraphInstantiate(&w->fwdExec, graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphDestroy(graph));
  }
  HIP_CHECK(hipGraphLaunch(w->fwdExec, st));
}

/* d prolongated trilinearly from the network grid to the finest, written
 * into `fine` (FP32 padded, face ghosts) or, when null, into lv[0].d */
void warmstartProlongCorrection(WarmStart *w, float *fine)
{
  for (int l = w->Ln; l > 0; l--)
    launchProlongSet32(l == 1 && fine ? fine : w->lv[l - 1].d, w->lv[l - 1].g, w->lv[l].d, w->lv[l].g);
}

const float *warmstartProlongToLevel1(WarmStart *w, Geo *g1)
{
  if (w->Ln < 1) return nullptr;
  for (int l = w->Ln; l > 1; l--) launchProlongSet32(w->lv[l - 1].d, w->lv[l - 1].g, w->lv[l].d, w->lv[l].g);
  *g1 = w->lv[1].g;
  return w->lv[1].d;
}

/* r - A d on the network grid (7-point Neumann Laplacian with that grid's
 * spacings; d carries its face ghosts), FP32 */
__global__ __launch_bounds__(256) void kCoarseDefect(const float *__restrict__ r, const float *__restrict__ d, Geo g,
    double cx, double cy, double cz, float *__restrict__ out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > g.ni || j > g.nj) return;
  long long c = g.at(i, j, k);
  double dc   = d[c];
  double lap  = ((double)d[c + 1] - 2.0 * dc + (double)d[c - 1]) * cx +
               ((double)d[c + g.sj] - 2.0 * dc + (double)d[c - g.sj]) * cy +
               ((double)d[c + g.sk] - 2.0 * dc + (double)d[c - g.sk]) * cz;
  out[c] = (float)((double)r[c] - lap);
}

/* The acceptance test of a mixed-precision solve, on coarse grids only: the
 * guess's residual estimated on the network grid as R r0 - A_n d (A_n the
 * network grid's operator, which stands in for R A P there), restricted to
 * the acceptance level, its sum of squares into scal[2]; compared with
 * warmstartReferenceNorm2 (R r0 there). No fine-grid pass. */
void warmstartCoarseAcceptLaunch(WarmStart *w)
{
  hipStream_t st = nusifStream();
  const WsLevel &N = w->lv[w->Ln];
  kCoarseDefect<<<cells(N.g), dim3(WTX, WTY), 0, st>>>(N.r, N.d, N.g, w->cn[0], w->cn[1], w->cn[2], N.b);
  if (w->Lc > w->Ln) {
    restrictToAcceptance(w, N.b, w->Ln, 2);
  } else {
    const dim3 g2 = sumGrid(N.g);
    kSumSq2<<<g2, dim3(WTX, WTY), 0, st>>>(N.b, N.b, N.g, w->partials);
    kFoldComp0<<<1, 256, 0, st>>>(w->partials, (int)(g2.x * g2.y * g2.z), w->scal + 2, w->scalHostDev + 2);
  }
}

/* The acceptance decision on the device: the guess is taken when its
 * restricted sum of squares (scal[2]) is below the previous pressure's
 * (scal[ref]), or always with force; 1.0 or 0.0 into scal[5] and its host
 * mirror. A rejected guess's correction on the network grid is zeroed, so
 * the solve that follows starts from e = 0 on either branch without the
 * host waiting for the network. */
__global__ void kAcceptDecide(double *scal, double *scalHost, int ref, int force)
{
  if (threadIdx.x == 0) {
    const double u = force || scal[2] < scal[ref] ? 1.0 : 0.0;
    scal[5]        = u;
    scalHost[5]    = u;
  }
}

__global__ __launch_bounds__(256) void kZeroRejected(float *__restrict__ d, long long n, const double *scal)
{
  if (scal[5] != 0.0) return;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long long)gridDim.x * blockDim.x)
    d[q] = 0.f;
}

void warmstartDecideLaunch(WarmStart *w, int force)
{
  hipStream_t st   = nusifStream();
  const WsLevel &N = w->lv[w->Ln];
  kAcceptDecide<<<1, 64, 0, st>>>(w->scal, w->scalHostDev, w->Lc > w->Ln ? 4 : 0, force);
  const long long n = (long long)N.g.n;
  kZeroRejected<<<(unsigned)std::min<long long>((n + 255) / 256, 1024), 256, 0, st>>>(N.d, n, w->scal);
}

int warmstartDecision(const WarmStart *w) { return w->scalHost[5] != 0.0 ? 1 : 0; }

float *warmstartResidualBuffer(WarmStart *w) { return w->lv[0].r; }

/* Sum of squares of R r on the network level, R = repeated 8-child
 * averaging, of a level-0 FP32 residual (any padded buffer on the finest
 * grid), into the device slot scal[2]; no host synchronisation. lv[l].r
 * (l >= 1) are reused as scratch. */
float *warmstartLevelResidual(WarmStart *w, int l, Geo *g)
{
  if (l < 1 || l > w->Ln) return nullptr;
  *g = w->lv[l].g;
  return w->lv[l].r;
}

void warmstartCoarseNormLaunch(WarmStart *w, const float *rFine, bool level1)
{
  hipStream_t st = nusifStream();
  const float *src = level1 ? w->lv[1].r : rFine;
  const int l0     = level1 ? 2 : 1;
  const WsLevel &L = w->lv[w->Lc];
  /* one component into scal[2]: scal[3] keeps the fine ||r0||^2 of the input */
  if (l0 <= w->Lc) { /* the last restriction also gives the sum of squares */
    restrictToAcceptance(w, src, l0 - 1, 2);
  } else {
    const dim3 g2 = sumGrid(L.g);
    kSumSq2<<<g2, dim3(WTX, WTY), 0, st>>>(src, src, L.g, w->partials);
    kFoldComp0<<<1, 256, 0, st>>>(w->partials, (int)(g2.x * g2.y * g2.z), w->scal + 2, w->scalHostDev + 2);
  }
}

/* scal[0..4] ([0] = sum (R r0)^2, [2] = the last warmstartCoarseNormLaunch)
 * in pinned host memory, which the folds write themselves: valid once the
 * stream has run them (after the caller's next synchronisation) */
const double *warmstartNormsAsync(WarmStart *w) { return w->scalHost; }

/* the previous pressure's restricted sum of squares the acceptance test
 * compares with the guess's (scal[2]): on the acceptance level */
double warmstartReferenceNorm2(const WarmStart *w) { return w->scalHost[w->Lc > w->Ln ? 4 : 0]; }

/* capture, part 1 (before the solve, after warmstartPrepare): the input and
 * a copy of p */
void warmstartCaptureBegin(WarmStart *w, const double *p, const Geo &g0)
{
  if (!w->capture) return;
  HIP_CHECK(hipMemcpyAsync(w->pPrev, p, (size_t)g0.n * sizeof(double), hipMemcpyDeviceToDevice, nusifStream()));
  uint16_t *x = nullptr;
  size_t nv   = netVoxels(w);
  HIP_CHECK(hipMalloc((void **)&x, nv * 8 * sizeof(uint16_t)));
  HIP_CHECK(hipMemcpyAsync(x, w->x, nv * 8 * sizeof(uint16_t), hipMemcpyDeviceToDevice, nusifStream()));
  w->dsX.push_back(x);
}

/* capture, part 2 (after the solve): target from p - pPrev */
void warmstartCaptureEnd(WarmStart *w, const double *p)
{
  if (!w->capture) return;
  hipStream_t st = nusifStream();
  kDiff<<<cells(w->lv[0].g), dim3(WTX, WTY), 0, st>>>(p, w->pPrev, w->lv[0].g, w->lv[0].d);
  for (int l = 1; l <= w->Ln; l++)
    kRestrictAvg<float><<<cells(w->lv[l].g), dim3(WTX, WTY), 0, st>>>(w->lv[l - 1].d, w->lv[l - 1].g, w->lv[l].d, w->lv[l].g);
  size_t nv = netVoxels(w);
  float *t  = nullptr;
  HIP_CHECK(hipMalloc((void **)&t, nv * sizeof(float)));
  kTarget<<<1024, 256, 0, st>>>(w->lv[w->Ln].d, w->lv[w->Ln].g, w->scal, w->h2, t);
  w->dsT.push_back(t);
}

size_t warmstartSamples(const WarmStart *w) { return w->dsT.size(); }

/* hand the captured pairs over (ownership moves to the caller) */
void warmstartTakeSamples(WarmStart *w, std::vector<uint16_t *> &xs, std::vector<float *> &ts)
{
  xs.insert(xs.end(), w->dsX.begin(), w->dsX.end());
  ts.insert(ts.end(), w->dsT.begin(), w->dsT.end());
  w->dsX.clear();
  w->dsT.clear();
}

/* ---- C ABI: training drivers ----------------------------------------------- */
static const int kUnetLevels = 4;
static const int kUnetChannels[4] = {16, 32, 64, 128};

/* xorshift64* uniform in [0, 1), seeded per training call */
struct Uniform {
  unsigned long long s;
  explicit Uniform(int seed) : s(0x243F6A8885A308D3ull ^ ((unsigned long long)seed * 0x9E3779B97F4A7C15ull)) {}
  double operator()()
  {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return ((s * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
  }
};

/* Training pairs captured from solver runs, all on one network grid. */
struct PairSet {
  std::vector<uint16_t *> xs;
  std::vector<float *> ts;
  int nx = 0, ny = 0, nz = 0;
  ~PairSet()
  {
    for (auto p : xs) HIP_CHECK(hipFree(p));
    for (auto p : ts) HIP_CHECK(hipFree(p));
  }
};

/* One solver run on `prm` (gmg, mixed precision, 1e-6, quiet): skipSteps
 * steps of spin-up, then stepsPerRun steps whose solves are captured as
 * pairs on the network grid the warm start would use (netExtent). */
static void capturePairs(Parameter *prm, int netExtent, int skipSteps, int stepsPerRun, PairSet &ps)
{
  char buf[64];
  nusifSetOption(prm, "solver", "gmg");
  nusifSetOption(prm, "precision", "mixed");
  nusifSetOption(prm, "mg_tol", "1e-6");
  nusifSetOption(prm, "verbose", "0");
  snprintf(buf, sizeof buf, "%d", netExtent);
  nusifSetOption(prm, "unet_extent", buf);
  Discretization d;
  Solver s;
  memset(&d, 0, sizeof d);
  commInit(&d.comm, 0, nullptr);
  commPartition(&d.comm, prm->kmax, prm->jmax, prm->imax);
  initDiscretization(&d, prm);
  initSolver(&s, &d, prm);
  NusifSolverCtx *sc = nusifSolverCtxOf(&s);
  if (skipSteps > 0) nusifRunSteps(&d, &s, skipSteps, nullptr);
  sc->ws = warmstartCreate(sc->dctx, nullptr, true, kUnetLevels);
  WarmStart *w = sc->ws;
  if (ps.nx && (w->nx != ps.nx || w->ny != ps.ny || w->nz != ps.nz))
    NUSIF_FATAL("unet training: grid %dx%dx%d gives network grid %dx%dx%d, earlier runs %dx%dx%d", prm->imax,
        prm->jmax, prm->kmax, w->nx, w->ny, w->nz, ps.nx, ps.ny, ps.nz);
  ps.nx = w->nx;
  ps.ny = w->ny;
  ps.nz = w->nz;
  nusifRunSteps(&d, &s, stepsPerRun, nullptr);
  warmstartTakeSamples(w, ps.xs, ps.ts);
  nusifFreeSolver(&s);
  nusifFreeDiscretization(&d);
}

/* sum of t^2 over a flat array: per-workgroup partials */
__global__ __launch_bounds__(256) void kSumSqFlat(const float *t, size_t n, double *partials)
{
  __shared__ double lds[4];
  double a = 0.0;
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    a += (double)t[q] * t[q];
  a = blockSum(a, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = a;
}

/* the loss of a network that predicts zero: the mean over the pairs of the
 * targets' mean square, the scale the training loss is read against */
static double zeroPredictionLoss(const PairSet &ps)
{
  const size_t nv = (size_t)ps.nx * ps.ny * ps.nz;
  const int nb    = 256;
  double *part = nullptr, *out = nullptr, *outDev = nullptr;
  HIP_CHECK(hipMalloc((void **)&part, nb * sizeof(double)));
  HIP_CHECK(hipMalloc((void **)&outDev, sizeof(double)));
  HIP_CHECK(hipHostMalloc((void **)&out, sizeof(double), hipHostMallocCoherent));
  double *outHostDev = nullptr;
  HIP_CHECK(hipHostGetDevicePointer((void **)&outHostDev, out, 0));
  hipStream_t st = nusifStream();
  double sum     = 0.0;
  for (float *t : ps.ts) {
    kSumSqFlat<<<nb, 256, 0, st>>>(t, nv, part);
    kFoldSum1<<<1, 256, 0, st>>>(part, nb, outDev, outHostDev);
    HIP_CHECK(hipStreamSynchronize(st));
    sum += *out / (double)nv;
  }
  HIP_CHECK(hipFree(part));
  HIP_CHECK(hipFree(outDev));
  HIP_CHECK(hipHostFree(out));
  return ps.ts.empty() ? 0.0 : sum / ps.ts.size();
}

/* Adam over the pairs: `epochs` shuffled passes, cosine-decayed lr, an
 * update per `batch` samples; from weightsIn when given (fine-tuning), else
 * from a seeded He initialisation. Writes weightsOut. */
static void trainOnPairs(PairSet &ps, const char *weightsIn, const char *weightsOut, int epochs, int batch,
    double lr, int seed, double *lossPerEpoch)
{
  Uniform uni(seed ^ 0x5bd1e995);
  UnetModel *m = nullptr;
  if (weightsIn && weightsIn[0]) {
    m = unetLoad(weightsIn, ps.nx, ps.ny, ps.nz);
    if (!m) NUSIF_FATAL("unet training: cannot read %s", weightsIn);
  } else {
    m = unetNew(kUnetLevels, kUnetChannels, 8, ps.nx, ps.ny, ps.nz);
    unetInit(m, (unsigned)seed);
    unetSyncWeights(m);
  }
  unetZeroGrads(m);
  const double zeroLoss = zeroPredictionLoss(ps);
  const double t0       = getTimeStamp();
  printf("[unet] %zu pairs on %dx%dx%d; loss of a zero prediction %.5g\n", ps.ts.size(), ps.nx, ps.ny, ps.nz,
      zeroLoss);
  std::vector<int> order(ps.ts.size());
  for (size_t q = 0; q < order.size(); q++) order[q] = (int)q;
  for (int e = 0; e < epochs; e++) {
    for (size_t q = order.size() - 1; q > 0; q--) std::swap(order[q], order[(size_t)(uni() * (q + 1)) % (q + 1)]);
    double sum = 0.0;
    float lrE  = (float)(lr * 0.5 * (1.0 + cos(M_PI * e / (double)epochs))); /* cosine decay */
    for (size_t q = 0; q < order.size(); q++) {
      unetForward(m, ps.xs[order[q]], true);
      bool last = (q + 1) % batch == 0 || q + 1 == order.size();
      sum += unetBackward(m, ps.xs[order[q]], ps.ts[order[q]], true);
      if (last) unetAdam(m, lrE, batch);
      if ((q + 1) % 1024 == 0) {
        printf("[unet] epoch %d: %zu/%zu samples, running loss %.5g\n", e + 1, q + 1, order.size(), sum / (q + 1));
        fflush(stdout);
      }
    }
    if (lossPerEpoch) lossPerEpoch[e] = sum / order.size();
    printf("[unet] epoch %d/%d: mean loss %.5g, %.4f of a zero prediction's (lr %.3g), %.0f s\n", e + 1,
        epochs, sum / order.size(), zeroLoss > 0.0 ? sum / order.size() / zeroLoss : 0.0, lrE,
        getTimeStamp() - t0);
    fflush(stdout);
    /* the weights after every epoch, so an interrupted run keeps its last one */
    if (unetSave(m, weightsOut) != 0) NUSIF_FATAL("cannot write %s", weightsOut);
  }
  unetFree(m);
}

/* the driver loop's parameters shared by the training runs */
static void trainingParameter(Parameter *prm, int imax, int jmax, int kmax, double re)
{
  initParameter(prm);
  prm->imax = imax;
  prm->jmax = jmax;
  prm->kmax = kmax;
  prm->re   = re;
  prm->tau  = 0.5;
  prm->te   = 1e9;
  prm->dt   = 0.02;
  prm->gx = prm->gy = prm->gz = 0.0;
  prm->u_init = prm->v_init = prm->w_init = prm->p_init = 0.0;
  prm->levels    = 6;
  prm->presmooth = prm->postsmooth = 2;
}

extern "C" int nusifUnetTrainCavity(const char *weightsOut, int grid, int runs, int stepsPerRun,
    int skipSteps, int epochs, int batch, double lr, int seed, double *lossPerEpoch)
{
  return nusifUnetTrainCavityGrids(weightsOut, &grid, 1, 64, runs, stepsPerRun, skipSteps, epochs, batch, lr,
      seed, lossPerEpoch);
}

/* as nusifUnetTrainCavity, run r on a grids[r % ngrids]^3 cavity, with the
 * network on the first multigrid level of extent <= netExtent (all grids
 * must give the same network grid) */
extern "C" int nusifUnetTrainCavityGrids(const char *weightsOut, const int *grids, int ngrids, int netExtent,
    int runs, int stepsPerRun, int skipSteps, int epochs, int batch, double lr, int seed, double *lossPerEpoch)
{
  if (ngrids < 1) NUSIF_FATAL("unet training: no grid given");
  const double tStart = getTimeStamp();
  Uniform uni(seed);
  PairSet ps;
  for (int r = 0; r < runs; r++) {
    const int grid = grids[r % ngrids];
    double re = 10.0 + 990.0 * uni(), lid = 0.5 + 1.5 * uni();
    Parameter prm;
    trainingParameter(&prm, grid, grid, grid, re);
    prm.xlength = prm.ylength = prm.zlength = 1.0;
    prm.name = (char *)"dcavity";
    prm.bcLeft = prm.bcRight = prm.bcBottom = prm.bcTop = prm.bcFront = prm.bcBack = NOSLIP;
    char buf[64];
    snprintf(buf, sizeof buf, "%.17g", lid);
    nusifSetOption(&prm, "lid", buf);
    capturePairs(&prm, netExtent, skipSteps, stepsPerRun, ps);
    printf("[unet] run %d/%d: %d^3, Re %.1f lid %.2f, %zu pairs so far, %.0f s\n", r + 1, runs, grid, re, lid,
        ps.ts.size(), getTimeStamp() - tStart);
    fflush(stdout);
  }
  if (ps.ts.empty()) return 0;
  trainOnPairs(ps, nullptr, weightsOut, epochs, batch, lr, seed, lossPerEpoch);
  return (int)ps.ts.size();
}

/* Channel pairs (the canal problem with uniform inflow u = 1 perturbed by a
 * seeded 1% uniform noise, zero-gradient outflow, no-slip walls, domain
 * xlength x 1 x 1; Re ~ U[100, 1000], an inflow seed per run; the rhs mean
 * removed), then training from weightsIn (fine-tuning; NULL or "" trains
 * from scratch). */
extern "C" int nusifUnetTrainChannel(const char *weightsIn, const char *weightsOut, int imax, int jmax, int kmax,
    double xlength, int netExtent, int runs, int stepsPerRun, int skipSteps, int epochs, int batch, double lr,
    int seed, double *lossPerEpoch)
{
  Uniform uni(seed);
  PairSet ps;
  for (int r = 0; r < runs; r++) {
    double re = 100.0 + 900.0 * uni();
    Parameter prm;
    trainingParameter(&prm, imax, jmax, kmax, re);
    prm.xlength = xlength;
    prm.ylength = prm.zlength = 1.0;
    prm.u_init  = 1.0;
    prm.gamma   = 0.9;
    prm.name    = (char *)"canal";
    prm.bcLeft = prm.bcRight = OUTFLOW;
    prm.bcBottom = prm.bcTop = prm.bcFront = prm.bcBack = NOSLIP;
    char buf[64];
    nusifSetOption(&prm, "u_inflow", "1.0");
    nusifSetOption(&prm, "inflow_noise", "0.01");
    snprintf(buf, sizeof buf, "%d", seed * 1000 + r + 1);
    nusifSetOption(&prm, "seed", buf);
    nusifSetOption(&prm, "mg_project", "1");
    capturePairs(&prm, netExtent, skipSteps, stepsPerRun, ps);
    printf("[unet] channel run %d/%d: %dx%dx%d, Re %.1f, %zu pairs so far\n", r + 1, runs, imax, jmax, kmax, re,
        ps.ts.size());
    fflush(stdout);
  }
  if (ps.ts.empty()) return 0;
  trainOnPairs(ps, weightsIn, weightsOut, epochs, batch, lr, seed, lossPerEpoch);
  return (int)ps.ts.size();
}

/* ---- C ABI: one forward/backward for gradient checks ------------------------ */
/* a model with the oracle's flat parameters and x (nx ny nz x cin,
 * channels-last) uploaded as bf16 */
static UnetModel *checkModel(const int *cfg, int nx, int ny, int nz, const double *params, const double *x,
    uint16_t **dx)
{
  UnetModel *m = unetNew(cfg[0], cfg + 1, cfg[5], nx, ny, nz);
  size_t np    = unetParamCount(m);
  std::vector<float> hp(np);
  for (size_t q = 0; q < np; q++) hp[q] = (float)params[q];
  unetSetParams(m, hp.data());
  unetSyncWeights(m);
  size_t nv = (size_t)nx * ny * nz;
  std::vector<uint16_t> hx(nv * cfg[5]);
  for (size_t q = 0; q < hx.size(); q++) hx[q] = f2bf((float)x[q]);
  HIP_CHECK(hipMalloc((void **)dx, hx.size() * sizeof(uint16_t)));
  HIP_CHECK(hipMemcpy(*dx, hx.data(), hx.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  return m;
}

static void readPrediction(UnetModel *m, size_t nv, double *pred)
{
  std::vector<float> hy(nv);
  /* the forward was queued on the (non-blocking) library stream */
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  HIP_CHECK(hipMemcpy(hy.data(), unetOutput(m), nv * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < nv; q++) pred[q] = hy[q];
}

/* cfg = {L, ch0, ch1, ch2, ch3, cin}; x: nx ny nz x cin (channels-last),
 * target nx ny nz; params in the oracle's flat layout. Returns the loss;
 * grad and pred (may be NULL) receive the gradient and the prediction. */
extern "C" double nusifUnetLossGrad(const int *cfg, int nx, int ny, int nz, const double *params, const double *x,
    const double *target, double *grad, double *pred)
{
  uint16_t *dx = nullptr;
  UnetModel *m = checkModel(cfg, nx, ny, nz, params, x, &dx);
  unetZeroGrads(m);
  size_t nv = (size_t)nx * ny * nz, np = unetParamCount(m);
  std::vector<float> ht(nv);
  for (size_t q = 0; q < nv; q++) ht[q] = (float)target[q];
  float *dt = nullptr;
  HIP_CHECK(hipMalloc((void **)&dt, nv * sizeof(float)));
  HIP_CHECK(hipMemcpy(dt, ht.data(), nv * sizeof(float), hipMemcpyHostToDevice));
  unetForward(m, dx, true);
  double loss = unetBackward(m, dx, dt, true);
  if (pred) readPrediction(m, nv, pred);
  if (grad) {
    std::vector<float> hg(np);
    unetGetGrads(m, hg.data());
    for (size_t q = 0; q < np; q++) grad[q] = hg[q];
  }
  HIP_CHECK(hipFree(dx));
  HIP_CHECK(hipFree(dt));
  unetFree(m);
  return loss;
}

/* The inference forward (the one the warm start runs: decoder inputs read
 * upsampled in place) on the same arguments; pred receives nx ny nz values. */
extern "C" void nusifUnetPredict(const int *cfg, int nx, int ny, int nz, const double *params, const double *x,
    double *pred)
{
  uint16_t *dx = nullptr;
  UnetModel *m = checkModel(cfg, nx, ny, nz, params, x, &dx);
  unetForward(m, dx, false);
  readPrediction(m, (size_t)nx * ny * nz, pred);
  HIP_CHECK(hipFree(dx));
  unetFree(m);
}
