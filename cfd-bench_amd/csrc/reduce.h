/* Wavefront / workgroup reductions for the two-kernel reduction pattern used
 * by every reduction: each workgroup writes one partial, then a second
 * one-workgroup kernel folds the partials in a fixed order, so results do
 * not depend on scheduling. */
#ifndef NUSIF_REDUCE_H
#define NUSIF_REDUCE_H

#include <hip/hip_runtime.h>

/* 64-lane butterfly sum / max over one wavefront */
__device__ inline double waveSum(double x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

__device__ inline double waveMax(double x)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    double y = __shfl_xor(x, o, 64);
    x        = x > y ? x : y;
  }
  return x;
}

/* Sum over the workgroup; the result is valid in thread 0. `lds` holds one
 * double per wavefront. */
__device__ inline double blockSum(double x, double *lds)
{
  int tid   = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  int nthr  = blockDim.x * blockDim.y * blockDim.z;
  int lane  = tid & 63, wave = tid >> 6, nwave = (nthr + 63) >> 6;
  x         = waveSum(x);
  if (lane == 0) lds[wave] = x;
  __syncthreads();
  double r = 0.0;
  if (tid == 0)
    for (int q = 0; q < nwave; q++) r += lds[q];
  __syncthreads();
  return r;
}

__device__ inline double blockMax(double x, double *lds)
{
  int tid   = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  int nthr  = blockDim.x * blockDim.y * blockDim.z;
  int lane  = tid & 63, wave = tid >> 6, nwave = (nthr + 63) >> 6;
  x         = waveMax(x);
  if (lane == 0) lds[wave] = x;
  __syncthreads();
  double r = 0.0;
  if (tid == 0) {
    r = lds[0];
    for (int q = 1; q < nwave; q++) r = r > lds[q] ? r : lds[q];
  }
  __syncthreads();
  return r;
}

/* Second stage of a two-kernel reduction: one workgroup folds `n` partials
 * (stride `stride`, component `comp`) in a fixed order, so the result does
 * not depend on scheduling. Kernel boundaries order the partial writes
 * before these reads; no device-scope atomics or fences are needed. */
__device__ inline double foldSum(const double *part, int n, int stride, int comp, double *lds)
{
  /* each thread adds its elements in ascending index order; four loads are
   * issued before their adds, so the fold is not one latency per element */
  double s     = 0.0;
  const int st = blockDim.x;
  int q        = threadIdx.x;
  for (; q + 3 * st < n; q += 4 * st) {
    double a0 = part[(long long)q * stride + comp], a1 = part[(long long)(q + st) * stride + comp];
    double a2 = part[(long long)(q + 2 * st) * stride + comp], a3 = part[(long long)(q + 3 * st) * stride + comp];
    s += a0;
    s += a1;
    s += a2;
    s += a3;
  }
  for (; q < n; q += st) s += part[(long long)q * stride + comp];
  return blockSum(s, lds);
}

/* Two interleaved components (stride 2) at once. Each thread adds its
 * elements in ascending index order, exactly as foldSum does, but issues the
 * loads of four elements before adding them, so the fold is not one memory
 * latency per element. */
__device__ inline void foldSum2(const double *part, int n, double *lds, double &s0, double &s1)
{
  double a = 0.0, b = 0.0;
  const int st = blockDim.x;
  int q        = threadIdx.x;
  for (; q + 3 * st < n; q += 4 * st) {
    double a0 = part[2LL * q], b0 = part[2LL * q + 1];
    double a1 = part[2LL * (q + st)], b1 = part[2LL * (q + st) + 1];
    double a2 = part[2LL * (q + 2 * st)], b2 = part[2LL * (q + 2 * st) + 1];
    double a3 = part[2LL * (q + 3 * st)], b3 = part[2LL * (q + 3 * st) + 1];
    a += a0;
    a += a1;
    a += a2;
    a += a3;
    b += b0;
    b += b1;
    b += b2;
    b += b3;
  }
  for (; q < n; q += st) {
    a += part[2LL * q];
    b += part[2LL * q + 1];
  }
  s0 = blockSum(a, lds);
  s1 = blockSum(b, lds);
}

__device__ inline double foldMax(const double *part, int n, int stride, int comp, double init,
    double *lds)
{
  double m = init;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    double x = part[(long long)q * stride + comp];
    m        = m > x ? m : x;
  }
  return blockMax(m, lds);
}

#endif
