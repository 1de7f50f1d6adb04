/* Kernels of the warm-start U-Net: BF16 implicit-GEMM 3x3x3 convolutions on
 * gfx950 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation), their data and
 * weight gradients, and the elementwise layers. Tensors are channels-last
 * ([z][y][x][c], c a multiple of 8), boxes of nx x ny x nz voxels, 'same'
 * zero padding.
 *
 * MFMA 16x16x32 lane layout (checked on the device): lane l holds
 *   A[row l%16][k 8*(l/16) .. +7],  B[k 8*(l/16) .. +7][col l%16],
 *   D[rows 4*(l/16) .. +3][col l%16].
 * Convolutions map rows to output channels and columns to voxels, so each
 * lane's four results are four consecutive channels of one voxel (one 8 B
 * store), and both operands are 16 B contiguous loads: weights [cout][27][cin]
 * along cin, activations along c. */
#ifndef NUSIF_UNET_KERNELS_H
#define NUSIF_UNET_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t f2bf(float f)
{
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u); /* nearest even */
  return (uint16_t)(u >> 16);
}

__device__ __host__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

/* 8 bf16 from 16 contiguous bytes (16 B aligned) */
__device__ inline bf16x8 ld8(const uint16_t *p) { return __builtin_bit_cast(bf16x8, *(const uint4 *)p); }

/* 8 floats (two 16 B loads) rounded to bf16 */
__device__ inline bf16x8 ld8(const float *p)
{
  float4 a = *(const float4 *)p, b = *(const float4 *)(p + 4);
  uint4 r;
  r.x = (uint32_t)f2bf(a.x) | ((uint32_t)f2bf(a.y) << 16);
  r.y = (uint32_t)f2bf(a.z) | ((uint32_t)f2bf(a.w) << 16);
  r.z = (uint32_t)f2bf(b.x) | ((uint32_t)f2bf(b.y) << 16);
  r.w = (uint32_t)f2bf(b.z) | ((uint32_t)f2bf(b.w) << 16);
  return __builtin_bit_cast(bf16x8, r);
}

__device__ inline bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

/* One convolution as C[rows][vox] = sum_k A[rows][k] B[k][vox]:
 *   forward:  rows = cout, A = W[cout][27][cin], B = in[vox + off(t)][cin]
 *   dgrad:    rows = cin,  A = Wt[cin][27][cout], B = dOut[vox - off(t)][cout]
 * The B operand may be the concatenation of two tensors (channels c0 + c1);
 * the result rows may be split over two outputs (r0 + r1 rows). */
template <typename InT, typename OutT> struct ConvArgs {
  const InT *in0, *in1;   /* B operand tensors, c0 and c1 channels (c1 = 0: none) */
  int c0, c1;
  const uint16_t *A;      /* [rows][27][c0 + c1] bf16 */
  const float *bias;      /* [rows] or null */
  OutT *out0, *out1;      /* [vox][r0], [vox][r1] */
  int r0, r1;
  int nx, ny, nz;         /* box extents (voxels) */
  int tapSign;            /* +1 forward, -1 data gradient */
  int relu;               /* ReLU on the result */
  int accumulate;         /* out += result (FP32 outputs) */
  int half0 = 0;          /* in0 is stored at extents (nx, ny, nz) / 2 and read
                             upsampled (nearest parent voxel); LDS forward conv only */
  /* Epilogue extras of the LDS forward conv (one K chunk set per workgroup,
   * no split-K): */
  uint16_t *pool = nullptr;     /* 2x2x2 average of the bf16 result at extents / 2,
                                   r0 channels (the next encoder level's input) */
  const float *headW = nullptr; /* 1x1 head on the bf16 result (all r0 <= 16 NB rows in
                                   one workgroup): y[v] = headW[r0] + sum_r headW[r] act[v][r] */
  float *headY = nullptr;       /* head output [vox] */
  float *corr = nullptr;        /* optional: y scaled into a padded (nx+2)(ny+2)(nz+2) FP32
                                   array with its face ghosts (the warm start's correction) */
  const double *corrScal = nullptr; /* corrScal[0] = sum of squares of the input's r channel */
  double corrH2 = 0.0;              /* correction = y * (float)(sqrt(corrScal[0] / voxels) * corrH2) */
  int storeAct = 1;             /* write out0 (off: only the head reads the result) */
  /* First layer only (8 input channels): instead of in0, stage channel 0
   * = packR * (1 / sqrt(packScal[0] / voxels)) and channel 1 = packB *
   * (1 / sqrt(packScal[1] / voxels)) (a zero norm gives 0), rounded to bf16,
   * channels 2..7 zero, read from padded (nx+2)(ny+2)(nz+2) FP32 arrays:
   * the network input without a packing pass */
  const float *packR = nullptr, *packB = nullptr;
  const double *packScal = nullptr;
  __host__ __device__ long long voxels() const { return (long long)nx * ny * nz; }
};

template <typename T> __device__ inline void storeOut(T *p, float v, int accumulate);
template <> __device__ inline void storeOut<float>(float *p, float v, int accumulate)
{
  *p = accumulate ? *p + v : v;
}
template <> __device__ inline void storeOut<uint16_t>(uint16_t *p, float v, int) { *p = f2bf(v); }

/* Workgroup: 4 waves; each wave owns MB blocks of 16 voxels and NB blocks of
 * 16 rows (grid.y walks the row tiles). */
template <typename InT, typename OutT, int NB, int MB>
__global__ __launch_bounds__(256) void kConvMfma(ConvArgs<InT, OutT> a)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nx = a.nx, ny = a.ny, nz = a.nz, nvox = nx * ny * nz;
  const int cin = a.c0 + a.c1, g8 = cin >> 3, G = 27 * g8;
  const int rows = a.r0 + a.r1;
  const int row0 = blockIdx.y * (16 * NB);
  const int vox0 = (blockIdx.x * 4 + wave) * (16 * MB);
  const int lr = lane & 15, lg = lane >> 4;

  int vz[MB], vy[MB], vx[MB];
  bool vin[MB];
#pragma unroll
  for (int mb = 0; mb < MB; mb++) {
    int v   = vox0 + mb * 16 + lr;
    vin[mb] = v < nvox;
    int vc  = vin[mb] ? v : 0;
    vx[mb]  = vc % nx;
    vy[mb]  = (vc / nx) % ny;
    vz[mb]  = vc / (nx * ny);
  }
  floatx4 acc[NB][MB];
#pragma unroll
  for (int nb = 0; nb < NB; nb++)
#pragma unroll
    for (int mb = 0; mb < MB; mb++) acc[nb][mb] = floatx4{0.f, 0.f, 0.f, 0.f};

  for (int s = 0; s < G; s += 4) {
    int gi  = s + lg;
    bool gv = gi < G;
    int t = gv ? gi / g8 : 0, c8 = gv ? gi - t * g8 : 0;
    int dz = (t / 9 - 1) * a.tapSign, dy = ((t / 3) % 3 - 1) * a.tapSign, dx = (t % 3 - 1) * a.tapSign;
    bf16x8 fa[NB], fb[MB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      int r  = row0 + nb * 16 + lr;
      fa[nb] = (gv && r < rows) ? ld8(a.A + ((size_t)r * 27 + t) * cin + c8 * 8) : zero8();
    }
#pragma unroll
    for (int mb = 0; mb < MB; mb++) {
      int z = vz[mb] + dz, y = vy[mb] + dy, x = vx[mb] + dx;
      bool ok = gv && vin[mb] && z >= 0 && z < nz && y >= 0 && y < ny && x >= 0 && x < nx;
      if (ok) {
        size_t vv = ((size_t)z * ny + y) * nx + x;
        int cc    = c8 * 8;
        fb[mb]    = cc < a.c0 ? ld8(a.in0 + vv * a.c0 + cc) : ld8(a.in1 + vv * a.c1 + (cc - a.c0));
      } else {
        fb[mb] = zero8();
      }
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
      for (int mb = 0; mb < MB; mb++)
        acc[nb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[nb], fb[mb], acc[nb][mb], 0, 0, 0);
  }
#pragma unroll
  for (int mb = 0; mb < MB; mb++) {
    int v = vox0 + mb * 16 + lr;
    if (!vin[mb]) continue;
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        int r = row0 + nb * 16 + 4 * lg + q;
        if (r >= rows) continue;
        float val = acc[nb][mb][q] + (a.bias ? a.bias[r] : 0.f);
        if (a.relu) val = val > 0.f ? val : 0.f;
        if (r < a.r0) storeOut<OutT>(a.out0 + (size_t)v * a.r0 + r, val, a.accumulate);
        else storeOut<OutT>(a.out1 + (size_t)v * a.r1 + (r - a.r0), val, a.accumulate);
      }
  }
}

/* Forward convolution with the input brick staged in LDS. A workgroup owns a
 * brick of 16 (x) x 4 (y) x 4 (z) output voxels; wave w owns z-slice w and
 * its four 16-voxel blocks are the brick's four y-rows. The input brick plus
 * its one-voxel halo (18 x 6 x 6 voxels) is loaded once per chunk of up to
 * 32 channels, and all 27 taps read their B fragments from LDS (one 16 B LDS
 * read per lane and block), instead of 27 global reads per voxel. K is
 * ordered (tap, 8-channel group) within a chunk; the A fragments (weights,
 * [cout][27][cin]) come from global memory through the caches. Voxel rows in
 * LDS are padded to 40 bf16 (80 B) to spread the 16 B reads over the banks. */
#define CL_BX 16
#define CL_BY 4
#define CL_BZ 4
#define CL_HX (CL_BX + 2)
#define CL_HY (CL_BY + 2)
#define CL_HZ (CL_BZ + 2)
#define CL_HV (CL_HX * CL_HY * CL_HZ)
#define CL_CC 32
#define CL_CSTR 40

/* One channel chunk of G8 8-channel groups (G8 = 4, 2 or 1): stage the
 * brick's halo into LDS, then K-steps of four (tap, group) pairs. G8 is a
 * compile-time constant so the (tap, group) split of the K index and the
 * tap's 3-D offset are shifts and constant multiplies, not divisions.
 * With KS K-groups of four waves (256 KS threads) all threads stage the
 * brick and group g runs the g-th of KS consecutive ranges of K-steps (a
 * chain of K-steps KS times shorter per wave; the caller adds the groups'
 * sums). */
template <int NB> struct ConvRing {
  static constexpr int D = NB == 1 ? 8 : (NB == 2 ? 4 : 2); /* K-steps the weights run ahead */
};

template <int NB, int G8, int CSTR, int KS, int KG>
__device__ inline void convSteps(const ConvArgs<uint16_t, uint16_t> &a, const uint16_t *sIn, const size_t (&aRow)[NB],
    const bool (&rowOk)[NB], const int (&hvBase)[4], int cin, int lg, floatx4 (&acc)[NB][4], int phase,
    bf16x8 (&ring)[ConvRing<NB>::D][NB])
{
  constexpr int G    = 27 * G8;
  constexpr int NIT  = (G + 3) / 4; /* K-steps of four (tap, group) pairs */
  constexpr int NITG = (NIT + KS - 1) / KS;
  constexpr int IB   = KG * NITG;
  constexpr int IE   = IB + NITG < NIT ? IB + NITG : NIT;
  constexpr int D    = ConvRing<NB>::D;
  auto loadA = [&](int s, bf16x8 *fa) {
    int gi  = s + lg;
    bool gv = gi < G;
    int t = gv ? gi / G8 : 0, c8 = gv ? gi % G8 : 0;
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
      fa[nb] = (gv && rowOk[nb]) ? ld8(a.A + aRow[nb] + (size_t)t * cin + c8 * 8) : zero8();
  };
  if (phase == 0) { /* the first D K-steps' weight fragments, before the staging */
#pragma unroll
    for (int p = 0; p < D; p++)
      if (IB + p < IE) loadA(4 * (IB + p), ring[p]);
    return;
  }
#pragma unroll
  for (int it = IB; it < IE; it++) {
    const int s = 4 * it;
    int gi      = s + lg;
    bool gv     = gi < G;
    int t = gv ? gi / G8 : 0, c8 = gv ? gi % G8 : 0;
    int off = ((t / 9) * CL_HY + (t / 3) % 3) * CL_HX + t % 3;
    bf16x8 fa[NB], fb[4];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) fa[nb] = ring[(it - IB) % D][nb];
    if (it + D < IE) loadA(4 * (it + D), ring[(it - IB) % D]);
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      uint4 raw = *(const uint4 *)&sIn[(hvBase[mb] + off) * CSTR + c8 * 8];
      fb[mb]    = gv ? __builtin_bit_cast(bf16x8, raw) : zero8();
    }
#pragma unroll
    for (int nb = 0; nb < NB; nb++)
#pragma unroll
      for (int mb = 0; mb < 4; mb++)
        acc[nb][mb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[nb], fb[mb], acc[nb][mb], 0, 0, 0);
  }
}

template <int NB, int G8, int CSTR = CL_CSTR, int KS = 1>
__device__ inline void convChunk(const ConvArgs<uint16_t, uint16_t> &a, uint16_t *sIn, int cc0, int cin, int rows,
    int row0, int x0, int y0, int z0, int wave, int lr, int lg, floatx4 (&acc)[NB][4], int tid, int kg)
{
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  /* per-lane bases: the A row offsets and the LDS voxel of each block at tap
   * offset (-1, -1, -1); a tap adds (dz * HY + dy) * HX + dx */
  size_t aRow[NB];
  bool rowOk[NB];
#pragma unroll
  for (int nb = 0; nb < NB; nb++) {
    int r     = row0 + nb * 16 + lr;
    rowOk[nb] = r < rows;
    aRow[nb]  = (size_t)(rowOk[nb] ? r : 0) * 27 * cin + cc0;
  }
  int hvBase[4];
#pragma unroll
  for (int mb = 0; mb < 4; mb++) hvBase[mb] = (wave * CL_HY + mb) * CL_HX + lr;
  /* weight fragments run D K-steps ahead of the MFMAs in a register ring, so
   * their global (L2) latency is paid once per chunk, not once per K-step;
   * the first D are requested before the staging */
  bf16x8 ring[ConvRing<NB>::D][NB];
  auto steps = [&](int phase) {
    if constexpr (KS == 1) {
      convSteps<NB, G8, CSTR, 1, 0>(a, sIn, aRow, rowOk, hvBase, cin, lg, acc, phase, ring);
    } else {
      static_assert(KS == 2, "two K-groups at most");
      if (kg == 0) convSteps<NB, G8, CSTR, 2, 0>(a, sIn, aRow, rowOk, hvBase, cin, lg, acc, phase, ring);
      else convSteps<NB, G8, CSTR, 2, 1>(a, sIn, aRow, rowOk, hvBase, cin, lg, acc, phase, ring);
    }
  };
  steps(0);
  /* stage the brick and its halo (this chunk's channels) into LDS: every
   * thread issues all its global loads before its first LDS store, so the
   * staging costs one memory latency, not one per element */
  constexpr int NT  = 256 * KS;
  constexpr int NST = (CL_HV * G8 + NT - 1) / NT;
  constexpr int SB  = 6; /* loads in flight per thread (register budget) */
  if constexpr (G8 == 1) {
    if (a.packR) { /* the network input straight from the FP32 level arrays */
      const double cnt = (double)a.voxels();
      const float sr = (float)sqrt(a.packScal[0] / cnt), sb = (float)sqrt(a.packScal[1] / cnt);
      const float ir = sr > 0.f ? 1.f / sr : 0.f, ib = sb > 0.f ? 1.f / sb : 0.f;
      const long long psj = nx + 2, psk = (long long)(nx + 2) * (ny + 2);
      float rv[NST], bv[NST];
#pragma unroll
      for (int r = 0; r < NST; r++) {
        const int v = tid + NT * r;
        rv[r] = bv[r] = 0.f;
        if (v < CL_HV) {
          int hx = v % CL_HX, hy = (v / CL_HX) % CL_HY, hz = v / (CL_HX * CL_HY);
          int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
          if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
            const long long c = (z + 1) * psk + (y + 1) * psj + x + 1;
            rv[r] = a.packR[c];
            bv[r] = a.packB[c];
          }
        }
      }
#pragma unroll
      for (int r = 0; r < NST; r++) {
        const int v = tid + NT * r;
        if (v < CL_HV) {
          uint4 o;
          o.x = (uint32_t)f2bf(rv[r] * ir) | ((uint32_t)f2bf(bv[r] * ib) << 16);
          o.y = o.z = o.w = 0;
          *(uint4 *)&sIn[v * CSTR] = o;
        }
      }
      __syncthreads();
      steps(1);
      __syncthreads();
      return;
    }
  }
#pragma unroll
  for (int r0 = 0; r0 < NST; r0 += SB) {
    uint4 val[SB];
#pragma unroll
    for (int r = 0; r < SB; r++) {
      const int u = tid + NT * (r0 + r);
      val[r]      = make_uint4(0, 0, 0, 0);
      if (r0 + r < NST && u < CL_HV * G8) {
        int v = u / G8, q = u - v * G8;
        int hx = v % CL_HX, hy = (v / CL_HX) % CL_HY, hz = v / (CL_HX * CL_HY);
        int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
        if (x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz) {
          size_t vv = ((size_t)z * ny + y) * nx + x;
          int c     = cc0 + q * 8;
          if (c < a.c0) {
            size_t v0 = a.half0 ? ((size_t)(z >> 1) * (ny >> 1) + (y >> 1)) * (nx >> 1) + (x >> 1) : vv;
            val[r]      = *(const uint4 *)(a.in0 + v0 * a.c0 + c);
          } else {
            val[r] = *(const uint4 *)(a.in1 + vv * a.c1 + (c - a.c0));
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < SB; r++) {
      const int u = tid + NT * (r0 + r);
      if (r0 + r < NST && u < CL_HV * G8) {
        int v = u / G8, q = u - v * G8;
        *(uint4 *)&sIn[v * CSTR + q * 8] = val[r];
      }
    }
  }
  __syncthreads();
  steps(1);
  __syncthreads();
}

/* GB: 8-channel groups of the last, partial chunk (cin % 32 / 8; 0 when
 * every chunk is full). Full chunks take four groups. Only the chunk shapes a
 * layer has are instantiated, which keeps the register count (and so the
 * number of workgroups per CU) at what one chunk loop needs. */
template <int NB, int GB, int CSTR = CL_CSTR, int KS = 1>
__device__ inline void convTail(const ConvArgs<uint16_t, uint16_t> &a, uint16_t *sIn, int cc0, int cin, int rows,
    int row0, int x0, int y0, int z0, int wave, int lr, int lg, floatx4 (&acc)[NB][4], int tid, int kg)
{
  if constexpr (GB == 3) { /* 24 channels: as groups of 2 and 1 */
    convChunk<NB, 2, CSTR, KS>(a, sIn, cc0, cin, rows, row0, x0, y0, z0, wave, lr, lg, acc, tid, kg);
    convChunk<NB, 1, CSTR, KS>(a, sIn, cc0 + 16, cin, rows, row0, x0, y0, z0, wave, lr, lg, acc, tid, kg);
  } else if constexpr (GB > 0) {
    convChunk<NB, GB, CSTR, KS>(a, sIn, cc0, cin, rows, row0, x0, y0, z0, wave, lr, lg, acc, tid, kg);
  }
}

/* KS: K-groups of four waves per workgroup (256 KS threads). All threads
 * stage each chunk; group g runs the g-th range of its K-steps; the groups'
 * sums are added in group order through LDS and group 0 writes the result.
 * For the coarse levels, where a layer has too few bricks to fill the
 * device: twice the waves per brick, each with half the chain of K-steps.
 *
 * Epilogue (one workgroup has all K, gridDim.z == 1): the brick's bf16
 * results also go to LDS, from where the workgroup forms the 2x2x2 average
 * pool (a.pool) or the 1x1 head (a.headW, with the optional scaled
 * correction a.corr) in the same order as the stand-alone kernels
 * (kPoolFwd, kHeadFwd, kUnpack), so the values are the same bits. */
template <typename OutT, int NB, int GB, bool FULL = true, int KS = 1>
__global__ __launch_bounds__(256 * KS) void kConvLds(ConvArgs<uint16_t, OutT> a, float *__restrict__ split)
{
  /* voxel stride in LDS (bf16): 32 channels + 8 of padding; a layer with
   * only a partial chunk of 8 or 16 channels stages 8 or 16 + 8 (16 B or 48 B:
   * 16 lanes' 16 B reads still fall in distinct banks), so more workgroups
   * fit a CU */
  constexpr int CSTR = FULL || GB == 3 ? CL_CSTR : (GB == 2 ? 24 : 8);
  constexpr int SIN  = CL_HV * CSTR;          /* the staging area (bf16) */
  constexpr int SA   = 16 * NB + 4;           /* bf16 per voxel of the epilogue's result brick */
  constexpr int SACT = 256 * SA;
  constexpr int SRED = (KS - 1) * 256 * 16 * NB * 2; /* the groups' FP32 sums, in bf16 units */
  constexpr int SL0  = SIN > SACT ? SIN : SACT;
  constexpr int SLDS = SL0 > SRED ? SL0 : SRED;
  __shared__ __attribute__((aligned(16))) uint16_t sIn[SLDS];
  const int tid = threadIdx.x & 255, kg = threadIdx.x >> 8, lane = tid & 63, wave = tid >> 6;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int nbx = (nx + CL_BX - 1) / CL_BX, nby = (ny + CL_BY - 1) / CL_BY;
  const int x0 = (blockIdx.x % nbx) * CL_BX;
  const int y0 = ((blockIdx.x / nbx) % nby) * CL_BY;
  const int z0 = (blockIdx.x / (nbx * nby)) * CL_BZ;
  const int cin = a.c0 + a.c1, rows = a.r0 + a.r1;
  const int row0 = blockIdx.y * (16 * NB);
  const int lr = lane & 15, lg = lane >> 4;

  floatx4 acc[NB][4];
#pragma unroll
  for (int nb = 0; nb < NB; nb++)
#pragma unroll
    for (int mb = 0; mb < 4; mb++) acc[nb][mb] = floatx4{0.f, 0.f, 0.f, 0.f};

  /* split-K: with gridDim.z > 1, workgroup z takes channel chunk z only and
   * writes its FP32 partial sums to split[z][vox][rows] (kConvSplitFinish
   * adds the chunks in order) */
  /* FULL: the layer has full 32-channel chunks (cin >= 32); without them
   * only the partial chunk's loop is compiled, which keeps the register count
   * of the narrow layers at what that loop needs */
  const int nFull = FULL ? cin / CL_CC : 0;
  const int tidAll = threadIdx.x; /* staging: all 256 KS threads */
  if (gridDim.z > 1) {
    const int cc0 = blockIdx.z * CL_CC;
    if (FULL && (int)blockIdx.z < nFull)
      convChunk<NB, 4, CL_CSTR, KS>(a, sIn, cc0, cin, rows, row0, x0, y0, z0, wave, lr, lg, acc, tidAll, kg);
    else
      convTail<NB, GB, CSTR, KS>(a, sIn, cc0, cin, rows, row0, x0, y0, z0, wave, lr, lg, acc, tidAll, kg);
  } else {
    if constexpr (FULL)
      for (int q = 0; q < nFull; q++)
        convChunk<NB, 4, CL_CSTR, KS>(a, sIn, q * CL_CC, cin, rows, row0, x0, y0, z0, wave, lr, lg, acc, tidAll, kg);
    convTail<NB, GB, CSTR, KS>(a, sIn, nFull * CL_CC, cin, rows, row0, x0, y0, z0, wave, lr, lg, acc, tidAll, kg);
  }
  if constexpr (KS > 1) {
    /* the chunk loops end at a barrier: the staging areas are free */
    float *sRed = (float *)sIn;
    if (kg > 0) {
      float *dst = sRed + ((size_t)(kg - 1) * 256 + tid) * (16 * NB);
#pragma unroll
      for (int nb = 0; nb < NB; nb++)
#pragma unroll
        for (int mb = 0; mb < 4; mb++) *(floatx4 *)(dst + (nb * 4 + mb) * 4) = acc[nb][mb];
    }
    __syncthreads();
    if (kg == 0)
      for (int g = 1; g < KS; g++) {
        const float *src = sRed + ((size_t)(g - 1) * 256 + tid) * (16 * NB);
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
          for (int mb = 0; mb < 4; mb++) acc[nb][mb] += *(const floatx4 *)(src + (nb * 4 + mb) * 4);
      }
    __syncthreads();
  }
  const int x = x0 + lr, z = z0 + wave;
  const bool own = kg == 0 && x < nx && z < nz; /* this lane's voxels exist and it holds their sums */
  /* each lane holds four consecutive rows (channels) of one voxel: rows are
   * multiples of 16, so they are one 16 B (FP32) or 8 B (bf16) store */
  if (gridDim.z > 1) {
    if (!own) return;
    float *sp = split + (size_t)blockIdx.z * a.voxels() * rows;
#pragma unroll
    for (int mb = 0; mb < 4; mb++) {
      int y = y0 + mb;
      if (y >= ny) continue;
      size_t v = ((size_t)z * ny + y) * nx + x;
#pragma unroll
      for (int nb = 0; nb < NB; nb++) {
        int r = row0 + nb * 16 + 4 * lg;
        if (r < rows) *(float4 *)(sp + v * rows + r) = make_float4(acc[nb][mb][0], acc[nb][mb][1], acc[nb][mb][2], acc[nb][mb][3]);
      }
    }
    return;
  }
  const bool epi  = a.pool != nullptr || a.headW != nullptr;
  uint16_t *sAct  = sIn; /* [brick voxel][SA] bf16 results for the epilogue */
#pragma unroll
  for (int mb = 0; mb < 4; mb++) {
    int y = y0 + mb;
    size_t v = ((size_t)z * ny + y) * nx + x;
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
      int r = row0 + nb * 16 + 4 * lg;
      if (kg != 0 || r >= rows) continue;
      float val[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        val[q] = acc[nb][mb][q] + (a.bias ? a.bias[r + q] : 0.f);
        if (a.relu) val[q] = val[q] > 0.f ? val[q] : 0.f;
      }
      if (sizeof(OutT) == 2 && !a.accumulate && r + 3 < a.r0) {
        uint2 pk;
        pk.x = (uint32_t)f2bf(val[0]) | ((uint32_t)f2bf(val[1]) << 16);
        pk.y = (uint32_t)f2bf(val[2]) | ((uint32_t)f2bf(val[3]) << 16);
        if (own && y < ny && a.storeAct) *(uint2 *)(a.out0 + v * a.r0 + r) = pk;
        if (epi) *(uint2 *)&sAct[((wave * 4 + mb) * 16 + lr) * SA + nb * 16 + 4 * lg] = pk;
      } else if (own && y < ny) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
          int rq = r + q;
          if (rq < a.r0) storeOut<OutT>(a.out0 + v * a.r0 + rq, val[q], a.accumulate);
          else storeOut<OutT>(a.out1 + v * a.r1 + (rq - a.r0), val[q], a.accumulate);
        }
      }
    }
  }
  if (!epi) return;
  __syncthreads();
  if (a.pool) {
    /* the brick's 8 x 2 x 2 parents (extents are even, so a parent's eight
     * children are all inside the box or all outside) */
    const int mx = nx >> 1, my = ny >> 1, mz = nz >> 1, R = 16 * NB;
    for (int e = threadIdx.x; e < 32 * R; e += blockDim.x) {
      const int row = e % R, pc = e / R;
      const int px = pc & 7, py = (pc >> 3) & 1, pz = pc >> 4;
      const int gx = (x0 >> 1) + px, gy = (y0 >> 1) + py, gz = (z0 >> 1) + pz, rr = row0 + row;
      if (gx >= mx || gy >= my || gz >= mz || rr >= a.r0) continue;
      float s = 0.f;
      for (int q = 0; q < 8; q++) {
        const int lx = 2 * px + (q & 1), ly = 2 * py + ((q >> 1) & 1), lz = 2 * pz + (q >> 2);
        s += bf2f(sAct[((lz * 4 + ly) * 16 + lx) * SA + row]);
      }
      a.pool[(((size_t)gz * my + gy) * mx + gx) * a.r0 + rr] = f2bf(0.125f * s);
    }
  }
  if (a.headW && threadIdx.x < 256) {
    /* every row of the voxel is in this workgroup (gridDim.y == 1) */
    const int t = threadIdx.x, gx = x0 + (t & 15), gy = y0 + ((t >> 4) & 3), gz = z0 + (t >> 6);
    if (gx < nx && gy < ny && gz < nz) {
      float s = a.headW[a.r0];
      for (int k = 0; k < a.r0; k++) s += a.headW[k] * bf2f(sAct[t * SA + k]);
      const size_t v = ((size_t)gz * ny + gy) * nx + gx;
      if (a.headY) a.headY[v] = s;
      if (a.corr) {
        const float sr  = (float)sqrt(a.corrScal[0] / (double)a.voxels());
        const float f   = (float)(sr * a.corrH2);
        const float val = s * f;
        const long long sj = nx + 2, sk = (long long)(nx + 2) * (ny + 2);
        const long long c  = (gz + 1) * sk + (gy + 1) * sj + gx + 1;
        a.corr[c] = val;
        if (gx == 0) a.corr[c - 1] = val;
        if (gx == nx - 1) a.corr[c + 1] = val;
        if (gy == 0) a.corr[c - sj] = val;
        if (gy == ny - 1) a.corr[c + sj] = val;
        if (gz == 0) a.corr[c - sk] = val;
        if (gz == nz - 1) a.corr[c + sk] = val;
      }
    }
  }
}

/* sum of the split-K chunks in chunk order, bias, ReLU, store */
template <typename OutT>
__global__ __launch_bounds__(256) void kConvSplitFinish(ConvArgs<uint16_t, OutT> a, const float *__restrict__ split,
    int chunks)
{
  const int rows = a.r0 + a.r1;
  const size_t total = (size_t)a.voxels() * rows;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    float val = 0.f;
    for (int z = 0; z < chunks; z++) val += split[(size_t)z * total + e];
    size_t v = e / rows;
    int r    = (int)(e - v * rows);
    val += a.bias ? a.bias[r] : 0.f;
    if (a.relu) val = val > 0.f ? val : 0.f;
    if (r < a.r0) storeOut<OutT>(a.out0 + v * a.r0 + r, val, a.accumulate);
    else storeOut<OutT>(a.out1 + v * a.r1 + (r - a.r0), val, a.accumulate);
  }
}

/* Weight gradient of one tap: dW[n][t][c] += sum_v dOut[v][n] * in[v + off(t)][c].
 * Workgroup: 32 output channels x 32 input channels of one tap over a chunk
 * of voxels; 32-voxel slices of both operands are staged transposed in LDS
 * so each MFMA fragment is one 16 B LDS read. Each wave owns one 16 x 16
 * block; partial sums go to dW with FP32 atomics (split over voxel chunks). */
template <typename GT>
__global__ __launch_bounds__(256) void kConvWgrad(const GT *__restrict__ dOut, int cout,
    const uint16_t *__restrict__ in0, const uint16_t *__restrict__ in1, int c0, int c1, int nx, int ny,
    int nz, float *__restrict__ dW, int voxPerChunk)
{
  __shared__ __attribute__((aligned(16))) uint16_t sA[32][40]; /* [cout][vox], padded rows */
  __shared__ __attribute__((aligned(16))) uint16_t sB[32][40]; /* [cin][vox] */
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cin = c0 + c1, nvox = nx * ny * nz;
  const int t   = blockIdx.y % 27;
  const int nT  = blockIdx.y / 27;          /* cout tile */
  const int cT  = blockIdx.x;               /* cin tile */
  const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
  const int v0 = blockIdx.z * voxPerChunk, v1 = min(v0 + voxPerChunk, nvox);
  const int nb = wave >> 1, cb = wave & 1;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int vs = v0; vs < v1; vs += 32) {
    /* stage: 32 vox x 32 channels of each operand, 4 elements per thread */
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int e = tid + 256 * q;       /* 0..1023 */
      int vv = e >> 5, ch = e & 31; /* voxel in slice, channel in tile */
      int v  = vs + vv;
      float ga = 0.f;
      int co = nT * 32 + ch;
      if (v < v1 && co < cout) ga = (float)dOut[(size_t)v * cout + co];
      sA[ch][vv] = f2bf(ga);
      uint16_t gb = 0;
      int ci = cT * 32 + ch;
      if (v < v1 && ci < cin) {
        int x = v % nx + dx, y = (v / nx) % ny + dy, z = v / (nx * ny) + dz;
        if (z >= 0 && z < nz && y >= 0 && y < ny && x >= 0 && x < nx) {
          size_t w = ((size_t)z * ny + y) * nx + x;
          gb       = ci < c0 ? in0[w * c0 + ci] : in1[w * c1 + (ci - c0)];
        }
      }
      sB[ch][vv] = gb;
    }
    __syncthreads();
    bf16x8 fa = ld8(&sA[nb * 16 + (lane & 15)][8 * (lane >> 4)]);
    bf16x8 fb = ld8(&sB[cb * 16 + (lane & 15)][8 * (lane >> 4)]);
    acc       = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa, fb, acc, 0, 0, 0);
    __syncthreads();
  }
  int ci = cT * 32 + cb * 16 + (lane & 15);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int co = nT * 32 + nb * 16 + 4 * (lane >> 4) + q;
    if (co < cout && ci < cin) atomicAdd(&dW[((size_t)co * 27 + t) * cin + ci], acc[q]);
  }
}

#endif
