/* Geometric multigrid pressure solve to a relative residual — the `gmg`
 * solver. Added form: the reference's src/solver-mg.c has the V-cycle's
 * skeleton (smooth / residual / restrict / prolongate / correct,
 * src/solver-mg.c:299-354) but its coarse correction never reaches p; this is
 * a working cell-centred V(nu1, nu2) cycle on the same discrete operator
 * (7-point Laplacian with Neumann ghost copy, src/solver-rb.c:66-69,129-148).
 *
 *   smoother      red-black Gauss-Seidel/SOR, odd cells first (as the
 *                 reference), relaxation mg_omega; face ghosts written by the
 *                 second colour pass itself (no extra launch)
 *   restriction   average of the 8 children of the fine residual, fused with
 *                 the residual: b_{l+1} = (1/8) sum (b - A x)
 *   prolongation  cell-centred trilinear (27/64, 9/64, 3/64, 1/64), coarse
 *                 values read at clamped indices (Neumann extension)
 *   coarsest      exact separable DCT solve in one workgroup (LDS); the
 *                 constant mode is dropped (all-Neumann compatibility)
 *   stop test     ||rhs - A p||_2 / ||rhs||_2 <= mg_tol in FP64 on the finest
 *                 grid, checked after every cycle
 *
 * precision fp64:  V-cycles in FP64 directly on p.
 * precision mixed: defect correction — FP64 residual r = rhs - A p (and its
 *                  norm) on the GPU, cast to FP32, one FP32 V-cycle for
 *                  A e = r from e = 0, then p += e in FP64.
 *
 * All level arrays use the padded layout of the fields (one ghost layer,
 * i fastest); level l has h_l = 2^l h_0. */
#include <math.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

#include "nusif_internal.h"
#include "reduce.h"
#include "warmstart.h"

#define MAXLEV 12
#define GTX 64
#define GTY 4
#define GKCH 8
#define COARSE_MAX_CELLS 2048 /* interior cells of the coarsest level (LDS solve) */

/* Set while cycles are enqueued speculatively (the device flag that makes
 * their kernels return at once), null otherwise. The library is
 * single-threaded, as the reference's API. */
static const int *gCycleStop = nullptr;

template <typename T> struct Coef {
  T idx2, idy2, idz2, factor;
};

template <typename T> struct LevelT {
  Geo g;
  T *x, *b;
  T *xAlt; /* ping-pong partner of x for the fused sweep */
  Coef<T> cf;
};

struct DctTables {
  int oCi, oCj, oCk, oDi, oDj, oDk, oLi, oLj, oLk, total;
};

static DctTables dctLayout(int ni, int nj, int nk)
{
  DctTables t;
  int o = 0;
  t.oCi = o; o += ni * ni;
  t.oCj = o; o += nj * nj;
  t.oCk = o; o += nk * nk;
  t.oDi = o; o += ni * ni;
  t.oDj = o; o += nj * nj;
  t.oDk = o; o += nk * nk;
  t.oLi = o; o += ni;
  t.oLj = o; o += nj;
  t.oLk = o; o += nk;
  t.total = o;
  return t;
}

/* One captured cycle: the buffer pointers it was captured with (the key)
 * and the pointer state the host-side swaps leave behind. */
struct CycleGraph {
  std::vector<void *> before, after;
  hipGraphExec_t exec;
};

#define CTL_RING 8 /* cycles in flight at most (pinned read-back slots) */
struct CycleCtl;

struct GmgState {
  std::vector<CycleGraph> graphs; /* replayable cycles, at most graphCache */
  int useGraphs;
  int graphCache;
  int speculate;  /* enqueue cycles ahead of the convergence test */
  int predicted;  /* cycles of the previous solve: the first batch */
  int inner;      /* cycles before a solve's first stop test; 0: nextGroup */
  int nextGroup;  /* mg_inner 0: the next solve's first group (starts at 1) */
  int lastGroup;  /* the last solve's first group */
  int groupCap;   /* mixed, mg_inner 0: longest useful first group (0: none) */
  int groupHeld;  /* the last solve needed no cycle after its first group */
  CycleCtl *dCtl; /* device cycle control */
  CycleCtl *hRing;/* pinned copies of dCtl after each enqueued cycle */
  CycleCtl *hRingDev; /* their device address (written by kFoldCycle) */
  CycleCtl *hInit;
  hipEvent_t ev[CTL_RING];
  hipEvent_t evInput; /* after the warm start's input pass (its norms) */
  DctTables dct;
  double *tab64;
  float *tab32;
  int nl;
  int nu1, nu2, maxCycles, project, mixed;
  int tail;     /* first level of the one-workgroup tail cycle, -1: none */
  int tileRows; /* fused-sweep tile height: 0 automatic, 4, 8 or 16 */
  int fuseProlong; /* prolongation inside the first post-smoothing sweep */
  /* the next finest-level cycle starts from P setXc (level setGc, the
   * finest halved) in its first sweep instead of from L32[0].x */
  const float *setXc;
  Geo setGc;
  double tol, omega;
  LevelT<double> L64[MAXLEV];
  LevelT<float> L32[MAXLEV];
  Coef<double> cf0; /* finest level, FP64 */
};

template <typename T> static Coef<T> coefFor(double dx, double dy, double dz, double omega)
{
  Coef<T> c;
  double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
  c.idx2   = (T)(1.0 / dx2);
  c.idy2   = (T)(1.0 / dy2);
  c.idz2   = (T)(1.0 / dz2);
  c.factor = (T)(omega * 0.5 * (dx2 * dy2 * dz2) / (dy2 * dz2 + dx2 * dz2 + dx2 * dy2));
  return c;
}

template <typename T>
__device__ inline T lap7(const T *x, long long c, long long sj, long long sk, const Coef<T> &cf)
{
  T xc = x[c];
  return (x[c + 1] - (T)2 * xc + x[c - 1]) * cf.idx2 + (x[c + sj] - (T)2 * xc + x[c - sj]) * cf.idy2 +
         (x[c + sk] - (T)2 * xc + x[c - sk]) * cf.idz2;
}

/* One colour pass. Thread x owns the cell pair (2x+1, 2x+2) of a row and
 * updates the one of colour `parity`. With `ghosts` (second pass), it also
 * writes the face ghosts of both cells of its pair: the other colour's
 * values are final by then. */
template <typename T>
__global__ __launch_bounds__(256) void kSmooth(T *__restrict__ x, const T *__restrict__ b, Geo g,
    Coef<T> cf, int parity, int ghosts, const int *__restrict__ stop)
{
  if (stop && *stop) return;
  int tx = blockIdx.x * blockDim.x + threadIdx.x;
  int j  = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k0 = blockIdx.z * GKCH + 1;
  if (j > g.nj) return;
  int i1 = 2 * tx + 1;
  if (i1 > g.ni) return;
  int k1 = min(k0 + GKCH, g.nk + 1);
  for (int k = k0; k < k1; k++) {
    int i   = (((1 + j + k) & 1) == parity) ? i1 : i1 + 1;
    int io  = (i == i1) ? i1 + 1 : i1; /* the pair's other cell */
    bool in = i <= g.ni;
    T v     = 0;
    long long c = g.at(i, j, k);
    if (in) {
      T r = b[c] - lap7(x, c, g.sj, g.sk, cf);
      v   = x[c] - cf.factor * r;
      x[c] = v;
    }
    if (ghosts) {
      if (in) faceGhosts(x, g, i, j, k, c, v);
      if (io <= g.ni) {
        long long co = g.at(io, j, k);
        faceGhosts(x, g, io, j, k, co, x[co]);
      }
    }
  }
}

/* Workgroup barrier that orders LDS traffic only: waits for this wave's
 * LDS operations, then s_barrier. Unlike __syncthreads() it does not wait
 * for outstanding global loads, so the prefetches below stay in flight
 * across it. The "memory" clobber keeps the compiler from moving memory
 * accesses across the barrier. */
__device__ inline void ldsBarrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

#define NUM_XCD 8 /* MI355X: 8 XCDs, each with its own L2 */
#define FT_I 64
#define FT_KC 32
#define FUSED_MIN_EXTENT 32 /* levels narrower than this use the two-pass sweep */

/* One full red-black sweep in one launch (ping-pong: reads xo, writes xn).
 * A workgroup of 32 * TJ threads owns a 64 x TJ (i, j) tile and marches a
 * chunk of kc planes in k. Plane k's red cells ((i + j + k) odd) are updated
 * from the old values of planes k-1..k+1 held in LDS (X), over the tile plus a
 * one-cell ring, into LDS (Y); then plane k-1's black cells are updated from
 * Y, one per thread. The ring and the planes k0-1 and k1 are computed
 * redundantly, so no workgroup reads another's new values. Every cell gets
 * exactly the values of the two-pass sweep (same operands, same order); face
 * ghosts are written with the new values. X and Y rotate over 4 plane slots:
 * loads of plane k+2 never touch a slot still being read. Each thread reads
 * the b of its black cell and of that cell's red partner (the column pair
 * 2p, 2p+1) in one load and writes the pair's two new values in one store;
 * the red ring cells of the tile are a second, partial round of tasks.
 *
 * LDS holds every plane split by the parity of (i + j): half h, row lj,
 * position p = li / 2 of the local column li (tile origin minus 2). On plane
 * k the red cells are exactly half (k + 1) & 1 and the black cells half k & 1,
 * and every neighbour a red or black update reads sits at position p - 1 + s,
 * p + s (i -/+ 1, other half) or p (j, k neighbours) of its row, s = li & 1
 * being uniform along a row. Consecutive lanes therefore read consecutive LDS
 * words with no stride-2 bank conflicts, and no lane computes a value that is
 * thrown away (apart from two corner positions of the red ring and lanes past
 * the task count of the last red round).
 *
 * X: old values of plane k over local columns 0..67 and rows 0..TJ+3, both
 * halves. Y: new red values over rows 1..TJ+2, positions 0..33 (the ghost and
 * out-of-box entries keep their old values). */
template <typename T, int TJ>
struct SweepRBShape {
  static constexpr int NT   = 32 * TJ;        /* threads: one black cell each */
  static constexpr int XW   = FT_I + 4;       /* local columns 0..67 */
  static constexpr int XH   = TJ + 4;         /* local rows 0..TJ+3 */
  static constexpr int RW   = XW / 2;         /* positions per half row */
  static constexpr int NXE  = XW * XH;        /* X elements per plane */
  static constexpr int NX   = (NXE + NT - 1) / NT;
  /* half stride = 16 (mod 32) words: a plane store's even and odd lanes land
   * in disjoint banks */
  static constexpr int HALF = XH * RW + ((16 - (XH * RW) % 32) + 32) % 32;
  static constexpr int NRT  = RW * (TJ + 2);  /* red tasks per plane */
  static constexpr int YS   = NRT;
};

/* Plane-relative addressing through buffer instructions: a buffer
 * descriptor (base = the first cell of a plane) is uniform and sits in
 * scalar registers, each lane passes a 32-bit byte offset within a plane
 * that stays fixed for the whole march, and the plane is a uniform scalar
 * offset. Global loads with 64-bit addresses cost two 64-bit VALU adds per
 * access here (about a third of the sweep's VALU instructions); buffer
 * accesses cost none, and one descriptor per array and workgroup (not one
 * per plane and access) saves the scalar work of a 64-bit base each time
 * (3.7% of the finest FP32 sweep at 256^3). Every offset is inside the
 * padded box (the callers clamp), so the descriptor's range (2^31 - 1
 * bytes) is no bound check. 0x00020000: the descriptor's fourth word for
 * gfx9-family GPUs (32-bit data format, no swizzle). */
template <typename T> __device__ inline __amdgpu_buffer_rsrc_t planeRsrc(const T *a, const Geo &g, int k)
{
  return __builtin_amdgcn_make_buffer_rsrc((void *)(a + (long long)k * g.sk), (short)0, 0x7fffffff, 0x00020000);
}
__device__ inline float ldP(float *, __amdgpu_buffer_rsrc_t r, unsigned off)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ inline double ldP(double *, __amdgpu_buffer_rsrc_t r, unsigned off)
{
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ inline void stP(__amdgpu_buffer_rsrc_t r, unsigned off, float v)
{
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}
__device__ inline void stP(__amdgpu_buffer_rsrc_t r, unsigned off, double v)
{
  using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, 0, 0);
}
/* the same with a scalar (uniform) byte offset `so`: one descriptor per
 * array serves every plane, so a plane step costs a multiply, not a new
 * 64-bit base and descriptor per access */
__device__ inline float ldPS(float *, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned so)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, so, 0));
}
__device__ inline double ldPS(double *, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned so)
{
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, so, 0));
}
__device__ inline void stPS(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned so, float v)
{
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, so, 0);
}
__device__ inline void stPS(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned so, double v)
{
  using V2 = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, off, so, 0);
}

/* two adjacent cells in one access (multi-dword buffer accesses need only
 * dword alignment): b under a black thread's column pair, and the pair's two
 * new values */
__device__ inline void ldPair(float *, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned so, float &a, float &b)
{
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, so, 0);
  /* through plain scalars: a bit cast of a vector element itself reads
   * element 0 with this compiler */
  const unsigned e0 = v[0], e1 = v[1];
  a = __builtin_bit_cast(float, e0);
  b = __builtin_bit_cast(float, e1);
}
__device__ inline void ldPair(double *, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned so, double &a, double &b)
{
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, off, so, 0);
  a = __builtin_bit_cast(double, (unsigned long long)v[0] | ((unsigned long long)v[1] << 32));
  b = __builtin_bit_cast(double, (unsigned long long)v[2] | ((unsigned long long)v[3] << 32));
}
__device__ inline void stPair(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned so, float a, float b)
{
  decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)) v;
  v[0] = __builtin_bit_cast(unsigned, a);
  v[1] = __builtin_bit_cast(unsigned, b);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, off, so, 0);
}
__device__ inline void stPair(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned so, double a, double b)
{
  decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0)) v;
  const unsigned long long ua = __builtin_bit_cast(unsigned long long, a), ub = __builtin_bit_cast(unsigned long long, b);
  v[0] = (unsigned)ua;
  v[1] = (unsigned)(ua >> 32);
  v[2] = (unsigned)ub;
  v[3] = (unsigned)(ub >> 32);
  __builtin_amdgcn_raw_buffer_store_b128(v, r, off, so, 0);
}

/* RES (the reference's red-black SOR iteration, src/solver-rb.c:55-148): the
 * sweep is iteration q of a batch; it returns at once when q >= *stopAt (the
 * device-side stop test) and adds the in-sweep r^2 of the cells its tile owns
 * (red and black, each r taken before that cell's update, as
 * src/solver-rb.c:66-72) into partials[tile]. */
struct SweepRes {
  double *partials  = nullptr;
  const int *stopAt = nullptr;
  int q             = 0;
};

/* PRO (the first post-smoothing sweep of a level): the sweep starts from
 * x + P xc (PRO 1) or from P xc alone (PRO 2, x not read) instead of x, P
 * the cell-centred trilinear prolongation of kProlongMarch below, evaluated
 * with the same operands in the same order; the face ghosts of the start
 * are those of its interior neighbours, as kProlongMarch leaves them. So the
 * sweep gives the bits of kProlongMarch followed by a plain sweep, without
 * the pass that writes and reads x + P xc once more. */
template <typename T> struct SweepPro {
  const T *xc = nullptr; /* the coarse level's x (face ghosts not read) */
  Geo gc{};
};

template <typename T, int TJ, int MINW, bool ZX, bool RES = false, int PRO = 0>
__global__ __launch_bounds__(32 * TJ, MINW) void kSweepRB(const T *__restrict__ xo, T *__restrict__ xn,
    const T *__restrict__ b, Geo g, Coef<T> cf, int kc, const int *__restrict__ stop,
    SweepRes rs = SweepRes(), SweepPro<T> pr = SweepPro<T>())
{
  if (stop && *stop) return;
  if (RES && rs.q >= *rs.stopAt) return;
  using S = SweepRBShape<T, TJ>;
  constexpr int NT = S::NT, XW = S::XW, RW = S::RW, HALF = S::HALF, NX = S::NX;
  /* red tasks: round 0 the red cell of this thread's black pair, round 1 the
   * ring (rows 1 and TJ+2, and positions 0 and RW-1 of the rows between) */
  constexpr int NR = 2, NRING = 2 * RW + 2 * TJ;
  static_assert(NRING <= NT, "ring tasks fit one round");
  static_assert(PRO == 0 || (!ZX && !RES), "the prolongation goes into a plain sweep only");
  __shared__ T X[4][2 * HALF];
  __shared__ T Y[4][S::YS];
  /* PRO: the coarse values under a tile (CW x CH cells of one coarse plane,
   * indices clamped into the coarse interior), two planes of them; and two
   * planes of S (below) at the X elements, each thread's own */
  constexpr int CW = FT_I / 2 + 4, CH = TJ / 2 + 4, CSZ = CW * CH, NCS = (CSZ + NT - 1) / NT;
  static_assert(4 * CSZ <= 8 * HALF, "the first coarse slabs fit the X planes");
  __shared__ T CSl[PRO ? 2 * CSZ : 1];
  __shared__ T SPw[PRO ? 2 : 1][PRO ? S::NXE : 1];
  const int tid = threadIdx.x, wave0 = tid & ~63;
  const int gxT = (g.ni + FT_I - 1) / FT_I, gyT = (g.nj + TJ - 1) / TJ;
  const int nTiles = gxT * gyT * ((g.nk + kc - 1) / kc);
  const int perXcd = (nTiles + NUM_XCD - 1) / NUM_XCD;
  const int t = (blockIdx.x % NUM_XCD) * perXcd + blockIdx.x / NUM_XCD;
  if (t >= nTiles) return;
  const int i0 = (t % gxT) * FT_I + 1, j0 = ((t / gxT) % gyT) * TJ + 1;
  const int k0 = (t / (gxT * gyT)) * kc + 1;
  const int k1 = min(k0 + kc, g.nk + 1); /* black planes [k0, k1) */
  const int nk = g.nk, sj = g.sj, ni = g.ni, nj = g.nj;
  const long long sk = g.sk;
  const int ib = i0 - 2, jb = j0 - 2; /* global (i, j) of local (0, 0) */

  /* X loads: element e = (li, lj) row-major, clamped into the padded box
   * (clamped entries are only read by updates whose results are not kept);
   * PRO: into the interior, a face ghost of the start being its interior
   * neighbour's value */
  constexpr unsigned ES = sizeof(T);
  constexpr int CLO = PRO ? 1 : 0;
  unsigned xOff[NX]; /* byte offsets within a plane */
  int xLds[NX];
#pragma unroll
  for (int q = 0; q < NX; q++) {
    int e  = min(tid + NT * q, S::NXE - 1);
    int li = e % XW, lj = e / XW;
    xOff[q] = (unsigned)(min(max(jb + lj, CLO), nj + 1 - CLO) * sj + min(max(ib + li, CLO), ni + 1 - CLO)) * ES;
    xLds[q] = ((li + lj) & 1) * HALF + lj * RW + (li >> 1);
  }
  /* black task: position p in 1..32, row lj in 2..TJ+1; black column 2p + s,
   * s = (k + lj) & 1, its red partner 2p + 1 - s. The pair's b comes in one
   * load and serves both. */
  const int bp = 1 + (tid & 31), blj = 2 + (tid >> 5);
  const int bj = jb + blj, bi = ib + 2 * bp; /* global i of column 2p */
  const unsigned bRow = (unsigned)(min(bj, nj + 1) * sj + min(bi, ni)) * ES; /* bytes */
  const unsigned sjB  = (unsigned)sj * ES;
  const bool bjIn = bj <= nj;
  const bool bIn0 = bjIn && bi <= ni, bIn1 = bjIn && bi + 1 <= ni;
  /* workgroup-uniform: the tile holds cells next to an i or j face, whose
   * Neumann ghosts its black pass writes */
  const bool faceTile = i0 == 1 || i0 + FT_I > ni || j0 == 1 || j0 + TJ > nj;
  /* red tasks (position p, row lj; the column is 2p + s with s = (k + 1 + lj)
   * & 1 on plane k; rIn bit s: that cell is interior in (i, j)). Round 0 is
   * (bp, blj); round 1 ring task q: row 1 (q < RW), row TJ+2 (q < 2 RW), then
   * positions 0 and RW-1 of rows 2..TJ+1. */
  int rP[NR], rLj[NR], rIn[NR];
  unsigned rOff[NR]; /* byte offset of column 2p within a plane */
#pragma unroll
  for (int r = 0; r < NR; r++) {
    int p, lj;
    if (r == 0) {
      p  = bp;
      lj = blj;
    } else {
      const int q = min(tid, NRING - 1);
      if (q < RW) {
        p  = q;
        lj = 1;
      } else if (q < 2 * RW) {
        p  = q - RW;
        lj = TJ + 2;
      } else {
        p  = ((q - 2 * RW) & 1) ? RW - 1 : 0;
        lj = 2 + ((q - 2 * RW) >> 1);
      }
    }
    int j = jb + lj, i = ib + 2 * p;
    rP[r] = p; rLj[r] = lj;
    rOff[r] = (unsigned)(min(max(j, 0), nj + 1) * sj + min(max(i, 0), ni)) * ES;
    bool jin = j >= 1 && j <= nj;
    rIn[r] = (jin && i >= 1 && i <= ni ? 1 : 0) | (jin && i + 1 >= 1 && i + 1 <= ni ? 2 : 0);
  }
  double acc = 0.0; /* RES: this thread's share of the sum of r^2 */
  /* one descriptor per array, based at the lowest plane this workgroup
   * touches (k0 - 2, clamped); plane k is the scalar offset (k - kBase)
   * planes, at most kc + 3 of them (the launchers check that it fits) */
  const int kBase = max(k0 - 2, 0);
  const __amdgpu_buffer_rsrc_t rXo = planeRsrc(xo, g, kBase);
  const __amdgpu_buffer_rsrc_t rB  = planeRsrc(b, g, kBase);
  const __amdgpu_buffer_rsrc_t rXn = planeRsrc(xn, g, kBase);
  const unsigned PB = (unsigned)sk * ES;
  auto soff = [&](int k) { return (unsigned)(k - kBase) * PB; };

  /* ---- PRO: the coarse correction ----------------------------------------
   * Fine plane k (clamped into 1..nk) lies in coarse plane K = (k + 1) / 2
   * and leans towards K2 (K - 1 for odd k, K + 1 for even, clamped), so it
   * takes P xc = a S_K + q S_K2 with S_P = a c1 + q c2 as in kProlongMarch:
   * S_P is the bilinear part in (i, j) on coarse plane P. The planes it
   * needs are hiC(k) - 1 and hiC(k), hiC(k) = min(k / 2 + 1, nkc), and each
   * thread needs S only at its own X elements, so it keeps S of plane P in
   * SPw[P & 1] at those elements (written and read by that thread alone:
   * no barrier orders them) and makes the next plane's when storeX reaches
   * an even plane (hiC grows by one; hi is the highest plane made). The new
   * plane's S comes from its coarse slab in LDS (CSl[P & 1]), put there
   * two plane steps before: an even step kk loads the slab of plane
   * kk / 2 + 3 into registers, the odd step after it puts it into LDS, and
   * the odd step two later makes its S. The first planes of S and the slabs
   * the steps would have put before the first one come from the prologue.
   * (S in registers instead of SPw took the 64 x 16 FP32 sweep from 61 to
   * 85 VGPRs.) */
  const int nkc = PRO ? pr.gc.nk : 1;
  auto kClamp = [&](int k) { return min(max(k, 1), nk); };
  auto hiC    = [&](int k) { return min((kClamp(k) >> 1) + 1, nkc); };
  const int Ib = ((ib + 1) >> 1) - 1, Jb = ((jb + 1) >> 1) - 1; /* coarse (i, j) of slab (0, 0) */
  const int pNeed  = hiC(k1 + 1);                /* the last coarse plane this chunk uses */
  const int pFirst = max(hiC(k0 - 2) - 1, 1);   /* the first */
  unsigned cOff[NCS]; /* byte offsets of this thread's slab cells within a coarse plane */
#pragma unroll
  for (int r = 0; r < NCS; r++) {
    const int s = min(tid + NT * r, CSZ - 1);
    const int I = min(max(Ib + s % CW, 1), pr.gc.ni), J = min(max(Jb + s / CW, 1), pr.gc.nj);
    cOff[r] = PRO ? (unsigned)(J * pr.gc.sj + I) * ES : 0u;
  }
  /* based at the chunk's first coarse plane, as the fine arrays' are */
  const __amdgpu_buffer_rsrc_t rXc = PRO ? planeRsrc(pr.xc, pr.gc, pFirst) : rXo;
  const unsigned PBc = PRO ? (unsigned)pr.gc.sk * ES : 0u;
  T slab[NCS]; /* a coarse slab in flight from an even plane step to the next */
  auto loadSlab = [&](int P, T(&v)[NCS]) {
#pragma unroll
    for (int r = 0; r < NCS; r++) v[r] = ldPS((T *)nullptr, rXc, cOff[r], (unsigned)(P - pFirst) * PBc);
  };
  auto putSlab = [&](T *dst, const T(&v)[NCS]) {
#pragma unroll
    for (int r = 0; r < NCS; r++)
      if (r < NCS - 1 || tid + NT * r < CSZ) dst[tid + NT * r] = v[r];
  };
  /* slab indices of each X element's coarse cells (I, J) and (I, J2), and
   * I2 - I (+-1) */
  int sO11[NX], sO12[NX], sDI[NX];
#pragma unroll
  for (int u = 0; u < NX; u++) {
    const int e  = min(tid + NT * u, S::NXE - 1);
    const int li = e % XW, lj = e / XW;
    const int fi = min(max(ib + li, 1), ni), fj = min(max(jb + lj, 1), nj);
    const int lI = ((fi + 1) >> 1) - Ib, lJ = ((fj + 1) >> 1) - Jb;
    sO11[u] = lJ * CW + lI;
    sO12[u] = (lJ + ((fj & 1) ? -1 : 1)) * CW + lI;
    sDI[u]  = (fi & 1) ? -1 : 1;
  }
  auto makeS = [&](int P, const T *cs) { /* S_P at this thread's X elements from its slab cs */
    const T a = (T)0.75, q = (T)0.25;
#pragma unroll
    for (int u = 0; u < NX; u++)
      if (u < NX - 1 || tid + NT * u < S::NXE) {
        const T c1 = a * cs[sO11[u]] + q * cs[sO11[u] + sDI[u]];
        const T c2 = a * cs[sO12[u]] + q * cs[sO12[u] + sDI[u]];
        SPw[P & 1][tid + NT * u] = a * c1 + q * c2;
      }
  };
  int hi = 0; /* the highest plane of S made */
  /* the slab ring's work at plane step kk (before its first barrier); the
   * parity of kk is a compile-time argument, so a step holds no branch on it
   * (a join of paths with and without a load in flight makes the compiler
   * drain every outstanding load before the next use of one) */
  auto proStep = [&](auto odd, int kk) {
    if constexpr (!decltype(odd)::value) loadSlab(min(kk / 2 + 3, pNeed), slab);
    else if ((kk + 5) / 2 <= pNeed) putSlab(CSl + (((kk + 5) / 2) & 1) * CSZ, slab);
  };

  T xr[NX];
  auto fetchX = [&](int k) {
    if (ZX || PRO == 2) {
#pragma unroll
      for (int q = 0; q < NX; q++) xr[q] = (T)0;
      return;
    }
    const unsigned so = soff(PRO ? kClamp(k) : min(max(k, 0), nk + 1));
#pragma unroll
    for (int q = 0; q < NX; q++)
      xr[q] = q < NX - 1 || wave0 + NT * q < S::NXE ? ldPS((T *)nullptr, rXo, xOff[q], so) : (T)0;
  };
  auto storeX = [&](int k) {
    if (ZX) return;
    if constexpr (PRO != 0) { /* SPw holds planes hiC(k) - 1 and hiC(k) */
      const int kq = kClamp(k), K = (kq + 1) >> 1;
      const int K2 = (kq & 1) ? max(K - 1, 1) : min(K + 1, nkc);
      const T *s0 = SPw[K & 1], *s2 = SPw[K2 & 1];
#pragma unroll
      for (int q = 0; q < NX; q++)
        if (q < NX - 1 || tid + NT * q < S::NXE) {
          const T c = (T)0.75 * s0[tid + NT * q] + (T)0.25 * s2[tid + NT * q];
          X[k & 3][xLds[q]] = PRO == 2 ? c : xr[q] + c;
        }
      return;
    }
#pragma unroll
    for (int q = 0; q < NX; q++)
      if (q < NX - 1 || tid + NT * q < S::NXE) X[k & 3][xLds[q]] = xr[q];
  };
  auto fetchPair = [&](int k, T &pa, T &pb) { /* b under this thread's pair of plane k */
    ldPair((T *)nullptr, rB, bRow, soff(min(max(k, 1), nk)), pa, pb);
  };
  auto fetchRing = [&](int k) { /* b under this thread's ring task of plane k */
    if (wave0 >= NRING) return (T)0;
    const int s0 = (k + 1) & 1;
    return ldPS((T *)nullptr, rB, rOff[1] + ((s0 + rLj[1]) & 1) * ES, soff(min(max(k, 1), nk)));
  };
  auto redPlane = [&](int k, const T *bred) {
    const int hr = (k + 1) & 1;
    const T *Xm = X[(k + 3) & 3] + hr * HALF, *Xc = X[k & 3] + hr * HALF, *Xp = X[(k + 1) & 3] + hr * HALF;
    const T *Xo = X[k & 3] + (1 - hr) * HALF; /* the other half of plane k */
    T *Yc = Y[k & 3];
    const bool kin = k >= 1 && k <= nk;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (r == 0 || tid < NRING) {
        const int s = (hr + rLj[r]) & 1;
        const int a = rLj[r] * RW + rP[r];
        T v, lap;
        if (ZX) {
          v   = (T)0;
          lap = (((T)0 - (T)2 * v + (T)0) * cf.idx2 + ((T)0 - (T)2 * v + (T)0) * cf.idy2) +
                ((T)0 - (T)2 * v + (T)0) * cf.idz2;
        } else {
          v   = Xc[a];
          lap = (Xo[a + s] - (T)2 * v + Xo[a - 1 + s]) * cf.idx2 +
                (Xo[a + RW] - (T)2 * v + Xo[a - RW]) * cf.idy2 +
                (Xp[a] - (T)2 * v + Xm[a]) * cf.idz2;
        }
        T rr   = bred[r] - lap;
        T vNew = v - cf.factor * rr;
        bool in = kin && ((rIn[r] >> s) & 1);
        Yc[a - RW] = in ? vNew : v;
        if (RES) { /* the tile's own red cells: columns 2..65, rows 2..TJ+1, planes [k0, k1) */
          const int li = 2 * rP[r] + s;
          if (in && k >= k0 && k < k1 && li >= 2 && li <= FT_I + 1 && rLj[r] >= 2 && rLj[r] <= TJ + 1)
            acc += (double)rr * (double)rr;
        }
      }
    }
  };
  auto blackPlane = [&](int k, T bc) {
    const int hb = k & 1, s = (hb + blj) & 1;
    const T *Yc = Y[k & 3], *Ym = Y[(k + 3) & 3], *Yp = Y[(k + 1) & 3];
    const int a = (blj - 1) * RW + bp; /* Y index of this row and position */
    T xc  = ZX ? (T)0 : X[k & 3][hb * HALF + blj * RW + bp];
    T lap = (Yc[a + s] - (T)2 * xc + Yc[a - 1 + s]) * cf.idx2 +
            (Yc[a + RW] - (T)2 * xc + Yc[a - RW]) * cf.idy2 +
            (Yp[a] - (T)2 * xc + Ym[a]) * cf.idz2;
    T rr   = bc - lap;
    T vB   = xc - cf.factor * rr;
    T vR   = Yc[a];
    if (RES && (s ? bIn1 : bIn0)) acc += (double)rr * (double)rr;
    /* column 2p holds the black value when s = 0, the red one when s = 1 */
    T v0 = s ? vR : vB, v1 = s ? vB : vR;
    const unsigned so = soff(k);
    const unsigned c0 = bRow, c1 = bRow + ES;
    auto put = [&](unsigned sp) {
      if (bIn0 && bIn1) {
        stPair(rXn, c0, sp, v0, v1);
      } else {
        if (bIn0) stPS(rXn, c0, sp, v0);
        if (bIn1) stPS(rXn, c1, sp, v1);
      }
    };
    put(so);
    if (faceTile) { /* Neumann ghosts of the i and j faces */
      if (bIn0) {
        if (bi == 1) stPS(rXn, c0 - ES, so, v0);
        if (bi == ni) stPS(rXn, c1, so, v0);
        if (bj == 1) stPS(rXn, c0 - sjB, so, v0);
        if (bj == nj) stPS(rXn, c0 + sjB, so, v0);
      }
      if (bIn1) {
        if (bi + 1 == ni) stPS(rXn, c1 + ES, so, v1);
        if (bj == 1) stPS(rXn, c1 - sjB, so, v1);
        if (bj == nj) stPS(rXn, c1 + sjB, so, v1);
      }
    }
    if (k == 1) put(soff(0)); /* the k-face ghost planes (k == 1: kBase is 0) */
    if (k == nk) put(soff(nk + 1));
  };
  /* one plane step with register set (xs, pa/pb, rg): store plane kk+1 to
   * LDS, issue the loads of plane kk+3 (x) and kk+2 (b of the pair and the
   * ring task), then red(kk) and black(kk-1). The pair of plane kk gives
   * red(kk) its b now and black(kk) its b at the next step (bbCarry). Loads
   * for planes past the chunk's last use (x: k1+1, b: k1) are skipped. */
  auto iterate = [&](auto odd, int kk, T(&xs)[NX], T &pa, T &pb, T &rg, T &bbCarry) {
    if constexpr (PRO != 0) {
      proStep(odd, kk);
      if constexpr (decltype(odd)::value) { /* kk + 1 is even: the window may move up */
        const int h = hiC(kk + 1);
        if (h > hi) {
          makeS(h, CSl + (h & 1) * CSZ);
          hi = h;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < NX; q++) xr[q] = xs[q];
    storeX(kk + 1);
    const bool sR = ((kk + 1 + blj) & 1) != 0; /* the red cell is column 2p + 1 */
    T bUse[NR];
    bUse[0]    = sR ? pb : pa;
    bUse[1]    = rg;
    T bbUse    = bbCarry;
    bbCarry    = sR ? pa : pb;
    if (kk + 3 <= k1 + 1) {
      fetchX(kk + 3);
#pragma unroll
      for (int q = 0; q < NX; q++) xs[q] = xr[q];
    }
    if (kk + 2 <= k1) {
      fetchPair(kk + 2, pa, pb);
      rg = fetchRing(kk + 2);
    }
    ldsBarrier();
    redPlane(kk, bUse);
    ldsBarrier();
    if (kk - 1 >= k0) blackPlane(kk - 1, bbUse);
  };

  /* the chunk's first loads all at once (planes k0-2 .. k0+1 of x, the b of
   * k0-1 and k0), so the prologue waits one memory latency, not two: it is
   * a large part of a short chunk's time on the coarser levels */
  T xP[NX], xQ[NX], xA[NX], xB[NX], paA, pbA, paB, pbB, rgA, rgB, bbC = (T)0;
  fetchX(k0 - 2);
#pragma unroll
  for (int q = 0; q < NX; q++) xP[q] = xr[q];
  fetchX(k0 - 1);
#pragma unroll
  for (int q = 0; q < NX; q++) xQ[q] = xr[q];
  fetchX(k0);
#pragma unroll
  for (int q = 0; q < NX; q++) xA[q] = xr[q];
  fetchX(k0 + 1);
#pragma unroll
  for (int q = 0; q < NX; q++) xB[q] = xr[q];
  fetchPair(k0 - 1, paA, pbA);
  rgA = fetchRing(k0 - 1);
  fetchPair(k0, paB, pbB);
  rgB = fetchRing(k0);
  if constexpr (PRO != 0) {
    /* S of planes lo = pFirst and h0 = hiC(k0 - 2) for storeX(k0 - 2), of
     * h1 = hiC(k0 - 1) for storeX(k0 - 1) (made after storeX(k0 - 2): it
     * takes the SPw slot of lo), the slab of the next plane P1 = (k0 + 3) / 2
     * in CSl[P1 & 1] (the first odd step makes its S) and, when k0 is even,
     * the slab of P1 + 1 in registers for the first step (odd) to put: at
     * most four slabs. Those of lo and h0 are staged in the X planes, free
     * until the first storeX; h1's goes to CSl[h1 & 1] (k0 odd: the first
     * step to write that buffer is k0, two barriers after its last read). */
    const int h0 = hiC(k0 - 2), h1 = hiC(k0 - 1), P1 = min((k0 + 3) / 2, pNeed);
    const int pLoad = min(P1 + ((k0 & 1) ? 0 : 1), pNeed);
    T pro[4][NCS];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (pFirst + u <= pLoad) loadSlab(pFirst + u, pro[u]);
    T *tmp = &X[0][0];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int P = pFirst + u;
      if (P <= h0) putSlab(tmp + u * CSZ, pro[u]);
      if ((P == h1 && h1 > h0) || (P == P1 && P1 > h1)) putSlab(CSl + (P & 1) * CSZ, pro[u]);
      if (P == P1 + 1 && !(k0 & 1))
#pragma unroll
        for (int r = 0; r < NCS; r++) slab[r] = pro[u][r];
    }
    ldsBarrier();
    makeS(pFirst, tmp);
    if (h0 > pFirst) makeS(h0, tmp + (h0 - pFirst) * CSZ);
    hi = h0;
    ldsBarrier(); /* the reads of the staged slabs before the X planes are written */
#pragma unroll
    for (int q = 0; q < NX; q++) xr[q] = xP[q];
    storeX(k0 - 2);
    if (h1 > h0) {
      makeS(h1, CSl + (h1 & 1) * CSZ);
      hi = h1;
    }
#pragma unroll
    for (int q = 0; q < NX; q++) xr[q] = xQ[q];
    storeX(k0 - 1);
  } else {
#pragma unroll
    for (int q = 0; q < NX; q++) xr[q] = xP[q];
    storeX(k0 - 2);
#pragma unroll
    for (int q = 0; q < NX; q++) xr[q] = xQ[q];
    storeX(k0 - 1);
  }
  using Odd  = std::integral_constant<bool, true>;
  using Even = std::integral_constant<bool, false>;
  if constexpr (PRO == 0) {
    iterate(Odd(), k0 - 1, xA, paA, pbA, rgA, bbC);
    for (int kk = k0; kk <= k1; kk += 2) {
      iterate(Even(), kk, xB, paB, pbB, rgB, bbC);
      if (kk + 1 <= k1) iterate(Odd(), kk + 1, xA, paA, pbA, rgA, bbC);
    }
  } else {
    /* an even step loads a coarse slab and the odd step after it puts it
     * into LDS, so every loop pass runs an even step first and the slab's
     * registers are not carried round the loop (a loop-carried value still
     * being loaded is copied at the loop's end, after a wait for every
     * outstanding load). With k0 odd the first two steps run before the
     * loop, and the register sets swap so that one loop serves both. */
    int ks = k0;
    if (k0 & 1) {
      iterate(Even(), k0 - 1, xA, paA, pbA, rgA, bbC);
      iterate(Odd(), k0, xB, paB, pbB, rgB, bbC); /* k0 < k1 */
      ks = k0 + 1;
#pragma unroll
      for (int q = 0; q < NX; q++) {
        const T t = xA[q];
        xA[q]     = xB[q];
        xB[q]     = t;
      }
      T t = paA;
      paA = paB;
      paB = t;
      t   = pbA;
      pbA = pbB;
      pbB = t;
      t   = rgA;
      rgA = rgB;
      rgB = t;
    } else {
      iterate(Odd(), k0 - 1, xA, paA, pbA, rgA, bbC);
    }
    for (int kk = ks; kk <= k1; kk += 2) {
      iterate(Even(), kk, xB, paB, pbB, rgB, bbC);
      if (kk + 1 <= k1) iterate(Odd(), kk + 1, xA, paA, pbA, rgA, bbC);
    }
  }
  if constexpr (RES) {
    __shared__ double lsum[NT / 64];
    const double sum = blockSum(acc, lsum);
    if (tid == 0) rs.partials[t] = sum;
  }
}

/* Launch shape of the fused sweep on a level, from a sweep of tile heights
 * 4/8/16 and k-chunks 1..32 on 32^3..256^3 cubes (FP32 and FP64). Levels with
 * nj >= 128 use 64 x 16 tiles (512 threads): fewer ring cells per interior
 * cell, 44 us for FP32 at 256^3 against 49 us for 64 x 4. Their k-chunk is the
 * longest (<= FT_KC planes, a power of two >= 4) that still gives >= 1024
 * workgroups for FP32, >= 256 for FP64 (FP32 256^3: 16 planes, 128^3: 4;
 * FP64 256^3: 32, 128^3: 8). The small levels are latency-bound: 64 x 8
 * tiles for nj >= 64, 64 x 4 below, and the k-chunk that gives >= 2048
 * workgroups, down to one plane (64^3 and 32^3: one plane, 6.1 and 5.1 us
 * for FP32 against 8.0 and 7.7 us with four). `rows` (4, 8 or 16) forces the
 * tile height; 0 picks it. */
struct SweepShape {
  int tj, kc;
};
static long long sweepTiles(const Geo &g, int tj, int kc)
{
  return (long long)((g.ni + FT_I - 1) / FT_I) * ((g.nj + tj - 1) / tj) * ((g.nk + kc - 1) / kc);
}
template <typename T> static SweepShape sweepShape(const Geo &g, int rows)
{
  const int tj = rows ? rows : (g.nj >= 128 ? 16 : (g.nj >= 64 ? 8 : 4));
  const long long want = tj == 16 ? (sizeof(T) == 4 ? 1024 : 256) : 2048;
  const int kmin       = tj == 16 ? 4 : 1;
  int kc = FT_KC;
  while (kc > kmin && sweepTiles(g, tj, kc) < want) kc /= 2;
  return {tj, kc};
}

/* 1-D grid: NUM_XCD * ceil(tiles / NUM_XCD) workgroups (surplus ids exit) */
static dim3 fusedGrid(const Geo &g, SweepShape s)
{
  long long n = sweepTiles(g, s.tj, s.kc);
  return dim3((unsigned)(((n + NUM_XCD - 1) / NUM_XCD) * NUM_XCD));
}

template <typename T, int TJ, int MINW>
static void launchSweepShape(const LevelT<T> &L, dim3 grd, int kc, bool timed, bool zx)
{
  /* the zero-start sweep moves 2/3 of a sweep's bytes (no x read): timed apart */
  const char *name = sizeof(T) == 8 ? (zx ? "gmg_smooth64z_l0" : "gmg_smooth64_l0")
                                    : (zx ? "gmg_smooth32z_l0" : "gmg_smooth32_l0");
  constexpr int NT = SweepRBShape<T, TJ>::NT;
  if (zx) {
    if (timed)
      NUSIF_TIMED(name, (kSweepRB<T, TJ, MINW, true><<<grd, NT, 0, nusifStream()>>>(L.x, L.xAlt, L.b, L.g, L.cf, kc, gCycleStop)));
    else
      kSweepRB<T, TJ, MINW, true><<<grd, NT, 0, nusifStream()>>>(L.x, L.xAlt, L.b, L.g, L.cf, kc, gCycleStop);
  } else {
    if (timed)
      NUSIF_TIMED(name, (kSweepRB<T, TJ, MINW, false><<<grd, NT, 0, nusifStream()>>>(L.x, L.xAlt, L.b, L.g, L.cf, kc, gCycleStop)));
    else
      kSweepRB<T, TJ, MINW, false><<<grd, NT, 0, nusifStream()>>>(L.x, L.xAlt, L.b, L.g, L.cf, kc, gCycleStop);
  }
}

/* PRO 1: the sweep from L.x + P C.x; PRO 2: from P C.x (L.x not read) */
template <typename T, int TJ, int MINW, int PRO>
static void launchSweepProShape(const LevelT<T> &L, const LevelT<T> &C, dim3 grd, int kc, bool timed)
{
  const char *name = sizeof(T) == 8 ? (PRO == 2 ? "gmg_smooth64s_l0" : "gmg_smooth64p_l0")
                                    : (PRO == 2 ? "gmg_smooth32s_l0" : "gmg_smooth32p_l0");
  constexpr int NT = SweepRBShape<T, TJ>::NT;
  SweepPro<T> pr;
  pr.xc = C.x;
  pr.gc = C.g;
  if (timed)
    NUSIF_TIMED(name, (kSweepRB<T, TJ, MINW, false, false, PRO><<<grd, NT, 0, nusifStream()>>>(
                          L.x, L.xAlt, L.b, L.g, L.cf, kc, gCycleStop, SweepRes(), pr)));
  else
    kSweepRB<T, TJ, MINW, false, false, PRO><<<grd, NT, 0, nusifStream()>>>(
        L.x, L.xAlt, L.b, L.g, L.cf, kc, gCycleStop, SweepRes(), pr);
}

/* Workgroups of the PRO sweep resident at once on the device (its LDS and
 * registers allow fewer per CU than the plain sweep's: 3 against 4 for the
 * 64 x 16 FP32 tile) */
template <typename T, int TJ, int MINW> static long long proResident()
{
  static long long n = 0;
  if (!n) {
    int per = 0, dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kSweepRB<T, TJ, MINW, false, false, 1>,
        SweepRBShape<T, TJ>::NT, 0));
    n = (long long)(per > 0 ? per : 1) * prop.multiProcessorCount;
  }
  return n;
}

template <typename T, int PRO>
static void launchSweepPro(const LevelT<T> &L, const LevelT<T> &C, SweepShape s, bool timed)
{
  /* the coarse level is the fine one halved (the fused sweep's slab of
   * 36 x (TJ / 2 + 4) coarse cells per tile assumes it) */
  if (C.g.ni * 2 != L.g.ni || C.g.nj * 2 != L.g.nj || C.g.nk * 2 != L.g.nk)
    NUSIF_FATAL("fused prolongation: level %dx%dx%d is not twice %dx%dx%d", L.g.ni, L.g.nj, L.g.nk, C.g.ni,
        C.g.nj, C.g.nk);
  /* longer k-chunks where the plain sweep's tiles would take more than one
   * round of resident workgroups: a second, partial round costs nearly as
   * long as a full one (256^3 FP32: 768 chunks of 22 planes, not 1024 of 16) */
  const long long res = s.tj == 16 ? proResident<T, 16, 4>() : (s.tj == 8 ? proResident<T, 8, 6>() : proResident<T, 4, 8>());
  while (s.kc < FT_KC && sweepTiles(L.g, s.tj, s.kc) > res) s.kc++;
  dim3 grd = fusedGrid(L.g, s);
  if (s.tj == 16)
    launchSweepProShape<T, 16, 4, PRO>(L, C, grd, s.kc, timed);
  else if (s.tj == 8)
    launchSweepProShape<T, 8, 6, PRO>(L, C, grd, s.kc, timed);
  else
    launchSweepProShape<T, 4, 8, PRO>(L, C, grd, s.kc, timed);
}

template <typename T> static void launchSweep(const LevelT<T> &L, SweepShape s, bool timed, bool zx)
{
  dim3 grd = fusedGrid(L.g, s);
  if (s.tj == 16)
    launchSweepShape<T, 16, 4>(L, grd, s.kc, timed, zx);
  else if (s.tj == 8)
    launchSweepShape<T, 8, 6>(L, grd, s.kc, timed, zx);
  else
    launchSweepShape<T, 4, 8>(L, grd, s.kc, timed, zx);
}

/* true when a level's smoother is the fused sweep */
/* the fused sweep addresses at most FT_KC + 3 planes past its base with
 * 32-bit scalar offsets (descriptor range 2^31 - 1 bytes) */
static bool fusedPlanesFit(const Geo &g, size_t elem) { return (size_t)(FT_KC + 4) * g.sk * elem < ((size_t)1 << 31); }
template <typename T> static bool fusedLevel(const LevelT<T> &L)
{
  return L.g.ni >= FUSED_MIN_EXTENT && L.g.nk >= 4 && fusedPlanesFit(L.g, sizeof(T));
}

/* ---- the reference's red-black SOR iteration as one fused sweep ----------
 * Both colour passes of iteration q (src/solver-rb.c:55-79: pass 0 = odd
 * cells first) and the Neumann ghost copy after them (src/solver-rb.c:129-148:
 * the sweep writes the face ghosts of the new values; during both passes the
 * stencil reads the old ghosts, as the reference does) in one launch, from
 * xo into xn. Same operands in the same order as the colour-pass kernel, so
 * the same bits; the in-sweep sum of r^2 goes to partials[tile]. */
bool rbFusedFits(const Geo &g) { return g.ni >= FUSED_MIN_EXTENT && g.nk >= 4 && fusedPlanesFit(g, sizeof(double)); }

int rbFusedTiles(const Geo &g)
{
  const SweepShape s = sweepShape<double>(g, 0);
  return (int)sweepTiles(g, s.tj, s.kc);
}

void launchRBFused(const double *xo, double *xn, const double *rhs, const Geo &g, const SweepCoef &sc,
    double *partials, const int *stopAt, int q)
{
  const Coef<double> cf = {sc.idx2, sc.idy2, sc.idz2, sc.factor};
  const SweepShape s    = sweepShape<double>(g, 0);
  const dim3 grd        = fusedGrid(g, s);
  if (!partials) { /* a sweep without the residual (the reference MG's smoother) */
    if (s.tj == 16)
      NUSIF_TIMED("rb_fused", (kSweepRB<double, 16, 4, false><<<grd, SweepRBShape<double, 16>::NT, 0,
                                  nusifStream()>>>(xo, xn, rhs, g, cf, s.kc, nullptr)));
    else if (s.tj == 8)
      NUSIF_TIMED("rb_fused", (kSweepRB<double, 8, 6, false><<<grd, SweepRBShape<double, 8>::NT, 0,
                                  nusifStream()>>>(xo, xn, rhs, g, cf, s.kc, nullptr)));
    else
      NUSIF_TIMED("rb_fused", (kSweepRB<double, 4, 8, false><<<grd, SweepRBShape<double, 4>::NT, 0,
                                  nusifStream()>>>(xo, xn, rhs, g, cf, s.kc, nullptr)));
    return;
  }
  SweepRes rs;
  rs.partials = partials;
  rs.stopAt   = stopAt;
  rs.q        = q;
  if (s.tj == 16)
    NUSIF_TIMED("rb_fused", (kSweepRB<double, 16, 4, false, true><<<grd, SweepRBShape<double, 16>::NT, 0,
                                nusifStream()>>>(xo, xn, rhs, g, cf, s.kc, nullptr, rs)));
  else if (s.tj == 8)
    NUSIF_TIMED("rb_fused", (kSweepRB<double, 8, 6, false, true><<<grd, SweepRBShape<double, 8>::NT, 0,
                                nusifStream()>>>(xo, xn, rhs, g, cf, s.kc, nullptr, rs)));
  else
    NUSIF_TIMED("rb_fused", (kSweepRB<double, 4, 8, false, true><<<grd, SweepRBShape<double, 4>::NT, 0,
                                nusifStream()>>>(xo, xn, rhs, g, cf, s.kc, nullptr, rs)));
}

/* b_c(I,J,K) = (1/8) sum over the 8 children of (b - A x); the coarse x is
 * zeroed (interior and face ghosts: the coarse cycle's initial guess). */
template <typename T>
__global__ __launch_bounds__(256) void kResidualRestrict(const T *__restrict__ x,
    const T *__restrict__ b, Geo gf, Coef<T> cf, T *__restrict__ bc, T *__restrict__ xc, Geo gc,
    int gridGx, int gridGy, const int *__restrict__ stop)
{
  if (stop && *stop) return;
  int bx, by, bz;
  if (!xcdBlock(gridGx, gridGy, gc.nk, bx, by, bz)) return;
  int I = bx * blockDim.x + threadIdx.x + 1;
  int J = by * blockDim.y + threadIdx.y + 1;
  int K = bz + 1;
  if (I > gc.ni || J > gc.nj) return;
  T s = 0;
#pragma unroll
  for (int dk = 0; dk < 2; dk++)
#pragma unroll
    for (int dj = 0; dj < 2; dj++)
#pragma unroll
      for (int di = 0; di < 2; di++) {
        long long c = gf.at(2 * I - 1 + di, 2 * J - 1 + dj, 2 * K - 1 + dk);
        s += b[c] - lap7(x, c, gf.sj, gf.sk, cf);
      }
  long long cc = gc.at(I, J, K);
  bc[cc]       = (T)0.125 * s;
  xc[cc]       = (T)0;
  faceGhosts(xc, gc, I, J, K, cc, (T)0);
}

template <typename T> __global__ __launch_bounds__(256) void kFill(T *__restrict__ x, long long n, T v)
{
  long long stride = (long long)gridDim.x * blockDim.x;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) x[q] = v;
}

/* x_f += P x_c, cell-centred trilinear (weights 3/4, 1/4 per direction;
 * coarse indices clamped to the interior, i.e. a Neumann extension); face
 * ghosts of x_f refreshed. Fine cell (i,j,k) lies in coarse cell
 * (I,J,K) = ((i+1)/2, ...) and leans towards I2 = I -/+ 1 for odd/even i.
 * With a = 3/4, q = 1/4 and, per coarse plane P,
 *   c1(P) = a x(I,J,P) + q x(I2,J,P),  c2(P) = a x(I,J2,P) + q x(I2,J2,P),
 *   S(P)  = a c1(P) + q c2(P),
 * the correction is e = a S(K) + q S(K2). A thread owns one (i, j) column
 * and marches PKC coarse planes (2 PKC fine planes): each S(P) is computed
 * once and used by the two fine planes next to it (odd plane 2K-1 takes
 * K2 = max(K-1, 1), even plane 2K takes K2 = min(K+1, nkc)), so coarse values
 * are read once per column rather than eight times per fine cell. The tail
 * cycle's tailProlong evaluates the same expression per cell. */
#define PKC 4
template <typename T, bool SET = false>
__global__ __launch_bounds__(256) void kProlongMarch(T *__restrict__ xf, Geo gf, const T *__restrict__ xc,
    Geo gc, int gridGx, int gridGy, int gridGz, const int *__restrict__ stop)
{
  if (stop && *stop) return;
  int bx, by, bz;
  if (!xcdBlock(gridGx, gridGy, gridGz, bx, by, bz)) return;
  int i = bx * blockDim.x + threadIdx.x + 1;
  int j = by * blockDim.y + threadIdx.y + 1;
  if (i > gf.ni || j > gf.nj) return;
  const int I = (i + 1) >> 1, J = (j + 1) >> 1;
  const int I2 = min(max(I + ((i & 1) ? -1 : 1), 1), gc.ni);
  const int J2 = min(max(J + ((j & 1) ? -1 : 1), 1), gc.nj);
  const T a = (T)0.75, q = (T)0.25;
  const long long o11 = (long long)J * gc.sj + I, o21 = (long long)J * gc.sj + I2;
  const long long o12 = (long long)J2 * gc.sj + I, o22 = (long long)J2 * gc.sj + I2;
  auto S = [&](int P) {
    const T *pl = xc + (long long)P * gc.sk;
    T c1 = a * pl[o11] + q * pl[o21];
    T c2 = a * pl[o12] + q * pl[o22];
    return a * c1 + q * c2;
  };
  const int K0 = bz * PKC + 1, K1 = min(K0 + PKC, gc.nk + 1);
  T sm = S(max(K0 - 1, 1)), s0 = S(K0);
  for (int K = K0; K < K1; K++) {
    T sp = S(min(K + 1, gc.nk));
    int k = 2 * K - 1;
    long long c = gf.at(i, j, k);
    /* SET: x = P x_c (no old value added) */
    T v = SET ? (a * s0 + q * sm) : xf[c] + (a * s0 + q * sm);
    xf[c] = v;
    faceGhosts(xf, gf, i, j, k, c, v);
    c += gf.sk;
    v     = SET ? (a * s0 + q * sp) : xf[c] + (a * s0 + q * sp);
    xf[c] = v;
    faceGhosts(xf, gf, i, j, k + 1, c, v);
    sm = s0;
    s0 = sp;
  }
}

/* the same solve for any element count per thread (the decompositions
 * recomputed in every phase) */
template <typename T>
__device__ void coarseDctGeneric(T *__restrict__ x, const T *__restrict__ b, const Geo &g,
    const T *__restrict__ tab, const DctTables &L, unsigned char *smem)
{
  const int ni = g.ni, nj = g.nj, nk = g.nk, N = ni * nj * nk;
  T *W1 = (T *)smem;
  T *W2 = W1 + N;
  const T *Ci = tab + L.oCi, *Cj = tab + L.oCj, *Ck = tab + L.oCk;
  const T *Di = tab + L.oDi, *Dj = tab + L.oDj, *Dk = tab + L.oDk;
  const T *Li = tab + L.oLi, *Lj = tab + L.oLj, *Lk = tab + L.oLk;
  int tid = threadIdx.x, nt = blockDim.x;
  for (int q = tid; q < N; q += nt) {
    int i = q % ni, j = (q / ni) % nj, k = q / (ni * nj);
    W1[q] = b[g.at(i + 1, j + 1, k + 1)];
  }
  __syncthreads();
  for (int q = tid; q < N; q += nt) { /* along i */
    int m = q % ni, row = q - m;
    T s = 0;
    for (int i = 0; i < ni; i++) s += Ci[m * ni + i] * W1[row + i];
    W2[q] = s;
  }
  __syncthreads();
  for (int q = tid; q < N; q += nt) { /* along j */
    int mi = q % ni, m = (q / ni) % nj, k = q / (ni * nj);
    T s = 0;
    for (int j = 0; j < nj; j++) s += Cj[m * nj + j] * W2[(k * nj + j) * ni + mi];
    W1[q] = s;
  }
  __syncthreads();
  for (int q = tid; q < N; q += nt) { /* along k, then divide */
    int mi = q % ni, mj = (q / ni) % nj, m = q / (ni * nj);
    T s = 0;
    for (int k = 0; k < nk; k++) s += Ck[m * nk + k] * W1[(k * nj + mj) * ni + mi];
    T lam = Li[mi] + Lj[mj] + Lk[m];
    W2[q] = (q == 0) ? (T)0 : s / lam;
  }
  __syncthreads();
  for (int q = tid; q < N; q += nt) { /* inverse along k */
    int mi = q % ni, mj = (q / ni) % nj, k = q / (ni * nj);
    T s = 0;
    for (int m = 0; m < nk; m++) s += Dk[m * nk + k] * W2[(m * nj + mj) * ni + mi];
    W1[q] = s;
  }
  __syncthreads();
  for (int q = tid; q < N; q += nt) { /* inverse along j */
    int mi = q % ni, j = (q / ni) % nj, k = q / (ni * nj);
    T s = 0;
    for (int m = 0; m < nj; m++) s += Dj[m * nj + j] * W1[(k * nj + m) * ni + mi];
    W2[q] = s;
  }
  __syncthreads();
  for (int q = tid; q < N; q += nt) { /* inverse along i, store with face ghosts */
    int i = q % ni, row = q - i, j = (q / ni) % nj, k = q / (ni * nj);
    T s = 0;
    for (int m = 0; m < ni; m++) s += Di[m * ni + i] * W2[row + m];
    long long c = g.at(i + 1, j + 1, k + 1);
    x[c]        = s;
    faceGhosts(x, g, i + 1, j + 1, k + 1, c, s);
  }
}

/* Coarsest level, exact: the Neumann (ghost-copy) Laplacian is diagonal in
 * the separable DCT-II basis C_n[m][i] = cos(pi m (i + 1/2) / n) with
 * eigenvalues -4/h^2 sin^2(pi m / (2n)), so x = A^+ b is three 1-D forward
 * transforms, a division by the eigenvalue sums (the constant mode, the
 * null space, set to 0 — this is the rhs-mean projection), and three inverse
 * transforms. One workgroup, both work arrays in LDS; every output element
 * is one thread's ascending-index sum. Tables (host-built):
 *   fwd  Ci[ni*ni] Cj[nj*nj] Ck[nk*nk]          C[m][i]
 *   inv  Di Dj Dk (same shapes)                  s_m C[m][i], s_0 = 1/n, s_m = 2/n
 *   lam  Li[ni] Lj[nj] Lk[nk]                    eigenvalues */
template <typename T>
__device__ void coarseDct(T *__restrict__ x, const T *__restrict__ b, const Geo &g,
    const T *__restrict__ tab, const DctTables &L, unsigned char *smem)
{
  const int ni = g.ni, nj = g.nj, nk = g.nk, N = ni * nj * nk;
  T *W1 = (T *)smem;
  T *W2 = W1 + N;
  const T *Ci = tab + L.oCi, *Cj = tab + L.oCj, *Ck = tab + L.oCk;
  const T *Di = tab + L.oDi, *Dj = tab + L.oDj, *Dk = tab + L.oDk;
  const T *Li = tab + L.oLi, *Lj = tab + L.oLj, *Lk = tab + L.oLk;
  const int tid = threadIdx.x, nt = blockDim.x;
  /* each element's (a, b, c) = (q % ni, q / ni % nj, q / (ni nj)), found
   * once (integer divisions by the runtime extents were most of the
   * phases' instructions); a thread's elements are q = tid + u nt */
  constexpr int MAXQ = 2; /* COARSE_MAX_CELLS / 1024 */
  if (N > MAXQ * nt) {
    coarseDctGeneric(x, b, g, tab, L, smem);
    return;
  }
  int qa[MAXQ], qb[MAXQ], qc[MAXQ];
#pragma unroll
  for (int u = 0; u < MAXQ; u++) {
    const int q = tid + u * nt;
    qa[u] = q % ni;
    qb[u] = (q / ni) % nj;
    qc[u] = q / (ni * nj);
  }
#define DCT_EACH(...)                                                                                           \
  _Pragma("unroll") for (int u = 0; u < MAXQ; u++)                                                              \
  {                                                                                                             \
    const int q = tid + u * nt;                                                                                 \
    if (q < N) {                                                                                                \
      const int ea = qa[u], eb = qb[u], ec = qc[u];                                                             \
      __VA_ARGS__                                                                                               \
    }                                                                                                           \
  }
  DCT_EACH(W1[q] = b[g.at(ea + 1, eb + 1, ec + 1)];)
  __syncthreads();
  DCT_EACH({ /* along i */
    const int m = ea, row = q - m;
    T s = 0;
    _Pragma("unroll 8")
    for (int i = 0; i < ni; i++) s += Ci[m * ni + i] * W1[row + i];
    W2[q] = s;
  })
  __syncthreads();
  DCT_EACH({ /* along j */
    const int mi = ea, m = eb, k = ec;
    T s = 0;
    _Pragma("unroll 8")
    for (int j = 0; j < nj; j++) s += Cj[m * nj + j] * W2[(k * nj + j) * ni + mi];
    W1[q] = s;
  })
  __syncthreads();
  DCT_EACH({ /* along k, then divide */
    const int mi = ea, mj = eb, m = ec;
    T s = 0;
    _Pragma("unroll 8")
    for (int k = 0; k < nk; k++) s += Ck[m * nk + k] * W1[(k * nj + mj) * ni + mi];
    T lam = Li[mi] + Lj[mj] + Lk[m];
    W2[q] = (q == 0) ? (T)0 : s / lam;
  })
  __syncthreads();
  DCT_EACH({ /* inverse along k */
    const int mi = ea, mj = eb, k = ec;
    T s = 0;
    _Pragma("unroll 8")
    for (int m = 0; m < nk; m++) s += Dk[m * nk + k] * W2[(m * nj + mj) * ni + mi];
    W1[q] = s;
  })
  __syncthreads();
  DCT_EACH({ /* inverse along j */
    const int mi = ea, j = eb, k = ec;
    T s = 0;
    _Pragma("unroll 8")
    for (int m = 0; m < nj; m++) s += Dj[m * nj + j] * W1[(k * nj + m) * ni + mi];
    W2[q] = s;
  })
  __syncthreads();
  DCT_EACH({ /* inverse along i, store with face ghosts */
    const int i = ea, row = q - i, j = eb, k = ec;
    T s = 0;
    _Pragma("unroll 8")
    for (int m = 0; m < ni; m++) s += Di[m * ni + i] * W2[row + m];
    long long c = g.at(i + 1, j + 1, k + 1);
    x[c]        = s;
    faceGhosts(x, g, i + 1, j + 1, k + 1, c, s);
  })
#undef DCT_EACH
}

template <typename T>
__global__ __launch_bounds__(1024) void kCoarseDct(T *__restrict__ x, const T *__restrict__ b, Geo g,
    const T *__restrict__ tab, DctTables L, const int *__restrict__ stop)
{
  if (stop && *stop) return;
  extern __shared__ unsigned char smem[];
  coarseDct(x, b, g, tab, L, smem);
}

/* ---- the coarse tail of the V-cycle in one workgroup ---------------------
 * Levels of at most TAIL_MAX_CELLS cells are too small to fill the chip;
 * launched one kernel per colour pass they cost a launch gap each. One
 * workgroup of 1024 threads runs the whole V-cycle below the first such
 * level instead: the same smoother passes, residual restriction, trilinear
 * prolongation and DCT coarse solve as the per-level kernels (same operands,
 * same order, so the same values), separated by workgroup barriers. The
 * level arrays stay in this XCD's L2. Loads of B cells are issued before
 * any of their stores: within a colour pass (and within a prolongation) no
 * cell reads a value another cell of the pass writes, so this is safe and
 * keeps B x 8 loads in flight per thread. */
#define TAIL_MAX_CELLS 4096
#define MAXTAIL 8
#define TAIL_THREADS 1024
/* cell pairs (and prolongation cells) per thread with loads in flight: 2,
 * as many as a first tail level of at most 4096 cells has per thread (more
 * slots are code that every launch fetches and no level uses) */
template <typename T> struct TailB { static constexpr int value = 2; };

template <typename T> struct TailArgs {
  int n;          /* levels in the tail; level 0 of the tail is the finest of them */
  int nu1, nu2;
  T *x[MAXTAIL];
  T *b[MAXTAIL];
  Geo g[MAXTAIL];
  Coef<T> cf[MAXTAIL];
  const T *tab;   /* DCT tables of the coarsest level */
  DctTables dct;
  const int *stop; /* skip the launch when *stop (speculative cycles) */
  /* LDS residency (the kernel's RES): byte offsets of each level's x (RES
   * >= 1) and b (RES == 2) in the dynamic LDS, and of the DCT workspace */
  unsigned xo[MAXTAIL], bo[MAXTAIL], wo;
  int tabLds;   /* the DCT tables copied to LDS at offset `to` (RES >= 1) */
  unsigned to;
};

__device__ inline void tailSync() { __syncthreads(); }

/* q / d for 0 <= q < 2^22 and 1 <= d < 2^12 (the tail's index ranges): a
 * float estimate, corrected once in each direction, so exact; a few
 * instructions instead of an integer division by a runtime extent */
__device__ inline int tdiv(int q, int d)
{
  int t = (int)((float)q * __builtin_amdgcn_rcpf((float)d));
  if (t * d > q) t--;
  if ((t + 1) * d <= q) t++;
  return t;
}

/* one colour pass over all cell pairs (the kSmooth work of every thread) */
template <typename T>
__device__ void tailPass(T *__restrict__ x, const T *__restrict__ b, const Geo &g, const Coef<T> &cf,
    int parity, int ghosts)
{
  constexpr int TAIL_B = TailB<T>::value;
  const int np = (g.ni + 1) >> 1, total = np * g.nj * g.nk;
  for (int base = threadIdx.x; base < total; base += TAIL_B * TAIL_THREADS) {
    T xc[TAIL_B], xm[TAIL_B], xp[TAIL_B], ym[TAIL_B], yp[TAIL_B], zm[TAIL_B], zp[TAIL_B], bb[TAIL_B];
    T xo[TAIL_B];
    long long cc[TAIL_B], co[TAIL_B];
    int ii[TAIL_B], io[TAIL_B], jj[TAIL_B], kk[TAIL_B];
    bool live[TAIL_B];
#pragma unroll
    for (int u = 0; u < TAIL_B; u++) {
      int q   = base + u * TAIL_THREADS;
      live[u] = q < total;
      int qq  = live[u] ? q : 0;
      int r = tdiv(qq, np), tx = qq - r * np;
      int k0 = tdiv(r, g.nj);
      int j = r - k0 * g.nj + 1, k = k0 + 1, i1 = 2 * tx + 1;
      int i  = (((1 + j + k) & 1) == parity) ? i1 : i1 + 1;
      ii[u]  = i;
      io[u]  = (i == i1) ? i1 + 1 : i1;
      jj[u]  = j;
      kk[u]  = k;
      /* a pair's missing second cell (odd ni) reads a valid dummy cell */
      long long c = g.at(min(i, g.ni), j, k);
      cc[u] = c;
      co[u] = g.at(min(io[u], g.ni), j, k);
      xc[u] = x[c];
      xm[u] = x[c - 1];
      xp[u] = x[c + 1];
      ym[u] = x[c - g.sj];
      yp[u] = x[c + g.sj];
      zm[u] = x[c - g.sk];
      zp[u] = x[c + g.sk];
      bb[u] = b[c];
      xo[u] = x[co[u]];
    }
#pragma unroll
    for (int u = 0; u < TAIL_B; u++) {
      if (!live[u]) continue;
      bool in = ii[u] <= g.ni;
      T v     = 0;
      if (in) {
        T lap = (xp[u] - (T)2 * xc[u] + xm[u]) * cf.idx2 + (yp[u] - (T)2 * xc[u] + ym[u]) * cf.idy2 +
                (zp[u] - (T)2 * xc[u] + zm[u]) * cf.idz2;
        T r   = bb[u] - lap;
        v     = xc[u] - cf.factor * r;
        x[cc[u]] = v;
      }
      if (ghosts) {
        if (in) faceGhosts(x, g, ii[u], jj[u], kk[u], cc[u], v);
        if (io[u] <= g.ni) faceGhosts(x, g, io[u], jj[u], kk[u], co[u], xo[u]);
      }
    }
  }
}

/* A thread's cell pairs on one tail level, decomposed once for all the
 * colour passes of a smoothing block (the decomposition's integer divisions
 * by runtime extents were the larger part of a pass): pair q = tid + u *
 * TAIL_THREADS holds cells i1 = 2 p + 1 and i1 + 1 of row (j, k). Valid when
 * the level has at most TailB * TAIL_THREADS pairs. */
template <typename T> struct TailPlan {
  static constexpr int B = TailB<T>::value;
  int c1[B]; /* cell index of (i1, j, k) */
  int i1[B], j[B], k[B];
  bool live[B];
};

template <typename T> __device__ bool tailPlan(const Geo &g, TailPlan<T> &P)
{
  constexpr int B = TailB<T>::value;
  const int np = (g.ni + 1) >> 1, total = np * g.nj * g.nk;
  if (total > B * TAIL_THREADS) return false;
#pragma unroll
  for (int u = 0; u < B; u++) {
    int q     = threadIdx.x + u * TAIL_THREADS;
    P.live[u] = q < total;
    int qq    = P.live[u] ? q : 0;
    int r = tdiv(qq, np), tx = qq - r * np;
    int k0  = tdiv(r, g.nj);
    P.j[u]  = r - k0 * g.nj + 1;
    P.k[u]  = k0 + 1;
    P.i1[u] = 2 * tx + 1;
    P.c1[u] = (int)g.at(P.i1[u], P.j[u], P.k[u]);
  }
  return true;
}

/* tailPass on a plan: the same cells, operands and order */
template <typename T>
__device__ void tailPassPlan(T *__restrict__ x, const T *__restrict__ b, const Geo &g, const Coef<T> &cf,
    int parity, int ghosts, const TailPlan<T> &P)
{
  constexpr int B = TailB<T>::value;
  T xc[B], xm[B], xp[B], ym[B], yp[B], zm[B], zp[B], bb[B], xo[B];
  long long cc[B], co[B];
  int ii[B], io[B];
#pragma unroll
  for (int u = 0; u < B; u++) {
    /* slots past the level's pairs load nothing (on a level with fewer
     * pairs than slots the whole workgroup skips them together) */
    if (!P.live[u]) continue;
    const int i1 = P.i1[u];
    const int i  = (((1 + P.j[u] + P.k[u]) & 1) == parity) ? i1 : i1 + 1;
    ii[u] = i;
    io[u] = (i == i1) ? i1 + 1 : i1;
    /* a pair's missing second cell (odd ni) reads a valid dummy cell */
    long long c = P.c1[u] + (i <= g.ni ? i - i1 : 0);
    cc[u] = c;
    co[u] = P.c1[u] + (io[u] <= g.ni ? io[u] - i1 : 0);
    xc[u] = x[c];
    xm[u] = x[c - 1];
    xp[u] = x[c + 1];
    ym[u] = x[c - g.sj];
    yp[u] = x[c + g.sj];
    zm[u] = x[c - g.sk];
    zp[u] = x[c + g.sk];
    bb[u] = b[c];
    xo[u] = x[co[u]];
  }
#pragma unroll
  for (int u = 0; u < B; u++) {
    if (!P.live[u]) continue;
    bool in = ii[u] <= g.ni;
    T v     = 0;
    if (in) {
      T lap = (xp[u] - (T)2 * xc[u] + xm[u]) * cf.idx2 + (yp[u] - (T)2 * xc[u] + ym[u]) * cf.idy2 +
              (zp[u] - (T)2 * xc[u] + zm[u]) * cf.idz2;
      T r   = bb[u] - lap;
      v     = xc[u] - cf.factor * r;
      x[cc[u]] = v;
    }
    if (ghosts) {
      if (in) faceGhosts(x, g, ii[u], P.j[u], P.k[u], cc[u], v);
      if (io[u] <= g.ni) faceGhosts(x, g, io[u], P.j[u], P.k[u], co[u], xo[u]);
    }
  }
}

template <typename T>
__device__ void tailRestrict(const T *__restrict__ x, const T *__restrict__ b, const Geo &gf,
    const Coef<T> &cf, T *__restrict__ bc, const Geo &gc)
{
  const int total = gc.ni * gc.nj * gc.nk;
  for (int q = threadIdx.x; q < total; q += TAIL_THREADS) {
    const int r = tdiv(q, gc.ni), k0 = tdiv(r, gc.nj);
    int I = q - r * gc.ni + 1, J = r - k0 * gc.nj + 1, K = k0 + 1;
    T s = 0;
#pragma unroll
    for (int dk = 0; dk < 2; dk++)
#pragma unroll
      for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int di = 0; di < 2; di++) {
          long long c = gf.at(2 * I - 1 + di, 2 * J - 1 + dj, 2 * K - 1 + dk);
          s += b[c] - lap7(x, c, gf.sj, gf.sk, cf);
        }
    bc[gc.at(I, J, K)] = (T)0.125 * s;
  }
}

template <typename T>
__device__ void tailProlong(T *__restrict__ xf, const Geo &gf, const T *__restrict__ xc, const Geo &gc)
{
  constexpr int TAIL_B = TailB<T>::value;
  const int total = gf.ni * gf.nj * gf.nk;
  const T a = (T)0.75, qq = (T)0.25;
  for (int base = threadIdx.x; base < total; base += TAIL_B * TAIL_THREADS) {
    T e[TAIL_B], old[TAIL_B];
    long long cf[TAIL_B];
    int ii[TAIL_B], jj[TAIL_B], kk[TAIL_B];
    bool live[TAIL_B];
#pragma unroll
    for (int u = 0; u < TAIL_B; u++) {
      int q   = base + u * TAIL_THREADS;
      live[u] = q < total;
      int qv  = live[u] ? q : 0;
      const int rr = tdiv(qv, gf.ni), k0 = tdiv(rr, gf.nj);
      int i = qv - rr * gf.ni + 1, j = rr - k0 * gf.nj + 1, k = k0 + 1;
      ii[u] = i;
      jj[u] = j;
      kk[u] = k;
      int I = (i + 1) >> 1, J = (j + 1) >> 1, K = (k + 1) >> 1;
      int I2 = min(max(I + ((i & 1) ? -1 : 1), 1), gc.ni);
      int J2 = min(max(J + ((j & 1) ? -1 : 1), 1), gc.nj);
      int K2 = min(max(K + ((k & 1) ? -1 : 1), 1), gc.nk);
      T c11 = a * xc[gc.at(I, J, K)] + qq * xc[gc.at(I2, J, K)];
      T c21 = a * xc[gc.at(I, J2, K)] + qq * xc[gc.at(I2, J2, K)];
      T c12 = a * xc[gc.at(I, J, K2)] + qq * xc[gc.at(I2, J, K2)];
      T c22 = a * xc[gc.at(I, J2, K2)] + qq * xc[gc.at(I2, J2, K2)];
      e[u]  = a * (a * c11 + qq * c21) + qq * (a * c12 + qq * c22);
      cf[u] = gf.at(i, j, k);
      old[u] = xf[cf[u]];
    }
#pragma unroll
    for (int u = 0; u < TAIL_B; u++) {
      if (!live[u]) continue;
      T v = old[u] + e[u];
      xf[cf[u]] = v;
      faceGhosts(xf, gf, ii[u], jj[u], kk[u], cf[u], v);
    }
  }
}

template <typename T> __device__ void tailZero(T *__restrict__ x, long long n)
{
  for (long long q = threadIdx.x; q < n; q += TAIL_THREADS) x[q] = (T)0;
}

/* dst[0..n) = src[0..n), global to LDS, eight loads per thread in flight */
template <typename T> __device__ void tailCopyToLds(T *__restrict__ dst, const T *__restrict__ src, long long n)
{
  constexpr int PER = 8;
  for (long long base = threadIdx.x; base < n; base += PER * TAIL_THREADS) {
    T v[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const long long q = base + (long long)u * TAIL_THREADS;
      v[u]              = q < n ? src[q] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const long long q = base + (long long)u * TAIL_THREADS;
      if (q < n) dst[q] = v[u];
    }
  }
}

/* RES: which level arrays live in LDS for the whole cycle (0: none, the
 * arrays stay in global memory; 1: every x; 2: every x and b). Every phase
 * then reads and writes LDS instead of L2, and a phase costs an LDS round
 * trip and a barrier, not an L2 one. The tail's first level takes its b
 * from global memory (the restriction above wrote it) and gives its x back
 * (the prolongation above reads it); the coarser levels exist only in LDS.
 * The passes are the same code on other pointers: same values. */
template <typename T, int RES>
__global__ __launch_bounds__(TAIL_THREADS) void kTailCycle(TailArgs<T> a)
{
  if (a.stop && *a.stop) return;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  auto X = [&](int l) -> T * { return RES >= 1 ? (T *)(smem + a.xo[l]) : a.x[l]; };
  auto B = [&](int l) -> T * { return RES == 2 ? (T *)(smem + a.bo[l]) : a.b[l]; };
  unsigned char *work = RES >= 1 ? smem + a.wo : smem;
  const T *tab        = a.tab;
  /* the global loads of the prologue (DCT tables, the first level's b) are
   * all issued before their LDS stores: one memory latency, not one per
   * element a thread copies */
  if (RES >= 1 && a.tabLds) {
    T *t = (T *)(smem + a.to);
    tailCopyToLds(t, a.tab, (long long)a.dct.total);
    tab = t;
  }
  for (int l = 0; l < a.n; l++) tailZero(X(l), a.g[l].n);
  if (RES == 2) tailCopyToLds(B(0), a.b[0], a.g[0].n);
  tailSync();
  /* nu colour-pass pairs on level l, on a plan when the level allows one */
  /* the first level's plan serves both its smoothing blocks */
  TailPlan<T> P0;
  const bool planned0 = tailPlan(a.g[0], P0);
  auto smoothLevel = [&](int l, int nu) {
    TailPlan<T> P;
    const bool planned = l == 0 ? planned0 : tailPlan(a.g[l], P);
    if (l == 0) P = P0;
    for (int s = 0; s < nu; s++) {
      if (planned) tailPassPlan(X(l), B(l), a.g[l], a.cf[l], 1, 0, P);
      else tailPass(X(l), B(l), a.g[l], a.cf[l], 1, 0);
      tailSync();
      if (planned) tailPassPlan(X(l), B(l), a.g[l], a.cf[l], 0, 1, P);
      else tailPass(X(l), B(l), a.g[l], a.cf[l], 0, 1);
      tailSync();
    }
  };
  /* down and up as one loop over the 2 (n - 1) smoothing blocks, so the
   * smoothing code exists once (the kernel is fetched into the instruction
   * cache anew at every launch: a second copy of a block costs its fetch
   * again): block s smooths level s on the way down, then restricts (and
   * after the last one solves the coarsest level); on the way up it first
   * prolongates onto the level */
  const int nb = 2 * (a.n - 1);
  if (nb == 0) {
    coarseDct(X(0), B(0), a.g[0], tab, a.dct, work);
    tailSync();
  }
  for (int s = 0; s < nb; s++) {
    const bool down = s < a.n - 1;
    const int l     = down ? s : nb - 1 - s;
    if (!down) {
      tailProlong(X(l), a.g[l], X(l + 1), a.g[l + 1]);
      tailSync();
    }
    smoothLevel(l, down ? a.nu1 : a.nu2);
    if (down) {
      tailRestrict(X(l), B(l), a.g[l], a.cf[l], B(l + 1), a.g[l + 1]);
      tailSync();
      if (l == a.n - 2) {
        coarseDct(X(a.n - 1), B(a.n - 1), a.g[a.n - 1], tab, a.dct, work);
        tailSync();
      }
    }
  }
  if (RES >= 1) /* the first level's x (every padded cell) back to global */
    for (long long q = threadIdx.x; q < a.g[0].n; q += TAIL_THREADS) a.x[0][q] = X(0)[q];
}

/* ---- FP64 finest-level kernels: norms, defect correction ---------------- */
/* Residual of the finest level with its norms, one workgroup per 64 x RTY
 * (i, j) tile marching NKC planes: the tile's values plus a one-cell ring
 * are staged in LDS per plane (three planes rotate), so the 7-point stencil
 * reads LDS instead of seven global loads per cell. (64 x 8 tiles, with
 * fewer ring loads per cell, measured no faster at 256^3: L2 catches the
 * ring's re-reads.)
 *   CORR = false: x = p;      r = rhs - A x
 *   CORR = true:  x = p + e;  pn = x (face ghosts refreshed), r = rhs - A x
 *                 (the mixed-precision defect correction and the next
 *                 residual in one pass; A x at a neighbour is p + e there:
 *                 both hold Neumann face ghosts, so at a ghost this is the
 *                 value the corrected p gets, and pn is bit-identical to
 *                 correcting first)
 * Optional outputs: r as FP32 (r32), an FP32 array zeroed with its face
 * ghosts (eZero: the next FP32 cycle's initial guess). Partial sums of r^2
 * and rhs^2 per workgroup (slots 2 b, 2 b + 1). pn, eZero must not alias
 * p or e. */
#define NKC 16
#define RTY 4 /* rows of a residual tile */
/* Optional second output of kResidualLds: the residual (rc) and rhs (bc,
 * may be null) averaged over the 8 children of each cell of the next
 * coarser level (extents halved), summed in the children's order (i, then
 * j, then k) as double, r rounded to FP32 first when roundR: the values of
 * a separate 8-child restriction of r (or of its FP32 copy) and rhs. The
 * tile (64 x RTY, k-chunks of 16 from an odd plane) holds whole children
 * groups when the extents are even. */
struct ResidualRestrict {
  float *rc = nullptr, *bc = nullptr;
  Geo gc;
  int roundR = 0;
};

template <bool CORR, bool RESTR = false>
__global__ __launch_bounds__(GTX * RTY, RESTR ? 6 : 8) void kResidualLds(const double *__restrict__ p, const float *__restrict__ e,
    const double *__restrict__ rhs, Geo g, Coef<double> cf, double *__restrict__ pn, float *__restrict__ r32,
    float *__restrict__ eZero, double *partials, int gridGx, int gridGy, int gridGz,
    const int *__restrict__ stop = nullptr, ResidualRestrict ro = ResidualRestrict())
{
  if (stop && *stop) return;
  constexpr int NT = GTX * RTY, HW = GTX + 2, HN = HW * (RTY + 2);
  static_assert(HN <= 2 * NT, "two ring elements per thread at most");
  __shared__ double lds[NT / 64];
  __shared__ double X[3][RTY + 2][HW];
  __shared__ double sR[RTY][GTX], sB[RTY][GTX]; /* this plane's r and rhs, for the restriction */
  int bx, by, bz;
  if (!xcdBlock(gridGx, gridGy, gridGz, bx, by, bz)) return;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + GTX * ty;
  const int i = bx * GTX + tx + 1, j = by * RTY + ty + 1;
  const int k0 = bz * NKC + 1, k1 = min(k0 + NKC, g.nk + 1);
  /* this thread's ring elements (two at most); positions past the padded box
   * load a clamped cell, used only by cells outside the interior */
  const int e1 = tid + NT;
  const bool has1 = e1 < HN;
  const int hi0 = tid % HW, hj0 = tid / HW, hi1 = e1 % HW, hj1 = has1 ? e1 / HW : 0;
  const long long o0 = (long long)min(by * RTY + hj0, g.nj + 1) * g.sj + min(bx * GTX + hi0, g.ni + 1);
  const long long o1 = (long long)min(by * RTY + hj1, g.nj + 1) * g.sj + min(bx * GTX + hi1, g.ni + 1);
  auto value = [&](long long q) {
    if constexpr (CORR) return p[q] + (double)e[q];
    else return p[q];
  };
  /* plane k + 2 is loaded into registers while plane k is computed and
   * stored to LDS at the next step, so no step waits on global memory */
  double r0 = 0.0, r1 = 0.0;
  auto fetch = [&](int k) {
    long long base = (long long)k * g.sk;
    r0 = value(base + o0);
    if (has1) r1 = value(base + o1);
  };
  auto store = [&](int slot) {
    X[slot][hj0][hi0] = r0;
    if (has1) X[slot][hj1][hi1] = r1;
  };
  fetch(k0 - 1);
  store((k0 - 1) % 3);
  fetch(k0);
  store(k0 % 3);
  fetch(k0 + 1);
  const bool in = i <= g.ni && j <= g.nj;
  /* rhs one plane ahead, issued with the p loads */
  double bNext = in ? rhs[g.at(i, j, k0)] : 0.0;
  double a0 = 0.0, a1 = 0.0;
  /* restriction: the thread of each 2 x 2 group's first cell (even tx, ty)
   * sums the group's 8 children over a plane pair */
  const bool rOwner = RESTR && ro.rc && (tx & 1) == 0 && (ty & 1) == 0;
  double accR = 0.0, accB = 0.0;
  for (int k = k0; k < k1; k++) {
    store((k + 1) % 3);
    if (k + 1 < k1) fetch(k + 2);
    const double bb = bNext;
    if (in && k + 1 < k1) bNext = rhs[g.at(i, j, k + 1)];
    __syncthreads();
    double rv = 0.0, bv = 0.0;
    if (in) {
      const int sm = (k + 2) % 3, sc = k % 3, sp = (k + 1) % 3;
      const long long c = g.at(i, j, k);
      double xc  = X[sc][ty + 1][tx + 1];
      double lap = (X[sc][ty + 1][tx + 2] - 2.0 * xc + X[sc][ty + 1][tx]) * cf.idx2 +
                   (X[sc][ty + 2][tx + 1] - 2.0 * xc + X[sc][ty][tx + 1]) * cf.idy2 +
                   (X[sp][ty + 1][tx + 1] - 2.0 * xc + X[sm][ty + 1][tx + 1]) * cf.idz2;
      double r  = bb - lap;
      if constexpr (CORR) {
        pn[c] = xc;
        faceGhosts(pn, g, i, j, k, c, xc);
      }
      if (r32) r32[c] = (float)r;
      if (eZero) {
        eZero[c] = 0.0f;
        faceGhosts(eZero, g, i, j, k, c, 0.0f);
      }
      a0 += r * r;
      a1 += bb * bb;
      if (RESTR) {
        rv = ro.roundR ? (double)(float)r : r;
        bv = bb;
      }
    }
    if (RESTR && ro.rc) {
      sR[ty][tx] = rv;
      sB[ty][tx] = bv;
    }
    __syncthreads();
    if (RESTR && rOwner) {
      if (k & 1) { /* first plane of the pair */
        accR = 0.0;
        accB = 0.0;
      }
      accR += sR[ty][tx];
      accR += sR[ty][tx + 1];
      accR += sR[ty + 1][tx];
      accR += sR[ty + 1][tx + 1];
      accB += sB[ty][tx];
      accB += sB[ty][tx + 1];
      accB += sB[ty + 1][tx];
      accB += sB[ty + 1][tx + 1];
      const int I = (i + 1) >> 1, J = (j + 1) >> 1, K = k >> 1;
      if (!(k & 1) && I <= ro.gc.ni && J <= ro.gc.nj && K <= ro.gc.nk) {
        const long long cc = ro.gc.at(I, J, K);
        ro.rc[cc] = (float)(0.125 * accR);
        if (ro.bc) ro.bc[cc] = (float)(0.125 * accB);
      }
    }
  }
  a0 = blockSum(a0, lds);
  a1 = blockSum(a1, lds);
  unsigned bid = bx + gridGx * (by + gridGy * bz);
  if (tid == 0) {
    partials[2 * bid]     = a0;
    partials[2 * bid + 1] = a1;
  }
}

/* out: device norm pair; outHost (optional): device address of a pinned,
 * host-coherent copy the host reads after the stream synchronises */
/* extra0 / extra0Host (optional): component 0 once more, e.g. the warm
 * start's copy of ||r0||^2 (the same partials, so no second fold) */
__global__ __launch_bounds__(256) void kFoldNorms(const double *partials, int nb, double *out, double *outHost,
    double *extra0 = nullptr, double *extra0Host = nullptr)
{
  __shared__ double lds[4];
  double s0, s1;
  foldSum2(partials, nb, lds, s0, s1);
  if (threadIdx.x == 0) {
    if (extra0) *extra0 = s0;
    if (extra0Host) *extra0Host = s0;
  }
  if (outHost && threadIdx.x == 0) {
    outHost[0] = s0;
    outHost[1] = s1;
    __threadfence_system();
  }
  if (threadIdx.x == 0) {
    out[0] = s0;
    out[1] = s1;
  }
}

/* Cycle control of speculatively enqueued cycles: the host enqueues cycles
 * ahead of knowing whether the previous one converged; every kernel of a
 * cycle returns at once when `stop` is set, and the cycle's last kernel
 * (kFoldCycle) sets it with the host loop's own test. */
struct CycleCtl {
  int stop;      /* no further cycle runs */
  int executed;  /* cycles that ran */
  int maxCycles;
  int pad;
  double tol;
  double r2, b2; /* ||r||^2, ||rhs||^2 after the last cycle that ran */
};

/* hostRing: device address of CTL_RING pinned, host-coherent CycleCtl
 * slots. A cycle that runs writes its result into slot (executed - 1) %
 * CTL_RING, the slot of its own cycle index, which the host reads after the
 * cycle's event. The slot comes from the device count, not from a launch
 * argument, so a replayed captured cycle writes the right one. Skipped
 * cycles write nothing: the host stops reading at the cycle that set stop. */
__global__ __launch_bounds__(256) void kFoldCycle(const double *partials, int nb, double *norm2, CycleCtl *ctl,
    CycleCtl *hostRing)
{
  __shared__ double lds[4];
  if (ctl->stop) return;
  double s0, s1;
  foldSum2(partials, nb, lds, s0, s1);
  if (threadIdx.x == 0) {
    norm2[0]      = s0;
    norm2[1]      = s1;
    ctl->r2       = s0;
    ctl->b2       = s1;
    int e         = ctl->executed + 1;
    ctl->executed = e;
    double rn = sqrt(s0), nb2 = sqrt(s1);
    ctl->stop = (nb2 > 0.0 && rn / nb2 > ctl->tol && e < ctl->maxCycles) ? 0 : 1;
    hostRing[(e - 1) % CTL_RING] = *ctl;
    __threadfence_system();
  }
}

/* sum of rhs over the interior -> norm2[2] */
__global__ __launch_bounds__(256) void kInteriorSumRhs(const double *__restrict__ rhs, Geo g, double *partials)
{
  __shared__ double lds[4];
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  double acc = 0.0;
  if (i <= g.ni && j <= g.nj)
    for (int k = 1; k <= g.nk; k++) acc += rhs[g.at(i, j, k)];
  acc = blockSum(acc, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[blockIdx.x + gridDim.x * blockIdx.y] = acc;
}

__global__ __launch_bounds__(256) void kFoldRhsSum(const double *partials, int nb, DevScalars *sc)
{
  __shared__ double lds[4];
  double s = foldSum(partials, nb, 1, 0, lds);
  if (threadIdx.x == 0) sc->norm2[2] = s;
}

__global__ __launch_bounds__(256) void kSubtractRhsMean(double *__restrict__ rhs, Geo g,
    const DevScalars *sc, double count)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  int k = blockIdx.z + 1;
  if (i > g.ni || j > g.nj) return;
  double m = sc->norm2[2] / count;
  long long c = g.at(i, j, k);
  rhs[c]      = rhs[c] - m;
}

/* ---- host orchestration -------------------------------------------------- */
static dim3 cellGrid(const Geo &g) { return dim3((g.ni + GTX - 1) / GTX, (g.nj + GTY - 1) / GTY, g.nk); }
static dim3 sweepGrid(const Geo &g)
{
  return dim3(((g.ni + 1) / 2 + GTX - 1) / GTX, (g.nj + GTY - 1) / GTY, (g.nk + GKCH - 1) / GKCH);
}
static dim3 normGrid(const Geo &g)
{
  return dim3((g.ni + GTX - 1) / GTX, (g.nj + RTY - 1) / RTY, (g.nk + NKC - 1) / NKC);
}

/* The warm start's input pass: r0 = rhs - A p (FP64) and rhs averaged over
 * the 8 children into the next level's rc, bc (FP32), and per-workgroup
 * partials of r0^2, rhs^2 (slots 2 b, 2 b + 1). Returns the partial count.
 * Extents must be even. */
int launchInputRestrict(const double *p, const double *rhs, const Geo &g, double idx2, double idy2, double idz2,
    float *rc, float *bc, const Geo &gc, double *partials, float *r32, float *eZero)
{
  const Coef<double> cf = {idx2, idy2, idz2, 0.0};
  ResidualRestrict ro;
  ro.rc   = rc;
  ro.bc   = bc;
  ro.gc   = gc;
  dim3 ng = normGrid(g);
  kResidualLds<false, true><<<xcdGrid(ng), dim3(GTX, RTY), 0, nusifStream()>>>(p, nullptr, rhs, g, cf, nullptr,
      r32, eZero, partials, ng.x, ng.y, ng.z, nullptr, ro);
  return (int)(ng.x * ng.y * ng.z);
}

void launchFoldNorms(const double *partials, int nb, double *out, double *outHost, double *extra0,
    double *extra0Host)
{
  kFoldNorms<<<1, 256, 0, nusifStream()>>>(partials, nb, out, outHost, extra0, extra0Host);
}

int inputRestrictPartials(const Geo &g)
{
  dim3 ng = normGrid(g);
  return (int)(ng.x * ng.y * ng.z);
}

/* x_f = P x_c (trilinear, cell-centred, face ghosts written), FP32 */
void launchProlongSet32(float *xf, const Geo &gf, const float *xc, const Geo &gc)
{
  dim3 pg((gf.ni + GTX - 1) / GTX, (gf.nj + GTY - 1) / GTY, (gc.nk + PKC - 1) / PKC);
  kProlongMarch<float, true><<<xcdGrid(pg), dim3(GTX, GTY), 0, nusifStream()>>>(xf, gf, xc, gc, pg.x, pg.y, pg.z,
      nullptr);
}

template <typename T> static void fill(T *x, long long n, T v)
{
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  kFill<T><<<(unsigned)b, 256, 0, nusifStream()>>>(x, n, v);
}

/* Timers cover the finest level only ("..._l0"), whose algorithmic bytes
 * nusifKernelBytes reports. Each sweep swaps x and xAlt. */
/* `zeroStart`: x is zero at the first sweep (for the fused sweep it is
 * taken as zero, so the caller may skip zeroing it) */
template <typename T> static void smooth(GmgState *G, LevelT<T> &L, int sweeps, bool fine, bool zeroStart)
{
  if (!fusedLevel(L)) { /* small levels: two colour passes */
    dim3 grd = sweepGrid(L.g);
    for (int s = 0; s < sweeps; s++) {
      kSmooth<T><<<grd, dim3(GTX, GTY), 0, nusifStream()>>>(L.x, L.b, L.g, L.cf, 1, 0, gCycleStop);
      kSmooth<T><<<grd, dim3(GTX, GTY), 0, nusifStream()>>>(L.x, L.b, L.g, L.cf, 0, 1, gCycleStop);
    }
    return;
  }
  const SweepShape shape = sweepShape<T>(L.g, G->tileRows);
  for (int s = 0; s < sweeps; s++) {
    launchSweep(L, shape, fine, zeroStart && s == 0);
    T *t   = L.x;
    L.x    = L.xAlt;
    L.xAlt = t;
  }
}

static const double *tables(GmgState *G, double *) { return G->tab64; }
static const float *tables(GmgState *G, float *) { return G->tab32; }

template <typename T> static void coarseSolve(GmgState *G, LevelT<T> &L)
{
  size_t lds = 2 * (size_t)L.g.ni * L.g.nj * L.g.nk * sizeof(T);
  kCoarseDct<T><<<1, 1024, lds, nusifStream()>>>(L.x, L.b, L.g, tables(G, (T *)nullptr), G->dct, gCycleStop);
}

/* DCT-II tables of the coarsest level (closed forms; host libm) */
static void buildDctTables(GmgState *G, const Geo &g, double hx, double hy, double hz)
{
  G->dct = dctLayout(g.ni, g.nj, g.nk);
  std::vector<double> t((size_t)G->dct.total);
  const int n[3] = {g.ni, g.nj, g.nk};
  const double h[3] = {hx, hy, hz};
  const int oC[3] = {G->dct.oCi, G->dct.oCj, G->dct.oCk};
  const int oD[3] = {G->dct.oDi, G->dct.oDj, G->dct.oDk};
  const int oL[3] = {G->dct.oLi, G->dct.oLj, G->dct.oLk};
  for (int d = 0; d < 3; d++) {
    int nn = n[d];
    for (int m = 0; m < nn; m++) {
      double sm = (m == 0 ? 1.0 : 2.0) / nn;
      for (int i = 0; i < nn; i++) {
        double cv = cos(M_PI * m * (i + 0.5) / nn);
        t[(size_t)oC[d] + (size_t)m * nn + i] = cv;
        t[(size_t)oD[d] + (size_t)m * nn + i] = sm * cv;
      }
      double sn = sin(M_PI * m / (2.0 * nn));
      t[(size_t)oL[d] + m] = -4.0 / (h[d] * h[d]) * sn * sn;
    }
  }
  std::vector<float> f(t.size());
  for (size_t q = 0; q < t.size(); q++) f[q] = (float)t[q];
  HIP_CHECK(hipMalloc((void **)&G->tab64, t.size() * sizeof(double)));
  HIP_CHECK(hipMalloc((void **)&G->tab32, f.size() * sizeof(float)));
  HIP_CHECK(hipMemcpy(G->tab64, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_CHECK(hipMemcpy(G->tab32, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
}

/* dynamic LDS a workgroup may take on this device (64 KiB if the query
 * fails) */
static size_t tailLdsCap()
{
  static size_t cap = 0;
  if (!cap) {
    int dev = 0, v = 0;
    cap = 64 * 1024;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      cap = (size_t)v;
  }
  return cap;
}

/* levels G->tail .. nl-1 in one workgroup */
template <typename T> static void tailCycle(GmgState *G, LevelT<T> *L)
{
  TailArgs<T> a;
  memset(&a, 0, sizeof(a));
  a.n   = G->nl - G->tail;
  a.nu1 = G->nu1;
  a.nu2 = G->nu2;
  for (int q = 0; q < a.n; q++) {
    a.x[q]  = L[G->tail + q].x;
    a.b[q]  = L[G->tail + q].b;
    a.g[q]  = L[G->tail + q].g;
    a.cf[q] = L[G->tail + q].cf;
  }
  a.tab  = tables(G, (T *)nullptr);
  a.dct  = G->dct;
  a.stop = gCycleStop;
  const Geo &gc = L[G->nl - 1].g;
  const size_t work = 2 * (size_t)gc.ni * gc.nj * gc.nk * sizeof(T);
  /* LDS residency: x and b of every tail level, or x only, when it fits
   * what a workgroup may take on this device; else none. The DCT tables go
   * to LDS too when there is room */
  auto align16 = [](size_t v) { return (v + 15) & ~(size_t)15; };
  size_t xs = 0, bs = 0;
  for (int q = 0; q < a.n; q++) {
    a.xo[q] = (unsigned)xs;
    xs += align16((size_t)a.g[q].n * sizeof(T));
  }
  for (int q = 0; q < a.n; q++) {
    a.bo[q] = (unsigned)(xs + bs);
    bs += align16((size_t)a.g[q].n * sizeof(T));
  }
  const size_t cap  = tailLdsCap();
  const size_t tabs = align16((size_t)G->dct.total * sizeof(T));
  if (xs + bs + work <= cap) {
    a.wo     = (unsigned)(xs + bs);
    a.to     = (unsigned)(xs + bs + work);
    a.tabLds = xs + bs + work + tabs <= cap;
    kTailCycle<T, 2><<<1, TAIL_THREADS, xs + bs + work + (a.tabLds ? tabs : 0), nusifStream()>>>(a);
  } else if (xs + work <= cap) {
    a.wo     = (unsigned)xs;
    a.to     = (unsigned)(xs + work);
    a.tabLds = xs + work + tabs <= cap;
    kTailCycle<T, 1><<<1, TAIL_THREADS, xs + work + (a.tabLds ? tabs : 0), nusifStream()>>>(a);
  } else {
    a.wo = 0;
    kTailCycle<T, 0><<<1, TAIL_THREADS, work, nusifStream()>>>(a);
  }
  HIP_CHECK(hipGetLastError());
}

/* `cont` (mixed mode, finest level): the cycle continues the correction e
 * of the previous cycle on the same defect instead of starting from 0 */
template <typename T> static void vcycle(GmgState *G, LevelT<T> *L, int l, bool cont = false)
{
  if (l == G->tail) {
    tailCycle(G, L);
    return;
  }
  if (l == G->nl - 1) {
    coarseSolve(G, L[l]);
    return;
  }
  bool fine = l == 0;
  /* the coarse levels start from x = 0 (zeroed by the restriction); the
   * finest one in mixed mode too (the error equation) unless the cycle
   * continues e, FP64 mode smooths p */
  bool fromSet = false;
  if constexpr (std::is_same<T, float>::value)
    if (fine && G->setXc) { /* the first sweep from P setXc: the prolongation's pass is not made */
      LevelT<float> S = L[1];
      S.x = const_cast<float *>(G->setXc);
      S.g = G->setGc;
      launchSweepPro<float, 2>(L[0], S, sweepShape<float>(L[0].g, G->tileRows), true);
      float *t   = L[0].x;
      L[0].x     = L[0].xAlt;
      L[0].xAlt  = t;
      G->setXc   = nullptr;
      fromSet    = true;
      smooth(G, L[l], G->nu1 - 1, fine, false);
    }
  if (!fromSet) smooth(G, L[l], G->nu1, fine, l > 0 || (G->mixed && !cont));
  LevelT<T> &C = L[l + 1];
  if (fine)
    NUSIF_TIMED(sizeof(T) == 8 ? "gmg_restrict64_l0" : "gmg_restrict32_l0",
        (kResidualRestrict<T><<<xcdGrid(cellGrid(C.g)), dim3(GTX, GTY), 0, nusifStream()>>>(
            L[l].x, L[l].b, L[l].g, L[l].cf, C.b, C.x, C.g, cellGrid(C.g).x, cellGrid(C.g).y, gCycleStop)));
  else
    kResidualRestrict<T><<<xcdGrid(cellGrid(C.g)), dim3(GTX, GTY), 0, nusifStream()>>>(
        L[l].x, L[l].b, L[l].g, L[l].cf, C.b, C.x, C.g, cellGrid(C.g).x, cellGrid(C.g).y, gCycleStop);
  vcycle(G, L, l + 1);
  if (G->fuseProlong && G->nu2 > 0 && fusedLevel(L[l])) {
    /* the prolongation inside the first post-smoothing sweep */
    launchSweepPro<T, 1>(L[l], C, sweepShape<T>(L[l].g, G->tileRows), fine);
    T *t      = L[l].x;
    L[l].x    = L[l].xAlt;
    L[l].xAlt = t;
    smooth(G, L[l], G->nu2 - 1, fine, false);
    return;
  }
  dim3 pg((L[l].g.ni + GTX - 1) / GTX, (L[l].g.nj + GTY - 1) / GTY, (C.g.nk + PKC - 1) / PKC);
  if (fine)
    NUSIF_TIMED(sizeof(T) == 8 ? "gmg_prolong64_l0" : "gmg_prolong32_l0",
        (kProlongMarch<T><<<xcdGrid(pg), dim3(GTX, GTY), 0, nusifStream()>>>(
            L[l].x, L[l].g, C.x, C.g, pg.x, pg.y, pg.z, gCycleStop)));
  else
    kProlongMarch<T><<<xcdGrid(pg), dim3(GTX, GTY), 0, nusifStream()>>>(
        L[l].x, L[l].g, C.x, C.g, pg.x, pg.y, pg.z, gCycleStop);
  smooth(G, L[l], G->nu2, fine, false);
}

template <typename T> static T *devAlloc(long long n)
{
  T *a = nullptr;
  HIP_CHECK(hipMalloc((void **)&a, (size_t)n * sizeof(T)));
  return a;
}

GmgState *gmgCreate(NusifSolverCtx *sc, const Parameter *prm)
{
  NusifCtx *c  = sc->dctx;
  const Grid *gr = &c->owner->grid;
  GmgState *G  = new GmgState();
  G->dct       = DctTables();
  G->tab64     = nullptr;
  G->tab32     = nullptr;
  G->nl        = 0;
  for (int l = 0; l < MAXLEV; l++) {
    G->L64[l] = LevelT<double>();
    G->L32[l] = LevelT<float>();
  }
  G->useGraphs = c->opt.mgGraph;
  G->graphCache = c->opt.mgGraphCache;
  G->speculate = c->opt.mgSpeculate;
  G->inner     = c->opt.mgInner;
  G->nextGroup = 1;
  G->lastGroup = 0;
  G->groupCap  = 0;
  G->groupHeld = 1;
  G->predicted = 1;
  HIP_CHECK(hipMalloc((void **)&G->dCtl, sizeof(CycleCtl)));
  HIP_CHECK(hipHostMalloc((void **)&G->hRing, CTL_RING * sizeof(CycleCtl), hipHostMallocCoherent));
  HIP_CHECK(hipHostGetDevicePointer((void **)&G->hRingDev, G->hRing, 0));
  HIP_CHECK(hipHostMalloc((void **)&G->hInit, sizeof(CycleCtl), hipHostMallocDefault));
  for (int q = 0; q < CTL_RING; q++) HIP_CHECK(hipEventCreateWithFlags(&G->ev[q], hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&G->evInput, hipEventDisableTiming));
  G->tileRows  = c->opt.mgTileRows;
  G->fuseProlong = c->opt.mgFuseProlong;
  G->setXc       = nullptr;
  G->nu1       = prm->presmooth;
  G->nu2       = prm->postsmooth;
  G->maxCycles = c->opt.mgMaxCycles;
  G->project   = c->opt.mgProject;
  G->mixed     = c->opt.mixed;
  G->tol       = c->opt.mgTol;
  G->omega     = c->opt.mgOmega;
  /* levels: halve while every extent is even and the coarse one stays >= 2 */
  int want = prm->levels < 1 ? 1 : prm->levels;
  if (want > MAXLEV) want = MAXLEV;
  int ni = c->geo.ni, nj = c->geo.nj, nk = c->geo.nk, nl = 1;
  while (nl < want && ni % 2 == 0 && nj % 2 == 0 && nk % 2 == 0 && ni >= 4 && nj >= 4 && nk >= 4) {
    ni /= 2;
    nj /= 2;
    nk /= 2;
    nl++;
  }
  /* the coarsest level must fit the LDS solve; add levels if it does not */
  while ((long long)ni * nj * nk > COARSE_MAX_CELLS) {
    if (ni % 2 || nj % 2 || nk % 2 || ni < 4 || nj < 4 || nk < 4 || nl >= MAXLEV)
      NUSIF_FATAL("gmg: grid %dx%dx%d cannot be coarsened to <= %d cells "
                  "(extents must stay even)", c->geo.ni, c->geo.nj, c->geo.nk, COARSE_MAX_CELLS);
    ni /= 2;
    nj /= 2;
    nk /= 2;
    nl++;
  }
  G->nl  = nl;
  G->cf0 = coefFor<double>(gr->dx, gr->dy, gr->dz, G->omega);
  /* the tail starts at the first level below the finest that is small
   * enough, when it covers at least one level above the coarsest */
  G->tail = -1;
  {
    int a = c->geo.ni, b2 = c->geo.nj, c3 = c->geo.nk;
    for (int l = 1; l < nl; l++) {
      a /= 2;
      b2 /= 2;
      c3 /= 2;
      if ((long long)a * b2 * c3 <= TAIL_MAX_CELLS) {
        if (nl - l >= 2 && nl - l <= MAXTAIL) G->tail = l;
        break;
      }
    }
    if (!c->opt.mgTail) G->tail = -1;
  }
  double dx = gr->dx, dy = gr->dy, dz = gr->dz;
  ni = c->geo.ni;
  nj = c->geo.nj;
  nk = c->geo.nk;
  for (int l = 0; l < nl; l++) {
    Geo g = makeGeo(ni, nj, nk);
    if (G->mixed) {
      LevelT<float> &L = G->L32[l];
      L.g              = g;
      L.x              = devAlloc<float>(g.n);
      L.xAlt           = devAlloc<float>(g.n);
      L.b              = devAlloc<float>(g.n);
      L.cf             = coefFor<float>(dx, dy, dz, G->omega);
      fill<float>(L.x, g.n, 0.0f);
      fill<float>(L.xAlt, g.n, 0.0f);
      fill<float>(L.b, g.n, 0.0f);
    } else {
      LevelT<double> &L = G->L64[l];
      L.g               = g;
      L.cf              = coefFor<double>(dx, dy, dz, G->omega);
      if (l > 0) {
        L.x    = devAlloc<double>(g.n);
        L.xAlt = devAlloc<double>(g.n);
        L.b    = devAlloc<double>(g.n);
        fill<double>(L.x, g.n, 0.0);
        fill<double>(L.xAlt, g.n, 0.0);
        fill<double>(L.b, g.n, 0.0);
      }
    }
    if (l == nl - 1) buildDctTables(G, g, dx, dy, dz);
    ni /= 2;
    nj /= 2;
    nk /= 2;
    dx *= 2;
    dy *= 2;
    dz *= 2;
  }
  dim3 ng = normGrid(c->geo);
  if ((size_t)2 * ng.x * ng.y * ng.z > c->nPartials) NUSIF_FATAL("partials buffer too small");
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  return G;
}

void gmgFree(GmgState *G)
{
  /* no graph is destroyed, and no buffer freed, while a launch may be queued */
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  HIP_CHECK(hipFree(G->dCtl));
  HIP_CHECK(hipHostFree(G->hRing));
  HIP_CHECK(hipHostFree(G->hInit));
  for (int q = 0; q < CTL_RING; q++) HIP_CHECK(hipEventDestroy(G->ev[q]));
  HIP_CHECK(hipEventDestroy(G->evInput));
  for (CycleGraph &e : G->graphs) HIP_CHECK(hipGraphExecDestroy(e.exec));
  G->graphs.clear();
  if (G->tab64) HIP_CHECK(hipFree(G->tab64));
  if (G->tab32) HIP_CHECK(hipFree(G->tab32));
  for (int l = 0; l < G->nl; l++) {
    if (G->L32[l].x) HIP_CHECK(hipFree(G->L32[l].x));
    if (G->L32[l].xAlt) HIP_CHECK(hipFree(G->L32[l].xAlt));
    if (G->L32[l].b) HIP_CHECK(hipFree(G->L32[l].b));
    if (l > 0 && G->L64[l].xAlt) HIP_CHECK(hipFree(G->L64[l].xAlt));
    if (l > 0 && G->L64[l].x) HIP_CHECK(hipFree(G->L64[l].x));
    if (l > 0 && G->L64[l].b) HIP_CHECK(hipFree(G->L64[l].b));
  }
  delete G;
}

/* ---- cycle replay through HIP graphs ---------------------------------------
 * A cycle is ~25 launches, many of them a few microseconds long on the
 * coarse levels; launched one by one, the host's launch rate sets the pace
 * there. The cycle body is therefore captured once per buffer arrangement
 * (the fused sweeps and the defect correction ping-pong their buffers, so a
 * cycle alternates between a few arrangements) and replayed with one
 * hipGraphLaunch. The body contains kernels only (no memset or copy nodes).
 * Timers put events around launches, so cycles run eagerly while any timer
 * is enabled. At most graphCache graphs are kept; the oldest is destroyed
 * to make room, only after the stream has drained, so no launch of it is
 * still queued or running when it goes (HIP does not document destroying an
 * executable graph with launches in flight as safe). */

static std::vector<void *> cycleState(NusifCtx *c, GmgState *G)
{
  std::vector<void *> v;
  v.push_back(c->p);
  v.push_back(c->pAlt);
  for (int l = 0; l < G->nl; l++) {
    if (G->mixed) {
      v.push_back(G->L32[l].x);
      v.push_back(G->L32[l].xAlt);
    } else {
      v.push_back(G->L64[l].x);
      v.push_back(G->L64[l].xAlt);
    }
  }
  /* the rhs buffer (computeRHS may swap it): part of a captured graph's
   * arguments, so part of its key; setCycleState does not read it */
  v.push_back(c->rhs);
  return v;
}

static void setCycleState(NusifCtx *c, GmgState *G, const std::vector<void *> &v)
{
  size_t q = 0;
  c->p    = (double *)v[q++];
  c->pAlt = (double *)v[q++];
  for (int l = 0; l < G->nl; l++) {
    if (G->mixed) {
      G->L32[l].x    = (float *)v[q++];
      G->L32[l].xAlt = (float *)v[q++];
    } else {
      G->L64[l].x    = (double *)v[q++];
      G->L64[l].xAlt = (double *)v[q++];
    }
  }
}

extern bool nusifTimersActive();

/* enqueues body() (one cycle group, `key` = its cycle count), from a
 * captured graph when graphs are on: the graphs are keyed by the buffer
 * arrangement before the group and its cycle count */
template <typename Body> static void runCycle(NusifCtx *c, GmgState *G, int key, Body body)
{
  if (!G->useGraphs || nusifTimersActive()) {
    body();
    return;
  }
  std::vector<void *> before = cycleState(c, G);
  before.push_back((void *)(intptr_t)key);
  for (CycleGraph &e : G->graphs)
    if (e.before == before) {
      HIP_CHECK(hipGraphLaunch(e.exec, nusifStream()));
      setCycleState(c, G, e.after);
      return;
    }
  if ((int)G->graphs.size() >= G->graphCache) {
    HIP_CHECK(hipStreamSynchronize(nusifStream()));
    HIP_CHECK(hipGraphExecDestroy(G->graphs.front().exec));
    G->graphs.erase(G->graphs.begin());
  }
  hipGraph_t graph;
  HIP_CHECK(hipStreamBeginCapture(nusifStream(), hipStreamCaptureModeThreadLocal));
  body();
  HIP_CHECK(hipStreamEndCapture(nusifStream(), &graph));
  CycleGraph e;
  e.before = before;
  e.after  = cycleState(c, G); /* (no key: setCycleState reads the buffers only) */
  HIP_CHECK(hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0));
  HIP_CHECK(hipGraphDestroy(graph));
  HIP_CHECK(hipGraphLaunch(e.exec, nusifStream()));
  G->graphs.push_back(e);
}

/* Mixed mode: whether the finest level's correction e must be zeroed in
 * memory before a cycle. Not when the cycle's first pre-smoothing sweep is
 * a fused sweep, which takes the old values as zero (ZX); a two-pass sweep,
 * or a cycle without pre-smoothing, reads them. */
static bool l0NeedsZero(const GmgState *G)
{
  return G->mixed && G->nl > 1 && (!fusedLevel(G->L32[0]) || G->nu1 == 0);
}

/* Residual of p on the finest level (FP64): partial sums of r^2 and rhs^2
 * folded into the device pair out[0..1]; optionally r as FP32 (r32) and an
 * FP32 array zeroed (eZero). No host synchronisation. */
static void launchResidual(NusifCtx *c, GmgState *G, const double *p, const double *rhs, float *r32,
    float *eZero, double *out)
{
  dim3 ng = normGrid(c->geo);
  NUSIF_TIMED("gmg_residual64", (kResidualLds<false><<<xcdGrid(ng), dim3(GTX, RTY), 0, nusifStream()>>>(
                                    p, nullptr, rhs, c->geo, G->cf0, nullptr, r32, eZero, c->partials, ng.x,
                                    ng.y, ng.z)));
  kFoldNorms<<<1, 256, 0, nusifStream()>>>(c->partials, (int)(ng.x * ng.y * ng.z), out,
      out == c->scal->norm2 ? c->scalHostDev->norm2 : (out == c->scal->norm2 + 2 ? c->scalHostDev->norm2 + 2 : nullptr));
}

/* read norm2[0..1] (||r||^2, ||rhs||^2), which the fold kernels also write
 * into the pinned mirror: returns ||r||, sets *nb */
static double readNorms(NusifCtx *c, double *nb)
{
  HIP_CHECK(hipStreamSynchronize(nusifStream()));
  *nb = sqrt(c->scalHost->norm2[1]);
  return sqrt(c->scalHost->norm2[0]);
}

/* Runs V-cycles until ||rhs - A p|| / ||rhs|| <= tol or maxCycles, from the
 * residual norms (rel, nb) of the starting p; updates them and returns the
 * cycle count. `body(m)` enqueues m cycles that end with the residual
 * partials of the new iterate (normGrid layout in c->partials).
 *
 * The first stop test of a solve comes after a group of G->inner cycles,
 * the later ones after every cycle. Skipped tests save the host round trip
 * and, in mixed mode, the FP64 correction + residual pass between the
 * cycles. G->inner = 0 sizes the group from the previous solve
 * (nextGroupAfter): a group that did not converge gives the exact count the
 * solve needed (in mixed precision at most the length the FP32 bound on one
 * correction leaves useful); one that converged shrinks by one when its own
 * contraction says one cycle fewer would have reached the tolerance. A solve that needs
 * fewer cycles than its group still runs the whole group.
 *
 * Speculative mode: the host enqueues cycles before it knows whether the
 * previous one converged (as many as the previous solve needed, then one
 * at a time), the device decides (kFoldCycle sets the stop flag every
 * cycle kernel checks first), and the host reads the decision through
 * pinned memory per cycle. Cycles enqueued past the stop do nothing. The
 * kernels that run, and so the result, are those of the one-at-a-time loop;
 * the buffer pointers are set back to their arrangement after the cycles
 * that ran. */
/* the next solve's first group under mg_inner 0, from this solve's: group
 * `first` took the relative residual from rel0 to relFirst; the solve ended
 * after `cycles` */
static int nextGroupAfter(GmgState *G, int first, int cycles, double rel0, double relFirst, double relEnd)
{
  int next = first;
  if (cycles > first) {
    next = cycles;
    /* Mixed precision: the group's cycles solve for one correction of one
     * FP64 defect in FP32, so the reduction a group can reach is bounded
     * (FP32 rounding of the correction, amplified by the operator: about
     * eps32 (n/pi)^2 for n cells along the longest side, a few percent at
     * 768 cells). A group that reduced the residual less per cycle than the
     * corrected single cycles after it had reached that bound: a longer one
     * buys nothing. Its useful length is the single cycles that give the
     * same reduction; it caps every later group (without the cap, each
     * solve would lengthen the group by the cycles the bound forced after
     * it, without end). */
    if (G->mixed && first > 1 && rel0 > 0.0 && relFirst > 0.0 && relEnd > 0.0 && relEnd < relFirst) {
      const double rhoGroup = pow(relFirst / rel0, 1.0 / first);
      const double rhoOne   = pow(relEnd / relFirst, 1.0 / (cycles - first));
      if (rhoGroup > rhoOne) {
        int m = (int)ceil(log(relFirst / rel0) / log(rhoOne));
        G->groupCap = std::max(1, std::min(m, first - 1));
      }
    }
    if (G->groupCap > 0) next = std::min(next, G->groupCap);
  } else if (first > 1 && rel0 > 0.0 && relFirst > 0.0) {
    /* mean contraction per cycle over the group; the residual one cycle
     * earlier was about relFirst / rho */
    double rho = pow(relFirst / rel0, 1.0 / first);
    if (rho < 1.0 && relFirst / rho <= G->tol) next = first - 1;
  }
  return std::max(1, std::min(next, G->maxCycles));
}

template <typename Body>
static int runCycles(NusifCtx *c, GmgState *G, NusifSolverCtx *sc, double &rel, double &nb, Body body)
{
  dim3 ng        = normGrid(c->geo);
  const int nbk  = (int)(ng.x * ng.y * ng.z);
  hipStream_t st = nusifStream();
  int cycles     = 0;
  if (!G->speculate || nusifTimersActive()) {
    const int first   = G->inner > 0 ? G->inner : G->nextGroup;
    const double rel0 = rel;
    double relFirst   = 0.0;
    G->lastGroup      = 0;
    while (nb > 0.0 && rel > G->tol && cycles < G->maxCycles) {
      int m = cycles == 0 ? first : 1;
      if (m > G->maxCycles - cycles) m = G->maxCycles - cycles;
      if (cycles == 0) G->lastGroup = m;
      runCycle(c, G, m, [&] {
        body(m);
        kFoldNorms<<<1, 256, 0, st>>>(c->partials, nbk, c->scal->norm2, c->scalHostDev->norm2);
      });
      cycles += m;
      double rn;
      if (sc->afterFirst && sc->afterFirstState == 0 && G->groupHeld) {
        /* the group's p is in c->p (the bodies keep it there): queue the
         * caller's work, then wait for the group's fold alone (it writes the
         * norms into the pinned mirror), not for that work. Only while the
         * solves finish within their first group: after one that did not,
         * the work would likely have to run again after the later cycles. */
        HIP_CHECK(hipEventRecord(G->ev[0], st));
        sc->afterFirst(sc->afterFirstArg);
        sc->afterFirstState = 1;
        HIP_CHECK(hipEventSynchronize(G->ev[0]));
        nb = sqrt(c->scalHost->norm2[1]);
        rn = sqrt(c->scalHost->norm2[0]);
      } else {
        if (sc->afterFirst) sc->afterFirstState = 2;
        rn = readNorms(c, &nb);
      }
      rel       = rn / nb;
      if (cycles == m) relFirst = rel;
    }
    if (G->inner == 0 && cycles > 0) G->nextGroup = nextGroupAfter(G, G->lastGroup, cycles, rel0, relFirst, rel);
    G->groupHeld = cycles <= G->lastGroup ? 1 : 0;
    return cycles;
  }
  if (!(nb > 0.0 && rel > G->tol && G->maxCycles > 0)) return 0;
  *G->hInit = CycleCtl{0, 0, G->maxCycles, 0, G->tol, 0.0, 0.0};
  HIP_CHECK(hipMemcpyAsync(G->dCtl, G->hInit, sizeof(CycleCtl), hipMemcpyHostToDevice, st));
  gCycleStop = &G->dCtl->stop;
  std::vector<std::vector<void *>> states;
  const int pred = G->predicted > 0 ? G->predicted : 1;
  int enq = 0, chk = 0;
  CycleCtl last = *G->hInit;
  for (;;) {
    int target = pred > chk + 1 ? pred : chk + 1;
    while (enq < G->maxCycles && enq < target && enq < chk + CTL_RING) {
      states.push_back(cycleState(c, G));
      runCycle(c, G, 1, [&] {
        body(1);
        kFoldCycle<<<1, 256, 0, st>>>(c->partials, nbk, c->scal->norm2, G->dCtl, G->hRingDev);
      });
      HIP_CHECK(hipEventRecord(G->ev[enq % CTL_RING], st));
      enq++;
    }
    HIP_CHECK(hipEventSynchronize(G->ev[chk % CTL_RING]));
    last = G->hRing[chk % CTL_RING];
    if (last.executed != chk + 1) /* every cycle up to chk ran: the slot is cycle chk's */
      NUSIF_FATAL("gmg: cycle control slot of cycle %d holds count %d", chk, last.executed);
    chk++;
    if (last.stop) break;
    if (chk >= enq && enq >= G->maxCycles)
      NUSIF_FATAL("gmg: cycle control: %d cycles ran without the stop flag", enq);
  }
  gCycleStop = nullptr;
  states.push_back(cycleState(c, G));
  if (last.executed < 1 || last.executed >= (int)states.size())
    NUSIF_FATAL("gmg: %d cycles reported for %d enqueued", last.executed, enq);
  setCycleState(c, G, states[last.executed]);
  nb           = sqrt(last.b2);
  rel          = sqrt(last.r2) / nb;
  G->predicted = last.executed;
  G->lastGroup = 1;
  return last.executed;
}

/* fold the partials of a normGrid launch into norm2[0..1] and read them */
static double foldNorms(NusifCtx *c, double *nb)
{
  dim3 ng = normGrid(c->geo);
  kFoldNorms<<<1, 256, 0, nusifStream()>>>(c->partials, (int)(ng.x * ng.y * ng.z), c->scal->norm2,
      c->scalHostDev->norm2);
  return readNorms(c, nb);
}

/* residual norms on the finest level: returns ||r||, sets *nb = ||rhs||;
 * optionally writes r as FP32 (r32) and zeroes an FP32 array (eZero) */
static double residualNorm(NusifCtx *c, GmgState *G, double *p, const double *rhs, float *r32, double *nb,
    float *eZero = nullptr)
{
  launchResidual(c, G, p, rhs, r32, eZero, c->scal->norm2);
  return readNorms(c, nb);
}

double gmgSolve(NusifSolverCtx *sc, double *p, const double *rhsIn)
{
  NusifCtx *c  = sc->dctx;
  GmgState *G  = sc->gmg;
  const Geo &g = c->geo;
  double *rhs  = (double *)rhsIn;
  if (G->project) {
    dim3 g2((g.ni + GTX - 1) / GTX, (g.nj + GTY - 1) / GTY);
    kInteriorSumRhs<<<g2, dim3(GTX, GTY), 0, nusifStream()>>>(rhs, g, c->partials);
    kFoldRhsSum<<<1, 256, 0, nusifStream()>>>(c->partials, (int)(g2.x * g2.y), c->scal);
    kSubtractRhsMean<<<cellGrid(g), dim3(GTX, GTY), 0, nusifStream()>>>(
        rhs, g, c->scal, (double)g.ni * g.nj * g.nk);
  }
  /* the Neumann face ghosts of the starting p (the last solve's already
   * holds them) */
  if (!c->pGhostsOk || p != c->p) launchPressureGhosts(g, p, nusifStream());
  double nb = 0.0;
  int cycles = 0;
  double rel = 0.0;
  WarmStart *ws = sc->ws;
  /* warmstart_skip: after a solve of at most that many cycles this one
   * starts from the previous-step pressure without the network */
  if (ws && warmstartNet(ws) && c->opt.warmstartSkip > 0 && sc->lastIterations > 0 &&
      sc->lastIterations <= c->opt.warmstartSkip) {
    ws           = nullptr;
    sc->warmUsed = 0;
  }
  bool haveInitial = false; /* mixed: residual of c->p already in L0.b, its norms in rnInit/nb */
  double rnInit = 0.0;
  bool fromGuess = false; /* mixed: the first cycle starts from the network's correction in L0.x */
  bool decisionPending = false; /* mixed: the device's acceptance decision is read after the cycles */
  if (ws && G->mixed && warmstartNet(ws) && warmstartLevel(ws) > 0) {
    /* Mixed precision: the network's correction is the starting value of
     * the FP32 error equation A e = r0, not a new FP64 iterate. The input
     * pass (r0 = rhs - A p_prev restricted for the network) also writes r0
     * as the solve's first FP32 defect with its norms, and the acceptance
     * test runs on the coarse grids; an accepted correction is prolongated
     * straight into the first cycle's e. No fine-grid pass beyond the one
     * the solve makes without a network. (With the network on the finest
     * grid, small problems, the general path below runs.) */
    LevelT<float> &L0 = G->L32[0];
    if (warmstartPrepare(ws, p, rhs, G->cf0.idx2, G->cf0.idy2, G->cf0.idz2, L0.b,
            l0NeedsZero(G) ? L0.x : nullptr, c->scal->norm2, c->scalHostDev->norm2)) {
      /* the host waits for the input pass's norms only; the network, the
       * acceptance decision and the first cycles are queued behind them.
       * The decision is made on the device: a rejected correction is zeroed
       * there, so both branches start the first cycle from the prolongated
       * correction (e = 0 when rejected, the same values as the solve from
       * the previous pressure). */
      HIP_CHECK(hipEventRecord(G->evInput, nusifStream()));
      warmstartCaptureBegin(ws, p, g);
      warmstartForward(ws);
      warmstartCoarseAcceptLaunch(ws);
      warmstartDecideLaunch(ws, c->opt.warmstart == 2);
      /* the last step of the correction's prolongation goes into the
       * first cycle's first sweep where that sweep is the fused one */
      Geo g1;
      const float *c1 = G->fuseProlong && G->nu1 > 0 && fusedLevel(L0) ? warmstartProlongToLevel1(ws, &g1) : nullptr;
      if (c1 && g1.ni * 2 == L0.g.ni && g1.nj * 2 == L0.g.nj && g1.nk * 2 == L0.g.nk) {
        G->setXc = c1;
        G->setGc = g1;
      } else {
        warmstartProlongCorrection(ws, L0.x);
      }
      fromGuess = true;
      HIP_CHECK(hipEventSynchronize(G->evInput));
      const double nb0 = sqrt(c->scalHost->norm2[1]);
      const double rn0 = sqrt(c->scalHost->norm2[0]);
      sc->wsRes[0] = rn0;
      sc->wsRes[1] = -1.0; /* the guess's fine-grid residual is not formed */
      sc->wsTried++;
      decisionPending = true;
      haveInitial     = true;
      rnInit          = rn0;
      nb              = nb0;
    }
  }
  if (ws && !haveInitial) {
    /* the network input: r0 = rhs - A p_prev and rhs, restricted to the
     * network grid, and their norms */
    warmstartPrepare(ws, p, rhs, G->cf0.idx2, G->cf0.idy2, G->cf0.idz2);
    warmstartCaptureBegin(ws, p, g);
    if (warmstartNet(ws)) {
      const float *d = warmstartCorrection(ws);
      /* the guess p + d into c->pAlt and its residual in one pass, in mixed
       * mode straight into the first defect buffer (L0.b) with the first
       * correction zeroed, so an accepted guess costs no further residual
       * pass; then its restriction to the network grid, and one read-back
       * for the decision */
      float *rg  = G->mixed ? G->L32[0].b : warmstartResidualBuffer(ws);
      float *e0  = l0NeedsZero(G) ? G->L32[0].x : nullptr;
      dim3 ng    = normGrid(g);
      /* the same pass also averages the FP32 residual onto level 1 (the
       * first step of its restriction to the network grid) */
      ResidualRestrict ro;
      ro.rc     = warmstartLevelResidual(ws, 1, &ro.gc);
      ro.roundR = 1;
      NUSIF_TIMED("ws_guess_residual",
          (kResidualLds<true, true><<<xcdGrid(ng), dim3(GTX, RTY), 0, nusifStream()>>>(
              p, d, rhs, g, G->cf0, c->pAlt, rg, e0, c->partials, ng.x, ng.y, ng.z, nullptr, ro)));
      kFoldNorms<<<1, 256, 0, nusifStream()>>>(c->partials, (int)(ng.x * ng.y * ng.z), c->scal->norm2,
          c->scalHostDev->norm2);
      warmstartCoarseNormLaunch(ws, rg, ro.rc != nullptr);
      const double *wn = warmstartNormsAsync(ws);
      double nbg       = 0.0;
      double rng       = readNorms(c, &nbg);
      double rc0 = sqrt(warmstartReferenceNorm2(ws)), rcg = sqrt(wn[2]);
      /* the starting residuals are compared on the network grid (restricted
       * fine residuals): a correction interpolated from that grid carries
       * interpolation error whose fine-grid residual the first smoothing
       * sweep removes, and which would otherwise hide any gain */
      sc->wsRes[0] = sqrt(wn[3]);
      sc->wsRes[1] = rng;
      sc->wsRes[2] = rc0;
      sc->wsRes[3] = rcg;
      sc->warmUsed = (c->opt.warmstart == 2 || rcg < rc0) ? 1 : 0;
      sc->wsTried++;
      sc->wsTaken += sc->warmUsed;
      if (sc->warmUsed) { /* start from the network's guess */
        double *t = c->p;
        c->p      = c->pAlt;
        c->pAlt   = t;
        p         = c->p;
        if (G->mixed) {
          haveInitial = true;
          rnInit      = rng;
          nb          = nbg;
        }
      }
    }
  }
  if (!G->mixed) {
    LevelT<double> &L0 = G->L64[0];
    L0.x    = p;
    L0.xAlt = c->pAlt;
    L0.b    = rhs;
    double rn = residualNorm(c, G, L0.x, rhs, nullptr, &nb);
    rel       = nb > 0.0 ? rn / nb : 0.0;
    dim3 ng   = normGrid(g);
    cycles    = runCycles(c, G, sc, rel, nb, [&](int m) {
      for (int q = 0; q < m; q++) vcycle<double>(G, G->L64, 0);
      NUSIF_TIMED("gmg_residual64", (kResidualLds<false><<<xcdGrid(ng), dim3(GTX, RTY), 0, nusifStream()>>>(
                                        L0.x, nullptr, rhs, g, G->cf0, nullptr, nullptr, nullptr, c->partials,
                                        ng.x, ng.y, ng.z, gCycleStop)));
      if (L0.x != c->p) { /* keep the current iterate in c->p */
        c->pAlt = c->p;
        c->p    = L0.x;
      }
    });
    if (L0.x != c->p) { /* the result sits in the partner buffer */
      c->pAlt = c->p;
      c->p    = L0.x;
    }
  } else {
    /* r (FP32) into L0.b (and e = 0 into L0.x where the first sweep reads
     * it); each cycle then ends with one pass that applies e to p (into the
     * partner buffer) and forms the next r */
    LevelT<float> &L0 = G->L32[0];
    double rn = haveInitial ? rnInit : residualNorm(c, G, p, rhs, L0.b, &nb, l0NeedsZero(G) ? L0.x : nullptr);
    rel       = nb > 0.0 ? rn / nb : 0.0;
    dim3 ng   = normGrid(g);
    cycles    = runCycles(c, G, sc, rel, nb, [&](int m) {
      for (int q = 0; q < m; q++) vcycle<float>(G, G->L32, 0, q > 0 || fromGuess);
      fromGuess = false;
      NUSIF_TIMED("gmg_correct_residual",
          (kResidualLds<true><<<xcdGrid(ng), dim3(GTX, RTY), 0, nusifStream()>>>(c->p, L0.x, rhs, g, G->cf0,
              c->pAlt, L0.b, l0NeedsZero(G) ? L0.xAlt : nullptr, c->partials, ng.x, ng.y, ng.z, gCycleStop)));
      float *te  = L0.x; /* the next cycle's (zeroed) e */
      L0.x       = L0.xAlt;
      L0.xAlt    = te;
      double *tp = c->p; /* the corrected p is in c->pAlt */
      c->p       = c->pAlt;
      c->pAlt    = tp;
    });
    p = c->p;
    if (decisionPending) {
      /* the cycles ended with a host synchronisation; without one (no
       * cycle) the stream is drained here, and a prolongated correction no
       * cycle used is dropped */
      if (cycles == 0) HIP_CHECK(hipStreamSynchronize(nusifStream()));
      G->setXc         = nullptr;
      const double *wn = warmstartNormsAsync(ws);
      sc->wsRes[2]     = sqrt(warmstartReferenceNorm2(ws));
      sc->wsRes[3]     = sqrt(wn[2]);
      sc->warmUsed     = warmstartDecision(ws);
      sc->wsTaken += sc->warmUsed;
    }
  }
  if (ws) warmstartCaptureEnd(ws, c->p);
  /* every path above leaves c->p with its face ghosts written: the last
   * smoothing sweep (FP64) or correction pass (mixed) writes them, and a
   * solve that ran no cycle keeps the ghosts set at its start */
  c->pGhostsOk       = true;
  sc->lastIterations = cycles;
  sc->lastRel        = rel;
  return rel;
}



/* ---- C ABI: checks of the warm start's grid transfers ---------------------
 * Host arrays in, host arrays out (padded layouts), for tests. */
extern "C" void nusifRestrictResidualCheck(int ni, int nj, int nk, double dx, double dy, double dz,
    const double *p, const double *rhs, float *rc, float *bc, double *sumsq)
{
  if (ni % 2 || nj % 2 || nk % 2) NUSIF_FATAL("restriction check: extents must be even");
  const Geo g = makeGeo(ni, nj, nk), gc = makeGeo(ni / 2, nj / 2, nk / 2);
  const Coef<double> cf = coefFor<double>(dx, dy, dz, 1.0);
  const int nb = inputRestrictPartials(g);
  double *dp = devAlloc<double>(g.n), *dr = devAlloc<double>(g.n), *dpart = devAlloc<double>(2LL * nb);
  float *drc = devAlloc<float>(gc.n), *dbc = devAlloc<float>(gc.n);
  hipStream_t st = nusifStream();
  HIP_CHECK(hipMemcpyAsync(dp, p, g.n * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(dr, rhs, g.n * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemsetAsync(drc, 0, gc.n * sizeof(float), st));
  HIP_CHECK(hipMemsetAsync(dbc, 0, gc.n * sizeof(float), st));
  launchInputRestrict(dp, dr, g, cf.idx2, cf.idy2, cf.idz2, drc, dbc, gc, dpart);
  std::vector<double> part(2 * (size_t)nb);
  HIP_CHECK(hipMemcpyAsync(rc, drc, gc.n * sizeof(float), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(bc, dbc, gc.n * sizeof(float), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipMemcpyAsync(part.data(), dpart, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  double sum = 0.0;
  for (int q = 0; q < nb; q++) sum += part[2 * (size_t)q];
  *sumsq = sum;
  HIP_CHECK(hipFree(dp));
  HIP_CHECK(hipFree(dr));
  HIP_CHECK(hipFree(dpart));
  HIP_CHECK(hipFree(drc));
  HIP_CHECK(hipFree(dbc));
}

extern "C" void nusifProlongSetCheck(int nic, int njc, int nkc, const float *xc, float *xf)
{
  const Geo gc = makeGeo(nic, njc, nkc), gf = makeGeo(2 * nic, 2 * njc, 2 * nkc);
  float *dc = devAlloc<float>(gc.n), *df = devAlloc<float>(gf.n);
  hipStream_t st = nusifStream();
  HIP_CHECK(hipMemcpyAsync(dc, xc, gc.n * sizeof(float), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemsetAsync(df, 0, gf.n * sizeof(float), st));
  launchProlongSet32(df, gf, dc, gc);
  HIP_CHECK(hipMemcpyAsync(xf, df, gf.n * sizeof(float), hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  HIP_CHECK(hipFree(dc));
  HIP_CHECK(hipFree(df));
}

extern "C" int nusifSolverFirstGroup(Solver *s)
{
  NusifSolverCtx *sc = nusifSolverCtxOf(s);
  return sc->kind == SOLVER_GMG && sc->gmg ? sc->gmg->lastGroup : 0;
}

/* Duration of the finest level's red-black sweep (the dominant kernel of a
 * gmg solve): `reps` back-to-back launches on the solver's own level-0
 * buffers, HIP events on the library stream around the batch only, so the
 * per-launch event packets of the kernel timers do not inflate it. The
 * sweeps change only the level-0 work arrays (in mixed mode the correction
 * e, in FP64 mode p and its partner: use it after the steps it measures).
 * reps is rounded up to an even number, which leaves x / xAlt as found.
 * Returns microseconds per launch, or -1 when the solver is not gmg or the
 * finest level does not use the fused sweep. */
extern "C" double nusifTimeFinestSweep(Solver *s, int reps)
{
  NusifSolverCtx *sc = nusifSolverCtxOf(s);
  if (sc->kind != SOLVER_GMG || !sc->gmg) return -1.0;
  GmgState *G = sc->gmg;
  NusifCtx *c = sc->dctx;
  reps        = reps < 2 ? 2 : reps + (reps & 1);
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  float ms = 0.0f;
  auto run = [&](auto &L) {
    if (!fusedLevel(L)) return false;
    auto shape = sweepShape<typename std::remove_pointer<decltype(L.x)>::type>(L.g, G->tileRows);
    launchSweep(L, shape, false, false); /* warm */
    std::swap(L.x, L.xAlt);
    launchSweep(L, shape, false, false);
    std::swap(L.x, L.xAlt);
    HIP_CHECK(hipEventRecord(e0, nusifStream()));
    for (int q = 0; q < reps; q++) {
      launchSweep(L, shape, false, false);
      std::swap(L.x, L.xAlt);
    }
    HIP_CHECK(hipEventRecord(e1, nusifStream()));
    HIP_CHECK(hipEventSynchronize(e1));
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    return true;
  };
  bool ok;
  if (G->mixed) {
    ok = run(G->L32[0]);
  } else {
    LevelT<double> &L0 = G->L64[0];
    L0.x    = c->p;
    L0.xAlt = c->pAlt;
    L0.b    = c->rhs;
    ok      = run(L0);
  }
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  return ok ? 1e3 * ms / reps : -1.0;
}
