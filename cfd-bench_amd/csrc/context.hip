/* Device contexts, the library stream, option side table and kernel timers.
 *
 * The reference keeps its fields in host memory owned by Discretization
 * (reference src/discretization.c:85-94). Here every Discretization gets a
 * NusifCtx holding the HBM copies; it is found again from the struct address
 * or from any of its host array pointers (solve() only receives p and rhs,
 * reference src/solver.h:35). */
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "nusif_internal.h"

static hipStream_t gStream = nullptr;
static std::vector<NusifCtx *> gCtx;
static std::vector<NusifSolverCtx *> gSolvers;

hipStream_t nusifStream(void)
{
  if (!gStream) {
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (n < 1) NUSIF_FATAL("no GPU visible to the HIP runtime");
    /* default priority on purpose: every stream of this library shares it */
    HIP_CHECK(hipStreamCreateWithFlags(&gStream, hipStreamNonBlocking));
  }
  return gStream;
}

void nusifRegisterCtx(NusifCtx *c) { gCtx.push_back(c); }

void nusifUnregisterCtx(NusifCtx *c)
{
  for (size_t q = 0; q < gCtx.size(); q++)
    if (gCtx[q] == c) {
      gCtx.erase(gCtx.begin() + q);
      return;
    }
}

NusifCtx *nusifCtxOf(const Discretization *d)
{
  for (NusifCtx *c : gCtx)
    if (c->owner == d) return c;
  /* a copied struct still carries the same host arrays */
  for (NusifCtx *c : gCtx)
    if (c->host[0] == d->p && c->host[5] == d->u) return c;
  NUSIF_FATAL("Discretization %p was not initialised by this library", (const void *)d);
  return nullptr;
}

NusifCtx *nusifCtxOfHost(const double *hostPtr, int *fieldIndex)
{
  for (NusifCtx *c : gCtx)
    for (int q = 0; q < 8; q++)
      if (c->host[q] == hostPtr) {
        if (fieldIndex) *fieldIndex = q;
        return c;
      }
  return nullptr;
}

void nusifRegisterSolver(NusifSolverCtx *s) { gSolvers.push_back(s); }

void nusifUnregisterSolver(NusifSolverCtx *s)
{
  for (size_t q = 0; q < gSolvers.size(); q++)
    if (gSolvers[q] == s) {
      gSolvers.erase(gSolvers.begin() + q);
      return;
    }
}

NusifSolverCtx *nusifSolverCtxOf(const Solver *s)
{
  for (NusifSolverCtx *c : gSolvers)
    if (c->owner == s) return c;
  for (NusifSolverCtx *c : gSolvers)
    if (&c->dctx->owner->grid == s->grid) return c;
  NUSIF_FATAL("Solver %p was not initialised by this library", (const void *)s);
  return nullptr;
}

/* ---- options ------------------------------------------------------------ */
/* Options are kept per Parameter address until initDiscretization consumes
 * them. A consumed entry is not applied again: a later Parameter that reuses
 * the address starts from the defaults (initParameter also clears it). */
struct OptEntry {
  const Parameter *prm;
  bool consumed;
  NusifOptions opt;
};
static std::vector<OptEntry> gOpts;

void nusifDefaultOptions(NusifOptions *o)
{
  memset(o, 0, sizeof(*o));
  o->nsteps      = 0;
  o->solver      = SOLVER_DEFAULT;
  o->mixed       = 0;
  o->mgTol       = 1e-6;
  o->mgMaxCycles = 100;
  o->mgOmega     = 1.0;
  o->mgProject   = 0;
  o->mgTail      = 1;
  o->mgGraph     = 0;
  o->mgGraphCache = 64;
  o->mgTileRows  = 0;
  o->mgFuseProlong = 1;
  o->mgSpeculate = 0;
  o->mgInner     = 1;
  o->unetExtent  = 64;
  o->unetAcceptLevel = 2;
  o->warmstartSkip   = 0;
  o->fgRhs       = 1;
  o->rbFused     = 1;
  o->bcFused     = 1;
  o->lid         = 1.0;
  o->uInflow     = 2.0;
  o->inflowNoise = 0.0;
  o->seed        = 1;
  o->warmstart   = 0;
  o->verbose     = NUSIF_VERBOSE_DEFAULT;
}

static OptEntry *findEntry(const Parameter *p)
{
  for (OptEntry &e : gOpts)
    if (e.prm == p) return &e;
  return nullptr;
}

static NusifOptions *optionsSlot(const Parameter *p)
{
  OptEntry *e = findEntry(p);
  if (e && e->consumed) {
    nusifDefaultOptions(&e->opt);
    e->consumed = false;
  }
  if (e) return &e->opt;
  OptEntry n;
  n.prm      = p;
  n.consumed = false;
  nusifDefaultOptions(&n.opt);
  gOpts.push_back(n);
  return &gOpts.back().opt;
}

extern "C" void nusifClearOptions(const Parameter *p)
{
  for (size_t q = 0; q < gOpts.size(); q++)
    if (gOpts[q].prm == p) {
      gOpts.erase(gOpts.begin() + q);
      return;
    }
}

/* Options for a new Discretization: the pending ones of `p`, or defaults. */
void nusifConsumeOptions(const Parameter *p, NusifOptions *out)
{
  OptEntry *e = findEntry(p);
  if (e && !e->consumed) {
    *out        = e->opt;
    e->consumed = true;
  } else {
    nusifDefaultOptions(out);
  }
}

const NusifOptions *nusifOptionsFor(const Parameter *p)
{
  OptEntry *e = findEntry(p);
  if (e) return &e->opt;
  static NusifOptions defaults;
  nusifDefaultOptions(&defaults);
  return &defaults;
}

static int parseInt(const char *v, int *out)
{
  char *end = nullptr;
  long x    = strtol(v, &end, 10);
  if (end == v) return -2;
  *out = (int)x;
  return 0;
}

static int parseReal(const char *v, double *out)
{
  char *end = nullptr;
  double x  = strtod(v, &end);
  if (end == v) return -2;
  *out = x;
  return 0;
}

/* value up to the first whitespace */
static std::string word(const char *v)
{
  std::string s(v);
  size_t b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return "";
  size_t e = s.find_first_of(" \t\r\n", b);
  return s.substr(b, e == std::string::npos ? std::string::npos : e - b);
}

extern "C" int nusifSetOption(Parameter *p, const char *key, const char *value)
{
  if (!p || !key || !value) return -2;
  NusifOptions *o = optionsSlot(p);
  std::string k(key), v = word(value);
  if (k == "nsteps") return parseInt(v.c_str(), &o->nsteps);
  if (k == "solver") {
    if (v == "rb") o->solver = SOLVER_RB;
    else if (v == "mg") o->solver = SOLVER_MG;
    else if (v == "gmg") o->solver = SOLVER_GMG;
    else return -2;
    return 0;
  }
  if (k == "precision") {
    if (v == "fp64") o->mixed = 0;
    else if (v == "mixed") o->mixed = 1;
    else return -2;
    return 0;
  }
  if (k == "mg_tol") return parseReal(v.c_str(), &o->mgTol);
  if (k == "mg_maxcycles") return parseInt(v.c_str(), &o->mgMaxCycles);
  if (k == "mg_omega") return parseReal(v.c_str(), &o->mgOmega);
  if (k == "mg_project") return parseInt(v.c_str(), &o->mgProject);
  if (k == "mg_tail") return parseInt(v.c_str(), &o->mgTail);
  if (k == "mg_fuse_prolong") return parseInt(v.c_str(), &o->mgFuseProlong);
  if (k == "mg_graph") return parseInt(v.c_str(), &o->mgGraph);
  if (k == "mg_graph_cache") {
    int r = parseInt(v.c_str(), &o->mgGraphCache);
    if (r == 0 && o->mgGraphCache < 1) return -2;
    return r;
  }
  if (k == "verbose") {
    int r = parseInt(v.c_str(), &o->verbose);
    if (r == 0 && (o->verbose < 0 || o->verbose > 2)) return -2;
    return r;
  }
  if (k == "unet_extent") return parseInt(v.c_str(), &o->unetExtent);
  if (k == "unet_graph") return parseInt(v.c_str(), &o->unetGraph);
  if (k == "unet_accept_level") {
    int r = parseInt(v.c_str(), &o->unetAcceptLevel);
    if (r == 0 && o->unetAcceptLevel < 0) return -2;
    return r;
  }
  if (k == "warmstart_skip") {
    int r = parseInt(v.c_str(), &o->warmstartSkip);
    if (r == 0 && o->warmstartSkip < 0) return -2;
    return r;
  }
  if (k == "fg_rhs") return parseInt(v.c_str(), &o->fgRhs);
  if (k == "rb_fused") return parseInt(v.c_str(), &o->rbFused);
  if (k == "bc_fused") return parseInt(v.c_str(), &o->bcFused);
  if (k == "mg_speculate") return parseInt(v.c_str(), &o->mgSpeculate);
  if (k == "mg_inner") {
    int r = parseInt(v.c_str(), &o->mgInner);
    if (r == 0 && o->mgInner < 0) return -2;
    return r;
  }
  if (k == "mg_tile_rows") {
    int r = parseInt(v.c_str(), &o->mgTileRows);
    if (r == 0 && o->mgTileRows != 0 && o->mgTileRows != 4 && o->mgTileRows != 8 && o->mgTileRows != 16) return -2;
    return r;
  }
  if (k == "lid") return parseReal(v.c_str(), &o->lid);
  if (k == "u_inflow") return parseReal(v.c_str(), &o->uInflow);
  if (k == "inflow_noise") return parseReal(v.c_str(), &o->inflowNoise);
  if (k == "seed") return parseInt(v.c_str(), &o->seed);
  if (k == "warmstart") {
    if (v == "none") o->warmstart = 0;
    else if (v == "unet") o->warmstart = 1;
    else if (v == "unet_force") o->warmstart = 2; /* diagnostic: always take the guess */
    else return -2;
    return 0;
  }
  if (k == "unet_weights") {
    if (v.size() >= sizeof(o->unetWeights)) return -2;
    strcpy(o->unetWeights, v.c_str());
    return 0;
  }
  return -1;
}

/* ---- device info ---------------------------------------------------------- */
extern "C" int nusifDeviceName(char *buf, int len)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  if (buf && len > 0) {
    buf[0] = '\0';
    if (n > 0) {
      hipDeviceProp_t prop;
      HIP_CHECK(hipGetDeviceProperties(&prop, 0));
      if (prop.name[0])
        snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
      else
        snprintf(buf, (size_t)len, "%s, %d CUs", prop.gcnArchName, prop.multiProcessorCount);
    }
  }
  return n;
}

extern "C" void nusifDeviceSynchronize(void) { HIP_CHECK(hipStreamSynchronize(nusifStream())); }

/* ---- kernel timers -------------------------------------------------------- */
struct KernelTimer {
  std::string name;
  std::vector<hipEvent_t> pool; /* start/stop pairs, recycled */
  size_t used;
  double ms;
  int launches;
};
static std::vector<KernelTimer *> gTimers;
int nusifTimerGeneration = 0;

KernelTimer *nusifTimerFor(const char *name)
{
  for (KernelTimer *t : gTimers)
    if (t->name == name) return t;
  return nullptr;
}

extern "C" int nusifTimerEnable(const char *kernel)
{
  if (nusifTimerFor(kernel)) return 0;
  KernelTimer *t = new KernelTimer();
  t->name        = kernel;
  t->used        = 0;
  t->ms          = 0.0;
  t->launches    = 0;
  gTimers.push_back(t);
  nusifTimerGeneration++;
  return 0;
}

static void drain(KernelTimer *t)
{
  for (size_t q = 0; q + 1 < t->used; q += 2) {
    float ms = 0.0f;
    HIP_CHECK(hipEventSynchronize(t->pool[q + 1]));
    HIP_CHECK(hipEventElapsedTime(&ms, t->pool[q], t->pool[q + 1]));
    t->ms += ms;
    t->launches++;
  }
  t->used = 0;
}

void nusifTimerBegin(KernelTimer *t)
{
  if (t->used + 2 > t->pool.size()) {
    if (t->pool.size() >= 4096) drain(t); /* bound the outstanding events */
    while (t->used + 2 > t->pool.size()) {
      hipEvent_t e;
      HIP_CHECK(hipEventCreate(&e));
      t->pool.push_back(e);
    }
  }
  HIP_CHECK(hipEventRecord(t->pool[t->used], nusifStream()));
}

void nusifTimerEnd(KernelTimer *t)
{
  HIP_CHECK(hipEventRecord(t->pool[t->used + 1], nusifStream()));
  t->used += 2;
}

bool nusifTimersActive() { return !gTimers.empty(); }

extern "C" void nusifTimerDisableAll(void)
{
  for (KernelTimer *t : gTimers) {
    drain(t);
    for (hipEvent_t e : t->pool) HIP_CHECK(hipEventDestroy(e));
    delete t;
  }
  gTimers.clear();
  nusifTimerGeneration++;
}

extern "C" double nusifTimerRead(const char *kernel, int *launches)
{
  KernelTimer *t = nusifTimerFor(kernel);
  if (!t) {
    if (launches) *launches = 0;
    return 0.0;
  }
  drain(t);
  if (launches) *launches = t->launches;
  return t->ms;
}

extern "C" void nusifTimerReset(void)
{
  for (KernelTimer *t : gTimers) {
    drain(t);
    t->ms       = 0.0;
    t->launches = 0;
  }
}

/* Used by commCollectResult (host_api.c): make the host copy of a library
 * field current before it is read on the host. */
extern "C" int nusifSyncHostIfOwned(const double *hostPtr)
{
  NusifCtx *c = nusifCtxOfHost(hostPtr, nullptr);
  if (!c) return 0;
  nusifSyncHost(c->owner);
  return 1;
}

extern "C" int nusifStepLimit(Parameter *p) { return nusifOptionsFor(p)->nsteps; }
extern "C" int nusifVerboseLevel(Parameter *p) { return nusifOptionsFor(p)->verbose; }

/* Algorithmic HBM bytes per launch: every array element the kernel must
 * touch, read once and/or written once, on the finest grid (N interior
 * cells, Npad padded cells). */
extern "C" double nusifKernelBytes(Discretization *d, const char *kernel)
{
  double N    = (double)d->grid.imax * d->grid.jmax * d->grid.kmax;
  double Npad = (double)(d->grid.imax + 2) * (d->grid.jmax + 2) * (d->grid.kmax + 2);
  struct {
    const char *name;
    double perCell;
    int padded;
  } table[] = {
      {"computeFG", 48.0, 0},        /* rd u v w, wr f g h */
      {"computeRHS", 32.0, 0},       /* rd f g h, wr rhs */
      {"computeFG_rhs", 56.0, 0},    /* rd u v w, wr f g h rhs */
      {"adaptUV", 56.0, 0},          /* rd f g h p, wr u v w */
      {"absmax", 24.0, 1},           /* rd u v w, padded */
      {"rb_pass", 16.0, 0},          /* one colour: rd p (all) + rd rhs / wr p (half) */
      {"rb_fused", 24.0, 0},         /* one full iteration: rd p, rd rhs, wr p */
      {"gmg_smooth64_l0", 24.0, 0},  /* one red-black sweep, FP64: rd x, rd b, wr x */
      {"gmg_smooth32_l0", 12.0, 0},  /* one red-black sweep, FP32 */
      {"gmg_smooth64z_l0", 16.0, 0}, /* first sweep from x = 0, FP64: rd b, wr x */
      {"gmg_smooth32z_l0", 8.0, 0},  /* first sweep from x = 0, FP32 */
      {"gmg_smooth64p_l0", 25.0, 0}, /* sweep from x + P x_coarse: + rd x_coarse (1/8) */
      {"gmg_smooth32p_l0", 12.5, 0},
      {"gmg_smooth64s_l0", 17.0, 0}, /* sweep from P x_coarse: rd b, x_coarse (1/8), wr x */
      {"gmg_smooth32s_l0", 8.5, 0},
      {"gmg_restrict64_l0", 17.0, 0},/* rd x, b; wr b_coarse (1/8) */
      {"gmg_restrict32_l0", 8.5, 0},
      {"gmg_prolong64_l0", 17.0, 0}, /* rd/wr x, rd x_coarse (1/8) */
      {"gmg_prolong32_l0", 8.5, 0},
      {"gmg_residual64", 16.0, 0},   /* rd p, rhs (+4 when the FP32 copy is written) */
      {"gmg_correct", 20.0, 0},      /* rd p, e32; wr p */
      {"gmg_correct_residual", 32.0, 0}, /* rd p, e32, rhs; wr p', r32 (the zero-start
                                            sweep needs no zeroed e) */
      {"ws_guess_residual", 32.5, 0}, /* warm-start guess: rd p, d32, rhs; wr p + d,
                                         r32, and r averaged onto level 1 (1/8 x 4 B) */
  };
  for (auto &e : table)
    if (strcmp(kernel, e.name) == 0) return e.perCell * (e.padded ? Npad : N);
  return 0.0;
}

/* Device-to-device copy of `bytes` bytes, `reps` times, on the library
 * stream: a plain copy of known size, used to calibrate HBM byte counters
 * and as the achievable-bandwidth reference. Returns ms per copy. */
extern "C" double nusifCopyProbe(size_t bytes, int reps)
{
  void *a = nullptr, *b = nullptr;
  HIP_CHECK(hipMalloc(&a, bytes));
  HIP_CHECK(hipMalloc(&b, bytes));
  HIP_CHECK(hipMemsetAsync(a, 1, bytes, nusifStream()));
  HIP_CHECK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, nusifStream()));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, nusifStream()));
  for (int r = 0; r < reps; r++)
    HIP_CHECK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, nusifStream()));
  HIP_CHECK(hipEventRecord(e1, nusifStream()));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  HIP_CHECK(hipFree(a));
  HIP_CHECK(hipFree(b));
  return reps > 0 ? ms / reps : 0.0;
}

/* ---- streaming probes: the HBM rate a plain kernel reaches ------------- */
/* Grid-stride loops over float4 (16 B per lane, one 1 KiB row per wave). */
__global__ __launch_bounds__(256) void kProbeCopy(const float4 *__restrict__ a, float4 *__restrict__ b, long long n)
{
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) b[i] = a[i];
}

/* two reads, one write: the access pattern of a smoother sweep (x, b -> x) */
__global__ __launch_bounds__(256) void kProbeAdd(const float4 *__restrict__ a, const float4 *__restrict__ c,
    float4 *__restrict__ b, long long n)
{
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    float4 x = a[i], y = c[i];
    b[i]     = make_float4(x.x + y.x, x.y + y.y, x.z + y.z, x.w + y.w);
  }
}

/* read only: the sum decides a store that never happens for finite input,
 * so every load is needed */
__global__ __launch_bounds__(256) void kProbeRead(const float4 *__restrict__ a, float *__restrict__ out, long long n)
{
  float s = 0.0f;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    float4 x = a[i];
    s += (x.x + x.y) + (x.z + x.w);
  }
  if (s == -1.0f) out[blockIdx.x * 256 + threadIdx.x] = s;
}

/* Milliseconds per launch of the three probes over arrays of `bytes` bytes
 * each (rounded down to 16 B), `reps` launches each: ms[0] copy (rd + wr),
 * ms[1] two reads and one write, ms[2] read only. The arrays start zeroed. */
extern "C" void nusifStreamProbe(size_t bytes, int reps, double *ms)
{
  const long long n = (long long)(bytes / 16);
  float4 *a = nullptr, *b = nullptr, *c = nullptr;
  float *out = nullptr;
  const int grid = (int)std::min<long long>((n + 255) / 256, 16384);
  HIP_CHECK(hipMalloc(&a, (size_t)n * 16));
  HIP_CHECK(hipMalloc(&b, (size_t)n * 16));
  HIP_CHECK(hipMalloc(&c, (size_t)n * 16));
  HIP_CHECK(hipMalloc(&out, (size_t)grid * 256 * sizeof(float)));
  HIP_CHECK(hipMemsetAsync(a, 0, (size_t)n * 16, nusifStream()));
  HIP_CHECK(hipMemsetAsync(b, 0, (size_t)n * 16, nusifStream()));
  HIP_CHECK(hipMemsetAsync(c, 0, (size_t)n * 16, nusifStream()));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  for (int v = 0; v < 3; v++) {
    for (int r = -1; r < reps; r++) { /* r = -1: untimed warm-up launch */
      if (r == 0) HIP_CHECK(hipEventRecord(e0, nusifStream()));
      if (v == 0) kProbeCopy<<<grid, 256, 0, nusifStream()>>>(a, b, n);
      if (v == 1) kProbeAdd<<<grid, 256, 0, nusifStream()>>>(a, c, b, n);
      if (v == 2) kProbeRead<<<grid, 256, 0, nusifStream()>>>(a, out, n);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(e1, nusifStream()));
    HIP_CHECK(hipEventSynchronize(e1));
    float t = 0.0f;
    HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
    ms[v] = reps > 0 ? t / reps : 0.0;
  }
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  HIP_CHECK(hipFree(a));
  HIP_CHECK(hipFree(b));
  HIP_CHECK(hipFree(c));
  HIP_CHECK(hipFree(out));
}
