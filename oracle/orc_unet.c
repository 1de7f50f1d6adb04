/* CPU ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * Naive restatement of the warm-start network of the GPU library (the U-Net
 * of cfd-bench_amd/csrc/unet.hip; no counterpart in the reference), forward
 * and backward in FP64, used to check the GPU's MFMA convolutions and their
 * gradients.
 *
 * Network (channels-last tensors [z][y][x][c], 'same' zero padding):
 *   enc l = 0..L-1:  conv3(cin_l -> ch_l) ReLU, conv3(ch_l -> ch_l) ReLU = s_l
 *                    (l < L-1: 2x2x2 average pool of s_l feeds level l+1)
 *   dec l = L-2..0:  up = nearest x2 of the level-(l+1) output,
 *                    conv3(concat(up, s_l) -> ch_l) ReLU, conv3(ch_l -> ch_l) ReLU
 *   head:            1x1 conv (ch_0 -> 1), no activation
 *   loss:            mean over voxels of (y - target)^2
 * Parameters, flat: for every 3x3x3 conv in the order enc0a enc0b enc1a ...
 * dec(L-2)a dec(L-2)b ... dec0a dec0b: W[cout][27][cin] then b[cout]; then
 * the head W[ch_0], b[1]. Tap t = (dz+1)*9 + (dy+1)*3 + (dx+1).
 *
 * With `bf16` set, weights and every stored activation are rounded to
 * bfloat16 (round to nearest even) where the GPU stores them, so forward
 * outputs match the GPU up to FP32 accumulation order.
 *
 * Tensors are boxes of nx x ny x nz voxels (x fastest), each extent
 * divisible by 2^(L-1).
 *
 * Pins (tests/test_oracle_unet.py): central finite differences of the loss
 * against the analytic gradient for every parameter block; the whole forward
 * at L = 1, 2 and 4 on cubes and boxes against an independent numpy
 * restatement (reshape-mean pooling, np.repeat upsampling, concatenation
 * order, zero-padded correlation); the pooling and upsampling backward passes
 * against the adjoint identities <P x, y> = <x, P^T y>.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 4

static double bf16r(double x)
{
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) {
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
  }
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

typedef struct {
  int nx, ny, nz, c; /* box extents, channels */
  double *v;
} Ten;

static size_t nvox(const Ten *t) { return (size_t)t->nx * t->ny * t->nz; }

static Ten ten(int nx, int ny, int nz, int c)
{
  Ten t;
  t.nx = nx;
  t.ny = ny;
  t.nz = nz;
  t.c  = c;
  t.v  = calloc((size_t)nx * ny * nz * c, sizeof(double));
  return t;
}

/* same extents as t, c channels */
static Ten like(const Ten *t, int c) { return ten(t->nx, t->ny, t->nz, c); }

/* half / double the extents of t, c channels */
static Ten half(const Ten *t, int c) { return ten(t->nx / 2, t->ny / 2, t->nz / 2, c); }
static Ten twice(const Ten *t, int c) { return ten(2 * t->nx, 2 * t->ny, 2 * t->nz, c); }

static void roundTen(Ten *t, int bf16)
{
  if (!bf16) return;
  size_t m = nvox(t) * t->c;
  for (size_t q = 0; q < m; q++) t->v[q] = bf16r(t->v[q]);
}

/* voxel index of (x, y, z) in t, -1 outside (the zero padding) */
static long vox(const Ten *t, int x, int y, int z)
{
  if (x < 0 || x >= t->nx || y < 0 || y >= t->ny || z < 0 || z >= t->nz) return -1;
  return ((long)z * t->ny + y) * t->nx + x;
}

/* out = conv3(in0 ++ in1) + b, optional ReLU. W[cout][27][cin0+cin1] */
static void convFwd(const Ten *in0, const Ten *in1, const double *W, const double *b, Ten *out,
    int relu)
{
  int c0 = in0->c, c1 = in1 ? in1->c : 0, cin = c0 + c1, co = out->c;
  for (int z = 0; z < out->nz; z++)
    for (int y = 0; y < out->ny; y++)
      for (int x = 0; x < out->nx; x++)
        for (int o = 0; o < co; o++) {
          double s = b[o];
          for (int t = 0; t < 27; t++) {
            long v = vox(out, x + t % 3 - 1, y + (t / 3) % 3 - 1, z + t / 9 - 1);
            if (v < 0) continue;
            const double *w = W + ((size_t)o * 27 + t) * cin;
            for (int c = 0; c < c0; c++) s += w[c] * in0->v[v * c0 + c];
            for (int c = 0; c < c1; c++) s += w[c0 + c] * in1->v[v * c1 + c];
          }
          if (relu && s < 0.0) s = 0.0;
          out->v[vox(out, x, y, z) * co + o] = s;
        }
}

/* given dOut (already through the ReLU mask), accumulate dIn0/dIn1, dW, db */
static void convBwd(const Ten *in0, const Ten *in1, const double *W, const Ten *dOut, Ten *dIn0,
    Ten *dIn1, double *dW, double *db)
{
  int c0 = in0->c, c1 = in1 ? in1->c : 0, cin = c0 + c1, co = dOut->c;
  for (int z = 0; z < dOut->nz; z++)
    for (int y = 0; y < dOut->ny; y++)
      for (int x = 0; x < dOut->nx; x++)
        for (int o = 0; o < co; o++) {
          double g = dOut->v[vox(dOut, x, y, z) * co + o];
          if (g == 0.0) continue;
          db[o] += g;
          for (int t = 0; t < 27; t++) {
            long v = vox(dOut, x + t % 3 - 1, y + (t / 3) % 3 - 1, z + t / 9 - 1);
            if (v < 0) continue;
            const double *w = W + ((size_t)o * 27 + t) * cin;
            double *dw      = dW + ((size_t)o * 27 + t) * cin;
            for (int c = 0; c < c0; c++) {
              dw[c] += g * in0->v[v * c0 + c];
              if (dIn0) dIn0->v[v * c0 + c] += g * w[c];
            }
            for (int c = 0; c < c1; c++) {
              dw[c0 + c] += g * in1->v[v * c1 + c];
              if (dIn1) dIn1->v[v * c1 + c] += g * w[c0 + c];
            }
          }
        }
}

static void reluMask(Ten *d, const Ten *act)
{
  size_t m = nvox(d) * d->c;
  for (size_t q = 0; q < m; q++)
    if (!(act->v[q] > 0.0)) d->v[q] = 0.0;
}

/* out = average of the 2x2x2 children in `in` (child q = (dz, dy, dx) bits) */
static void poolFwd(const Ten *in, Ten *out)
{
  int c = in->c;
  for (int z = 0; z < out->nz; z++)
    for (int y = 0; y < out->ny; y++)
      for (int x = 0; x < out->nx; x++)
        for (int k = 0; k < c; k++) {
          double s = 0.0;
          for (int q = 0; q < 8; q++)
            s += in->v[vox(in, 2 * x + (q & 1), 2 * y + ((q >> 1) & 1), 2 * z + (q >> 2)) * c + k];
          out->v[vox(out, x, y, z) * c + k] = 0.125 * s;
        }
}

static void poolBwd(const Ten *dOut, Ten *dIn)
{
  int c = dOut->c;
  for (int z = 0; z < dIn->nz; z++)
    for (int y = 0; y < dIn->ny; y++)
      for (int x = 0; x < dIn->nx; x++)
        for (int k = 0; k < c; k++)
          dIn->v[vox(dIn, x, y, z) * c + k] += 0.125 * dOut->v[vox(dOut, x / 2, y / 2, z / 2) * c + k];
}

/* out = nearest x2 copy of in (each voxel takes its parent's value) */
static void upFwd(const Ten *in, Ten *out)
{
  int c = in->c;
  for (int z = 0; z < out->nz; z++)
    for (int y = 0; y < out->ny; y++)
      for (int x = 0; x < out->nx; x++)
        for (int k = 0; k < c; k++)
          out->v[vox(out, x, y, z) * c + k] = in->v[vox(in, x / 2, y / 2, z / 2) * c + k];
}

static void upBwd(const Ten *dOut, Ten *dIn)
{
  int c = dOut->c;
  for (int z = 0; z < dOut->nz; z++)
    for (int y = 0; y < dOut->ny; y++)
      for (int x = 0; x < dOut->nx; x++)
        for (int k = 0; k < c; k++)
          dIn->v[vox(dIn, x / 2, y / 2, z / 2) * c + k] += dOut->v[vox(dOut, x, y, z) * c + k];
}

/* cfg = {L, ch0, ch1, ch2, ch3, cinPad} */
typedef struct {
  int L, ch[MAXL], cin;
  int nconv;
  int cinOf[2 * 2 * MAXL], coutOf[2 * 2 * MAXL];
  size_t wOff[2 * 2 * MAXL], bOff[2 * 2 * MAXL], headW, headB, total;
} Net;

static Net netOf(const int *cfg)
{
  Net N;
  memset(&N, 0, sizeof N);
  N.L   = cfg[0];
  N.cin = cfg[5];
  for (int l = 0; l < N.L; l++) N.ch[l] = cfg[1 + l];
  int q = 0;
  for (int l = 0; l < N.L; l++) { /* encoder */
    N.cinOf[q]  = l == 0 ? N.cin : N.ch[l - 1];
    N.coutOf[q] = N.ch[l];
    q++;
    N.cinOf[q]  = N.ch[l];
    N.coutOf[q] = N.ch[l];
    q++;
  }
  for (int l = N.L - 2; l >= 0; l--) { /* decoder */
    N.cinOf[q]  = N.ch[l + 1] + N.ch[l];
    N.coutOf[q] = N.ch[l];
    q++;
    N.cinOf[q]  = N.ch[l];
    N.coutOf[q] = N.ch[l];
    q++;
  }
  N.nconv = q;
  size_t off = 0;
  for (int k = 0; k < q; k++) {
    N.wOff[k] = off;
    off += (size_t)N.coutOf[k] * 27 * N.cinOf[k];
    N.bOff[k] = off;
    off += (size_t)N.coutOf[k];
  }
  N.headW = off;
  off += (size_t)N.ch[0];
  N.headB = off;
  off += 1;
  N.total = off;
  return N;
}

long orcUnetParamCount(const int *cfg) { return (long)netOf(cfg).total; }

/* Loss of one sample on an nx x ny x nz box (each extent divisible by
 * 2^(L-1)); grad (may be NULL) receives d loss / d params (added to, not
 * overwritten); pred (may be NULL) receives the nx ny nz prediction. */
double orcUnetLossBox(const int *cfg, int nx, int ny, int nz, const double *params, const double *x,
    const double *target, double *grad, double *pred, int bf16)
{
  Net N = netOf(cfg);
  int L = N.L;
  double *P = malloc(N.total * sizeof(double));
  for (size_t q = 0; q < N.total; q++) P[q] = bf16 ? bf16r(params[q]) : params[q];
  /* head weights stay FP32 on the GPU */
  if (bf16)
    for (int c = 0; c <= N.ch[0]; c++) P[N.headW + c] = (double)(float)params[N.headW + c];

  Ten xin = ten(nx, ny, nz, N.cin);
  memcpy(xin.v, x, nvox(&xin) * N.cin * sizeof(double));
  roundTen(&xin, bf16);
  Ten a[MAXL], s[MAXL], pl[MAXL], up[MAXL], da[MAXL], d[MAXL];
  int q = 0;
  const Ten *cur = &xin;
  for (int l = 0; l < L; l++) {
    a[l] = like(cur, N.ch[l]);
    convFwd(cur, NULL, P + N.wOff[q], P + N.bOff[q], &a[l], 1);
    roundTen(&a[l], bf16);
    q++;
    s[l] = like(cur, N.ch[l]);
    convFwd(&a[l], NULL, P + N.wOff[q], P + N.bOff[q], &s[l], 1);
    roundTen(&s[l], bf16);
    q++;
    if (l < L - 1) {
      pl[l] = half(&s[l], N.ch[l]);
      poolFwd(&s[l], &pl[l]);
      roundTen(&pl[l], bf16);
      cur = &pl[l];
    }
  }
  const Ten *prev = &s[L - 1];
  for (int l = L - 2; l >= 0; l--) {
    up[l] = twice(prev, N.ch[l + 1]);
    upFwd(prev, &up[l]);
    da[l] = like(&s[l], N.ch[l]);
    convFwd(&up[l], &s[l], P + N.wOff[q], P + N.bOff[q], &da[l], 1);
    roundTen(&da[l], bf16);
    q++;
    d[l] = like(&s[l], N.ch[l]);
    convFwd(&da[l], NULL, P + N.wOff[q], P + N.bOff[q], &d[l], 1);
    roundTen(&d[l], bf16);
    q++;
    prev = &d[l];
  }
  const Ten *top = L > 1 ? &d[0] : &s[0];
  size_t nv = nvox(&xin);
  double loss = 0.0;
  double *y = malloc(nv * sizeof(double));
  for (size_t v = 0; v < nv; v++) {
    double acc = P[N.headB];
    for (int c = 0; c < N.ch[0]; c++) acc += P[N.headW + c] * top->v[v * N.ch[0] + c];
    y[v] = acc;
    double e = acc - target[v];
    loss += e * e;
  }
  loss /= (double)nv;
  if (pred) memcpy(pred, y, nv * sizeof(double));

  if (grad) {
    Ten dtop = like(&xin, N.ch[0]);
    for (size_t v = 0; v < nv; v++) {
      double g = 2.0 * (y[v] - target[v]) / (double)nv;
      grad[N.headB] += g;
      for (int c = 0; c < N.ch[0]; c++) {
        grad[N.headW + c] += g * top->v[v * N.ch[0] + c];
        dtop.v[v * N.ch[0] + c] = g * P[N.headW + c];
      }
    }
    /* decoder, reverse order */
    Ten dprev = dtop; /* gradient w.r.t. the current decoder output */
    Ten dS[MAXL];
    for (int l = 0; l < L; l++) dS[l] = like(&s[l], s[l].c);
    int qq = N.nconv;
    if (L == 1) {
      for (size_t m = 0; m < nv * N.ch[0]; m++) dS[0].v[m] = dtop.v[m];
    }
    for (int l = 0; l <= L - 2; l++) {
      /* layers of decoder level l are conv indices qq-2 (a), qq-1 (b) */
      int ib = qq - 1, ia = qq - 2;
      qq -= 2;
      reluMask(&dprev, &d[l]);
      Ten dda = like(&da[l], da[l].c);
      convBwd(&da[l], NULL, P + N.wOff[ib], &dprev, &dda, NULL, grad + N.wOff[ib], grad + N.bOff[ib]);
      reluMask(&dda, &da[l]);
      Ten dup = like(&up[l], up[l].c);
      convBwd(&up[l], &s[l], P + N.wOff[ia], &dda, &dup, &dS[l], grad + N.wOff[ia], grad + N.bOff[ia]);
      free(dda.v);
      free(dprev.v);
      /* up is the nearest-x2 copy of the level-(l+1) output */
      Ten dnext = half(&up[l], up[l].c);
      upBwd(&dup, &dnext);
      free(dup.v);
      if (l + 1 == L - 1) {
        for (size_t m = 0; m < nvox(&dnext) * dnext.c; m++) dS[L - 1].v[m] += dnext.v[m];
        free(dnext.v);
      } else {
        dprev = dnext; /* gradient of d[l+1] */
      }
      /* decoder levels are visited l = 0, 1, ...: d[l] depends on d[l+1], so
       * the loop proceeds outward from the top, which is the reverse of the
       * forward order (L-2 first). */
    }
    /* encoder, deepest level first */
    for (int l = L - 1; l >= 0; l--) {
      int ia = 2 * l, ib = 2 * l + 1;
      reluMask(&dS[l], &s[l]);
      Ten dA = like(&a[l], a[l].c);
      convBwd(&a[l], NULL, P + N.wOff[ib], &dS[l], &dA, NULL, grad + N.wOff[ib], grad + N.bOff[ib]);
      reluMask(&dA, &a[l]);
      const Ten *inl = l == 0 ? &xin : &pl[l - 1];
      Ten dIn = like(inl, inl->c);
      convBwd(inl, NULL, P + N.wOff[ia], &dA, l == 0 ? NULL : &dIn, NULL, grad + N.wOff[ia],
          grad + N.bOff[ia]);
      free(dA.v);
      if (l > 0) poolBwd(&dIn, &dS[l - 1]);
      free(dIn.v);
    }
    for (int l = 0; l < L; l++) free(dS[l].v);
  }

  free(y);
  free(P);
  free(xin.v);
  for (int l = 0; l < L; l++) {
    free(a[l].v);
    free(s[l].v);
    if (l < L - 1) free(pl[l].v);
  }
  for (int l = 0; l <= L - 2; l++) {
    free(up[l].v);
    free(da[l].v);
    free(d[l].v);
  }
  return loss;
}

/* the same on an n^3 cube */
double orcUnetLoss(const int *cfg, int n, const double *params, const double *x,
    const double *target, double *grad, double *pred, int bf16)
{
  return orcUnetLossBox(cfg, n, n, n, params, x, target, grad, pred, bf16);
}

/* test hooks: the pooling and upsampling layers alone, on channels-last
 * boxes; `fine` has extents (2 nx, 2 ny, 2 nz), `coarse` (nx, ny, nz).
 * dir 0: coarse = pool(fine); 1: fine += pool^T(coarse);
 *     2: fine = up(coarse);   3: coarse += up^T(fine). */
void orcUnetTransfer(int dir, int nx, int ny, int nz, int c, double *fine, double *coarse)
{
  Ten f = {2 * nx, 2 * ny, 2 * nz, c, fine}, g = {nx, ny, nz, c, coarse};
  switch (dir) {
  case 0: poolFwd(&f, &g); break;
  case 1: poolBwd(&g, &f); break;
  case 2: upFwd(&g, &f); break;
  default: upBwd(&f, &g); break;
  }
}

/* round-to-nearest-even bfloat16 of x, as a double (test hook) */
double orcBf16(double x) { return bf16r(x); }
