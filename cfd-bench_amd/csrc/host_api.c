/* Host-side support API of the NuSiF driver (reference src/comm.c sequential
 * branches, src/parameter.c, src/progress.c, src/profiler.c, src/timing.c,
 * src/allocate.c, src/vtkWriter-seq.c), restated for one GPU:
 * communication calls are single-domain no-ops and the only data movement is
 * the device->host download inside commCollectResult. */
#define _GNU_SOURCE
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "allocate.h"
#include "comm.h"
#include "nusif.h"
#include "parameter.h"
#include "profiler.h"
#include "progress.h"
#include "timing.h"
#include "vtkWriter.h"

/* context.hip: download a Discretization's device fields when `hostPtr` is
 * one of its host arrays; returns 1 when it did. */
extern int nusifSyncHostIfOwned(const double *hostPtr);

/* ---- allocate / timing (src/allocate.c:12-36, src/timing.c:7-19) -------- */
void *allocate(size_t alignment, size_t bytesize)
{
  void *ptr = NULL;
  int rc    = posix_memalign(&ptr, alignment, bytesize);
  if (rc == EINVAL) {
    fprintf(stderr, "Error: Alignment parameter is not a power of two\n");
    exit(EXIT_FAILURE);
  }
  if (rc == ENOMEM || ptr == NULL) {
    fprintf(stderr, "Error: Insufficient memory to fulfill the request\n");
    exit(EXIT_FAILURE);
  }
  return ptr;
}

double getTimeStamp(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + (double)ts.tv_nsec * 1.e-9;
}

double getTimeResolution(void)
{
  struct timespec ts;
  clock_getres(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + (double)ts.tv_nsec * 1.e-9;
}

/* ---- communication: one domain ------------------------------------------ */
void commInit(CommType *c, int argc, char **argv)
{
  (void)argc;
  (void)argv;
  c->rank = 0;
  c->size = 1;
}

void commPartition(CommType *c, int kmax, int jmax, int imax)
{
  c->imaxLocal = imax;
  c->jmaxLocal = jmax;
  c->kmaxLocal = kmax;
  for (int d = 0; d < NDIMS; d++) {
    c->dims[d]   = 1;
    c->coords[d] = 0;
  }
  for (int d = 0; d < NDIRS; d++) c->neighbours[d] = -1;
}

void commFinalize(CommType *c) { (void)c; }
void commPrintConfig(CommType *c) { (void)c; }
void commExchange(CommType *c, double *grid) { (void)c; (void)grid; }
void commShift(CommType *c, double *f, double *g, double *h)
{
  (void)c; (void)f; (void)g; (void)h;
}
void commReduceAll(double *v, int op) { (void)v; (void)op; }
void commReduce(double *v, double *o, int count, int op)
{
  (void)op;
  for (int i = 0; i < count; i++) o[i] = v[i];
}
int commIsBoundary(CommType *c, DirectionType direction)
{
  (void)c;
  (void)direction;
  return 1;
}
void commGetOffsets(CommType *c, int offsets[], int kmax, int jmax, int imax)
{
  (void)c; (void)offsets; (void)kmax; (void)jmax; (void)imax;
}
void commFreeCommunicator(CommType *comm) { (void)comm; }
void commUpdateDatatypes(
    CommType *oldcomm, CommType *newcomm, int imaxLocal, int jmaxLocal, int kmaxLocal)
{
  (void)oldcomm;
  newcomm->imaxLocal = imaxLocal;
  newcomm->jmaxLocal = jmaxLocal;
  newcomm->kmaxLocal = kmaxLocal;
}

/* src/comm.c:436-476: interior p, face velocities averaged to cell centres. */
void commCollectResult(CommType *c, double *ug, double *vg, double *wg, double *pg, double *u,
    double *v, double *w, double *p, int kmax, int jmax, int imax)
{
  (void)kmax; (void)jmax; (void)imax;
  nusifSyncHostIfOwned(p);
  if (u != p) nusifSyncHostIfOwned(u);
  int ni = c->imaxLocal, nj = c->jmaxLocal, nk = c->kmaxLocal;
  size_t sj = (size_t)ni + 2, sk = sj * ((size_t)nj + 2), q = 0;
  for (int k = 1; k <= nk; k++)
    for (int j = 1; j <= nj; j++)
      for (int i = 1; i <= ni; i++, q++) {
        size_t x = (size_t)k * sk + (size_t)j * sj + (size_t)i;
        pg[q]    = p[x];
        ug[q]    = (u[x] + u[x - 1]) / 2.0;
        vg[q]    = (v[x] + v[x - sj]) / 2.0;
        wg[q]    = (w[x] + w[x - sk]) / 2.0;
      }
}

/* ---- parameters (src/parameter.c:13-133) ------------------------------- */
extern void nusifClearOptions(const Parameter *p);

void initParameter(Parameter *param)
{
  nusifClearOptions(param);
  param->xlength    = 1.0;
  param->ylength    = 1.0;
  param->zlength    = 1.0;
  param->imax       = 100;
  param->jmax       = 100;
  param->kmax       = 100;
  param->itermax    = 1000;
  param->eps        = 0.0001;
  param->omg        = 1.7;
  param->re         = 100.0;
  param->gamma      = 0.9;
  param->tau        = 0.5;
  param->levels     = 5;
  param->presmooth  = 5;
  param->postsmooth = 5;
}

/* strtok(line, " ")-style scanning: a token ends at the next ' ' only, so a
 * value keeps any tab or newline that follows it (reference src/parameter.c:53-54). */
static char *nextToken(char **cursor)
{
  char *s = *cursor;
  while (*s == ' ') s++;
  if (*s == '\0') {
    *cursor = s;
    return NULL;
  }
  char *tok = s;
  while (*s != '\0' && *s != ' ') s++;
  if (*s == ' ') *s++ = '\0';
  *cursor = s;
  return tok;
}

/* Key tables: the reference matches a key when the token STARTS with it and
 * tries every key (no else), src/parameter.c:56-95. */
typedef enum { K_INT, K_REAL, K_STR } KeyKind;
typedef struct {
  const char *key;
  KeyKind kind;
  size_t offset;
} KeySpec;

#define KEY(f, kind) {#f, kind, offsetof(Parameter, f)}
static const KeySpec kKeys[] = {
    KEY(xlength, K_REAL), KEY(ylength, K_REAL), KEY(zlength, K_REAL), KEY(imax, K_INT),
    KEY(jmax, K_INT), KEY(kmax, K_INT), KEY(itermax, K_INT), KEY(levels, K_INT),
    KEY(presmooth, K_INT), KEY(postsmooth, K_INT), KEY(eps, K_REAL), KEY(omg, K_REAL),
    KEY(re, K_REAL), KEY(tau, K_REAL), KEY(gamma, K_REAL), KEY(dt, K_REAL), KEY(te, K_REAL),
    KEY(gx, K_REAL), KEY(gy, K_REAL), KEY(gz, K_REAL), KEY(name, K_STR), KEY(bcLeft, K_INT),
    KEY(bcRight, K_INT), KEY(bcBottom, K_INT), KEY(bcTop, K_INT), KEY(bcFront, K_INT),
    KEY(bcBack, K_INT), KEY(u_init, K_REAL), KEY(v_init, K_REAL), KEY(w_init, K_REAL),
    KEY(p_init, K_REAL),
};
#undef KEY

/* Keys this library adds; matched exactly. None of the reference keys is a
 * prefix of any of them, so the reference's own keys never fire on them. */
static const char *kExtraKeys[] = {"nsteps", "solver", "precision", "mg_tol", "mg_maxcycles",
    "mg_omega", "mg_project", "mg_tail", "mg_fuse_prolong", "bc_fused", "mg_graph", "mg_tile_rows", "mg_speculate", "mg_inner", "lid", "u_inflow", "inflow_noise", "seed",
    "warmstart", "unet_weights", "unet_extent", "mg_graph_cache", "verbose", "unet_accept_level", "unet_graph", "warmstart_skip"};

void readParameter(Parameter *param, const char *filename)
{
  FILE *fp = fopen(filename, "r");
  char line[4096];
  if (!fp) {
    fprintf(stderr, "Could not open parameter file: %s\n", filename);
    exit(EXIT_FAILURE);
  }
  while (!feof(fp)) {
    line[0] = '\0';
    if (fgets(line, sizeof(line), fp) == NULL && ferror(fp) != 0)
      fprintf(stderr, "Error in fgets function\n");
    char *hash = strchr(line, '#');
    if (hash) *hash = '\0';
    char *cursor = line;
    char *tok    = nextToken(&cursor);
    char *val    = tok ? nextToken(&cursor) : NULL;
    if (tok == NULL || val == NULL) continue;
    for (size_t q = 0; q < sizeof(kKeys) / sizeof(kKeys[0]); q++) {
      const KeySpec *k = &kKeys[q];
      if (strncmp(tok, k->key, strlen(k->key)) != 0) continue;
      char *field = (char *)param + k->offset;
      if (k->kind == K_INT) *(int *)field = atoi(val);
      else if (k->kind == K_REAL) *(double *)field = atof(val);
      else *(char **)field = strdup(val);
    }
    for (size_t q = 0; q < sizeof(kExtraKeys) / sizeof(kExtraKeys[0]); q++)
      if (strcmp(tok, kExtraKeys[q]) == 0 && nusifSetOption(param, tok, val) != 0) {
        fprintf(stderr, "Bad value for %s: %s\n", tok, val);
        exit(EXIT_FAILURE);
      }
  }
  fclose(fp);
}

void printParameter(Parameter *param)
{
  printf("Parameters for %s\n", param->name);
  printf("Boundary conditions Left:%d Right:%d Bottom:%d Top:%d Front:%d "
         "Back:%d\n",
      param->bcLeft, param->bcRight, param->bcBottom, param->bcTop, param->bcFront,
      param->bcBack);
  printf("\tReynolds number: %.2f\n", param->re);
  printf("\tInit arrays: U:%.2f V:%.2f W:%.2f P:%.2f\n", param->u_init, param->v_init,
      param->w_init, param->p_init);
  printf("Geometry data:\n");
  printf("\tDomain box size (x, y, z): %.2f, %.2f, %.2f\n", param->xlength, param->ylength,
      param->zlength);
  printf("\tCells (x, y, z): %d, %d, %d\n", param->imax, param->jmax, param->kmax);
  printf("Timestep parameters:\n");
  printf("\tDefault stepsize: %.2f, Final time %.2f\n", param->dt, param->te);
  printf("\tTau factor: %.2f\n", param->tau);
  printf("Iterative solver parameters:\n");
  printf("\tMax iterations: %d\n", param->itermax);
  printf("\tepsilon (stopping tolerance) : %f\n", param->eps);
  printf("\tgamma (stopping tolerance) : %f\n", param->gamma);
  printf("\tomega (SOR relaxation): %f\n", param->omg);
}

/* ---- progress + residual trace (src/progress.c:14-66) ------------------- */
static double gEnd;
static int gTicks;

void initProgress(double end)
{
  gEnd   = end;
  gTicks = 0;
  printf("[          ]");
  fflush(stdout);
}

void printProgress(double current)
{
  int ticks = (int)rint((current / gEnd) * 10.0);
  if (ticks > gTicks) {
    char bar[11];
    gTicks = ticks;
    for (int i = 0; i < 10; i++) bar[i] = (i < gTicks) ? '#' : ' ';
    bar[10] = '\0';
    printf("\r[%s]", bar);
  }
  fflush(stdout);
}

void stopProgress(void)
{
  printf("\n");
  fflush(stdout);
}

FILE *initResidualWriter(void)
{
  FILE *fp = fopen("residual.dat", "w");
  if (fp == NULL) {
    printf("Error!\n");
    exit(EXIT_FAILURE);
  }
  return fp;
}

void writeResidual(FILE *fp, double ts, double res) { fprintf(fp, "%f, %f\n", ts, res); }

/* ---- profiler (src/profiler.c:15-109), sequential ----------------------- */
double T[NUMREGIONS];
size_t C[NUMREGIONS];
static double gUpdates = 0.0, gDataset = 0.0;

void initProfiler(CommType *c)
{
  double points = (double)(c->kmaxLocal + 2) * (c->jmaxLocal + 2) * (c->imaxLocal + 2);
  gUpdates      = (double)c->kmaxLocal * c->jmaxLocal * c->imaxLocal;
  gDataset      = points * 16.0 * 1.E-6;
  for (int i = 0; i < NUMREGIONS; i++) {
    T[i] = 0.0;
    C[i] = 0;
  }
}

/* Solver MLUPS line: iterations x interior cells / solver time
 * (src/profiler.c:61-74). */
void printProfile(CommType *c, int it)
{
  if (commIsMaster(c) && T[SOLVER] > 0.0) {
    double mlups = (it * gUpdates) * 1.E-6;
    printf("Solver Time: %.2f s Dataset size: %.2f MB Performance: %.2f MLUPS/s "
           "Performance total: %.2f MLUPS/s\n",
        T[SOLVER], gDataset, mlups / T[SOLVER], (mlups * c->size) / T[SOLVER]);
  }
  for (int i = 0; i < NUMREGIONS; i++) {
    T[i] = 0.0;
    C[i] = 0;
  }
}

void finalizeProfiler(void) {}

/* ---- legacy VTK, sequential (src/vtkWriter-seq.c:31-146) ---------------- */
static double swapBytes(double x)
{
  unsigned char in[8], out[8];
  memcpy(in, &x, 8);
  for (int q = 0; q < 8; q++) out[q] = in[7 - q];
  memcpy(&x, out, 8);
  return x;
}

void vtkOpen(VtkOptions *o, char *problem)
{
  char filename[50];
  snprintf(filename, sizeof(filename), "%s.vtk", problem);
  o->fh = fopen(filename, "w");
  if (!o->fh) {
    fprintf(stderr, "Could not open %s\n", filename);
    exit(EXIT_FAILURE);
  }
  Grid *g = o->grid;
  fprintf(o->fh, "# vtk DataFile Version 3.0\n");
  fprintf(o->fh, "PAMPI cfd solver output\n");
  fprintf(o->fh, o->fmt == BINARY ? "BINARY\n" : "ASCII\n");
  fprintf(o->fh, "DATASET STRUCTURED_POINTS\n");
  fprintf(o->fh, "DIMENSIONS %d %d %d\n", g->imax, g->jmax, g->kmax);
  fprintf(o->fh, "ORIGIN %f %f %f\n", g->dx * 0.5, g->dy * 0.5, g->dz * 0.5);
  fprintf(o->fh, "SPACING %f %f %f\n", g->dx, g->dy, g->dz);
  fprintf(o->fh, "POINT_DATA %d\n", g->imax * g->jmax * g->kmax);
  printf("Writing VTK output for %s\n", problem);
}

static int vtkReady(VtkOptions *o)
{
  if (o->fh == NULL) {
    printf("vtkWriter not initialize! Call vtkOpen first!\n");
    return 0;
  }
  return 1;
}

void vtkScalar(VtkOptions *o, char *name, double *s)
{
  printf("Register scalar %s\n", name);
  if (!vtkReady(o)) return;
  fprintf(o->fh, "SCALARS %s double 1\n", name);
  fprintf(o->fh, "LOOKUP_TABLE default\n");
  size_t n = (size_t)o->grid->imax * o->grid->jmax * o->grid->kmax;
  for (size_t q = 0; q < n; q++) {
    if (o->fmt == BINARY) {
      double b = swapBytes(s[q]);
      fwrite(&b, sizeof(double), 1, o->fh);
    } else {
      fprintf(o->fh, "%f\n", s[q]);
    }
  }
  if (o->fmt == BINARY) fprintf(o->fh, "\n");
}

void vtkVector(VtkOptions *o, char *name, VtkVector vec)
{
  printf("Register vector %s\n", name);
  if (!vtkReady(o)) return;
  fprintf(o->fh, "VECTORS %s double\n", name);
  size_t n = (size_t)o->grid->imax * o->grid->jmax * o->grid->kmax;
  for (size_t q = 0; q < n; q++) {
    if (o->fmt == BINARY) {
      double b[3] = {swapBytes(vec.u[q]), swapBytes(vec.v[q]), swapBytes(vec.w[q])};
      fwrite(b, sizeof(double), 3, o->fh);
    } else {
      fprintf(o->fh, "%f %f %f\n", vec.u[q], vec.v[q], vec.w[q]);
    }
  }
  if (o->fmt == BINARY) fprintf(o->fh, "\n");
}

void vtkClose(VtkOptions *o)
{
  fclose(o->fh);
  o->fh = NULL;
}
