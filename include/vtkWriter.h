/* Legacy-VTK writer for cell-centred results (reference src/vtkWriter.h:8-33,
 * sequential form src/vtkWriter-seq.c). */
#ifndef NUSIF_VTKWRITER_H
#define NUSIF_VTKWRITER_H
#include <stdio.h>
#include "comm.h"
#include "grid.h"
#ifdef __cplusplus
extern "C" {
#endif
typedef enum VtkFormat { ASCII = 0, BINARY } VtkFormat;

typedef struct VtkOptions {
  Grid *grid;
  FILE *fh;
  VtkFormat fmt;
  CommType comm;
} VtkOptions;

typedef struct VtkVector {
  double *u, *v, *w;
} VtkVector;

extern void vtkOpen(VtkOptions *opts, char *filename);
extern void vtkVector(VtkOptions *opts, char *name, VtkVector vec);
extern void vtkScalar(VtkOptions *opts, char *name, double *p);
extern void vtkClose(VtkOptions *opts);
#ifdef __cplusplus
}
#endif
#endif
