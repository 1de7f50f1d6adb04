/* Grid geometry of one NuSiF domain.
 * Layout-identical to the reference's Grid (src/grid.h:8-12): a caller compiled
 * against either header sees the same struct. */
#ifndef NUSIF_GRID_H
#define NUSIF_GRID_H

typedef struct {
  double dx, dy, dz;
  int imax, jmax, kmax;
  double xlength, ylength, zlength;
} Grid;

#endif
