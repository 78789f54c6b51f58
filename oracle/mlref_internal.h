/* Internal layout helpers of the oracle (TEST INFRASTRUCTURE ONLY). */
#ifndef MLREF_INTERNAL_H
#define MLREF_INTERNAL_H

#include "mlref.h"

typedef struct {
  int dim;
  int nz;           /* valid cells in k (1 in 2D) */
  int gz;           /* ghost width in k (0 in 2D) */
  long sy, sz;      /* cell strides */
  long ncell;       /* doubles per component */
  long fn[3][3];    /* face array dims [dir][axis] */
  long nface[3];
} mlr_lay;

static inline void mlr_layout(const mlr_grid* g, mlr_lay* L) {
  L->dim = g->dim;
  L->nz = g->dim == 3 ? g->n[2] : 1;
  L->gz = g->dim == 3 ? 1 : 0;
  L->sy = g->n[0] + 2;
  L->sz = (long)(g->n[0] + 2) * (g->n[1] + 2);
  L->ncell = L->sz * (g->dim == 3 ? g->n[2] + 2 : 1);
  for (int d = 0; d < 3; ++d) {
    L->fn[d][0] = g->n[0] + (d == 0);
    L->fn[d][1] = g->n[1] + (d == 1);
    L->fn[d][2] = g->dim == 3 ? g->n[2] + (d == 2) : 1;
    L->nface[d] = L->fn[d][0] * L->fn[d][1] * L->fn[d][2];
  }
}

static inline long mlr_ci(const mlr_lay* L, int i, int j, int k) {
  return (long)(i + 1) + L->sy * (long)(j + 1) + L->sz * (long)(k + L->gz);
}

static inline long mlr_fi(const mlr_lay* L, int d, int i, int j, int k) {
  return (long)i + L->fn[d][0] * ((long)j + L->fn[d][1] * (long)k);
}

#endif
