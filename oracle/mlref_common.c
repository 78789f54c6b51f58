/*
 * mlref_common.c — layout helpers, ghost fills, transfers, norms and
 * coefficient averaging shared by every operator of the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see mlref.h).  Build with -ffp-contract=off so
 * that every a*b+c below is two rounded operations, as written.
 */
#include "mlref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mlref_internal.h"

long mlr_ncell(const mlr_grid* g) {
  long nz = g->dim == 3 ? g->n[2] + 2 : 1;
  return (long)(g->n[0] + 2) * (long)(g->n[1] + 2) * nz;
}

long mlr_nface(const mlr_grid* g, int d) {
  long r = 1;
  for (int e = 0; e < g->dim; ++e) r *= (long)(g->n[e] + (e == d ? 1 : 0));
  return r;
}

/* MG depth rule: depth l exists while every box length is divisible by
 * 2^l * s_maxCoarse with s_maxCoarse = 2, i.e. the coarsest level keeps at
 * least 2 cells per direction (src/AMRNonLinearMultiCompOp.cpp:1316-1325;
 * s_maxCoarse is Chombo's AMRPoissonOp default 2 — build spec). */
int mlr_mg_depth(const mlr_grid* g) {
  int depth = 1;
  long coarsening = 2;
  for (;;) {
    int ok = 1;
    for (int d = 0; d < g->dim; ++d)
      if (g->n[d] % (coarsening * 2) != 0) ok = 0;
    if (!ok) break;
    depth++;
    coarsening *= 2;
  }
  return depth;
}

void mlr_coarsen_grid(const mlr_grid* f, int ratio, mlr_grid* c) {
  *c = *f;
  for (int d = 0; d < f->dim; ++d) c->n[d] = f->n[d] / ratio;
  if (f->dim == 2) c->n[2] = 1;
  c->dx = f->dx * (double)ratio;
}

/* Periodic ghost fill of a single-box level: Chombo LevelData::exchange with
 * a periodic ProblemDomain (halo one cell, src/AMRNonLinearMultiCompOp.cpp
 * :1235-1236).  Ghosts on non-periodic sides are left untouched. */
void mlr_exchange(const mlr_grid* g, int ncomp, double* x) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int c = 0; c < ncomp; ++c) {
    double* xc = x + c * L.ncell;
    for (int d = 0; d < g->dim; ++d) {
      if (!g->periodic[d]) continue;
      int nd = g->n[d];
      /* loop over the valid range of the other directions */
      int lo[3] = {0, 0, 0}, hi[3] = {g->n[0] - 1, g->n[1] - 1, g->n[2] - 1};
      if (g->dim == 2) hi[2] = 0;
      /* transverse directions already filled periodically extend the copy
       * to their ghosts, so corners of doubly-periodic domains get filled */
      for (int e = 0; e < d; ++e)
        if (g->periodic[e]) { lo[e] = -1; hi[e] = g->n[e]; }
      lo[d] = hi[d] = 0;
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i) {
            int iv[3] = {i, j, k};
            int a[3] = {i, j, k}, b[3] = {i, j, k};
            a[d] = -1;      b[d] = nd - 1;
            xc[mlr_ci(&L, a[0], a[1], a[2])] = xc[mlr_ci(&L, b[0], b[1], b[2])];
            a[d] = nd;      b[d] = 0;
            xc[mlr_ci(&L, a[0], a[1], a[2])] = xc[mlr_ci(&L, b[0], b[1], b[2])];
            (void)iv;
          }
    }
  }
}

/* NoFluxBCFunction: homogeneous Neumann on every non-periodic domain side,
 * ghost = adjacent interior value over adjCellBox(valid, dir, side, 1)
 * (BCutil/PhysBCUtil.cpp:2295-2365; Chombo NeumBC with value 0). */
void mlr_noflux_bc(const mlr_grid* g, int ncomp, double* x) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int c = 0; c < ncomp; ++c) {
    double* xc = x + c * L.ncell;
    for (int d = 0; d < g->dim; ++d) {
      if (g->periodic[d]) continue;
      int hi[3] = {g->n[0] - 1, g->n[1] - 1, g->dim == 3 ? g->n[2] - 1 : 0};
      for (int side = 0; side < 2; ++side) {
        int lo2[3] = {0, 0, 0}, hi2[3] = {hi[0], hi[1], hi[2]};
        int gidx = side ? g->n[d] : -1, iidx = side ? g->n[d] - 1 : 0;
        lo2[d] = hi2[d] = 0;
        for (int k = lo2[2]; k <= hi2[2]; ++k)
          for (int j = lo2[1]; j <= hi2[1]; ++j)
            for (int i = lo2[0]; i <= hi2[0]; ++i) {
              int a[3] = {i, j, k}, b[3] = {i, j, k};
              a[d] = gidx;
              b[d] = iidx;
              xc[mlr_ci(&L, a[0], a[1], a[2])] = xc[mlr_ci(&L, b[0], b[1], b[2])];
            }
      }
    }
  }
}

/* Dirichlet (order 1) / Neumann ghost fill per side and component:
 *   Dirichlet  g = 2*v - x1                 (BCutil/BCFuncF.ChF:40-46)
 *   Neumann    g = x1 + isign*dx*v          (BCutil/BCFuncF.ChF:80-88)
 * over adjCellBox(domain, dir, side, 1) (BCutil/BCFunctions.cpp:101-161),
 * values zeroed when homogeneous. */
void mlr_scalar_bc(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                   int homogeneous, double* x) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int d = 0; d < g->dim; ++d) {
    if (g->periodic[d]) continue;
    for (int side = 0; side < 2; ++side) {
      int isign = side ? 1 : -1;
      for (int c = 0; c < ncomp; ++c) {
        double* xc = x + c * L.ncell;
        int type = bc->type[d][side][c];
        double v = homogeneous ? 0.0 : bc->val[d][side][c];
        int lo2[3] = {0, 0, 0};
        int hi2[3] = {g->n[0] - 1, g->n[1] - 1, g->dim == 3 ? g->n[2] - 1 : 0};
        int gidx = side ? g->n[d] : -1, iidx = side ? g->n[d] - 1 : 0;
        lo2[d] = hi2[d] = 0;
        for (int k = lo2[2]; k <= hi2[2]; ++k)
          for (int j = lo2[1]; j <= hi2[1]; ++j)
            for (int i = lo2[0]; i <= hi2[0]; ++i) {
              int a[3] = {i, j, k}, b[3] = {i, j, k};
              a[d] = gidx;
              b[d] = iidx;
              double x1 = xc[mlr_ci(&L, b[0], b[1], b[2])];
              double gv;
              if (type == 0)
                gv = 2.0 * v - x1;
              else
                gv = x1 + ((double)isign * g->dx) * v;
              xc[mlr_ci(&L, a[0], a[1], a[2])] = gv;
            }
      }
    }
  }
}

/* RESTRICT (src/AMRNonLinearMultiCompOpF.ChF:555-585) after phiCoarse.setVal(0)
 * (src/AMRNonLinearMultiCompOp.cpp:492-514): phi_c(i/2) += phi_f / 2^d, fine
 * cells visited in Fortran order (i fastest). */
void mlr_restrict_r(const mlr_grid* gf, int ncomp, const double* phif,
                    double* phic) {
  mlr_grid gc;
  mlr_coarsen_grid(gf, 2, &gc);
  mlr_lay Lf, Lc;
  mlr_layout(gf, &Lf);
  mlr_layout(&gc, &Lc);
  double denom = gf->dim == 3 ? 8.0 : 4.0;
  for (long q = 0; q < ncomp * Lc.ncell; ++q) phic[q] = 0.0;
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < Lf.nz; ++k)
      for (int j = 0; j < gf->n[1]; ++j)
        for (int i = 0; i < gf->n[0]; ++i) {
          long ic = mlr_ci(&Lc, i / 2, j / 2, gf->dim == 3 ? k / 2 : 0) + n * Lc.ncell;
          phic[ic] = phic[ic] + phif[mlr_ci(&Lf, i, j, k) + n * Lf.ncell] / denom;
        }
}

/* Fill every ghost (corners included) with the valid value at the
 * per-direction periodic image or, on walls, the nearest interior index:
 * the no-flux rule applied direction by direction. */
void mlr_fill_ghosts_clamp(const mlr_grid* g, int ncomp, double* x) {
  mlr_lay L;
  mlr_layout(g, &L);
  int lo2 = g->dim == 3 ? -1 : 0, hi2 = g->dim == 3 ? g->n[2] : 0;
  for (int c = 0; c < ncomp; ++c) {
    double* xc = x + c * L.ncell;
    for (int k = lo2; k <= hi2; ++k)
      for (int j = -1; j <= g->n[1]; ++j)
        for (int i = -1; i <= g->n[0]; ++i) {
          int iv[3] = {i, j, k}, src[3];
          int ghost = 0;
          for (int d = 0; d < 3; ++d) {
            int v = iv[d], nd = d < g->dim ? g->n[d] : 1;
            if (d >= g->dim) { src[d] = 0; continue; }
            if (v < 0 || v >= nd) {
              ghost = 1;
              if (g->periodic[d]) v = ((v % nd) + nd) % nd;
              else v = v < 0 ? 0 : nd - 1;
            }
            src[d] = v;
          }
          if (ghost) xc[mlr_ci(&L, i, j, k)] = xc[mlr_ci(&L, src[0], src[1], src[2])];
        }
  }
}

/* Linear prolongation of a coarse correction, ratio 2 (build spec taken from
 * PROLONG_JP, src/AMRProjectionOpF.ChF:15-77; it stands in for Chombo's
 * FORT_PROLONG_2 called at src/AMRNonLinearMultiCompOp.cpp:478).  The coarse
 * correction's ghosts are filled first by mlr_fill_ghosts_clamp (periodic
 * image, otherwise the homogeneous no-flux rule ghost = interior). */
void mlr_prolong_linear(const mlr_grid* gf, int ncomp, double* phif,
                        double* corrc) {
  mlr_grid gc;
  mlr_coarsen_grid(gf, 2, &gc);
  mlr_fill_ghosts_clamp(&gc, ncomp, corrc);
  mlr_lay Lf, Lc;
  mlr_layout(gf, &Lf);
  mlr_layout(&gc, &Lc);
  const int D = gf->dim;
  const double den = 1.0 / (D == 3 ? 64.0 : 16.0);
  const double fx1 = 3.0 * den, fx2 = 9.0 * den, fx3 = 27.0 * den;
  const double f0 = 1.0 * den;
  for (int k = 0; k < Lf.nz; ++k)
    for (int j = 0; j < gf->n[1]; ++j)
      for (int i = 0; i < gf->n[0]; ++i) {
        int ic = i / 2, jc = j / 2, kc = D == 3 ? k / 2 : 0;
        int ox = 2 * (i % 2) - 1, oy = 2 * (j % 2) - 1, oz = D == 3 ? 2 * (k % 2) - 1 : 0;
        for (int n = 0; n < ncomp; ++n) {
          const double* c = corrc + n * Lc.ncell;
          double* p = phif + n * Lf.ncell + mlr_ci(&Lf, i, j, k);
          if (D == 2) {
            *p = *p + fx2 * c[mlr_ci(&Lc, ic, jc, 0)] + f0 * c[mlr_ci(&Lc, ic + ox, jc + oy, 0)];
            *p = *p + fx1 * (c[mlr_ci(&Lc, ic + ox, jc, 0)] + c[mlr_ci(&Lc, ic, jc + oy, 0)]);
          } else {
            *p = *p + fx3 * c[mlr_ci(&Lc, ic, jc, kc)] +
                 f0 * c[mlr_ci(&Lc, ic + ox, jc + oy, kc + oz)];
            *p = *p + fx2 * (c[mlr_ci(&Lc, ic + ox, jc, kc)] + c[mlr_ci(&Lc, ic, jc + oy, kc)] +
                             c[mlr_ci(&Lc, ic, jc, kc + oz)]);
            *p = *p + fx1 * (c[mlr_ci(&Lc, ic + ox, jc + oy, kc)] +
                             c[mlr_ci(&Lc, ic, jc + oy, kc + oz)] +
                             c[mlr_ci(&Lc, ic + ox, jc, kc + oz)]);
          }
        }
      }
}

/* Piecewise-constant prolongation phi_f += e_c(i/2) (Chombo FORT_PROLONG as
 * called from src/AMRProjectionOp.cpp:91-95; build spec). */
void mlr_prolong_constant(const mlr_grid* gf, int ncomp, double* phif,
                          const double* corrc) {
  mlr_grid gc;
  mlr_coarsen_grid(gf, 2, &gc);
  mlr_lay Lf, Lc;
  mlr_layout(gf, &Lf);
  mlr_layout(&gc, &Lc);
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < Lf.nz; ++k)
      for (int j = 0; j < gf->n[1]; ++j)
        for (int i = 0; i < gf->n[0]; ++i) {
          long f = mlr_ci(&Lf, i, j, k) + n * Lf.ncell;
          phif[f] = phif[f] + corrc[mlr_ci(&Lc, i / 2, j / 2, gf->dim == 3 ? k / 2 : 0) + n * Lc.ncell];
        }
}

/* Arithmetic CoarseAverage of a cell field by `ratio` (used for aCoef at
 * src/AMRNonLinearMultiCompOp.cpp:1448-1458 and for aCoef/cCoef at
 * src/DarcyBrinkmanOp.cpp:977-1008).  Build spec for the order: each row of
 * fine cells (fixed j, k) is summed i fastest, the row sums are added in
 * (j, k) order, then multiplied by 1/ratio^d. */
void mlr_average_cells(const mlr_grid* gf, int ncomp, const double* af,
                       int ratio, double* ac) {
  mlr_grid gc;
  mlr_coarsen_grid(gf, ratio, &gc);
  mlr_lay Lf, Lc;
  mlr_layout(gf, &Lf);
  mlr_layout(&gc, &Lc);
  int rz = gf->dim == 3 ? ratio : 1;
  double scale = 1.0 / (double)(ratio * ratio * rz);
  for (int n = 0; n < ncomp; ++n)
    for (int K = 0; K < Lc.nz; ++K)
      for (int J = 0; J < gc.n[1]; ++J)
        for (int I = 0; I < gc.n[0]; ++I) {
          double s = 0.0;
          for (int kk = 0; kk < rz; ++kk)
            for (int jj = 0; jj < ratio; ++jj) {
              double rs = 0.0;
              for (int ii = 0; ii < ratio; ++ii)
                rs = rs + af[mlr_ci(&Lf, I * ratio + ii, J * ratio + jj, K * rz + kk) + n * Lf.ncell];
              s = s + rs;
            }
          ac[mlr_ci(&Lc, I, J, K) + n * Lc.ncell] = s * scale;
        }
}

/* Arithmetic CoarseAverageFace by `ratio` directly from the finest level
 * (src/AMRNonLinearMultiCompOp.cpp:1450-1469; build spec for the order: the
 * fine sub-faces of a coarse face summed with the lowest transverse index
 * fastest, then multiplied by 1/ratio^(d-1)). */
void mlr_average_faces(const mlr_grid* gf, int ncomp, const double* const* bf,
                       int ratio, double* const* bc) {
  mlr_grid gc;
  mlr_coarsen_grid(gf, ratio, &gc);
  mlr_lay Lf, Lc;
  mlr_layout(gf, &Lf);
  mlr_layout(&gc, &Lc);
  for (int d = 0; d < gf->dim; ++d) {
    int r[3] = {ratio, ratio, gf->dim == 3 ? ratio : 1};
    r[d] = 1;
    double scale = 1.0 / (double)(r[0] * r[1] * r[2]);
    int fn[3] = {gc.n[0], gc.n[1], gf->dim == 3 ? gc.n[2] : 1};
    fn[d] += 1;
    for (int n = 0; n < ncomp; ++n)
      for (int K = 0; K < fn[2]; ++K)
        for (int J = 0; J < fn[1]; ++J)
          for (int I = 0; I < fn[0]; ++I) {
            int base[3] = {I * (d == 0 ? ratio : r[0]), J * (d == 1 ? ratio : r[1]),
                           K * (d == 2 ? ratio : r[2])};
            double s = 0.0;
            for (int kk = 0; kk < r[2]; ++kk)
              for (int jj = 0; jj < r[1]; ++jj)
                for (int ii = 0; ii < r[0]; ++ii)
                  s = s + bf[d][mlr_fi(&Lf, d, base[0] + ii, base[1] + jj, base[2] + kk) + n * Lf.nface[d]];
            bc[d][mlr_fi(&Lc, d, I, J, K) + n * Lc.nface[d]] = s * scale;
          }
  }
}

/* Max norm over valid cells and all components (Chombo norm(x, 0); build
 * spec for the FAS stopping rule). */
double mlr_norm_max(const mlr_grid* g, int ncomp, const double* x) {
  mlr_lay L;
  mlr_layout(g, &L);
  double m = 0.0;
  int nan = 0;
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          double v = fabs(x[mlr_ci(&L, i, j, k) + n * L.ncell]);
          if (v != v) nan = 1;
          else if (v > m) m = v;
        }
  return nan ? NAN : m;
}

/* lambda = 1/(alpha*a [- beta*c] + sum_d scale*beta*(b_hi + b_lo)),
 * scale = 1/dx^2: AMRNonLinearMultiCompOp::resetLambda
 * (src/AMRNonLinearMultiCompOp.cpp:540-573) with SUMFACESNLVC
 * (src/AMRNonLinearMultiCompOpF.ChF:598-625); with c != NULL the
 * DarcyBrinkmanOp::resetLambda form (src/DarcyBrinkmanOp.cpp:265-296,
 * lambdaFab.plus(cCoef, -beta)).  a == NULL means aCoef == 1. */
void mlr_lambda(const mlr_grid* g, int ncomp, double alpha, const double* a,
                double beta, const double* const* b, const double* c,
                double* lambda) {
  mlr_lay L;
  mlr_layout(g, &L);
  double scale = 1.0 / (g->dx * g->dx);
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          double lam = (a ? a[q] : 1.0) * alpha;
          if (c) lam = lam + (-beta) * c[q];
          for (int d = 0; d < g->dim; ++d) {
            int ih = i + (d == 0), jh = j + (d == 1), kh = k + (d == 2);
            double s = b[d][mlr_fi(&L, d, ih, jh, kh) + n * L.nface[d]] +
                       b[d][mlr_fi(&L, d, i, j, k) + n * L.nface[d]];
            lam = lam + scale * beta * s;
          }
          lambda[q] = 1.0 / lam;
        }
}

/* CELLTOEDGEGEOMETRIC (src/CellToEdge2F.ChF:20-46):
 * face = sqrt(|chi_lo * chi_hi|), selected for porosity in fillScalarFace
 * (srcSubcycle/AMRLevelMushyLayerFill.cpp:454-458).  Build spec for domain
 * faces: chi ghosts are the periodic image or, on walls, the interior value
 * (zero-gradient porosity). */
void mlr_geometric_faces(const mlr_grid* g, double* chi, double* const* chif) {
  mlr_lay L;
  mlr_layout(g, &L);
  mlr_exchange(g, 1, chi);
  mlr_noflux_bc(g, 1, chi);
  for (int d = 0; d < g->dim; ++d) {
    int fn[3] = {g->n[0], g->n[1], L.nz};
    fn[d] += 1;
    for (int k = 0; k < fn[2]; ++k)
      for (int j = 0; j < fn[1]; ++j)
        for (int i = 0; i < fn[0]; ++i) {
          int il = i - (d == 0), jl = j - (d == 1), kl = k - (d == 2);
          double lo = chi[mlr_ci(&L, il, jl, kl)], hi = chi[mlr_ci(&L, i, j, k)];
          chif[d][mlr_fi(&L, d, i, j, k)] = sqrt(fabs(lo * hi));
        }
  }
}
