/*
 * mlref_hc.c — the nonlinear (H, Theta) operator of AMRNonLinearMultiCompOp,
 * its phase-diagram closure, the FAS V-cycle and the backward-Euler step.
 *
 * TEST INFRASTRUCTURE ONLY (see mlref.h).  Arithmetic is written in the
 * reference's evaluation order (Fortran left-to-right, no contraction).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mlref.h"
#include "mlref_internal.h"

/* ---- point closure ----------------------------------------------------- */

/* CALCULATE_BOUNDING_ENERGY (src/EnthalpyVariablesF.ChF:657-698). */
void mlr_bounding_energy(double H, double C, const mlr_params* p, double* Hs,
                         double* He, double* Hl) {
  (void)H;
  const double CR = p->compositionRatio, pc = p->waterDistributionCoeff;
  const double c = p->specificHeatRatio, St = p->stefan;
  const double te = p->thetaEutectic, Te = p->ThetaEutectic;
  double hs;
  if (C < -CR)
    hs = 0.0;
  else
    hs = c * ((-C - CR) / pc);
  if (hs < 0.0) hs = 0.0;
  hs = hs + te * c;
  double chie = (CR + C) / (CR + Te);
  *Hs = hs;
  *He = chie * (St + te * (1.0 - c)) + c * te;
  *Hl = St - C + te + Te;
}

/* CALCULATEPOROSITYMUSH (src/EnthalpyVariablesF.ChF:103-150): the mush root
 * of the porosity quadratic, Theta clamped to >= -CR inside the quadratic. */
double mlr_porosity_mush(double C, double H, const mlr_params* p) {
  const double CR = p->compositionRatio, pc = p->waterDistributionCoeff;
  const double c = p->specificHeatRatio, St = p->stefan;
  double validC = (C >= -CR) ? C : -CR;
  if (St == 0.0) return 1.0;
  double a = CR * (c - 1.0) + St * (pc - 1.0);
  double b = CR * (1.0 - 2.0 * c) + H * (1.0 - pc) - validC * (c - 1.0) - pc * St;
  double q = (CR + validC) * c + pc * H;
  return (-b - sqrt(b * b - 4.0 * a * q)) / (2.0 * a);
}

/* CALCULATE_T_CL (src/EnthalpyVariablesF.ChF:419-467): four regimes.  Returns
 * NaN when no branch matches (only for NaN inputs). */
void mlr_closure_point(double H, double C, const mlr_params* p, double* theta,
                       double* Cl) {
  double Hs, He, Hl;
  mlr_bounding_energy(H, C, p, &Hs, &He, &Hl);
  const double CR = p->compositionRatio, pc = p->waterDistributionCoeff;
  if (H <= Hs) {
    *theta = H / p->specificHeatRatio;
    *Cl = p->ThetaEutectic;
  } else if (H > Hs && H <= He) {
    *theta = p->thetaEutectic;
    *Cl = p->ThetaEutectic;
  } else if (H > He && H < Hl) {
    double chi = mlr_porosity_mush(C, H, p);
    *theta = -(C + CR * (1.0 - chi)) / (chi + pc * (1.0 - chi));
    *Cl = (C + CR * (1.0 - chi)) / (chi + pc * (1.0 - chi));
  } else if (H >= Hl) {
    *theta = H - p->stefan;
    *Cl = C;
  } else {
    *theta = NAN;
    *Cl = NAN;
  }
}

/* CALCULATEPOROSITY_MULTICOMP (src/EnthalpyVariablesF.ChF:56-94). */
double mlr_porosity_point(double H, double C, const mlr_params* p) {
  double Hs, He, Hl;
  mlr_bounding_energy(H, C, p, &Hs, &He, &Hl);
  const double c = p->specificHeatRatio, te = p->thetaEutectic;
  if (H <= Hs) return 0.0;
  if (H > Hs && H <= He) return (H - te * c) / (p->stefan + te * (1.0 - c));
  if (H > He && H < Hl) return mlr_porosity_mush(C, H, p);
  if (H >= Hl) return 1.0;
  return NAN;
}

/* ---- level operators ---------------------------------------------------- */

/* Closure over the whole grown box (valid + 1 ghost): region =
 * derivedVar.box() & Hs.box() (src/AMRNonLinearMultiCompOp.cpp:57-90). */
static void closure_box(const mlr_grid* g, const mlr_params* p,
                        const double* phi, double* phil) {
  mlr_lay L;
  mlr_layout(g, &L);
#pragma omp parallel for schedule(static) if (L.ncell > 4096)
  for (long q = 0; q < L.ncell; ++q) {
    double t, cl;
    mlr_closure_point(phi[q], phi[q + L.ncell], p, &t, &cl);
    phil[q] = t;
    phil[q + L.ncell] = cl;
  }
}

/* FAB form of computeDiffusedVar (src/AMRNonLinearMultiCompOp.cpp:57-98):
 * closure then the derived-variable BC (temperatureLiquidSalinityBC,
 * BCutil/PhysBCUtil.cpp:3741-3787, order 1). */
static void diffused_var_fab(const mlr_grid* g, const mlr_params* p,
                             const mlr_bc* bc, const double* phi, double* phil,
                             int homogeneous) {
  closure_box(g, p, phi, phil);
  if (bc->apply) mlr_scalar_bc(g, 2, bc, homogeneous, phil);
}

/* LevelData form (src/AMRNonLinearMultiCompOp.cpp:36-55): per box, then an
 * exchange of the derived variable. */
void mlr_hc_diffused_var(const mlr_grid* g, const mlr_params* p,
                         const mlr_bc* bc, double* phi, double* phil,
                         int homogeneous) {
  diffused_var_fab(g, p, bc, phi, phil, homogeneous);
  mlr_exchange(g, 2, phil);
}

/* The common L(phi) bracket of NLVCCOMPUTEOP / NLVCCOMPUTERES /
 * RESTRICTRESNLVC / NONLINEARSMOOTHING (src/AMRNonLinearMultiCompOpF.ChF
 * :125-160, :285-307, :407-429, :509-535):
 *   lofphi = alpha*a*phi - beta*(sum of b-weighted phi_l differences)*dxinv */
static inline double hc_lofphi(const mlr_lay* L, int dim, int n, int i, int j,
                               int k, double alpha, const double* a, double beta,
                               const double* const* b, const double* phi,
                               const double* phil, double dxinv) {
  long q = mlr_ci(L, i, j, k) + n * L->ncell;
  const double* pl = phil + n * L->ncell;
  double p0 = pl[mlr_ci(L, i, j, k)];
  double p1 = pl[mlr_ci(L, i + 1, j, k)];
  double p2 = pl[mlr_ci(L, i - 1, j, k)];
  double p3 = pl[mlr_ci(L, i, j + 1, k)];
  double p4 = pl[mlr_ci(L, i, j - 1, k)];
  const double* b0 = b[0] + n * L->nface[0];
  const double* b1 = b[1] + n * L->nface[1];
  double s = b0[mlr_fi(L, 0, i + 1, j, k)] * (p1 - p0) - b0[mlr_fi(L, 0, i, j, k)] * (p0 - p2) +
             b1[mlr_fi(L, 1, i, j + 1, k)] * (p3 - p0) - b1[mlr_fi(L, 1, i, j, k)] * (p0 - p4);
  if (dim == 3) {
    double p5 = pl[mlr_ci(L, i, j, k + 1)];
    double p6 = pl[mlr_ci(L, i, j, k - 1)];
    const double* b2 = b[2] + n * L->nface[2];
    s = s + b2[mlr_fi(L, 2, i, j, k + 1)] * (p5 - p0) - b2[mlr_fi(L, 2, i, j, k)] * (p0 - p6);
  }
  double aq = a ? a[q] : 1.0;
  return alpha * aq * phi[q] - beta * s * dxinv;
}

/* levelGSRB (src/AMRNonLinearMultiCompOp.cpp:815-949) for one iteration:
 * for pass in {0, 1}: exchange(phi); NoFlux BC; computeDiffusedVar (FAB form,
 * inhomogeneous); NONLINEARSMOOTHING pass (src/AMRNonLinearMultiCompOpF.ChF
 * :125-195): cells with (i+j+k+pass) even, both components, phi_l frozen
 * within the pass.  `scratch` holds 2*ncell doubles for phi_l. */
void mlr_hc_gsrb(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                 double alpha, const double* a, double beta,
                 const double* const* b, const double* lambda, double* phi,
                 const double* rhs, double* scratch) {
  mlr_lay L;
  mlr_layout(g, &L);
  double dxinv = 1.0 / (g->dx * g->dx);
  for (int pass = 0; pass <= 1; ++pass) {
    mlr_exchange(g, 2, phi);
    mlr_noflux_bc(g, 2, phi);
    diffused_var_fab(g, p, bc, phi, scratch, 0);
    /* the cells of one colour are independent: rows in parallel (identical
       results for any thread count) */
    for (int n = 0; n < 2; ++n)
      for (int k = 0; k < L.nz; ++k)
#pragma omp parallel for schedule(static) if (g->n[0] * g->n[1] > 4096)
        for (int j = 0; j < g->n[1]; ++j) {
          int imin = abs((j + k + pass) % 2);
          for (int i = imin; i < g->n[0]; i += 2) {
            double lof = hc_lofphi(&L, g->dim, n, i, j, k, alpha, a, beta, b, phi, scratch, dxinv);
            long q = mlr_ci(&L, i, j, k) + n * L.ncell;
            phi[q] = phi[q] - lambda[q] * (lof - rhs[q]);
          }
        }
  }
}

/* residualI (src/AMRNonLinearMultiCompOp.cpp:100-166) + NLVCCOMPUTERES
 * (src/AMRNonLinearMultiCompOpF.ChF:337-434): res = rhs - L(phi). */
void mlr_hc_residual(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                     double alpha, const double* a, double beta,
                     const double* const* b, double* phi, const double* rhs,
                     double* res, double* scratch) {
  mlr_lay L;
  mlr_layout(g, &L);
  double dxinv = 1.0 / (g->dx * g->dx);
  mlr_noflux_bc(g, 2, phi);
  mlr_exchange(g, 2, phi);
  diffused_var_fab(g, p, bc, phi, scratch, 0);
  for (int n = 0; n < 2; ++n)
    for (int k = 0; k < L.nz; ++k)
#pragma omp parallel for schedule(static) if (g->n[0] * g->n[1] > 4096)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          res[q] = rhs[q] - hc_lofphi(&L, g->dim, n, i, j, k, alpha, a, beta, b, phi, scratch, dxinv);
        }
}

/* applyOpI -> applyOpNoBoundary2 (src/AMRNonLinearMultiCompOp.cpp:228-300) +
 * NLVCCOMPUTEOP (src/AMRNonLinearMultiCompOpF.ChF:221-313). */
void mlr_hc_apply_op(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                     double alpha, const double* a, double beta,
                     const double* const* b, double* phi, double* lhs,
                     double* scratch) {
  mlr_lay L;
  mlr_layout(g, &L);
  double dxinv = 1.0 / (g->dx * g->dx);
  mlr_noflux_bc(g, 2, phi);
  mlr_exchange(g, 2, phi);
  mlr_hc_diffused_var(g, p, bc, phi, scratch, 0);
  for (int n = 0; n < 2; ++n)
    for (int k = 0; k < L.nz; ++k)
#pragma omp parallel for schedule(static) if (g->n[0] * g->n[1] > 4096)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          lhs[q] = hc_lofphi(&L, g->dim, n, i, j, k, alpha, a, beta, b, phi, scratch, dxinv);
        }
}

/* restrictResidual, FAS form (src/AMRNonLinearMultiCompOp.cpp:373-451) +
 * RESTRICTRESNLVC (src/AMRNonLinearMultiCompOpF.ChF:452-542):
 * res_c = 0; res_c(i/2) += (rhs - L(phi))/2^d, fine cells in Fortran order.
 * Inhomogeneous derived BC (build spec: the FAS coarse problem keeps the
 * real boundary values). */
void mlr_hc_restrict_residual(const mlr_grid* gf, const mlr_params* p,
                              const mlr_bc* bc, double alpha, const double* a,
                              double beta, const double* const* b, double* phi,
                              const double* rhs, double* resc, double* scratch) {
  mlr_grid gc;
  mlr_coarsen_grid(gf, 2, &gc);
  mlr_lay L, Lc;
  mlr_layout(gf, &L);
  mlr_layout(&gc, &Lc);
  double dxinv = 1.0 / (gf->dx * gf->dx);
  double denom = gf->dim == 3 ? 8.0 : 4.0;
  mlr_noflux_bc(gf, 2, phi);
  mlr_exchange(gf, 2, phi);
  mlr_hc_diffused_var(gf, p, bc, phi, scratch, 0);
  for (long q = 0; q < 2 * Lc.ncell; ++q) resc[q] = 0.0;
  /* a coarse row gathers fine rows 2jc and 2jc+1 only: coarse rows in
     parallel keep every coarse cell's summation order */
  for (int n = 0; n < 2; ++n)
    for (int k = 0; k < L.nz; ++k)
#pragma omp parallel for schedule(static) if (gf->n[0] * gf->n[1] > 4096)
      for (int jc = 0; jc < (gf->n[1] + 1) / 2; ++jc)
      for (int j = 2 * jc; j < 2 * jc + 2 && j < gf->n[1]; ++j)
        for (int i = 0; i < gf->n[0]; ++i) {
          long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          double lof = hc_lofphi(&L, gf->dim, n, i, j, k, alpha, a, beta, b, phi, scratch, dxinv);
          long qc = mlr_ci(&Lc, i / 2, j / 2, gf->dim == 3 ? k / 2 : 0) + n * Lc.ncell;
          resc[qc] = resc[qc] + (rhs[q] - lof) / denom;
        }
}

/* levelJacobi (src/AMRNonLinearMultiCompOp.cpp:1077-1101), one sweep:
 * resid = rhs - L(phi) (residualI, inhomogeneous); resid *= lambda;
 * phi += 0.5 * resid (incr with weight 0.5). */
void mlr_hc_jacobi(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc, double alpha, const double* a,
                   double beta, const double* const* b, const double* lambda, double* phi, const double* rhs,
                   double* res, double* scratch) {
  mlr_lay L;
  mlr_layout(g, &L);
  mlr_hc_residual(g, p, bc, alpha, a, beta, b, phi, rhs, res, scratch);
  for (int n = 0; n < 2; ++n)
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          double r = res[q] * lambda[q];
          phi[q] = phi[q] + 0.5 * r;
        }
}

/* preCond (src/AMRNonLinearMultiCompOp.cpp:172-206): phi = rhs * lambda,
 * then relax(phi, rhs, 2) (two levelGSRB iterations). */
void mlr_hc_precond(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc, double alpha, const double* a,
                    double beta, const double* const* b, const double* lambda, double* phi, const double* rhs,
                    double* scratch) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int n = 0; n < 2; ++n)
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          phi[q] = rhs[q] * lambda[q];
        }
  for (int it = 0; it < 2; ++it) mlr_hc_gsrb(g, p, bc, alpha, a, beta, b, lambda, phi, rhs, scratch);
}

/* getFlux (src/AMRNonLinearMultiCompOp.cpp:1103-1162) on every face of the
 * level, direction d, both components: flux = -b * ((phil_hi - phil_lo) * scale)
 * with scale = beta * ref / dx (ref = 1) and phil the diffused variable
 * (computeDiffusedVar with the inhomogeneous derived-variable BC, periodic
 * images across periodic sides). */
void mlr_hc_get_flux(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc, double beta,
                     const double* const* b, double* phi, double* const* flux, double* scratch) {
  mlr_lay L;
  mlr_layout(g, &L);
  mlr_noflux_bc(g, 2, phi);
  mlr_exchange(g, 2, phi);
  mlr_hc_diffused_var(g, p, bc, phi, scratch, 0);
  const double scale = beta * 1 / g->dx;
  for (int d = 0; d < g->dim; ++d)
    for (int n = 0; n < 2; ++n)
      for (int k = 0; k < (int)L.fn[d][2]; ++k)
        for (int j = 0; j < (int)L.fn[d][1]; ++j)
          for (int i = 0; i < (int)L.fn[d][0]; ++i) {
            int lo[3] = {i - (d == 0), j - (d == 1), k - (d == 2)};
            double phihi = scratch[mlr_ci(&L, i, j, k) + n * L.ncell];
            double philo = scratch[mlr_ci(&L, lo[0], lo[1], lo[2]) + n * L.ncell];
            double gradphi = (phihi - philo) * scale;
            long f = mlr_fi(&L, d, i, j, k) + n * L.nface[d];
            flux[d][f] = -b[d][f] * gradphi;
          }
}

/* ---- coefficient assembly ---------------------------------------------- */

/* Porosity of every valid cell (CALCULATEPOROSITY_MULTICOMP as used by
 * updateEnthalpyVariables, src/phaseDiagram.cpp:92-177). */
void mlr_porosity_field(const mlr_grid* g, const mlr_params* p,
                        const double* phi, double* chi) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        long q = mlr_ci(&L, i, j, k);
        chi[q] = mlr_porosity_point(phi[q], phi[q + L.ncell], p);
      }
}

/* HC bCoef assembly (srcSubcycle/AMRLevelMushyLayerInit.cpp:914-953):
 *   b_Theta = 1*chi_f;  b_H = ((1 - chi_f)*k + chi_f);  then
 *   b_H *= -D_H,  b_Theta *= -D_Theta,
 * with chi_f from CELLTOEDGEGEOMETRIC.  `chi` is a 1-component cell field
 * (ghosts are filled here); b[d] holds 2 components. */
void mlr_hc_bcoef(const mlr_grid* g, const mlr_params* p, double* chi,
                  double* const* b) {
  mlr_lay L;
  mlr_layout(g, &L);
  double* chif[3] = {0, 0, 0};
  for (int d = 0; d < g->dim; ++d) chif[d] = (double*)malloc(sizeof(double) * L.nface[d]);
  mlr_geometric_faces(g, chi, chif);
  for (int d = 0; d < g->dim; ++d) {
    for (long f = 0; f < L.nface[d]; ++f) {
      double cf = chif[d][f];
      double bC = 1.0 * cf;
      double bH = 1.0 - cf;
      bH = bH * p->heatConductivityRatio;
      bH = bH + cf;
      bH = bH * (-p->heatDiffusionCoeff);
      bC = bC * (-p->saltDiffusionCoeff);
      b[d][f] = bH;
      b[d][f + L.nface[d]] = bC;
    }
    free(chif[d]);
  }
}

/* ---- FAS solve (build spec for Chombo AMRFASMultiGrid) ----------------- */

void mlr_default_opts(mlr_opts* o) {
  o->nu1 = 1;
  o->nu2 = 4;
  o->bottom = 40;
  o->max_iter = 10;
  o->eps = 1e-10;
  o->hang = 1e-10;
  o->norm_thresh = 1e-10;
  o->bottom_solver = MLR_BOTTOM_RELAX;
  o->bottom_max_iter = 80;
}

typedef struct {
  mlr_grid g;
  long ncell;
  double *phi, *rhs, *phi0, *tmp, *scratch, *lambda, *a, *b[3];
  int own_phi;
} hc_level;

typedef struct {
  int nlev;
  hc_level lev[32];
  const mlr_params* p;
  const mlr_bc* bc;
  double alpha, beta;
  const mlr_opts* o;
} hc_hier;

static void hier_build(hc_hier* h, const mlr_grid* g, const double* a,
                       const double* const* b) {
  h->nlev = mlr_mg_depth(g);
  for (int l = 0; l < h->nlev; ++l) {
    hc_level* L = &h->lev[l];
    memset(L, 0, sizeof(*L));
    mlr_coarsen_grid(g, 1 << l, &L->g);
    L->ncell = mlr_ncell(&L->g);
    size_t nb = sizeof(double) * 2 * L->ncell;
    L->rhs = (double*)calloc(1, nb);
    L->phi0 = (double*)calloc(1, nb);
    L->tmp = (double*)calloc(1, nb);
    L->scratch = (double*)calloc(1, nb);
    L->lambda = (double*)calloc(1, nb);
    if (l > 0) {
      L->phi = (double*)calloc(1, nb);
      L->own_phi = 1;
    }
    for (int d = 0; d < g->dim; ++d)
      L->b[d] = (double*)calloc(1, sizeof(double) * 2 * mlr_nface(&L->g, d));
    if (l == 0) {
      for (int d = 0; d < g->dim; ++d)
        memcpy(L->b[d], b[d], sizeof(double) * 2 * mlr_nface(g, d));
      if (a) {
        L->a = (double*)calloc(1, nb);
        memcpy(L->a, a, nb);
      }
    } else {
      mlr_average_faces(g, 2, b, 1 << l, L->b);
      if (a) {
        L->a = (double*)calloc(1, nb);
        mlr_average_cells(g, 2, a, 1 << l, L->a);
      }
    }
    mlr_lambda(&L->g, 2, h->alpha, L->a, h->beta, (const double* const*)L->b, NULL, L->lambda);
  }
}

static void hier_free(hc_hier* h) {
  for (int l = 0; l < h->nlev; ++l) {
    hc_level* L = &h->lev[l];
    free(L->rhs);
    free(L->phi0);
    free(L->tmp);
    free(L->scratch);
    free(L->lambda);
    free(L->a);
    if (L->own_phi) free(L->phi);
    for (int d = 0; d < 3; ++d) free(L->b[d]);
  }
}

static void relax(hc_hier* h, int l, int iters) {
  hc_level* L = &h->lev[l];
  for (int it = 0; it < iters; ++it)
    mlr_hc_gsrb(&L->g, h->p, h->bc, h->alpha, L->a, h->beta, (const double* const*)L->b,
                L->lambda, L->phi, L->rhs, L->scratch);
}

/* FAS V-cycle (build spec of Chombo's AMRFASMultiGrid V-cycle on one level):
 * relax nu1; phi_c = R phi (kept as phi_c0); rhs_c = R(rhs - L phi) + L_c(phi_c);
 * recurse; phi += P(phi_c - phi_c0); relax nu2.  Coarsest: `bottom` GSRB
 * iterations (RelaxSolver with m_imax = 40,
 * srcSubcycle/AMRLevelMushyLayerInit.cpp:788,996-1000). */
static void vcycle(hc_hier* h, int l) {
  hc_level* L = &h->lev[l];
  if (l == h->nlev - 1) {
    relax(h, l, h->o->bottom);
    return;
  }
  hc_level* C = &h->lev[l + 1];
  relax(h, l, h->o->nu1);
  mlr_restrict_r(&L->g, 2, L->phi, C->phi);
  memcpy(C->phi0, C->phi, sizeof(double) * 2 * C->ncell);
  mlr_hc_restrict_residual(&L->g, h->p, h->bc, h->alpha, L->a, h->beta,
                           (const double* const*)L->b, L->phi, L->rhs, C->rhs, L->scratch);
  mlr_hc_apply_op(&C->g, h->p, h->bc, h->alpha, C->a, h->beta, (const double* const*)C->b,
                  C->phi, C->tmp, C->scratch);
  {
    mlr_lay Lc;
    mlr_layout(&C->g, &Lc);
    for (int n = 0; n < 2; ++n)
      for (int k = 0; k < Lc.nz; ++k)
        for (int j = 0; j < C->g.n[1]; ++j)
          for (int i = 0; i < C->g.n[0]; ++i) {
            long q = mlr_ci(&Lc, i, j, k) + n * Lc.ncell;
            C->rhs[q] = C->rhs[q] + C->tmp[q];
          }
  }
  vcycle(h, l + 1);
  {
    mlr_lay Lc;
    mlr_layout(&C->g, &Lc);
    for (long q = 0; q < 2 * Lc.ncell; ++q) C->tmp[q] = 0.0;
    for (int n = 0; n < 2; ++n)
      for (int k = 0; k < Lc.nz; ++k)
        for (int j = 0; j < C->g.n[1]; ++j)
          for (int i = 0; i < C->g.n[0]; ++i) {
            long q = mlr_ci(&Lc, i, j, k) + n * Lc.ncell;
            C->tmp[q] = C->phi[q] - C->phi0[q];
          }
  }
  mlr_prolong_linear(&L->g, 2, L->phi, C->tmp);
  relax(h, l, h->o->nu2);
}

/* Solve L(phi) = rhs by FAS V-cycles (build spec for AMRFASMultiGrid::solve).
 * r0 = |rhs - L phi|_inf.  Exit status is a bit set in the form Chombo's
 * multigrid reports it and the reference inspects
 * (srcSubcycle/AMRLevelMushyLayer.cpp:715-749, failure = 2, 4 or 6):
 *   1 |r| <= eps*r0,  2 iteration max reached unconverged,
 *   4 hang (|r| > (1-hang)*|r_prev|),  8 |r| <= norm_thresh. */
int mlr_hc_fas_solve(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                     double alpha, const double* a, double beta,
                     const double* const* b, double* phi, const double* rhs,
                     const mlr_opts* o, mlr_stats* st) {
  hc_hier h;
  h.p = p;
  h.bc = bc;
  h.alpha = alpha;
  h.beta = beta;
  h.o = o;
  hier_build(&h, g, a, b);
  hc_level* F = &h.lev[0];
  F->phi = phi;
  long nb = 2 * F->ncell;
  memcpy(F->rhs, rhs, sizeof(double) * nb);

  mlr_hc_residual(g, p, bc, alpha, F->a, beta, (const double* const*)F->b, phi, F->rhs, F->tmp, F->scratch);
  double r0 = mlr_norm_max(g, 2, F->tmp);
  double norm = r0, prev = r0;
  int iter = 0, status = 0;
  if (r0 <= o->norm_thresh) status = 8;
  while (status == 0 && iter < o->max_iter) {
    vcycle(&h, 0);
    iter++;
    mlr_hc_residual(g, p, bc, alpha, F->a, beta, (const double* const*)F->b, phi, F->rhs, F->tmp, F->scratch);
    norm = mlr_norm_max(g, 2, F->tmp);
    if (norm <= o->eps * r0) status |= 1;
    if (norm <= o->norm_thresh) status |= 8;
    if (status == 0 && norm > (1.0 - o->hang) * prev) status |= 4;
    prev = norm;
  }
  if (status == 0) status = 2;
  if (st) {
    st->exit_status = status;
    st->iterations = iter;
    st->initial_residual = r0;
    st->final_residual = norm;
  }
  F->phi = NULL;
  hier_free(&h);
  return status;
}

/* Backward-Euler step (build spec for Chombo LevelBackwardEuler::updateSoln
 * as driven from srcSubcycle/AMRLevelMushyLayer.cpp:1775-1778):
 * alpha = 1, beta = -dt; rhs = a*phi_old (diagonalScale,
 * src/AMRNonLinearMultiCompOp.H:75-83) + dt*src; solve from the caller's
 * phi_new (zeroPhi = false). */
int mlr_hc_be_step(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                   const double* a, const double* const* b, double dt,
                   double* phi_new, const double* phi_old, const double* src,
                   const mlr_opts* o, mlr_stats* st) {
  long nb = 2 * mlr_ncell(g);
  double* rhs = (double*)malloc(sizeof(double) * nb);
  for (long q = 0; q < nb; ++q) {
    double r = phi_old[q] * (a ? a[q] : 1.0);
    if (src) r = r + dt * src[q];
    rhs[q] = r;
  }
  int s = mlr_hc_fas_solve(g, p, bc, 1.0, a, -dt, b, phi_new, rhs, o, st);
  free(rhs);
  return s;
}

/* OpenMP threads of the parallel loops above (test/baseline control; the
 * results do not depend on it).  0 leaves the current setting. */
#ifdef _OPENMP
#include <omp.h>
#endif
int mlr_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
