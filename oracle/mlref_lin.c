/*
 * mlref_lin.c — TEST INFRASTRUCTURE ONLY.  CPU restatement of the linear
 * variable-coefficient operators on the path:
 *
 *   L u = alpha*a*u - beta*c*u - beta*div(b grad u)
 *
 *   DarcyBrinkmanOp (src/DarcyBrinkmanOp.cpp, src/DarcyBrinkmanOpF.ChF):
 *     a = 1, b = -Pr, c = (Pr/Da) chi/Pi, BE alpha = 1, beta = -dt, one
 *     solve per velocity component;
 *   AMRProjectionOp over VCAMRPoissonOp2 (src/AMRProjectionOp.cpp:31-98):
 *     c absent, alpha = 0, beta = 1, b = -S_face; piecewise-constant
 *     prolongation.
 *
 * The operator arithmetic follows the DarcyBrinkmanOp Fortran line by line
 * (left-to-right evaluation, -ffp-contract=off).  With c == NULL the c term
 * is omitted, which is VCAMRPoissonOp2's form (Chombo, not in the reference
 * tree: build spec, parity unpinned against Chombo).  The linear multigrid
 * driver (Chombo AMRMultiGrid + MultiGrid::cycle) is build spec as well.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mlref_internal.h"

/* Ghost fill of a linear-op field: periodic exchange, then the scalar BC
 * (DarcyBrinkmanOp::residualI / levelGSRB order: m_bc then exchange,
 * src/DarcyBrinkmanOp.cpp:40-47, :526-539; the two commute on a single box
 * since BC ghosts and periodic ghosts are disjoint). */
static void lin_ghosts(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                       int homogeneous, double* phi) {
  mlr_scalar_bc(g, ncomp, bc, homogeneous, phi);
  mlr_exchange(g, ncomp, phi);
}

/* lofphi of DBCOMPUTEOP / GSRBHELMHOLTZDB (src/DarcyBrinkmanOpF.ChF:125-147,
 * :262-286): alpha*a*phi - beta*c*phi - beta*(S)*dxinv, S summed in the
 * Fortran order, c term omitted when c == NULL. */
static double lin_lof(const mlr_lay* L, int dim, int n, int i, int j, int k,
                      double alpha, const double* a, double beta,
                      const double* c, const double* const* b,
                      const double* phi, double dxinv) {
  const long q = mlr_ci(L, i, j, k) + n * L->ncell;
  const double p0 = phi[q];
  const double* b0 = b[0] + n * L->nface[0];
  const double* b1 = b[1] + n * L->nface[1];
  double s = b0[mlr_fi(L, 0, i + 1, j, k)] * (phi[mlr_ci(L, i + 1, j, k) + n * L->ncell] - p0) -
             b0[mlr_fi(L, 0, i, j, k)] * (p0 - phi[mlr_ci(L, i - 1, j, k) + n * L->ncell]) +
             b1[mlr_fi(L, 1, i, j + 1, k)] * (phi[mlr_ci(L, i, j + 1, k) + n * L->ncell] - p0) -
             b1[mlr_fi(L, 1, i, j, k)] * (p0 - phi[mlr_ci(L, i, j - 1, k) + n * L->ncell]);
  if (dim == 3) {
    const double* b2 = b[2] + n * L->nface[2];
    s = s + b2[mlr_fi(L, 2, i, j, k + 1)] * (phi[mlr_ci(L, i, j, k + 1) + n * L->ncell] - p0) -
        b2[mlr_fi(L, 2, i, j, k)] * (p0 - phi[mlr_ci(L, i, j, k - 1) + n * L->ncell]);
  }
  const double aq = a ? a[q] : 1.0;
  double lof = alpha * aq * p0;
  if (c) lof = lof - beta * c[q] * p0;
  return lof - beta * s * dxinv;
}

/* DarcyBrinkmanOp::levelGSRB (src/DarcyBrinkmanOp.cpp:503-576): for pass in
 * {0, 1}: exchange, homogeneous BC, GSRBHELMHOLTZDB on the cells with
 * (i+j+k+pass) even (src/DarcyBrinkmanOpF.ChF:110-153). */
void mlr_lin_gsrb(const mlr_grid* g, int ncomp, const mlr_bc* bc, double alpha,
                  const double* a, double beta, const double* c,
                  const double* const* b, const double* lambda, double* phi,
                  const double* rhs) {
  mlr_lay L;
  mlr_layout(g, &L);
  const double dxinv = 1.0 / (g->dx * g->dx);
  for (int pass = 0; pass <= 1; ++pass) {
    lin_ghosts(g, ncomp, bc, 1, phi);
    /* the cells of one colour are independent: rows in parallel */
    for (int n = 0; n < ncomp; ++n)
      for (int k = 0; k < L.nz; ++k)
#pragma omp parallel for schedule(static) if (g->n[0] * g->n[1] > 4096)
        for (int j = 0; j < g->n[1]; ++j) {
          const int imin = abs((j + k + pass) % 2);
          for (int i = imin; i < g->n[0]; i += 2) {
            const long q = mlr_ci(&L, i, j, k) + n * L.ncell;
            const double lof = lin_lof(&L, g->dim, n, i, j, k, alpha, a, beta, c, b, phi, dxinv);
            phi[q] = phi[q] - lambda[q] * (lof - rhs[q]);
          }
        }
  }
}

/* DarcyBrinkmanOp::residualI (src/DarcyBrinkmanOp.cpp:31-83) with
 * DBCOMPUTERES (src/DarcyBrinkmanOpF.ChF:290-400): res = rhs - L(phi). */
void mlr_lin_residual(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                      double alpha, const double* a, double beta,
                      const double* c, const double* const* b, double* phi,
                      const double* rhs, double* res, int homogeneous) {
  mlr_lay L;
  mlr_layout(g, &L);
  const double dxinv = 1.0 / (g->dx * g->dx);
  lin_ghosts(g, ncomp, bc, homogeneous, phi);
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < L.nz; ++k)
#pragma omp parallel for schedule(static) if (g->n[0] * g->n[1] > 4096)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          const long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          res[q] = rhs[q] - lin_lof(&L, g->dim, n, i, j, k, alpha, a, beta, c, b, phi, dxinv);
        }
}

/* DarcyBrinkmanOp::applyOpI (src/DarcyBrinkmanOp.cpp:123-137) with
 * DBCOMPUTEOP (src/DarcyBrinkmanOpF.ChF:160-288): lhs = L(phi). */
void mlr_lin_apply_op(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                      double alpha, const double* a, double beta,
                      const double* c, const double* const* b, double* phi,
                      double* lhs, int homogeneous) {
  mlr_lay L;
  mlr_layout(g, &L);
  const double dxinv = 1.0 / (g->dx * g->dx);
  lin_ghosts(g, ncomp, bc, homogeneous, phi);
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < L.nz; ++k)
#pragma omp parallel for schedule(static) if (g->n[0] * g->n[1] > 4096)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          const long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          lhs[q] = lin_lof(&L, g->dim, n, i, j, k, alpha, a, beta, c, b, phi, dxinv);
        }
}

/* DarcyBrinkmanOp::restrictResidual (src/DarcyBrinkmanOp.cpp:183-240) with
 * RESTRICTRESDB (src/DarcyBrinkmanOpF.ChF:402-509): homogeneous BC,
 * res_c zeroed, res_c(i/2) += (rhs - lofphi)/denom, denom = 2^d, fine cells
 * in Fortran order (i fastest). */
void mlr_lin_restrict_residual(const mlr_grid* gf, int ncomp,
                               const mlr_bc* bc, double alpha, const double* a,
                               double beta, const double* c,
                               const double* const* b, double* phi,
                               const double* rhs, double* resc) {
  mlr_grid gc;
  mlr_coarsen_grid(gf, 2, &gc);
  mlr_lay Lf, Lc;
  mlr_layout(gf, &Lf);
  mlr_layout(&gc, &Lc);
  const double dxinv = 1.0 / (gf->dx * gf->dx);
  const double denom = gf->dim == 3 ? 8.0 : 4.0;
  lin_ghosts(gf, ncomp, bc, 1, phi);
  for (long q = 0; q < ncomp * Lc.ncell; ++q) resc[q] = 0.0;
  /* a coarse row gathers fine rows 2jc and 2jc+1 only: coarse rows in
     parallel keep every coarse cell's summation order */
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < Lf.nz; ++k)
#pragma omp parallel for schedule(static) if (gf->n[0] * gf->n[1] > 4096)
      for (int jc = 0; jc < (gf->n[1] + 1) / 2; ++jc)
      for (int j = 2 * jc; j < 2 * jc + 2 && j < gf->n[1]; ++j)
        for (int i = 0; i < gf->n[0]; ++i) {
          const long q = mlr_ci(&Lf, i, j, k) + n * Lf.ncell;
          const double lof = lin_lof(&Lf, gf->dim, n, i, j, k, alpha, a, beta, c, b, phi, dxinv);
          const long qc = mlr_ci(&Lc, i / 2, j / 2, gf->dim == 3 ? k / 2 : 0) + n * Lc.ncell;
          resc[qc] = resc[qc] + (rhs[q] - lof) / denom;
        }
}

/* ---- BiCGStab bottom solver --------------------------------------------
 * The reference's bottom solver of the projection (Projector::
 * makeBottomSolvers, srcSubcycle/Projector.cpp:3004-3023: BiCGStab unless
 * projection.relax_bottom_solver, :58, with m_imax = bottomSolveMaxIter = 20,
 * :63) and of the U* solves (s_botSolverUStar,
 * srcSubcycle/AMRLevelMushyLayerInit.cpp:739, Chombo's default m_imax 80).
 * Chombo's BiCGStabSolver is not in the reference tree; build spec: van der
 * Vorst's preconditioned BiCGStab (SIAM J. Sci. Stat. Comput. 13 (1992)
 * 631-644) with restarts,
 *   r = rhs - L(phi) (homogeneous), r~ = r, e = 0, |.| the L2 norm, r0 = |r|
 *   while i < imax and |r| > eps |r_prev| and |r_prev| > 0:
 *     rho = (r~, r); rho == 0: stop
 *     p = r (after a (re)start) or p = beta p + (-beta omega) v + r,
 *         beta = (rho / rho_prev) (alpha / omega)
 *     p~ = preCond(p); v = L p~; m = (r~, v); alpha = rho / m
 *     |m| > small |rho|: r += (-alpha) v, e += alpha p~;  else r = 0
 *     |r| > eps r0 and |r| > reps r0: s~ = preCond(r); t = L s~;
 *         omega = (t, r) / (t, t); e += omega s~; r += (-omega) t
 *     |r| <= eps r0 or |r| <= reps r0: converged
 *     omega == 0 or |r| > (1 - hang) |r_prev|: every second time, phi += e
 *         and restart from the true residual (status 3 after 5 restarts)
 *   phi += e
 * with eps 1e-6, reps 1e-12, hang 1e-8, small 1e-30.  Vector updates are
 * y = y + s*x (one rounding each), scaling y = s*y. */
enum { DOT_LANES = 64 };

/* (x, y) over the valid cells (compact index c = cell + n*nvalid, x
 * fastest): 64 interleaved partial sums s_t = sum of x_c*y_c over c = t,
 * t+64, ... in increasing c, then the pairwise tree s_t = s_t + s_{t+h} for
 * h = P/2, ..., 2, 1, P the smallest power of two >= min(cells, 64) (build
 * spec: the order of the GPU's one-wavefront reduction; Chombo sums box by
 * box). */
double mlr_lin_dot(const mlr_grid* g, int ncomp, const double* x, const double* y) {
  mlr_lay L;
  mlr_layout(g, &L);
  double s[DOT_LANES];
  for (int t = 0; t < DOT_LANES; ++t) s[t] = 0.0;
  long c = 0;
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i, ++c) {
          const long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          s[c % DOT_LANES] = s[c % DOT_LANES] + x[q] * y[q];
        }
  int P = 1;
  while (P < DOT_LANES && P < c) P *= 2;
  for (int h = P / 2; h >= 1; h /= 2)
    for (int t = 0; t < h; ++t) s[t] = s[t] + s[t + h];
  return s[0];
}

/* DarcyBrinkmanOp::preCond (src/DarcyBrinkmanOp.cpp:89-121; VCAMRPoissonOp2's
 * has the same form): phi = rhs*lambda on the valid cells, then relax(phi,
 * rhs, 2) */
void mlr_lin_precond(const mlr_grid* g, int ncomp, const mlr_bc* bc, double alpha,
                     const double* a, double beta, const double* c,
                     const double* const* b, const double* lambda, double* phi,
                     const double* rhs) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int n = 0; n < ncomp; ++n)
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          const long q = mlr_ci(&L, i, j, k) + n * L.ncell;
          phi[q] = rhs[q] * lambda[q];
        }
  for (int it = 0; it < 2; ++it) mlr_lin_gsrb(g, ncomp, bc, alpha, a, beta, c, b, lambda, phi, rhs);
}

/* y = y + s*x and y = s*y over every stored value */
static void lin_incr(long n, double* y, const double* x, double s) {
  for (long q = 0; q < n; ++q) y[q] = y[q] + s * x[q];
}
static void lin_scale(long n, double* y, double s) {
  for (long q = 0; q < n; ++q) y[q] = s * y[q];
}

int mlr_lin_bicgstab(const mlr_grid* g, int ncomp, const mlr_bc* bc, double alpha,
                     const double* a, double beta, const double* c,
                     const double* const* b, const double* lambda, double* phi,
                     const double* rhs, int imax, int* iters, double* final_norm) {
  const double eps = 1e-6, reps = 1e-12, hang = 1e-8, small = 1e-30;
  const int max_restarts = 5;
  const long n = ncomp * mlr_ncell(g);
  double* w = (double*)calloc((size_t)(8 * n), sizeof(double));
  double *r = w, *rt = w + n, *e = w + 2 * n, *p = w + 3 * n, *pt = w + 4 * n, *st = w + 5 * n, *t = w + 6 * n,
         *v = w + 7 * n;
  mlr_lin_residual(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs, r, 1);
  memcpy(rt, r, sizeof(double) * n);
  double norm0 = sqrt(mlr_lin_dot(g, ncomp, r, r)), norm1 = norm0;
  const double r0 = norm0;
  double rho0 = 0, rho1 = 0, alpha0 = 0, alpha1 = 0, omega0 = 0, omega1 = 0;
  int init = 1, restarts = 0, recount = 0, i = 0, status = 0;
  while (i < imax && norm0 > eps * norm1 && norm1 > 0) {
    ++i;
    norm1 = norm0;
    alpha1 = alpha0;
    omega1 = omega0;
    rho1 = rho0;
    rho0 = mlr_lin_dot(g, ncomp, rt, r);
    if (rho0 == 0.0) {
      status = 2;
      break;
    }
    if (init) {
      memcpy(p, r, sizeof(double) * n);
      init = 0;
    } else {
      const double bt = (rho0 / rho1) * (alpha1 / omega1);
      lin_scale(n, p, bt);
      lin_incr(n, p, v, -bt * omega1);
      lin_incr(n, p, r, 1.0);
    }
    mlr_lin_precond(g, ncomp, bc, alpha, a, beta, c, b, lambda, pt, p);
    mlr_lin_apply_op(g, ncomp, bc, alpha, a, beta, c, b, pt, v, 1);
    const double m = mlr_lin_dot(g, ncomp, rt, v);
    alpha0 = rho0 / m;
    if (fabs(m) > small * fabs(rho0)) {
      lin_incr(n, r, v, -alpha0);
      norm0 = sqrt(mlr_lin_dot(g, ncomp, r, r));
      lin_incr(n, e, pt, alpha0);
    } else {
      memset(r, 0, sizeof(double) * n);
      norm0 = 0.0;
    }
    if (norm0 > eps * r0 && norm0 > reps * r0) {
      mlr_lin_precond(g, ncomp, bc, alpha, a, beta, c, b, lambda, st, r);
      mlr_lin_apply_op(g, ncomp, bc, alpha, a, beta, c, b, st, t, 1);
      omega0 = mlr_lin_dot(g, ncomp, t, r) / mlr_lin_dot(g, ncomp, t, t);
      lin_incr(n, e, st, omega0);
      lin_incr(n, r, t, -omega0);
      norm0 = sqrt(mlr_lin_dot(g, ncomp, r, r));
    }
    if (norm0 <= eps * r0 || norm0 <= reps * r0) {
      status = 1;
      break;
    }
    if (omega0 == 0.0 || norm0 > (1.0 - hang) * norm1) {
      if (recount == 0) {
        recount = 1;
      } else {
        recount = 0;
        lin_incr(n, phi, e, 1.0);
        if (restarts == max_restarts) {
          status = 3;
          memset(e, 0, sizeof(double) * n);
          break;
        }
        mlr_lin_residual(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs, r, 1);
        norm0 = sqrt(mlr_lin_dot(g, ncomp, r, r));
        rho0 = alpha0 = omega0 = 0.0;
        memcpy(rt, r, sizeof(double) * n);
        memset(e, 0, sizeof(double) * n);
        ++restarts;
        init = 1;
      }
    }
  }
  if (status == 0) status = 2;
  lin_incr(n, phi, e, 1.0);
  if (iters) *iters = i;
  if (final_norm) *final_norm = norm0;
  free(w);
  return status;
}

/* ---- linear multigrid (build spec for Chombo AMRMultiGrid::solve on one
 * AMR level and MultiGrid::cycle):
 *   res = rhs - L(phi) (inhomogeneous), r0 = |res|_inf
 *   per iteration: e = 0; V-cycle on L e = res with homogeneous BCs; phi += e;
 *                  res = rhs - L(phi); stop as the FAS solve does
 *   V-cycle on level l: relax(e, r, nu1); if l is not the coarsest:
 *     r_c = restrictResidual(e, r); e_c = 0; V-cycle(l+1); e += P e_c;
 *     relax(e, r, nu2);  else relax(e, r, bottom), then (bottom_solver
 *     BiCGStab) BiCGStab on L e = r from that e (MultiGrid's m_bottom
 *     smoothings before the bottom solve, srcSubcycle/Projector.cpp:3031-3032)
 *   Coefficients of level l are arithmetic averages of the finest level's
 *   (DarcyBrinkmanOpFactory::MGnewOp, src/DarcyBrinkmanOp.cpp:977-1008).
 */
typedef struct {
  mlr_grid g;
  double *a, *c, *b[3], *lam;
  double *e, *r;
} lin_level;

static void lin_vcycle(lin_level* lev, int l, int nlev, int ncomp,
                       const mlr_bc* bc, double alpha, double beta,
                       int prolong_linear, const mlr_opts* o) {
  lin_level* L = &lev[l];
  const double* const* bb = (const double* const*)L->b;
  if (l == nlev - 1) {
    for (int it = 0; it < o->bottom; ++it)
      mlr_lin_gsrb(&L->g, ncomp, bc, alpha, L->a, beta, L->c, bb, L->lam, L->e, L->r);
    if (o->bottom_solver == MLR_BOTTOM_BICGSTAB)
      mlr_lin_bicgstab(&L->g, ncomp, bc, alpha, L->a, beta, L->c, bb, L->lam, L->e, L->r, o->bottom_max_iter, NULL, NULL);
    return;
  }
  for (int it = 0; it < o->nu1; ++it)
    mlr_lin_gsrb(&L->g, ncomp, bc, alpha, L->a, beta, L->c, bb, L->lam, L->e, L->r);
  lin_level* C = &lev[l + 1];
  mlr_lin_restrict_residual(&L->g, ncomp, bc, alpha, L->a, beta, L->c, bb, L->e, L->r, C->r);
  memset(C->e, 0, sizeof(double) * ncomp * mlr_ncell(&C->g));
  lin_vcycle(lev, l + 1, nlev, ncomp, bc, alpha, beta, prolong_linear, o);
  if (prolong_linear)
    mlr_prolong_linear(&L->g, ncomp, L->e, C->e);
  else
    mlr_prolong_constant(&L->g, ncomp, L->e, C->e);
  for (int it = 0; it < o->nu2; ++it)
    mlr_lin_gsrb(&L->g, ncomp, bc, alpha, L->a, beta, L->c, bb, L->lam, L->e, L->r);
}

int mlr_lin_mg_solve(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                     double alpha, const double* a, double beta,
                     const double* c, const double* const* b, double* phi,
                     const double* rhs, int prolong_linear, const mlr_opts* o,
                     mlr_stats* st) {
  const int nlev = mlr_mg_depth(g);
  lin_level* lev = (lin_level*)calloc((size_t)nlev, sizeof(lin_level));
  for (int l = 0; l < nlev; ++l) {
    lin_level* L = &lev[l];
    if (l == 0)
      L->g = *g;
    else
      mlr_coarsen_grid(g, 1 << l, &L->g);
    const long nc = ncomp * mlr_ncell(&L->g);
    L->e = (double*)calloc((size_t)nc, sizeof(double));
    L->r = (double*)calloc((size_t)nc, sizeof(double));
    L->lam = (double*)calloc((size_t)nc, sizeof(double));
    if (l == 0) {
      L->a = (double*)a;
      L->c = (double*)c;
      for (int d = 0; d < 3; ++d) L->b[d] = d < g->dim ? (double*)b[d] : NULL;
    } else {
      L->a = a ? (double*)calloc((size_t)nc, sizeof(double)) : NULL;
      L->c = c ? (double*)calloc((size_t)nc, sizeof(double)) : NULL;
      if (a) mlr_average_cells(g, ncomp, a, 1 << l, L->a);
      if (c) mlr_average_cells(g, ncomp, c, 1 << l, L->c);
      for (int d = 0; d < 3; ++d)
        L->b[d] = d < g->dim ? (double*)calloc((size_t)(ncomp * mlr_nface(&L->g, d)), sizeof(double)) : NULL;
      mlr_average_faces(g, ncomp, b, 1 << l, L->b);
    }
    mlr_lambda(&L->g, ncomp, alpha, L->a, beta, (const double* const*)L->b, L->c, L->lam);
  }
  lin_level* F = &lev[0];
  const long n0 = ncomp * mlr_ncell(g);
  mlr_lin_residual(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs, F->r, 0);
  const double r0 = mlr_norm_max(g, ncomp, F->r);
  double norm = r0, prev = r0;
  int iter = 0, status = 0;
  if (r0 <= o->norm_thresh) status = 8;
  while (status == 0 && iter < o->max_iter) {
    memset(F->e, 0, sizeof(double) * n0);
    lin_vcycle(lev, 0, nlev, ncomp, bc, alpha, beta, prolong_linear, o);
    mlr_lay L;
    mlr_layout(g, &L);
    for (int n = 0; n < ncomp; ++n)
      for (int k = 0; k < L.nz; ++k)
        for (int j = 0; j < g->n[1]; ++j)
          for (int i = 0; i < g->n[0]; ++i) {
            const long q = mlr_ci(&L, i, j, k) + n * L.ncell;
            phi[q] = phi[q] + F->e[q];
          }
    ++iter;
    mlr_lin_residual(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs, F->r, 0);
    norm = mlr_norm_max(g, ncomp, F->r);
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
  for (int l = 0; l < nlev; ++l) {
    lin_level* L = &lev[l];
    free(L->e);
    free(L->r);
    free(L->lam);
    if (l > 0) {
      free(L->a);
      free(L->c);
      for (int d = 0; d < 3; ++d) free(L->b[d]);
    }
  }
  free(lev);
  return status;
}
