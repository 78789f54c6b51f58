/* CPU oracle (TEST INFRASTRUCTURE ONLY): the advective source of the
 * (H, Theta) equations on one single-box level, the step before the
 * implicit solve (multiCompAdvectDiffuse -> computeScalarAdvectiveSrcHC,
 * srcSubcycle/AMRLevelMushyLayer.cpp:1645-1672, 2074-2104).
 *
 *   mlr_edge_to_cell   cell velocity from face velocities (EdgeToCell)
 *   mlr_advect_src     time-centred face values of W = (theta, Theta_l) by
 *                      the unsplit PLM Godunov predictor, fluxes
 *                      W_half * u_face * advectionCoeff, and
 *                      src = src - div(flux)
 *
 * What the reference tree holds:
 *   AdvectionPhysics (srcSubcycle/AdvectionPhysics.cpp): primitive =
 *     conserved (consToPrim :229-240), characteristic analysis/synthesis
 *     identities (:129-154), every wave speed the cell velocity in the sweep
 *     direction (charValues :156-178), no source increment (:180-189), the
 *     Riemann problem and quasilinearUpdate through RIEMANNF /
 *     QUASILINEARUPDATE (:191-227, :243-267), getFlux = W_half*u (:99-118);
 *   AdvectIBC::primBC (BCutil/AdvectIBC.cpp:130-260) with SOLIDBCF,
 *     NEUMANN, SOLIDEXTRAPBCF and ADVINFLOWOUTFLOW (BCutil/SolidF.ChF:14-109)
 *     for the faces on a physical wall; setBdrySlopes does nothing;
 *   the options (srcSubcycle/AMRLevelMushyLayerInit.cpp:497-537,
 *     srcSubcycle/MushyLayerSubcycleUtils.cpp:613,700-739): PLM normal
 *     predictor (HCNormalPredOrder 1), second-order slopes, van Leer
 *     limiting of the primitive variables, no flattening, no artificial
 *     viscosity, no higher-order limiter;
 *   computeScalarAdvectiveFlux (srcSubcycle/AMRLevelMushyLayer.cpp
 *     :1165-1198): the face values times the face velocity and
 *     m_advectionCoeff; the component loop multiplies every component by
 *     the coefficient once per component, so with two components the flux
 *     of component 0 is ((W*u)*coeff)*coeff and of component 1
 *     ((W*coeff)*u)*coeff (kept as the reference computes it);
 *   Divergence::levelDivergenceMACMultiComp (util/Divergence.cpp:19-60,
 *     util/DivergenceF.ChF:22-47) and a_src.plus(div, -1)
 *     (srcSubcycle/AMRLevelMushyLayer.cpp:2094-2101).
 *
 * Build spec restating Chombo (not in the reference tree; parity unpinned
 * against Chombo itself): PatchGodunov::computeWHalf (Chombo 3.2,
 * lib/src/AMRTimeDependent) is the corner-transport-upwind method of
 * Colella (J. Comput. Phys. 87, 1990) in 2D and of Saltzman (J. Comput.
 * Phys. 115, 1994) in 3D:
 *   slopes (GodunovUtilities::vanLeerSlopes): centred dW = 0.5*(dWr + dWl)
 *     with dWr = W[i+1]-W[i], dWl = W[i]-W[i-1], limited to
 *     sign(dW)*min(|dW|, 2|dWl|, 2|dWr|) and 0 where dWl*dWr < 0; one-sided
 *     unlimited differences in the first and last cell next to a wall;
 *   normal predictor (PLMNORMALPREDF): with lambda the cell velocity,
 *     lmin = min(lambda, 0), lmax = max(lambda, 0):
 *       lambda > 0: dW- = -0.5*(1 + dt/dx*lmin)*dW, dW+ = 0.5*(1 - dt/dx*lambda)*dW
 *       lambda < 0: dW- = -0.5*(1 + dt/dx*lambda)*dW, dW+ = 0.5*(1 - dt/dx*lmax)*dW
 *       else:       dW- = -0.5*(1 + dt/dx*lmin)*dW, dW+ = 0.5*(1 - dt/dx*lmax)*dW
 *     W-/W+ = (dW-/+ + W) + S*(0.5*dt) (the source added after the predictor);
 *   Riemann (RIEMANNF): the left state where u_face > 0, the right state
 *     where u_face < 0, 0.5*(left + right) where u_face == 0;
 *   quasilinear update (QUASILINEARUPDATE):
 *     AdWdx = (scale*u_cell)*(W_face[hi] - W_face[lo]);
 *   2D: W-/+[d1] += AdWdx(WHalf1[d2], scale = -(1/2)*dt/dx), then the final
 *     Riemann problem;
 *   3D: WHalf2[d1][d2] = Riemann(W-/+[d1] + AdWdx(WHalf1[d2], -(1/3)*dt/dx));
 *     W-/+[d1] += AdWdx(WHalf2[d2][d3], -(1/2)*dt/dx) for each d2 != d1 in
 *     increasing order (d3 the third direction), then the final Riemann
 *     problem;
 *   EdgeToCell: u_cell = 0.5*(u_face[lo] + u_face[hi]).
 * Every face of the domain is computed; in a periodic direction cell and
 * face indices wrap, which gives the values the reference computes on its
 * periodic ghost cells.
 *
 * Fields: W, S (diffusive source, may be NULL) and src are 2-component cell
 * arrays (ghosts are not read); u_face[d] are 1-component face arrays;
 * flux[d] and whalf[d] (optional outputs, NULL to skip) are 2-component
 * face arrays.  */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mlref_internal.h"

enum { ADV_DIRICHLET = 0, ADV_EXTRAP = 1, ADV_INFLOW_OUTFLOW = 2, ADV_PLUME = 3, ADV_NEUMANN = 4 };

typedef struct {
  const mlr_grid* g;
  mlr_lay L;
  int nv;       /* valid cells */
  double dt, dx, dtbydx;
  const double* W;
  const double* const* uf;
  const mlr_advect_bc* bc;
  double* vc[3];              /* cell velocity, compact valid-cell arrays */
  double* pred[3][2];         /* [d][0 = W-, 1 = W+] 2 comps, compact */
  double* h1[3];              /* WHalf1[d], 2 comps, face arrays */
  double* h2[3][3];           /* WHalf2[d1][d2] (3D) */
} adv_t;

/* compact valid-cell index with wrap in periodic directions */
static inline long cidx(const adv_t* A, int i, int j, int k) {
  const mlr_grid* g = A->g;
  int c[3] = {i, j, k};
  for (int d = 0; d < g->dim; ++d) {
    if (c[d] < 0) c[d] += g->n[d];
    if (c[d] >= g->n[d]) c[d] -= g->n[d];
  }
  return (long)c[0] + (long)g->n[0] * ((long)c[1] + (long)g->n[1] * (long)c[2]);
}

static inline long wcell(const adv_t* A, int i, int j, int k) {
  const mlr_grid* g = A->g;
  int c[3] = {i, j, k};
  for (int d = 0; d < g->dim; ++d) {
    if (c[d] < 0) c[d] += g->n[d];
    if (c[d] >= g->n[d]) c[d] -= g->n[d];
  }
  return mlr_ci(&A->L, c[0], c[1], c[2]);
}

/* face of direction d at face coordinate (i,j,k) (i..k of the face's lower
 * cell in the transverse directions may wrap) */
static inline long fidx(const adv_t* A, int d, int i, int j, int k) {
  const mlr_grid* g = A->g;
  int c[3] = {i, j, k};
  for (int e = 0; e < g->dim; ++e) {
    const int lim = g->n[e] + (e == d ? 1 : 0);
    if (e == d) {
      if (g->periodic[e]) {
        if (c[e] < 0) c[e] += g->n[e];
        if (c[e] > g->n[e]) c[e] -= g->n[e];
      }
    } else {
      if (c[e] < 0) c[e] += g->n[e];
      if (c[e] >= lim) c[e] -= g->n[e];
    }
  }
  return mlr_fi(&A->L, d, c[0], c[1], c[2]);
}

/* EdgeToCell: u_cell = 0.5*(u_face[lo] + u_face[hi]) */
void mlr_edge_to_cell(const mlr_grid* g, const double* const* uf, double* const* vc) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int d = 0; d < g->dim; ++d)
    for (int k = 0; k < L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          int h[3] = {i, j, k};
          h[d] += 1;
          vc[d][mlr_ci(&L, i, j, k)] = 0.5 * (uf[d][mlr_fi(&L, d, i, j, k)] + uf[d][mlr_fi(&L, d, h[0], h[1], h[2])]);
        }
}

/* limited (or one-sided) slope of component c in direction d */
static double slope(const adv_t* A, int d, int c, int i, int j, int k) {
  const mlr_grid* g = A->g;
  const double* W = A->W + (long)c * A->L.ncell;
  int p[3] = {i, j, k}, m[3] = {i, j, k};
  p[d] += 1;
  m[d] -= 1;
  const double w0 = W[wcell(A, i, j, k)];
  const int at = (d == 0 ? i : d == 1 ? j : k);
  if (!g->periodic[d] && at == 0) return W[wcell(A, p[0], p[1], p[2])] - w0;
  if (!g->periodic[d] && at == g->n[d] - 1) return w0 - W[wcell(A, m[0], m[1], m[2])];
  const double dWr = W[wcell(A, p[0], p[1], p[2])] - w0;
  const double dWl = w0 - W[wcell(A, m[0], m[1], m[2])];
  const double dWc = 0.5 * (dWr + dWl);
  double lim = fmin(2.0 * fabs(dWl), 2.0 * fabs(dWr));
  lim = fmin(lim, fabs(dWc));
  if (dWl * dWr < 0.0) lim = 0.0;
  return copysign(lim, dWc);
}

static void predict(adv_t* A, const double* S) {
  const mlr_grid* g = A->g;
  const double hdt = 0.5 * A->dt;
  for (int k = 0; k < A->L.nz; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        const long q = cidx(A, i, j, k), w = mlr_ci(&A->L, i, j, k);
        for (int d = 0; d < g->dim; ++d) {
          const double lam = A->vc[d][q];
          const double lmin = fmin(lam, 0.0), lmax = fmax(lam, 0.0);
          for (int c = 0; c < 2; ++c) {
            const double dW = slope(A, d, c, i, j, k);
            double dm, dp;
            if (lam > 0.0) {
              dm = -0.5 * (1.0 + A->dtbydx * lmin) * dW;
              dp = 0.5 * (1.0 - A->dtbydx * lam) * dW;
            } else if (lam < 0.0) {
              dm = -0.5 * (1.0 + A->dtbydx * lam) * dW;
              dp = 0.5 * (1.0 - A->dtbydx * lmax) * dW;
            } else {
              dm = -0.5 * (1.0 + A->dtbydx * lmin) * dW;
              dp = 0.5 * (1.0 - A->dtbydx * lmax) * dW;
            }
            const double wc = A->W[w + (long)c * A->L.ncell];
            double wm = dm + wc, wp = dp + wc;
            if (S) {
              const double sl = S[w + (long)c * A->L.ncell] * hdt;
              wm = wm + sl;
              wp = wp + sl;
            }
            A->pred[d][0][q + (long)c * A->nv] = wm;
            A->pred[d][1][q + (long)c * A->nv] = wp;
          }
        }
      }
}

/* the states either side of the faces of direction d, corrected by the
 * quasilinear updates of the stage */
typedef struct {
  int stage;      /* 1 WHalf1, 2 WHalf2[d][d2], 3 final */
  int d, d2;
} stage_t;

static double qlu(const adv_t* A, const double* wh, int e, double scale, int c, int i, int j, int k) {
  int h[3] = {i, j, k};
  h[e] += 1;
  const double diff = wh[fidx(A, e, h[0], h[1], h[2]) + (long)c * A->L.nface[e]] -
                      wh[fidx(A, e, i, j, k) + (long)c * A->L.nface[e]];
  return (scale * A->vc[e][cidx(A, i, j, k)]) * diff;
}

/* state side s (0: W-, 1: W+) of cell (i,j,k), component c, for stage st */
static double state(const adv_t* A, const stage_t* st, int s, int c, int i, int j, int k) {
  const int d = st->d;
  double w = A->pred[d][s][cidx(A, i, j, k) + (long)c * A->nv];
  if (st->stage == 1) return w;
  if (st->stage == 2) {
    const double s3 = -(1.0 / 3.0) * A->dt / A->dx;
    return w + qlu(A, A->h1[st->d2], st->d2, s3, c, i, j, k);
  }
  const double s2 = -(1.0 / 2.0) * A->dt / A->dx;
  if (A->g->dim == 2) {
    const int e = 1 - d;
    return w + qlu(A, A->h1[e], e, s2, c, i, j, k);
  }
  for (int e = 0; e < 3; ++e) {
    if (e == d) continue;
    const int e3 = 3 - d - e;
    w = w + qlu(A, A->h2[e][e3], e, s2, c, i, j, k);
  }
  return w;
}

static double riemann(double wl, double wr, double u) {
  if (u > 0.0) return wl;
  if (u < 0.0) return wr;
  return 0.5 * (wl + wr);
}

/* Godunov value of component c on face (i,j,k) of direction d, interior
 * formula (the face's lower cell is i-1 in direction d) */
static double face_interior(const adv_t* A, const stage_t* st, int c, int i, int j, int k) {
  const int d = st->d;
  int l[3] = {i, j, k};
  l[d] -= 1;
  const double wl = state(A, st, 1, c, l[0], l[1], l[2]);
  const double wr = state(A, st, 0, c, i, j, k);
  return riemann(wl, wr, A->uf[d][fidx(A, d, i, j, k)]);
}

static double face_value(const adv_t* A, const stage_t* st, int c, int i, int j, int k) {
  const mlr_grid* g = A->g;
  const int d = st->d;
  const int at = (d == 0 ? i : d == 1 ? j : k);
  if (g->periodic[d] || (at > 0 && at < g->n[d])) return face_interior(A, st, c, i, j, k);
  const int side = at == 0 ? 0 : 1;
  const int lohisign = side == 0 ? -1 : 1;
  const int type = A->bc->type[d][side][c];
  const double val = A->bc->val[d][side][c];
  int cell[3] = {i, j, k};
  if (side == 1) cell[d] -= 1;
  /* the extrapolated state: W- of the first cell, W+ of the last */
  const double wext = state(A, st, side == 0 ? 0 : 1, c, cell[0], cell[1], cell[2]);
  if (type == ADV_DIRICHLET) return val;
  if (type == ADV_EXTRAP) return wext;
  if (type == ADV_NEUMANN) {
    int in[3] = {i, j, k};
    in[d] -= lohisign;
    const double phi1 = face_interior(A, st, c, in[0], in[1], in[2]);
    return phi1 + ((double)lohisign * A->dx) * val;
  }
  /* ADV_INFLOW_OUTFLOW */
  const double inOut = A->uf[d][fidx(A, d, i, j, k)] * (double)lohisign;
  return inOut > 0.0 ? wext : val;
}

static void face_stage(const adv_t* A, const stage_t* st, double* out) {
  const mlr_grid* g = A->g;
  const int d = st->d;
  const long nf = A->L.nface[d];
  for (int k = 0; k < (int)A->L.fn[d][2]; ++k)
    for (int j = 0; j < (int)A->L.fn[d][1]; ++j)
      for (int i = 0; i < (int)A->L.fn[d][0]; ++i)
        for (int c = 0; c < 2; ++c) out[mlr_fi(&A->L, d, i, j, k) + c * nf] = face_value(A, st, c, i, j, k);
  (void)g;
}

int mlr_advect_src(const mlr_grid* g, const mlr_advect_bc* bc, double dt, double advCoeff, const double* W,
                   const double* const* uf, const double* S, double* src, double* const* flux,
                   double* const* whalf) {
  for (int d = 0; d < g->dim; ++d)
    for (int s = 0; s < 2; ++s)
      for (int c = 0; c < 2; ++c)
        if (!g->periodic[d] && (bc->type[d][s][c] < 0 || bc->type[d][s][c] > 4 || bc->type[d][s][c] == ADV_PLUME))
          return -1;
  adv_t A;
  memset(&A, 0, sizeof(A));
  A.g = g;
  mlr_layout(g, &A.L);
  A.nv = g->n[0] * g->n[1] * g->n[2];
  A.dt = dt;
  A.dx = g->dx;
  A.dtbydx = dt / g->dx;
  A.W = W;
  A.uf = uf;
  A.bc = bc;
  double* vcc[3] = {NULL, NULL, NULL};
  for (int d = 0; d < g->dim; ++d) {
    vcc[d] = (double*)malloc(sizeof(double) * A.L.ncell);
    A.vc[d] = (double*)malloc(sizeof(double) * A.nv);
    A.pred[d][0] = (double*)malloc(sizeof(double) * 2 * A.nv);
    A.pred[d][1] = (double*)malloc(sizeof(double) * 2 * A.nv);
    A.h1[d] = (double*)malloc(sizeof(double) * 2 * A.L.nface[d]);
  }
  mlr_edge_to_cell(g, uf, vcc);
  for (int d = 0; d < g->dim; ++d)
    for (int k = 0; k < A.L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) A.vc[d][cidx(&A, i, j, k)] = vcc[d][mlr_ci(&A.L, i, j, k)];
  predict(&A, S);
  for (int d = 0; d < g->dim; ++d) {
    stage_t st = {1, d, -1};
    face_stage(&A, &st, A.h1[d]);
  }
  if (g->dim == 3) {
    for (int d = 0; d < 3; ++d)
      for (int e = 0; e < 3; ++e) {
        if (e == d) continue;
        A.h2[d][e] = (double*)malloc(sizeof(double) * 2 * A.L.nface[d]);
        stage_t st = {2, d, e};
        face_stage(&A, &st, A.h2[d][e]);
      }
  }
  double* F[3] = {NULL, NULL, NULL};
  for (int d = 0; d < g->dim; ++d) {
    F[d] = (double*)malloc(sizeof(double) * 2 * A.L.nface[d]);
    stage_t st = {3, d, -1};
    face_stage(&A, &st, F[d]);
    const long nf = A.L.nface[d];
    if (whalf && whalf[d]) memcpy(whalf[d], F[d], sizeof(double) * 2 * nf);
    for (long f = 0; f < nf; ++f) {
      const double u = uf[d][f];
      F[d][f] = ((F[d][f] * u) * advCoeff) * advCoeff;
      F[d][f + nf] = ((F[d][f + nf] * advCoeff) * u) * advCoeff;
    }
    if (flux && flux[d]) memcpy(flux[d], F[d], sizeof(double) * 2 * nf);
  }
  const double one_on_dx = 1.0 / g->dx;
  for (int c = 0; c < 2; ++c)
    for (int k = 0; k < A.L.nz; ++k)
      for (int j = 0; j < g->n[1]; ++j)
        for (int i = 0; i < g->n[0]; ++i) {
          double div = 0.0;
          for (int d = 0; d < g->dim; ++d) {
            int h[3] = {i, j, k};
            h[d] += 1;
            const long nf = A.L.nface[d];
            div = div + one_on_dx * (F[d][mlr_fi(&A.L, d, h[0], h[1], h[2]) + c * nf] -
                                     F[d][mlr_fi(&A.L, d, i, j, k) + c * nf]);
          }
          const long q = mlr_ci(&A.L, i, j, k) + (long)c * A.L.ncell;
          src[q] = src[q] + (-1.0 * div);
        }
  for (int d = 0; d < 3; ++d) {
    free(vcc[d]);
    free(A.vc[d]);
    free(A.pred[d][0]);
    free(A.pred[d][1]);
    free(A.h1[d]);
    free(F[d]);
    for (int e = 0; e < 3; ++e) free(A.h2[d][e]);
  }
  return 0;
}

/* CellToEdge of a no-slip cell-centred velocity (1-component arrays):
 * interior and periodic faces 0.5*(u_lo + u_hi), wall faces 0 (the wall
 * ghost 2*0 - u1 averages to zero).  The advection velocity used by the
 * stepper in place of the reference's MAC-projected predictor. */
void mlr_face_velocity(const mlr_grid* g, const double* const* u, double* const* uf) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int d = 0; d < g->dim; ++d)
    for (int k = 0; k < (int)L.fn[d][2]; ++k)
      for (int j = 0; j < (int)L.fn[d][1]; ++j)
        for (int i = 0; i < (int)L.fn[d][0]; ++i) {
          int c[3] = {i, j, k}, lo[3] = {i, j, k};
          lo[d] -= 1;
          const long f = mlr_fi(&L, d, i, j, k);
          const int at = c[d];
          if (g->periodic[d]) {
            if (lo[d] < 0) lo[d] += g->n[d];
            if (c[d] >= g->n[d]) c[d] -= g->n[d];
            uf[d][f] = 0.5 * (u[d][mlr_ci(&L, lo[0], lo[1], lo[2])] + u[d][mlr_ci(&L, c[0], c[1], c[2])]);
          } else if (at == 0 || at == g->n[d]) {
            uf[d][f] = 0.0;
          } else {
            uf[d][f] = 0.5 * (u[d][mlr_ci(&L, lo[0], lo[1], lo[2])] + u[d][mlr_ci(&L, c[0], c[1], c[2])]);
          }
        }
}
