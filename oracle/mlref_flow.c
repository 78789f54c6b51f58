/* CPU oracle (TEST INFRASTRUCTURE ONLY): the state update either side of the
 * implicit solves and the cell-centred velocity closure of a
 * Darcy-Brinkman step on one single-box level.
 *
 *   mlr_enthalpy_vars   updateEnthalpyVariables (src/phaseDiagram.cpp:92-177)
 *   mlr_ustar_src       computeUstarSrc (srcSubcycle/AMRLevelMushyLayerVel.cpp
 *                       :1724-1853) with fillBuoyancy
 *                       (srcSubcycle/AMRLevelMushyLayerFill.cpp:870-885) and
 *                       fillPressureSrcTerm (:787-807)
 *   mlr_cc_project_rhs  the velocity before the cell-centred projection and the
 *                       projection's right-hand side (computeCCvelocity,
 *                       srcSubcycle/AMRLevelMushyLayerVel.cpp:673-917:
 *                       removeGradPressure :654-671, setVelZero, then
 *                       Projector::levelProject's div(u)/dt,
 *                       srcSubcycle/Projector.cpp:1379-1397)
 *   mlr_pressure_faces  the projection's face coefficient (arithmetic face
 *                       mean of the porosity at the half time,
 *                       srcSubcycle/AMRLevelMushyLayerVel.cpp:717-731)
 *   mlr_cc_correct      applyCCcorrection (srcSubcycle/Projector.cpp:1632-1683)
 *                       and the final setVelZero
 *
 * Fields: `vars` holds 4 components (theta, Theta_l, Theta_s, porosity);
 * velocity components and sources are separate 1-component arrays; the
 * pressure `pi` is 1 component whose ghosts are filled here.
 *
 * Build spec (Chombo, not in the reference tree; parity unpinned against it):
 * the time-centred value of a field at t^n + dt/2 is 0.5*x^n + 0.5*x^{n+1}
 * (Chombo timeInterp), and CellToEdge's arithmetic face value is
 * 0.5*(x_lo + x_hi). */
#include <math.h>
#include <stdlib.h>

#include "mlref_internal.h"

/* CALCULATE_BOUNDING_ENERGY (src/EnthalpyVariablesF.ChF:657-698), then per
 * regime CALCULATEPOROSITY (:13-50), CALCULATECL (:156-196), CALCULATECS
 * (:201-241), CALCULATET (:246-287).  NaN when no regime matches (a NaN
 * input), where the Fortran leaves the value unset. */
void mlr_enthalpy_vars_point(double H, double C, const mlr_params* p, double* out) {
  double Hs, He, Hl;
  mlr_bounding_energy(H, C, p, &Hs, &He, &Hl);
  const double CR = p->compositionRatio, pc = p->waterDistributionCoeff, c = p->specificHeatRatio;
  const double St = p->stefan, te = p->thetaEutectic, Te = p->ThetaEutectic;
  double T = NAN, Cl = NAN, Cs = NAN, chi = NAN;
  if (H <= Hs) {
    chi = 0;
    Cl = Te;
    Cs = C;
    T = H / c;
  } else if (H > Hs && H <= He) {
    chi = (H - te * c) / (St + te * (1 - c));
    Cl = Te;
    double pm = mlr_porosity_mush(C, H, p);
    Cs = C / (1 - pm);
    T = te;
  } else if (H > He && H < Hl) {
    double pm = mlr_porosity_mush(C, H, p);
    chi = pm;
    Cl = (C + CR * (1 - pm)) / (pm + pc * (1 - pm));
    Cs = (C * pc - CR * pm) / (pm + pc * (1 - pm));
    T = -(C + CR * (1 - pm)) / (pm + pc * (1 - pm));
  } else if (H >= Hl) {
    chi = 1;
    Cl = C;
    Cs = 0;
    T = H - St;
  }
  out[0] = T;
  out[1] = Cl;
  out[2] = Cs;
  out[3] = chi;
}

void mlr_enthalpy_vars(const mlr_grid* g, const mlr_params* p, const double* phi, double* vars) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        long q = mlr_ci(&L, i, j, k);
        double o[4];
        mlr_enthalpy_vars_point(phi[q], phi[q + L.ncell], p, o);
        for (int c = 0; c < 4; ++c) vars[q + c * L.ncell] = o[c];
      }
}

/* ---- ghost fills -------------------------------------------------------- */

/* Pressure ghosts: periodic exchange, then on every wall the second-order
 * extrapolation 3 p1 - 3 p2 + p3 (DomainExtrapBCFunction /
 * BasicGradPressureBCFunction with order 2, BCutil/PhysBCUtil.cpp:2524-2536,
 * :2628-2633; EXTRAPBC, BCutil/BCFuncF.ChF:97-156).  Corners are not used. */
static void extrap_bc(const mlr_grid* g, double* x) {
  mlr_lay L;
  mlr_layout(g, &L);
  mlr_exchange(g, 1, x);
  for (int d = 0; d < g->dim; ++d) {
    if (g->periodic[d]) continue;
    int hi[3] = {g->n[0] - 1, g->n[1] - 1, g->dim == 3 ? g->n[2] - 1 : 0};
    for (int side = 0; side < 2; ++side) {
      int lo2[3] = {0, 0, 0}, hi2[3] = {hi[0], hi[1], hi[2]};
      int isign = side ? 1 : -1;
      int gi = side ? g->n[d] : -1;
      lo2[d] = hi2[d] = 0;
      for (int k = lo2[2]; k <= hi2[2]; ++k)
        for (int j = lo2[1]; j <= hi2[1]; ++j)
          for (int i = lo2[0]; i <= hi2[0]; ++i) {
            int a[3] = {i, j, k};
            a[d] = gi;
            int b1[3] = {a[0], a[1], a[2]}, b2[3] = {a[0], a[1], a[2]}, b3[3] = {a[0], a[1], a[2]};
            b1[d] -= isign;
            b2[d] -= 2 * isign;
            b3[d] -= 3 * isign;
            double p1 = x[mlr_ci(&L, b1[0], b1[1], b1[2])];
            double p2 = x[mlr_ci(&L, b2[0], b2[1], b2[2])];
            double p3 = x[mlr_ci(&L, b3[0], b3[1], b3[2])];
            x[mlr_ci(&L, a[0], a[1], a[2])] = 3 * p1 - 3 * p2 + p3;
          }
    }
  }
}

/* Velocity ghosts: periodic exchange, then on every wall (SolidWall, viscous)
 * Dirichlet 0 of order 1 for every component, g = 2*0 - u1
 * (BasicCCVelBCFunction, BCutil/PhysBCUtil.cpp:1577-1611). */
static void vel_bc(const mlr_grid* g, double* u) {
  mlr_bc bc;
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      for (int c = 0; c < 2; ++c) {
        bc.type[d][s][c] = 0;
        bc.val[d][s][c] = 0.0;
      }
  bc.apply = 1;
  mlr_exchange(g, 1, u);
  mlr_scalar_bc(g, 1, &bc, 0, u);
}

/* GRADCC (util/GradientF.ChF:194-219): (phi(i+e_d) - phi(i-e_d)) * (half/dx) */
static inline double grad_cc(const mlr_lay* L, const double* x, double factor, int d, int i, int j, int k) {
  int a[3] = {i, j, k}, b[3] = {i, j, k};
  a[d] += 1;
  b[d] -= 1;
  return factor * (x[mlr_ci(L, a[0], a[1], a[2])] - x[mlr_ci(L, b[0], b[1], b[2])]);
}

static inline double centred(double xo, double xn) { return 0.5 * xo + 0.5 * xn; }

/* ---- the U* source ------------------------------------------------------- */

/* src_d = (0 + buoyancy_d) - P_src_d at every valid cell, with the fields at
 * the half time (CCVelSrcTermCentering = 0.5):
 *   buoyancy_z = chi*(buoyTCoeff*theta - buoySCoeff*Theta_l + body force 0),
 *                zero in the other directions (fillBuoyancy);
 *   P_src_d   = GRADCC(pi)_d, times chi when scaleP (fillPressureSrcTerm,
 *                computeGradP srcSubcycle/AMRLevelMushyLayer.cpp:1041-1077
 *                with the extrapolated pressure ghosts).
 * Advection (u.grad u) does not enter (no advective source on this path). */
void mlr_ustar_src(const mlr_grid* g, double buoyT, double buoyS, int scaleP, const double* vold,
                   const double* vnew, double* pi, double* const* src) {
  mlr_lay L;
  mlr_layout(g, &L);
  extrap_bc(g, pi);
  const double factor = 0.5 / g->dx;
  const int gd = g->dim - 1;
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        long q = mlr_ci(&L, i, j, k);
        double th = centred(vold[q], vnew[q]);
        double thl = centred(vold[q + L.ncell], vnew[q + L.ncell]);
        double chi = centred(vold[q + 3 * L.ncell], vnew[q + 3 * L.ncell]);
        for (int d = 0; d < g->dim; ++d) {
          double buoy = 0.0;
          if (d == gd) {
            buoy = buoy + buoyT * th;
            buoy = buoy + (-buoyS) * thl;
            buoy = buoy + 1.0 * 0.0;
            buoy = buoy * chi;
          }
          double gp = grad_cc(&L, pi, factor, d, i, j, k);
          if (scaleP) gp = gp * chi;
          double s = 0.0;
          s = s + buoy;
          s = s - gp;
          src[d][q] = s;
        }
      }
}

/* ---- the cell-centred projection ---------------------------------------- */

/* u_d = ustar_d + (P_src_d * dt) (removeGradPressure); u = 0 where the new
 * porosity is <= limit (setVelZero, SETZEROVELOCITY src/utils.ChF:73-96);
 * velocity ghosts; rhs = div(u) * (1/dt) with div the MAC divergence
 * (DIVERGENCE, util/DivergenceF.ChF:22-49, summed over directions in order)
 * of the arithmetic face velocity (levelDivergenceCC,
 * util/Divergence.cpp:349-390).  u is written with its ghosts filled. */
void mlr_cc_project_rhs(const mlr_grid* g, double dt, double limit, int scaleP, const double* vold,
                        const double* vnew, double* pi, const double* const* ustar, double* const* u,
                        double* rhs) {
  mlr_lay L;
  mlr_layout(g, &L);
  extrap_bc(g, pi);
  const double factor = 0.5 / g->dx;
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        long q = mlr_ci(&L, i, j, k);
        double chi = centred(vold[q + 3 * L.ncell], vnew[q + 3 * L.ncell]);
        for (int d = 0; d < g->dim; ++d) {
          double gp = grad_cc(&L, pi, factor, d, i, j, k);
          if (scaleP) gp = gp * chi;
          gp = gp * dt;
          u[d][q] = ustar[d][q] + gp;
        }
        if (vnew[q + 3 * L.ncell] <= limit)
          for (int d = 0; d < g->dim; ++d) u[d][q] = 0;
      }
  for (int d = 0; d < g->dim; ++d) vel_bc(g, u[d]);
  const double one_on_dx = 1.0 / g->dx;
  const double dtScale = 1.0 / dt;
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        long q = mlr_ci(&L, i, j, k);
        double div = 0.0;
        for (int d = 0; d < g->dim; ++d) {
          int a[3] = {i, j, k}, b[3] = {i, j, k};
          a[d] += 1;
          b[d] -= 1;
          double ehi = 0.5 * (u[d][q] + u[d][mlr_ci(&L, a[0], a[1], a[2])]);
          double elo = 0.5 * (u[d][mlr_ci(&L, b[0], b[1], b[2])] + u[d][q]);
          div = div + one_on_dx * (ehi - elo);
        }
        rhs[q] = div * dtScale;
      }
}

/* b_d = -(face porosity at the half time): 0.5*(chi_lo + chi_hi) inside and
 * across periodic sides, the adjacent cell's value on walls (the porosity's
 * ghost equals the interior cell, and the face value is re-set from it). */
void mlr_pressure_faces(const mlr_grid* g, const double* vold, const double* vnew, double* const* b) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int d = 0; d < g->dim; ++d)
    for (int k = 0; k < (int)L.fn[d][2]; ++k)
      for (int j = 0; j < (int)L.fn[d][1]; ++j)
        for (int i = 0; i < (int)L.fn[d][0]; ++i) {
          int lo[3] = {i, j, k}, hi[3] = {i, j, k};
          lo[d] -= 1;
          int nd = g->n[d];
          double v;
          if (lo[d] < 0 && !g->periodic[d]) {
            long q = mlr_ci(&L, hi[0], hi[1], hi[2]);
            v = centred(vold[q + 3 * L.ncell], vnew[q + 3 * L.ncell]);
          } else if (hi[d] >= nd && !g->periodic[d]) {
            long q = mlr_ci(&L, lo[0], lo[1], lo[2]);
            v = centred(vold[q + 3 * L.ncell], vnew[q + 3 * L.ncell]);
          } else {
            if (lo[d] < 0) lo[d] += nd;
            if (hi[d] >= nd) hi[d] -= nd;
            long ql = mlr_ci(&L, lo[0], lo[1], lo[2]), qh = mlr_ci(&L, hi[0], hi[1], hi[2]);
            double cl = centred(vold[ql + 3 * L.ncell], vnew[ql + 3 * L.ncell]);
            double ch = centred(vold[qh + 3 * L.ncell], vnew[qh + 3 * L.ncell]);
            v = 0.5 * (cl + ch);
          }
          b[d][mlr_fi(&L, d, i, j, k)] = -v;
        }
}

/* u_d -= (GRADCC(pi)_d * chi) * (1/(1/dt)) (applyCCcorrection with
 * correctScale = 1/dtScale, srcSubcycle/Projector.cpp:1479-1481), then u = 0
 * where the new porosity is <= limit, then the velocity ghosts. */
void mlr_cc_correct(const mlr_grid* g, double dt, double limit, int scaleP, const double* vold,
                    const double* vnew, double* pi, double* const* u) {
  mlr_lay L;
  mlr_layout(g, &L);
  extrap_bc(g, pi);
  const double factor = 0.5 / g->dx;
  const double dtScale = 1.0 / dt;
  const double scale = 1 / dtScale;
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        long q = mlr_ci(&L, i, j, k);
        double chi = centred(vold[q + 3 * L.ncell], vnew[q + 3 * L.ncell]);
        for (int d = 0; d < g->dim; ++d) {
          double gp = grad_cc(&L, pi, factor, d, i, j, k);
          if (scaleP) gp = gp * chi;
          gp = gp * scale;
          u[d][q] = u[d][q] - gp;
        }
        if (vnew[q + 3 * L.ncell] <= limit)
          for (int d = 0; d < g->dim; ++d) u[d][q] = 0;
      }
  for (int d = 0; d < g->dim; ++d) vel_bc(g, u[d]);
}

/* x^3 rounded once (x*x = p + e and p*x = h + l exactly by FMA remainders),
 * the value of a correctly rounded pow(x, 3.0) */
static double cube_rn(double x) {
  const double p = x * x;
  const double e = fma(x, x, -p);
  const double h = p * x;
  const double l = fma(p, x, -h);
  return h + (l + e * x);
}

/* U* Darcy coefficient c = (chi / Pi(chi)) * coeff
 * (srcSubcycle/AMRLevelMushyLayerInit.cpp:665-689) with
 * MushyLayerParams::calculatePermeability (src/MushyLayerParams.cpp:1031-1090):
 * law 0 pure fluid 1, 1 cubic chi^3, 2 Kozeny-Carman chi^3/(1-chi)^2, 3 log
 * -chi^2 log(1-chi), 4 porosity chi; hele_shaw > 0: 1/(1/hele_shaw + 1/Pi).
 * pow(x, 3.0) and pow(x, 2.0) are the once-rounded cube and square. */
void mlr_darcy_ccoef(const mlr_grid* g, const double* chi, double coeff, int law, double hele_shaw, double* c) {
  mlr_lay L;
  mlr_layout(g, &L);
  for (int k = 0; k < L.nz; ++k)
    for (int j = 0; j < g->n[1]; ++j)
      for (int i = 0; i < g->n[0]; ++i) {
        long q = mlr_ci(&L, i, j, k);
        double liquid = chi[q], solid = 1 - liquid, perm;
        switch (law) {
          case 0: perm = 1; break;
          case 1: perm = cube_rn(liquid); break;
          case 3: perm = -pow(liquid, 2.0) * log(solid); break;
          case 4: perm = liquid; break;
          default: perm = cube_rn(liquid) / (solid * solid); break;
        }
        if (hele_shaw > 0) perm = 1 / (1 / hele_shaw + 1 / perm);
        double v = liquid;
        v = v / perm;
        c[q] = v * coeff;
      }
}
