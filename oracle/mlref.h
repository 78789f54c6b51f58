/*
 * mlref — CPU oracle for the single-level implicit mushy-layer solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's arithmetic for the implicit (H, Theta) FAS solve and the
 * variable-coefficient Helmholtz / Poisson solves.  It is the checker the
 * GPU path is compared with; it is linked only by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py.  The product library
 * (libmlhip.so) never calls it.
 *
 * Layout (Chombo FAB order, one ghost layer, component outermost):
 *   idx(i,j,k,c) = (i+1) + (n0+2)*((j+1) + (n1+2)*(k+gz)) + c*ncell
 * with gz = 1 in 3D and 0 in 2D (2D arrays have a single k plane).
 * Face arrays of direction d live on surroundingNodes(valid box, d), no
 * ghosts: dims n + e_d, face index i in direction d sits between cells
 * i-1 and i, component outermost.
 *
 * Every function cites the reference file:line it restates.  Parts of the
 * solver that live in Chombo (not in the reference tree) follow the build's
 * own written specification; those are marked "build spec" and are
 * "parity unpinned" against Chombo itself.
 */
#ifndef MLREF_H
#define MLREF_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int dim;          /* 2 or 3 */
  int n[3];         /* valid cells per direction; n[2] == 1 in 2D */
  int periodic[3];  /* 1 if periodic in that direction */
  double dx;        /* cell size (isotropic) */
} mlr_grid;

/* Physical parameters used by the closure and the coefficient assembly
 * (src/MushyLayerParams.h; defaults src/MushyLayerParams.cpp:266-270,
 * :340-365, :476-484, :736-753). */
typedef struct {
  double stefan;                 /* St */
  double compositionRatio;       /* CR */
  double waterDistributionCoeff; /* pc */
  double specificHeatRatio;      /* c */
  double heatConductivityRatio;  /* k */
  double thetaEutectic;          /* theta_e */
  double ThetaEutectic;          /* Theta_e */
  double heatDiffusionCoeff;     /* D_H */
  double saltDiffusionCoeff;     /* D_Theta = 1/Le */
} mlr_params;

/* Boundary conditions of a cell-centred field, per [dir][side][comp]
 * (side 0 = lo, 1 = hi; at most 2 components).  type 0 = Dirichlet
 * (order 1, BCutil/BCFuncF.ChF:14-56), 1 = Neumann (BCutil/BCFuncF.ChF:63-96).
 * Periodic directions ignore the entry.  For the HC operator this is the
 * derived-variable (theta, Theta_l) BC; `apply` mirrors
 * m_apply_bcs_to_diagnostic_var (src/AMRNonLinearMultiCompOp.cpp:92-96). */
typedef struct {
  int type[3][2][2];
  double val[3][2][2];
  int apply;
} mlr_bc;

/* Solver options (srcSubcycle/MushyLayerSubcycleUtils.cpp:664-674 defaults:
 * nu1 = 1, nu2 = 4, bottom = 40, max_iter = 10, eps = hang = norm_thresh =
 * 1e-10). */
typedef struct {
  int nu1, nu2, bottom, max_iter;
  double eps, hang, norm_thresh;
  /* linear MG only: after the `bottom` relaxations the coarsest level is
   * solved by BiCGStab (bottom_solver MLR_BOTTOM_BICGSTAB, at most
   * bottom_max_iter iterations) or not (MLR_BOTTOM_RELAX) */
  int bottom_solver, bottom_max_iter;
} mlr_opts;
enum { MLR_BOTTOM_RELAX = 0, MLR_BOTTOM_BICGSTAB = 1 };

typedef struct {
  int exit_status;       /* bit set: 1 |r| <= eps*r0, 2 iteration max,
                            4 hang, 8 |r| <= norm_thresh */
  int iterations;
  double initial_residual;
  double final_residual;
} mlr_stats;

/* ---- layout helpers ---------------------------------------------------- */
long mlr_ncell(const mlr_grid* g);            /* doubles per component */
long mlr_nface(const mlr_grid* g, int d);     /* faces per component */
int mlr_mg_depth(const mlr_grid* g);          /* number of MG levels */
void mlr_coarsen_grid(const mlr_grid* f, int ratio, mlr_grid* c);

/* ---- point closure ----------------------------------------------------- */
void mlr_bounding_energy(double H, double C, const mlr_params* p,
                         double* Hs, double* He, double* Hl);
double mlr_porosity_mush(double C, double H, const mlr_params* p);
void mlr_closure_point(double H, double C, const mlr_params* p,
                       double* theta, double* Cl);
double mlr_porosity_point(double H, double C, const mlr_params* p);

/* ---- ghost fills ------------------------------------------------------- */
void mlr_exchange(const mlr_grid* g, int ncomp, double* x);
void mlr_noflux_bc(const mlr_grid* g, int ncomp, double* x);
void mlr_scalar_bc(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                   int homogeneous, double* x);
void mlr_fill_ghosts_clamp(const mlr_grid* g, int ncomp, double* x);

/* ---- HC (nonlinear multi-component) level operators -------------------- */
void mlr_hc_diffused_var(const mlr_grid* g, const mlr_params* p,
                         const mlr_bc* bc, double* phi, double* phil,
                         int homogeneous);
void mlr_lambda(const mlr_grid* g, int ncomp, double alpha, const double* a,
                double beta, const double* const* b, const double* c,
                double* lambda);
void mlr_hc_gsrb(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                 double alpha, const double* a, double beta,
                 const double* const* b, const double* lambda, double* phi,
                 const double* rhs, double* scratch);
void mlr_hc_residual(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                     double alpha, const double* a, double beta,
                     const double* const* b, double* phi, const double* rhs,
                     double* res, double* scratch);
void mlr_hc_apply_op(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                     double alpha, const double* a, double beta,
                     const double* const* b, double* phi, double* lhs,
                     double* scratch);
void mlr_hc_restrict_residual(const mlr_grid* gf, const mlr_params* p,
                              const mlr_bc* bc, double alpha, const double* a,
                              double beta, const double* const* b,
                              double* phi, const double* rhs, double* resc,
                              double* scratch);

/* ---- transfers ---------------------------------------------------------- */
void mlr_restrict_r(const mlr_grid* gf, int ncomp, const double* phif,
                    double* phic);
void mlr_prolong_linear(const mlr_grid* gf, int ncomp, double* phif,
                        double* corrc);
void mlr_prolong_constant(const mlr_grid* gf, int ncomp, double* phif,
                          const double* corrc);
void mlr_average_cells(const mlr_grid* gf, int ncomp, const double* af,
                       int ratio, double* ac);
void mlr_average_faces(const mlr_grid* gf, int ncomp, const double* const* bf,
                       int ratio, double* const* bc);
double mlr_norm_max(const mlr_grid* g, int ncomp, const double* x);

/* ---- coefficient assembly ---------------------------------------------- */
void mlr_porosity_field(const mlr_grid* g, const mlr_params* p,
                        const double* phi, double* chi);
void mlr_geometric_faces(const mlr_grid* g, double* chi, double* const* chif);
void mlr_hc_bcoef(const mlr_grid* g, const mlr_params* p, double* chi,
                  double* const* b);

/* ---- FAS solve and backward-Euler step --------------------------------- */
void mlr_default_opts(mlr_opts* o);
int mlr_hc_fas_solve(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                     double alpha, const double* a, double beta,
                     const double* const* b, double* phi, const double* rhs,
                     const mlr_opts* o, mlr_stats* st);
int mlr_hc_be_step(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc,
                   const double* a, const double* const* b, double dt,
                   double* phi_new, const double* phi_old, const double* src,
                   const mlr_opts* o, mlr_stats* st);

/* ---- linear variable-coefficient Helmholtz / Poisson (DarcyBrinkmanOp,
 * AMRProjectionOp): L u = alpha*a*u - beta*c*u - beta*div(b grad u);
 * a, c nullable (a == NULL: aCoef == 1; c == NULL: no c term). ----------- */
void mlr_lin_gsrb(const mlr_grid* g, int ncomp, const mlr_bc* bc, double alpha,
                  const double* a, double beta, const double* c,
                  const double* const* b, const double* lambda, double* phi,
                  const double* rhs);
void mlr_lin_residual(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                      double alpha, const double* a, double beta,
                      const double* c, const double* const* b, double* phi,
                      const double* rhs, double* res, int homogeneous);
void mlr_lin_apply_op(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                      double alpha, const double* a, double beta,
                      const double* c, const double* const* b, double* phi,
                      double* lhs, int homogeneous);
void mlr_lin_restrict_residual(const mlr_grid* gf, int ncomp,
                               const mlr_bc* bc, double alpha, const double* a,
                               double beta, const double* c,
                               const double* const* b, double* phi,
                               const double* rhs, double* resc);
/* (x, y) over the valid cells in the fixed order of mlref_lin.c */
double mlr_lin_dot(const mlr_grid* g, int ncomp, const double* x, const double* y);
/* preCond: phi = lambda*rhs, then 2 GSRB iterations (homogeneous BCs) */
void mlr_lin_precond(const mlr_grid* g, int ncomp, const mlr_bc* bc, double alpha,
                     const double* a, double beta, const double* c,
                     const double* const* b, const double* lambda, double* phi,
                     const double* rhs);
/* BiCGStab on L phi = rhs with homogeneous BCs from the given phi; returns
 * the exit status (1 converged, 2 breakdown or iteration limit, 3 restart
 * limit), *iters and *final_norm (L2) when non-NULL */
int mlr_lin_bicgstab(const mlr_grid* g, int ncomp, const mlr_bc* bc, double alpha,
                     const double* a, double beta, const double* c,
                     const double* const* b, const double* lambda, double* phi,
                     const double* rhs, int imax, int* iters, double* final_norm);
int mlr_lin_mg_solve(const mlr_grid* g, int ncomp, const mlr_bc* bc,
                     double alpha, const double* a, double beta,
                     const double* c, const double* const* b, double* phi,
                     const double* rhs, int prolong_linear, const mlr_opts* o,
                     mlr_stats* st);

/* OpenMP threads of the oracle's parallel loops (0: unchanged); returns the
 * count in use.  Results do not depend on it. */
int mlr_set_threads(int n);

/* levelJacobi, preCond, getFlux of the (H, Theta) operator (mlref_hc.c) */
void mlr_hc_jacobi(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc, double alpha, const double* a,
                   double beta, const double* const* b, const double* lambda, double* phi, const double* rhs,
                   double* res, double* scratch);
void mlr_hc_precond(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc, double alpha, const double* a,
                    double beta, const double* const* b, const double* lambda, double* phi, const double* rhs,
                    double* scratch);
void mlr_hc_get_flux(const mlr_grid* g, const mlr_params* p, const mlr_bc* bc, double beta,
                     const double* const* b, double* phi, double* const* flux, double* scratch);

/* ---- state update and velocity closure of a Darcy-Brinkman step
 * (mlref_flow.c) ---------------------------------------------------------- */
void mlr_enthalpy_vars_point(double H, double C, const mlr_params* p, double* out);
void mlr_enthalpy_vars(const mlr_grid* g, const mlr_params* p, const double* phi, double* vars);
void mlr_ustar_src(const mlr_grid* g, double buoyT, double buoyS, int scaleP, const double* vold,
                   const double* vnew, double* pi, double* const* src);
void mlr_cc_project_rhs(const mlr_grid* g, double dt, double limit, int scaleP, const double* vold,
                        const double* vnew, double* pi, const double* const* ustar, double* const* u,
                        double* rhs);
void mlr_pressure_faces(const mlr_grid* g, const double* vold, const double* vnew, double* const* b);
void mlr_cc_correct(const mlr_grid* g, double dt, double limit, int scaleP, const double* vold,
                    const double* vnew, double* pi, double* const* u);
void mlr_darcy_ccoef(const mlr_grid* g, const double* chi, double coeff, int law, double hele_shaw, double* c);

/* ---- advective source of the (H, Theta) equations (mlref_advect.c) ------
 * Advection boundary conditions per [dir][side][comp] with AdvectIBC's
 * numbering (BCutil/AdvectIBC.H:137-144): 0 Dirichlet, 1 extrapolation,
 * 2 inflow/outflow, 4 Neumann (3, the plume inflow, is not supported). */
typedef struct {
  int type[3][2][2];
  double val[3][2][2];
} mlr_advect_bc;
void mlr_edge_to_cell(const mlr_grid* g, const double* const* uf, double* const* vc);
void mlr_face_velocity(const mlr_grid* g, const double* const* u, double* const* uf);
int mlr_advect_src(const mlr_grid* g, const mlr_advect_bc* bc, double dt, double advCoeff, const double* W,
                   const double* const* uf, const double* S, double* src, double* const* flux,
                   double* const* whalf);

#ifdef __cplusplus
}
#endif
#endif
