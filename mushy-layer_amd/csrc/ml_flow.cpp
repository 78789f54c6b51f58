// ml_flow.cpp — C ABI of the level-wide state update (ml_hc_enthalpy_vars)
// and of the cell-centred velocity closure of a Darcy-Brinkman step
// (ml_vc_ustar_src, ml_vc_cc_project_rhs, ml_vc_pressure_faces,
// ml_vc_cc_correct, ml_vc_darcy_ccoef_law); kernels in ml_flow_kernels.hip.
#include "ml_internal.h"

namespace {

bool fields_ok(int dim, const double* const* f) {
  if (!f) return false;
  for (int d = 0; d < dim; ++d)
    if (!f[d]) return false;
  return true;
}

// the pressure stencils extrapolate from three cells inside a wall
int check_flow(ml_vc_op op, const ml_flow_params* fp, const char* who) {
  if (!fp) return ml::fail(op->ctx, ML_EINVAL, std::string(who) + ": null flow parameters");
  const ml::Geom& g = op->lev[0].g;
  for (int d = 0; d < g.dim; ++d)
    if (!g.periodic[d] && g.n[d] < 3)
      return ml::fail(op->ctx, ML_ENOTSUP, std::string(who) + ": a walled direction needs at least 3 cells");
  return ML_OK;
}

}  // namespace

extern "C" {

int ml_hc_enthalpy_vars(ml_hc_op op, const double* phi, double* vars) {
  if (!op || !phi || !vars) return ML_EINVAL;
  ML_HIP(op->ctx, ml::enthalpy_vars(op->lev[0].g, op->cp, phi, vars, op->ctx->stream));
  return ML_OK;
}

int ml_vc_ustar_src(ml_vc_op op, const ml_flow_params* fp, const double* vars_old, const double* vars_new,
                    const double* pi, double* const src[]) {
  if (!op || !vars_old || !vars_new || !pi || !fields_ok(op->dim, src)) return ML_EINVAL;
  int rc = check_flow(op, fp, "ml_vc_ustar_src");
  if (rc) return rc;
  ML_HIP(op->ctx, ml::ustar_src(op->lev[0].g, fp->buoyancyTCoeff, fp->buoyancySCoeff, fp->scalePressure, vars_old,
                                vars_new, pi, src, op->ctx->stream));
  return ML_OK;
}

int ml_vc_cc_project_rhs(ml_vc_op op, const ml_flow_params* fp, double dt, const double* vars_old,
                         const double* vars_new, const double* pi, const double* const ustar[], double* const u[],
                         double* rhs) {
  if (!op || !vars_old || !vars_new || !pi || !rhs || !fields_ok(op->dim, ustar) || !fields_ok(op->dim, u) ||
      !(dt > 0))
    return ML_EINVAL;
  for (int d = 0; d < op->dim; ++d)
    if (ustar[d] == u[d]) return ml::fail(op->ctx, ML_EINVAL, "ml_vc_cc_project_rhs: u must not alias ustar");
  int rc = check_flow(op, fp, "ml_vc_cc_project_rhs");
  if (rc) return rc;
  ML_HIP(op->ctx, ml::cc_project_rhs(op->lev[0].g, dt, fp->porosityLimit, fp->scalePressure, vars_old, vars_new, pi,
                                     ustar, u, rhs, op->ctx->stream));
  return ML_OK;
}

int ml_vc_pressure_faces(ml_vc_op op, const double* vars_old, const double* vars_new, double* const bCoef[]) {
  if (!op || !vars_old || !vars_new || !fields_ok(op->dim, bCoef)) return ML_EINVAL;
  ML_HIP(op->ctx, ml::pressure_faces(op->lev[0].g, vars_old, vars_new, bCoef, op->ctx->stream));
  return ML_OK;
}

int ml_vc_cc_correct(ml_vc_op op, const ml_flow_params* fp, double dt, const double* vars_old,
                     const double* vars_new, const double* pi, double* const u[]) {
  if (!op || !vars_old || !vars_new || !pi || !fields_ok(op->dim, u) || !(dt > 0)) return ML_EINVAL;
  int rc = check_flow(op, fp, "ml_vc_cc_correct");
  if (rc) return rc;
  ML_HIP(op->ctx, ml::cc_correct(op->lev[0].g, dt, fp->porosityLimit, fp->scalePressure, vars_old, vars_new, pi, u,
                                 op->ctx->stream));
  return ML_OK;
}

int ml_vc_darcy_ccoef_law(ml_vc_op op, const double* chi, double darcy_coeff, int law, double hele_shaw,
                          double* cCoef) {
  if (!op || !chi || !cCoef) return ML_EINVAL;
  if (law < ML_PERM_PURE_FLUID || law > ML_PERM_POROSITY)
    return ml::fail(op->ctx, ML_EINVAL, "ml_vc_darcy_ccoef_law: unknown permeability law");
  ML_HIP(op->ctx, ml::darcy_ccoef_law(op->lev[0].g, darcy_coeff, law, hele_shaw, chi, cCoef, op->ctx->stream));
  return ML_OK;
}

}  // extern "C"
