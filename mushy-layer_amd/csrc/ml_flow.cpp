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

// Batched Darcy-Brinkman step stages: one FlowJob per operator, workgroup
// maps over their cells and faces, staged as one table, then one launch per
// stage in the per-operator order.
struct FlowBatch {
  std::vector<ml::FlowJob> jobs;
  std::vector<unsigned> cmap, fmap[3];
  std::vector<ml::WordsJob> words;
  std::vector<unsigned> wmap;
  int dim = 0;

  ml::FlowJob& add(const ml::Geom& g) {
    ml::FlowJob J{};
    J.g = g;
    J.nbc = (g.nvalid() + 255) / 256;
    const unsigned hi = unsigned(jobs.size()) << ml::kBatchTileBits;
    for (long long b = 0; b < J.nbc; ++b) cmap.push_back(hi | unsigned(b));
    for (int d = 0; d < g.dim; ++d) {
      J.nbf[d] = (g.nface[d] + 255) / 256;
      for (long long b = 0; b < J.nbf[d]; ++b) fmap[d].push_back(hi | unsigned(b));
    }
    jobs.push_back(J);
    return jobs.back();
  }
  void copy(const double* src, double* dst, long long n) {
    const unsigned j = unsigned(words.size());
    words.push_back(ml::WordsJob{reinterpret_cast<const unsigned long long*>(src),
                                 reinterpret_cast<unsigned long long*>(dst), n});
    for (long long c = 0; c * ml::kWordsChunk < n; ++c) wmap.push_back((j << ml::kBatchTileBits) | unsigned(c));
  }
  // stages < 100: flow_batch stages; 100 + d: HC face coefficients of
  // direction d; 200: the copies / zero fills
  int run(ml_ctx ctx, const std::vector<int>& stages) {
    if (jobs.size() > (size_t(1) << (32 - ml::kBatchTileBits)) || words.size() > (size_t(1) << (32 - ml::kBatchTileBits)))
      return ml::fail(ctx, ML_ENOTSUP, "batched Darcy-Brinkman stage: more than 4096 operators");
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    const std::vector<std::pair<const void*, size_t>> parts = {
        {jobs.data(), jobs.size() * sizeof(ml::FlowJob)},   {cmap.data(), cmap.size() * sizeof(unsigned)},
        {fmap[0].data(), fmap[0].size() * sizeof(unsigned)}, {fmap[1].data(), fmap[1].size() * sizeof(unsigned)},
        {fmap[2].data(), fmap[2].size() * sizeof(unsigned)}, {words.data(), words.size() * sizeof(ml::WordsJob)},
        {wmap.data(), wmap.size() * sizeof(unsigned)}};
    std::vector<size_t> off(parts.size() + 1, 0);
    for (size_t k = 0; k < parts.size(); ++k) off[k + 1] = al(off[k] + parts[k].second);
    std::vector<unsigned char> blob(off.back());
    for (size_t k = 0; k < parts.size(); ++k)
      if (parts[k].second) std::memcpy(blob.data() + off[k], parts[k].first, parts[k].second);
    if (blob.empty()) return ML_OK;
    void* dv = nullptr;
    ML_HIP(ctx, ctx->tables.put(blob.data(), blob.size(), ctx->stream, &dv));
    const unsigned char* d = static_cast<const unsigned char*>(dv);
    const ml::FlowJob* dj = reinterpret_cast<const ml::FlowJob*>(d + off[0]);
    auto mp = [&](int k) { return reinterpret_cast<const unsigned*>(d + off[k]); };
    for (int st : stages) {
      if (st == 200) {
        ML_HIP(ctx, ml::words_batch(reinterpret_cast<const ml::WordsJob*>(d + off[5]), mp(6), int(wmap.size()),
                                    ctx->stream));
      } else if (st >= 100) {
        const int dd = st - 100;
        ML_HIP(ctx, ml::hc_bcoef_batch(dim, dd, dj, mp(2 + dd), int(fmap[dd].size()), ctx->stream));
      } else if (st >= ml::kFlowPressFaces && st < ml::kFlowPressFaces + 3) {
        const int dd = st - ml::kFlowPressFaces;
        ML_HIP(ctx, ml::flow_batch(dim, st, dj, mp(2 + dd), int(fmap[dd].size()), ctx->stream));
      } else {
        ML_HIP(ctx, ml::flow_batch(dim, st, dj, mp(1), int(cmap.size()), ctx->stream));
      }
    }
    return ML_OK;
  }
};

template <class Op>
int many_check(ml_ctx ctx, int n, const Op* ops, const char* who) {
  for (int i = 0; i < n; ++i) {
    if (!ops[i] || ops[i]->ctx != ctx || ops[i]->dim != ops[0]->dim)
      return ml::fail(ctx, ML_EINVAL, std::string(who) + ": operators of one context and dimension");
    for (int j = 0; j < i; ++j)
      if (ops[j] == ops[i]) return ml::fail(ctx, ML_EINVAL, std::string(who) + ": an operator appears twice");
  }
  return ML_OK;
}

}  // namespace

extern "C" {

int ml_hc_db_coefs_many(ml_ctx ctx, int n, const ml_hc_op ops[], const double* const phi[], double* const vars[],
                        double* const bCoef[]) {
  if (!ctx || n < 0 || (n > 0 && (!ops || !phi || !vars || !bCoef))) return ML_EINVAL;
  if (n == 0) return ML_OK;
  if (int rc = many_check(ctx, n, ops, "ml_hc_db_coefs_many")) return rc;
  FlowBatch B;
  B.dim = ops[0]->dim;
  for (int i = 0; i < n; ++i) {
    ml_hc_op op = ops[i];
    if (!phi[i] || !vars[i] || !fields_ok(op->dim, bCoef + i * op->dim)) return ML_EINVAL;
    ml::FlowJob& J = B.add(op->lev[0].g);
    J.cp = op->cp;
    J.phi = phi[i];
    J.vars = vars[i];
    J.chi = vars[i] + 3 * J.g.ncell;   // the porosity component of the enthalpy variables
    J.kcond = op->params.heatConductivityRatio;
    J.DH = op->params.heatDiffusionCoeff;
    J.DC = op->params.saltDiffusionCoeff;
    for (int d = 0; d < op->dim; ++d) J.b[d] = bCoef[i * op->dim + d];
  }
  std::vector<int> stages = {ml::kFlowEnthalpy};
  for (int d = 0; d < B.dim; ++d) stages.push_back(100 + d);
  return B.run(ctx, stages);
}

int ml_vc_db_ustar_many(ml_ctx ctx, int n, const ml_hc_op hops[], const ml_vc_op uops[],
                        const ml_flow_params* const fp[], const double darcy_coeff[], const int law[],
                        const double hele_shaw[], const double* const phi[], const double* const vars_old[],
                        double* const vars_new[], const double* const pi[], double* const cCoef[],
                        double* const src[], const double* const u[], double* const u_old[]) {
  if (!ctx || n < 0 ||
      (n > 0 && (!hops || !uops || !fp || !darcy_coeff || !law || !hele_shaw || !phi || !vars_old || !vars_new ||
                 !pi || !cCoef || !src || !u || !u_old)))
    return ML_EINVAL;
  if (n == 0) return ML_OK;
  if (int rc = many_check(ctx, n, uops, "ml_vc_db_ustar_many")) return rc;
  if (int rc = many_check(ctx, n, hops, "ml_vc_db_ustar_many")) return rc;
  FlowBatch B;
  B.dim = uops[0]->dim;
  for (int i = 0; i < n; ++i) {
    ml_vc_op op = uops[i];
    const int dim = op->dim;
    if (hops[i]->dim != dim || !phi[i] || !vars_old[i] || !vars_new[i] || !pi[i] || !cCoef[i] ||
        !fields_ok(dim, src + i * dim) || !fields_ok(dim, u + i * dim) || !fields_ok(dim, u_old + i * dim))
      return ML_EINVAL;
    if (law[i] < ML_PERM_PURE_FLUID || law[i] > ML_PERM_POROSITY)
      return ml::fail(ctx, ML_EINVAL, "ml_vc_db_ustar_many: unknown permeability law");
    if (int rc = check_flow(op, fp[i], "ml_vc_db_ustar_many")) return rc;
    ml::FlowJob& J = B.add(op->lev[0].g);
    J.cp = hops[i]->cp;
    J.phi = phi[i];
    J.vars = vars_new[i];
    J.vold = vars_old[i];
    J.vnew = vars_new[i];
    J.pi = pi[i];
    J.chi = vars_old[i] + 3 * J.g.ncell;
    J.c = cCoef[i];
    J.coeff = darcy_coeff[i];
    J.law = law[i];
    J.hele_shaw = hele_shaw[i];
    J.buoyT = fp[i]->buoyancyTCoeff;
    J.buoyS = fp[i]->buoyancySCoeff;
    J.scaleP = fp[i]->scalePressure;
    for (int d = 0; d < dim; ++d) {
      J.src[d] = src[i * dim + d];
      B.copy(u[i * dim + d], u_old[i * dim + d], J.g.ncell);
    }
  }
  // enthalpy variables at t^{n+1}, Darcy coefficient, U* source, u^n saved
  return B.run(ctx, {ml::kFlowEnthalpy, ml::kFlowDarcy, ml::kFlowUstarSrc, 200});
}

int ml_vc_db_project_many(ml_ctx ctx, int n, const ml_vc_op uops[], const ml_flow_params* const fp[], double dt,
                          const double* const vars_old[], const double* const vars_new[], double* const pi[],
                          const double* const ustar[], double* const u[], double* const rhs[],
                          double* const bCoef[]) {
  if (!ctx || n < 0 || (n > 0 && (!uops || !fp || !vars_old || !vars_new || !pi || !ustar || !u || !rhs || !bCoef)) ||
      !(dt > 0))
    return ML_EINVAL;
  if (n == 0) return ML_OK;
  if (int rc = many_check(ctx, n, uops, "ml_vc_db_project_many")) return rc;
  FlowBatch B;
  B.dim = uops[0]->dim;
  for (int i = 0; i < n; ++i) {
    ml_vc_op op = uops[i];
    const int dim = op->dim;
    if (!vars_old[i] || !vars_new[i] || !pi[i] || !rhs[i] || !fields_ok(dim, ustar + i * dim) ||
        !fields_ok(dim, u + i * dim) || !fields_ok(dim, bCoef + i * dim))
      return ML_EINVAL;
    for (int d = 0; d < dim; ++d)
      if (ustar[i * dim + d] == u[i * dim + d])
        return ml::fail(ctx, ML_EINVAL, "ml_vc_db_project_many: u must not alias ustar");
    if (int rc = check_flow(op, fp[i], "ml_vc_db_project_many")) return rc;
    ml::FlowJob& J = B.add(op->lev[0].g);
    J.vold = vars_old[i];
    J.vnew = vars_new[i];
    J.pi = pi[i];
    J.dt = dt;
    J.limit = fp[i]->porosityLimit;
    J.scaleP = fp[i]->scalePressure;
    J.rhs = rhs[i];
    for (int d = 0; d < dim; ++d) {
      J.ustar[d] = ustar[i * dim + d];
      J.u[d] = u[i * dim + d];
      J.b[d] = bCoef[i * dim + d];
    }
    B.copy(nullptr, pi[i], J.g.ncell);
  }
  // u = ustar + grad pi, rhs = div(u)/dt, pressure faces, then pi = 0
  std::vector<int> stages = {ml::kFlowVelocity, ml::kFlowDivergence};
  for (int d = 0; d < B.dim; ++d) stages.push_back(ml::kFlowPressFaces + d);
  stages.push_back(200);
  return B.run(ctx, stages);
}

int ml_vc_db_correct_many(ml_ctx ctx, int n, const ml_vc_op uops[], const ml_flow_params* const fp[], double dt,
                          const double* const vars_old[], const double* const vars_new[], const double* const pi[],
                          double* const u[]) {
  if (!ctx || n < 0 || (n > 0 && (!uops || !fp || !vars_old || !vars_new || !pi || !u)) || !(dt > 0))
    return ML_EINVAL;
  if (n == 0) return ML_OK;
  if (int rc = many_check(ctx, n, uops, "ml_vc_db_correct_many")) return rc;
  FlowBatch B;
  B.dim = uops[0]->dim;
  for (int i = 0; i < n; ++i) {
    ml_vc_op op = uops[i];
    const int dim = op->dim;
    if (!vars_old[i] || !vars_new[i] || !pi[i] || !fields_ok(dim, u + i * dim)) return ML_EINVAL;
    if (int rc = check_flow(op, fp[i], "ml_vc_db_correct_many")) return rc;
    ml::FlowJob& J = B.add(op->lev[0].g);
    J.vold = vars_old[i];
    J.vnew = vars_new[i];
    J.pi = pi[i];
    J.dt = dt;
    J.limit = fp[i]->porosityLimit;
    J.scaleP = fp[i]->scalePressure;
    for (int d = 0; d < dim; ++d) J.u[d] = u[i * dim + d];
  }
  return B.run(ctx, {ml::kFlowCorrect});
}


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
