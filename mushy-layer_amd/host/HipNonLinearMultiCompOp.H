// HipNonLinearMultiCompOp.H — the reference's AMRNonLinearMultiCompOp API
// (src/AMRNonLinearMultiCompOp.H:29-352, the AMRLevelOp/MGLevelOp methods a
// Chombo multigrid driver calls) over the device LevelData, forwarding every
// method to the C ABI of libmlhip.so.  One object per MG depth, sharing one
// ml_hc_op (the factory, AMRNonLinearMultiCompOpFactory,
// src/AMRNonLinearMultiCompOp.cpp:1201-1596).
#pragma once

#include <vector>

#include "LevelData.H"
#include "mlhip.h"

namespace mlhost {

class HipNonLinearMultiCompOp {
 public:
  HipNonLinearMultiCompOp(ml_hc_op op, ml_ctx ctx, int depth) : m_op(op), m_ctx(ctx), m_depth(depth) {
    double dx = 0;
    check(ml_hc_level_info(op, depth, m_n, &dx), ctx, "ml_hc_level_info");
    m_dx = dx;
  }

  int depth() const { return m_depth; }
  const int* n() const { return m_n; }
  double dx() const { return m_dx; }

  // ---- LinearOp / AMRPoissonOp vector functions
  void create(LevelData& lhs, const LevelData& rhs) const { lhs.define(m_ctx, rhs.dim(), rhs.n(), rhs.nComp()); }
  // createCoarser (src/AMRNonLinearMultiCompOp.cpp:355-371)
  void createCoarser(LevelData& coarse, const LevelData& fine) const {
    int nc[3] = {fine.n()[0] / 2, fine.n()[1] / 2, fine.dim() == 3 ? fine.n()[2] / 2 : 1};
    coarse.define(m_ctx, fine.dim(), nc, fine.nComp());
  }
  void assign(LevelData& lhs, const LevelData& rhs) const { lhs.copyFrom(rhs); }
  void incr(LevelData& lhs, const LevelData& x, double scale) const {
    check(ml_vec_incr(m_ctx, lhs.size(), lhs.data(), x.data(), scale), m_ctx, "incr");
  }
  void axby(LevelData& lhs, const LevelData& x, const LevelData& y, double a, double b) const {
    check(ml_vec_axby(m_ctx, lhs.size(), lhs.data(), x.data(), y.data(), a, b), m_ctx, "axby");
  }
  void scale(LevelData& lhs, double a) const {
    check(ml_vec_axby(m_ctx, lhs.size(), lhs.data(), lhs.data(), nullptr, a, 0.0), m_ctx, "scale");
  }
  void setToZero(LevelData& lhs) const { check(ml_vec_set(m_ctx, lhs.size(), lhs.data(), 0.0), m_ctx, "setToZero"); }
  // norm(x, 0): max norm over valid cells and components
  double norm(const LevelData& x, int /*ord*/ = 0) const {
    double v = 0;
    check(ml_norm_max(m_ctx, x.dim(), x.n(), x.nComp(), x.data(), &v), m_ctx, "norm");
    return v;
  }

  // ---- operator methods (reference: src/AMRNonLinearMultiCompOp.cpp)
  void setAlphaAndBeta(double alpha, double beta) {  // :515-523
    check(ml_hc_set_alpha_beta(m_op, alpha, beta), m_ctx, "setAlphaAndBeta");
  }
  // diagonalScale (src/AMRNonLinearMultiCompOp.H:75-83): rhs *= aCoef
  void diagonalScale(LevelData& rhs) const {
    check(ml_hc_diagonal_scale(m_op, m_depth, rhs.data()), m_ctx, "diagonalScale");
  }
  void residual(LevelData& lhs, LevelData& phi, const LevelData& rhs, bool homogeneous = false) {  // :100-166
    check(ml_hc_residual(m_op, m_depth, lhs.data(), phi.data(), rhs.data(), homogeneous), m_ctx, "residual");
  }
  void applyOp(LevelData& lhs, LevelData& phi, bool homogeneous = false) {  // :228-244
    check(ml_hc_apply_op(m_op, m_depth, lhs.data(), phi.data(), homogeneous), m_ctx, "applyOp");
  }
  void relax(LevelData& phi, const LevelData& rhs, int iters) {  // levelGSRB :815-949
    check(ml_hc_relax(m_op, m_depth, phi.data(), rhs.data(), iters), m_ctx, "relax");
  }
  void computeDiffusedVar(LevelData& out, const LevelData& phi, bool homogeneous = false) {  // :36-98
    check(ml_hc_compute_diffused_var(m_op, m_depth, phi.data(), out.data(), homogeneous), m_ctx,
          "computeDiffusedVar");
  }
  void restrictResidual(LevelData& resCoarse, LevelData& phiFine, const LevelData& rhsFine) {  // :373-451
    check(ml_hc_restrict_residual(m_op, m_depth, resCoarse.data(), phiFine.data(), rhsFine.data()), m_ctx,
          "restrictResidual");
  }
  void restrictR(LevelData& phiCoarse, const LevelData& phiFine) {  // :492-514
    check(ml_hc_restrict_r(m_op, m_depth, phiCoarse.data(), phiFine.data()), m_ctx, "restrictR");
  }
  void prolongIncrement(LevelData& phiFine, const LevelData& corrCoarse) {  // :454-482
    check(ml_hc_prolong_increment(m_op, m_depth, phiFine.data(), corrCoarse.data()), m_ctx, "prolongIncrement");
  }

 private:
  ml_hc_op m_op;
  ml_ctx m_ctx;
  int m_depth;
  int m_n[3] = {1, 1, 1};
  double m_dx = 0;
};

}  // namespace mlhost
