// HipVCOps.H — the reference's linear operator APIs over the device
// LevelData, forwarding every method to the C ABI of libmlhip.so:
//
//   HipDarcyBrinkmanOp  DarcyBrinkmanOp (src/DarcyBrinkmanOp.H:27-269): the U*
//                       Helmholtz operator  alpha a u - beta c u - beta div(b grad u)
//   HipProjectionOp     AMRProjectionOp (src/AMRProjectionOp.h:21-45) over
//                       VCAMRPoissonOp2: the projection's  -beta div(b grad p),
//                       alpha = 0, no c, piecewise-constant prolongation
//
// One object per MG depth, sharing one ml_vc_op (the factory:
// DarcyBrinkmanOpFactory, src/DarcyBrinkmanOp.cpp:801-1017).  The methods
// are those a Chombo AMRMultiGrid driver calls on a level (AMRLevelOp /
// MGLevelOp / LinearOp); LinearMultiGrid.H is such a driver.
#pragma once

#include "LevelData.H"
#include "mlhip.h"

namespace mlhost {

class HipVCLevelOp {
 public:
  HipVCLevelOp(ml_vc_op op, ml_ctx ctx, int depth, int dim, const int n0[3]) : m_op(op), m_ctx(ctx), m_depth(depth) {
    for (int d = 0; d < 3; ++d) m_n[d] = d < dim ? (n0[d] >> depth) : 1;
    m_dim = dim;
  }
  virtual ~HipVCLevelOp() = default;

  int depth() const { return m_depth; }
  const int* n() const { return m_n; }

  // ---- LinearOp vector functions (AMRPoissonOp, inherited by the reference's ops)
  void create(LevelData& lhs, const LevelData& rhs) const { lhs.define(m_ctx, rhs.dim(), rhs.n(), rhs.nComp()); }
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
  double norm(const LevelData& x, int /*ord*/ = 0) const {
    double v = 0;
    check(ml_norm_max(m_ctx, x.dim(), x.n(), x.nComp(), x.data(), &v), m_ctx, "norm");
    return v;
  }

  // ---- operator methods
  // setAlphaAndBeta (src/DarcyBrinkmanOp.cpp:242-248)
  void setAlphaAndBeta(double alpha, double beta) {
    check(ml_vc_set_alpha_beta(m_op, alpha, beta), m_ctx, "setAlphaAndBeta");
  }
  // residualI (src/DarcyBrinkmanOp.cpp:31-83): lhs = rhs - L(phi)
  void residual(LevelData& lhs, LevelData& phi, const LevelData& rhs, bool homogeneous = false) {
    check(ml_vc_residual(m_op, m_depth, lhs.data(), phi.data(), rhs.data(), homogeneous), m_ctx, "residual");
  }
  // applyOpI (src/DarcyBrinkmanOp.cpp:123-137)
  void applyOp(LevelData& lhs, LevelData& phi, bool homogeneous = false) {
    check(ml_vc_apply_op(m_op, m_depth, lhs.data(), phi.data(), homogeneous), m_ctx, "applyOp");
  }
  // preCond (src/DarcyBrinkmanOp.cpp:89-121): lambda*rhs, then relax(phi, rhs, 2)
  void preCond(LevelData& phi, const LevelData& rhs) {
    check(ml_vc_precond(m_op, m_depth, phi.data(), rhs.data()), m_ctx, "preCond");
  }
  // dotProduct over the valid cells (in the BiCGStab bottom's fixed order)
  double dotProduct(const LevelData& x, const LevelData& y) const {
    double v = 0;
    check(ml_vc_dot(m_op, m_depth, x.data(), y.data(), &v), m_ctx, "dotProduct");
    return v;
  }
  // relax -> levelGSRB x iters, homogeneous BCs (src/DarcyBrinkmanOp.cpp:503-576)
  void relax(LevelData& e, const LevelData& r, int iters) {
    check(ml_vc_relax(m_op, m_depth, e.data(), r.data(), iters), m_ctx, "relax");
  }
  // restrictResidual (src/DarcyBrinkmanOp.cpp:183-240): resCoarse = R(r - L e), homogeneous
  void restrictResidual(LevelData& resCoarse, LevelData& eFine, const LevelData& rFine) {
    check(ml_vc_restrict_residual(m_op, m_depth, resCoarse.data(), eFine.data(), rFine.data()), m_ctx,
          "restrictResidual");
  }
  // prolongIncrement (AMRProjectionOp::prolongIncrement, src/AMRProjectionOp.cpp:67-98):
  // piecewise constant
  void prolongIncrement(LevelData& eFine, const LevelData& eCoarse) {
    check(ml_vc_prolong_increment(m_op, m_depth, eFine.data(), eCoarse.data()), m_ctx, "prolongIncrement");
  }
  // diagonalScale: rhs *= aCoef (aCoef == 1 in both of the reference's uses)
  void diagonalScale(LevelData& /*rhs*/) const {}

 protected:
  ml_vc_op m_op;
  ml_ctx m_ctx;
  int m_depth;
  int m_dim = 2;
  int m_n[3] = {1, 1, 1};
};

// DarcyBrinkmanOp: backward-Euler U* solves set alpha = 1, beta = -dt
class HipDarcyBrinkmanOp : public HipVCLevelOp {
 public:
  using HipVCLevelOp::HipVCLevelOp;
};

// AMRProjectionOp: the projection's operator, alpha = 0, beta = 1
// (Projector::defineSolverMGLevelVCOp, srcSubcycle/Projector.cpp:2850-2991)
class HipProjectionOp : public HipVCLevelOp {
 public:
  HipProjectionOp(ml_vc_op op, ml_ctx ctx, int depth, int dim, const int n0[3]) : HipVCLevelOp(op, ctx, depth, dim, n0) {
    if (depth == 0) setAlphaAndBeta(0.0, 1.0);
  }
};

}  // namespace mlhost
