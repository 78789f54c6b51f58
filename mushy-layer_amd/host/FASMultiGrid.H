// FASMultiGrid.H — the nonlinear FAS V-cycle driver that Chombo's
// AMRFASMultiGrid provides to the reference (configured at
// srcSubcycle/AMRLevelMushyLayerInit.cpp:989-1022), written against the
// AMRLevelOp-style method set of an operator class (one operator per MG
// depth).  Chombo's driver is not part of the reference tree, so this is the
// project's own specification of it:
//
//   V(l):  relax nu1;  phi_c = R phi (kept as phi_c0);
//          rhs_c = R(rhs - L phi) + L_c(phi_c);  V(l+1);
//          phi += P(phi_c - phi_c0);  relax nu2
//   coarsest level: `bottom` relaxations (RelaxSolver, m_imax = 40,
//          srcSubcycle/AMRLevelMushyLayerInit.cpp:788,996-1000)
//   stop:  |r| <= eps*|r0| (status 1) or |r| <= norm_thresh (8); hang when
//          |r| > (1-hang)|r_prev| (4); iteration max (2) -- max norms.
#pragma once

#include <memory>
#include <vector>

#include "LevelData.H"

namespace mlhost {

struct FASParams {
  int nu1 = 1, nu2 = 4, bottom = 40, max_iter = 10;
  double eps = 1e-10, hang = 1e-10, norm_thresh = 1e-10;
  // linear MG: BiCGStab on the coarsest level after the `bottom` relaxations
  bool bicgstab_bottom = false;
  int bottom_imax = 80;
};

struct SolveStats {
  int exit_status = 0;
  int iterations = 0;
  double initial_residual = 0;
  double final_residual = 0;
};

template <class Op>
class FASMultiGrid {
 public:
  // ops[l] operates on MG depth l (0 = finest); the driver does not own them.
  void define(std::vector<Op*> ops, const LevelData& prototype) {
    m_ops = std::move(ops);
    const int L = int(m_ops.size());
    m_phi.resize(L);
    m_rhs.resize(L);
    m_phi0.resize(L);
    m_tmp.resize(L);
    m_ops[0]->create(m_rhs[0], prototype);
    m_ops[0]->create(m_tmp[0], prototype);
    for (int l = 1; l < L; ++l) {
      m_ops[l - 1]->createCoarser(m_phi[l], l == 1 ? prototype : m_phi[l - 1]);
      m_ops[l]->create(m_rhs[l], m_phi[l]);
      m_ops[l]->create(m_phi0[l], m_phi[l]);
      m_ops[l]->create(m_tmp[l], m_phi[l]);
    }
  }

  FASParams& params() { return m_params; }

  int solve(LevelData& phi, const LevelData& rhs, SolveStats* st = nullptr) {
    Op& op = *m_ops[0];
    m_fine = &phi;
    op.assign(m_rhs[0], rhs);
    op.residual(m_tmp[0], phi, m_rhs[0], false);
    const double r0 = op.norm(m_tmp[0], 0);
    double norm = r0, prev = r0;
    int iter = 0, status = 0;
    if (r0 <= m_params.norm_thresh) status = 8;
    while (status == 0 && iter < m_params.max_iter) {
      vcycle(0);
      ++iter;
      op.residual(m_tmp[0], phi, m_rhs[0], false);
      norm = op.norm(m_tmp[0], 0);
      if (norm <= m_params.eps * r0) status |= 1;
      if (norm <= m_params.norm_thresh) status |= 8;
      if (status == 0 && norm > (1.0 - m_params.hang) * prev) status |= 4;
      prev = norm;
    }
    if (status == 0) status = 2;
    if (st) {
      st->exit_status = status;
      st->iterations = iter;
      st->initial_residual = r0;
      st->final_residual = norm;
    }
    m_fine = nullptr;
    return status;
  }

 private:
  LevelData& phiAt(int l) { return l == 0 ? *m_fine : m_phi[l]; }

  void vcycle(int l) {
    Op& op = *m_ops[l];
    const int L = int(m_ops.size());
    if (l == L - 1) {
      op.relax(phiAt(l), m_rhs[l], m_params.bottom);
      return;
    }
    Op& opc = *m_ops[l + 1];
    op.relax(phiAt(l), m_rhs[l], m_params.nu1);
    op.restrictR(m_phi[l + 1], phiAt(l));
    opc.assign(m_phi0[l + 1], m_phi[l + 1]);
    op.restrictResidual(m_rhs[l + 1], phiAt(l), m_rhs[l]);
    opc.applyOp(m_tmp[l + 1], m_phi[l + 1], false);
    opc.incr(m_rhs[l + 1], m_tmp[l + 1], 1.0);
    vcycle(l + 1);
    opc.axby(m_tmp[l + 1], m_phi[l + 1], m_phi0[l + 1], 1.0, -1.0);
    op.prolongIncrement(phiAt(l), m_tmp[l + 1]);
    op.relax(phiAt(l), m_rhs[l], m_params.nu2);
  }

  std::vector<Op*> m_ops;
  std::vector<LevelData> m_phi, m_rhs, m_phi0, m_tmp;
  LevelData* m_fine = nullptr;
  FASParams m_params;
};

// LevelBackwardEuler::updateSoln (Chombo; driven from
// srcSubcycle/AMRLevelMushyLayer.cpp:1775-1778 with zeroPhi = false):
// alpha = 1, beta = -dt on every depth; rhs = diagonalScale(phi_old) + dt*src;
// FAS solve from phi_new.
template <class Op>
int backwardEulerStep(std::vector<Op*>& ops, FASMultiGrid<Op>& solver, LevelData& phiNew,
                      const LevelData& phiOld, const LevelData* src, double dt, LevelData& rhsScratch,
                      SolveStats* st = nullptr) {
  ops[0]->setAlphaAndBeta(1.0, -dt);
  ops[0]->assign(rhsScratch, phiOld);
  ops[0]->diagonalScale(rhsScratch);
  if (src) ops[0]->incr(rhsScratch, *src, dt);
  return solver.solve(phiNew, rhsScratch, st);
}

}  // namespace mlhost
