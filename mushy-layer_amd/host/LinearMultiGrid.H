// LinearMultiGrid.H — the linear multigrid driver that Chombo's AMRMultiGrid
// provides to the reference's U* solves (VelocityMultigrid,
// srcSubcycle/AMRLevelMushyLayerVel.cpp:1155-1427) and projection
// (Projector::solveMGlevel, srcSubcycle/Projector.cpp:3150-3221), on one AMR
// level, written against the AMRLevelOp-style method set of an operator class
// (one operator per MG depth).  Chombo's driver is not part of the reference
// tree; this is the project's own specification of it (residual-correction
// form, as the oracle's mlr_lin_mg_solve):
//
//   res = rhs - L(phi), r0 = |res|_inf
//   per iteration: e = 0; V(0) on L e = res (homogeneous BCs); phi += e;
//                  res = rhs - L(phi); stop rule
//   V(l): relax(e, r, nu1); r_c = restrictResidual(e, r); e_c = 0; V(l+1);
//         e += P e_c; relax(e, r, nu2).  Coarsest: relax(e, r, bottom), then
//         (bicgstab_bottom) BiCGStabSolver on L e = r from that e
//   stop:  |r| <= eps*|r0| (1), |r| <= norm_thresh (8), hang (4), max (2)
#pragma once

#include <vector>

#include "BiCGStabSolver.H"
#include "FASMultiGrid.H"
#include "LevelData.H"

namespace mlhost {

template <class Op>
class LinearMultiGrid {
 public:
  void define(std::vector<Op*> ops, const LevelData& prototype) {
    m_ops = std::move(ops);
    const int L = int(m_ops.size());
    m_e.resize(L);
    m_r.resize(L);
    m_ops[0]->create(m_e[0], prototype);
    m_ops[0]->create(m_r[0], prototype);
    for (int l = 1; l < L; ++l) {
      m_ops[l - 1]->createCoarser(m_e[l], m_e[l - 1]);
      m_ops[l]->create(m_r[l], m_e[l]);
    }
  }

  FASParams& params() { return m_params; }

  int solve(LevelData& phi, const LevelData& rhs, SolveStats* st = nullptr) {
    Op& op = *m_ops[0];
    op.residual(m_r[0], phi, rhs, false);
    const double r0 = op.norm(m_r[0], 0);
    double norm = r0, prev = r0;
    int iter = 0, status = 0;
    if (r0 <= m_params.norm_thresh) status = 8;
    while (status == 0 && iter < m_params.max_iter) {
      op.setToZero(m_e[0]);
      vcycle(0);
      op.incr(phi, m_e[0], 1.0);
      ++iter;
      op.residual(m_r[0], phi, rhs, false);
      norm = op.norm(m_r[0], 0);
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
    return status;
  }

 private:
  void vcycle(int l) {
    Op& op = *m_ops[l];
    const int L = int(m_ops.size());
    if (l == L - 1) {
      op.relax(m_e[l], m_r[l], m_params.bottom);
      if (m_params.bicgstab_bottom) {
        m_bottom.m_imax = m_params.bottom_imax;
        m_bottom.solve(op, m_e[l], m_r[l]);
      }
      return;
    }
    Op& opc = *m_ops[l + 1];
    op.relax(m_e[l], m_r[l], m_params.nu1);
    op.restrictResidual(m_r[l + 1], m_e[l], m_r[l]);
    opc.setToZero(m_e[l + 1]);
    vcycle(l + 1);
    op.prolongIncrement(m_e[l], m_e[l + 1]);
    op.relax(m_e[l], m_r[l], m_params.nu2);
  }

  std::vector<Op*> m_ops;
  std::vector<LevelData> m_e, m_r;
  FASParams m_params;
  BiCGStabSolver<Op> m_bottom;
};

// LevelBackwardEuler::updateSoln with a linear operator (the U* solve,
// srcSubcycle/AMRLevelMushyLayerVel.cpp:1368-1373): alpha = 1, beta = -dt;
// rhs = diagonalScale(phi_old) + dt*src; solve from phi_new.
template <class Op>
int backwardEulerStep(std::vector<Op*>& ops, LinearMultiGrid<Op>& solver, LevelData& phiNew,
                      const LevelData& phiOld, const LevelData* src, double dt, LevelData& rhsScratch,
                      SolveStats* st = nullptr) {
  ops[0]->setAlphaAndBeta(1.0, -dt);
  ops[0]->assign(rhsScratch, phiOld);
  ops[0]->diagonalScale(rhsScratch);
  if (src) ops[0]->incr(rhsScratch, *src, dt);
  return solver.solve(phiNew, rhsScratch, st);
}

}  // namespace mlhost
