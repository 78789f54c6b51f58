// BiCGStabSolver.H — the bottom solver the reference gives the linear
// multigrid of its projection (Projector::makeBottomSolvers,
// srcSubcycle/Projector.cpp:3004-3023, m_imax = bottomSolveMaxIter = 20) and
// of its U* solves (s_botSolverUStar, srcSubcycle/AMRLevelMushyLayerInit.cpp:739).
// Chombo's BiCGStabSolver is not part of the reference tree; this is the
// project's own specification of it, written against the LinearOp method set
// of an operator class (residual, applyOp, preCond, dotProduct, norm, incr,
// scale, assign, setToZero, create): van der Vorst's preconditioned BiCGStab
// (SIAM J. Sci. Stat. Comput. 13 (1992) 631-644) with restarts,
//
//   r = rhs - L(phi) (homogeneous), r~ = r, e = 0, |.| the L2 norm, r0 = |r|
//   while i < imax, |r| > eps |r_prev|, |r_prev| > 0:
//     rho = (r~, r); rho == 0: stop
//     p = r (after a (re)start), else p = beta p + (-beta omega) v + r
//     p~ = preCond(p); v = L p~; m = (r~, v); alpha = rho / m
//     |m| > small |rho|: r += (-alpha) v, e += alpha p~;  else r = 0
//     |r| > eps r0 and > reps r0: s~ = preCond(r); t = L s~;
//         omega = (t, r) / (t, t); e += omega s~; r += (-omega) t
//     |r| <= eps r0 or <= reps r0: converged
//     omega == 0 or |r| > (1 - hang) |r_prev|: every second time phi += e and
//         restart from the true residual (status 3 after m_numRestarts)
//   phi += e
//
// Bit-identical to the BiCGStab inside the fused engine's coarse V-cycle
// when the operator's dotProduct uses the same summation order.
#pragma once

#include <cmath>

#include "LevelData.H"

namespace mlhost {

template <class Op>
class BiCGStabSolver {
 public:
  int m_imax = 80;
  double m_eps = 1e-6, m_reps = 1e-12, m_hang = 1e-8, m_small = 1e-30;
  int m_numRestarts = 5;
  int m_exitStatus = -1;
  int m_iterations = 0;

  // solve L phi = rhs with homogeneous BCs, starting from phi
  void solve(Op& op, LevelData& phi, const LevelData& rhs) {
    LevelData r, rt, e, p, pt, st, t, v;
    for (LevelData* x : {&r, &rt, &e, &p, &pt, &st, &t, &v}) {
      op.create(*x, rhs);
      op.setToZero(*x);
    }
    op.residual(r, phi, rhs, true);
    op.assign(rt, r);
    double norm0 = std::sqrt(op.dotProduct(r, r)), norm1 = norm0;
    const double r0 = norm0;
    double rho0 = 0, rho1 = 0, alpha0 = 0, alpha1 = 0, omega0 = 0, omega1 = 0;
    bool init = true;
    int restarts = 0, recount = 0, i = 0;
    m_exitStatus = 0;
    while (i < m_imax && norm0 > m_eps * norm1 && norm1 > 0) {
      ++i;
      norm1 = norm0;
      alpha1 = alpha0;
      omega1 = omega0;
      rho1 = rho0;
      rho0 = op.dotProduct(rt, r);
      if (rho0 == 0.0) {
        m_exitStatus = 2;
        break;
      }
      if (init) {
        op.assign(p, r);
        init = false;
      } else {
        const double bt = (rho0 / rho1) * (alpha1 / omega1);
        op.scale(p, bt);
        op.incr(p, v, -bt * omega1);
        op.incr(p, r, 1.0);
      }
      op.preCond(pt, p);
      op.applyOp(v, pt, true);
      const double m = op.dotProduct(rt, v);
      alpha0 = rho0 / m;
      if (std::fabs(m) > m_small * std::fabs(rho0)) {
        op.incr(r, v, -alpha0);
        norm0 = std::sqrt(op.dotProduct(r, r));
        op.incr(e, pt, alpha0);
      } else {
        op.setToZero(r);
        norm0 = 0.0;
      }
      if (norm0 > m_eps * r0 && norm0 > m_reps * r0) {
        op.preCond(st, r);
        op.applyOp(t, st, true);
        omega0 = op.dotProduct(t, r) / op.dotProduct(t, t);
        op.incr(e, st, omega0);
        op.incr(r, t, -omega0);
        norm0 = std::sqrt(op.dotProduct(r, r));
      }
      if (norm0 <= m_eps * r0 || norm0 <= m_reps * r0) {
        m_exitStatus = 1;
        break;
      }
      if (omega0 == 0.0 || norm0 > (1.0 - m_hang) * norm1) {
        if (recount == 0) {
          recount = 1;
        } else {
          recount = 0;
          op.incr(phi, e, 1.0);
          if (restarts == m_numRestarts) {
            m_exitStatus = 3;
            op.setToZero(e);
            break;
          }
          op.residual(r, phi, rhs, true);
          norm0 = std::sqrt(op.dotProduct(r, r));
          rho0 = alpha0 = omega0 = 0.0;
          op.assign(rt, r);
          op.setToZero(e);
          ++restarts;
          init = true;
        }
      }
    }
    if (m_exitStatus == 0) m_exitStatus = 2;
    m_iterations = i;
    op.incr(phi, e, 1.0);
  }
};

}  // namespace mlhost
