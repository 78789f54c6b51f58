// ml_hc_engine.cpp — the fused FAS engine behind ml_hc_solve / ml_hc_solve_be.
//
// It performs exactly the operation sequence of the FAS V-cycle the
// reference configures (AMRFASMultiGrid driving AMRNonLinearMultiCompOp,
// srcSubcycle/AMRLevelMushyLayerInit.cpp:989-1022; driver specification in
// host/FASMultiGrid.H), with kernels fused per level:
//   tiled levels:   relax(nu1) | restrict phi, phi0 and rhs-L(phi) in one
//                   kernel | rhs_c += L_c(phi_c) | ... | prolong fused into
//                   the first post-smoothing launch | relax(nu2)
//   coarse levels:  one single-workgroup kernel for the whole sub-cycle,
//                   including the 40-iteration relax bottom solve.
// Norms accumulate into per-iteration slots zeroed once per solve.
//
// Iteration control runs on the device (SolveCtl): the host enqueues a batch
// of iterations, each followed by a one-thread check of the stopping rule;
// once the rule stops the solve, the remaining kernels of the batch return
// at once.  The host synchronises once per batch instead of once per
// iteration; the iterations that run are those of a host-checked loop.
#include <algorithm>
#include <cstring>

#include "ml_internal.h"

namespace ml {

namespace {
constexpr int kSmallCells = 1024;
constexpr int kMaxHalvings = 6;   // FP32-refined step-length backtracking
}

HCEngine::HCEngine(ml_hc_op op) : m_op(op) { m_ok = init() == ML_OK; }

HCEngine::~HCEngine() {
  for (float* p : m_lin_owned) hipFree(p);
  if (m_lindesc) hipFree(m_lindesc);
  if (m_hlindesc) hipHostFree(m_hlindesc);
  m_stage_lin.release();
  m_stage_fas.release();
  for (size_t l = 1; l < m_phi.size(); ++l) {
    if (m_phi[l]) hipFree(m_phi[l]);
    if (m_rhs[l]) hipFree(m_rhs[l]);
    if (m_phi0[l]) hipFree(m_phi0[l]);
  }
  for (double* q : m_alt)
    if (q) hipFree(q);
  if (m_rhs_be) hipFree(m_rhs_be);
  if (m_slots) hipFree(m_slots);
  if (m_hslots) hipHostFree(m_hslots);
  if (m_desc) hipFree(m_desc);
  if (m_hdesc) hipHostFree(m_hdesc);
  if (m_ctl) hipFree(m_ctl);
  if (m_arrive) hipFree(m_arrive);
  if (m_hctl) hipHostFree(m_hctl);
}

int HCEngine::init() {
  ml_ctx ctx = m_op->ctx;
  m_L = int(m_op->lev.size());
  m_lc = m_L;
  for (int l = 0; l < m_L; ++l)
    if (m_op->lev[l].g.nvalid() <= kSmallCells) {
      m_lc = l;
      break;
    }
  m_phi.assign(m_L, nullptr);
  m_rhs.assign(m_L, nullptr);
  m_phi0.assign(m_L, nullptr);
  for (int l = 1; l < m_L; ++l) {
    const size_t bytes = sizeof(double) * 2 * m_op->lev[l].g.ncell;
    void* p = nullptr;
    ML_HIP(ctx, hipMalloc(&p, bytes));
    m_phi[l] = static_cast<double*>(p);
    ML_HIP(ctx, hipMalloc(&p, bytes));
    m_rhs[l] = static_cast<double*>(p);
    ML_HIP(ctx, hipMalloc(&p, bytes));
    m_phi0[l] = static_cast<double*>(p);
    ML_HIP(ctx, vec_set(2 * m_op->lev[l].g.ncell, m_phi[l], 0.0, ctx->stream));
    ML_HIP(ctx, vec_set(2 * m_op->lev[l].g.ncell, m_rhs[l], 0.0, ctx->stream));
    ML_HIP(ctx, vec_set(2 * m_op->lev[l].g.ncell, m_phi0[l], 0.0, ctx->stream));
  }
  // relax ping-pong partners: every level relaxed by the tiled kernels
  // (above m_lc, or all levels when none fits one workgroup) and level 0
  m_alt.assign(m_L, nullptr);
  m_cur.assign(m_L, 0);
  for (int l = 0; l < m_L; ++l) {
    if (l > 0 && l >= m_lc) break;
    void* q = nullptr;
    ML_HIP(ctx, hipMalloc(&q, sizeof(double) * 2 * m_op->lev[l].g.ncell));
    m_alt[l] = static_cast<double*>(q);
    ML_HIP(ctx, vec_set(2 * m_op->lev[l].g.ncell, m_alt[l], 0.0, ctx->stream));
  }
  void* p = nullptr;
  ML_HIP(ctx, hipMalloc(&p, sizeof(double) * 2 * m_op->lev[0].g.ncell));
  m_rhs_be = static_cast<double*>(p);
  m_nslots = 1024;
  ML_HIP(ctx, hipMalloc(&p, sizeof(unsigned long long) * m_nslots));
  m_slots = static_cast<unsigned long long*>(p);
  ML_HIP(ctx, hipHostMalloc(&p, sizeof(unsigned long long) * m_nslots, hipHostMallocDefault));
  m_hslots = static_cast<unsigned long long*>(p);
  ML_HIP(ctx, hipMalloc(&p, sizeof(CoarseDesc<double>)));
  m_desc = static_cast<CoarseDesc<double>*>(p);
  ML_HIP(ctx, hipHostMalloc(&p, sizeof(CoarseDesc<double>), hipHostMallocDefault));
  m_hdesc = static_cast<CoarseDesc<double>*>(p);
  ML_HIP(ctx, hipMalloc(&p, sizeof(SolveCtl)));
  m_ctl = static_cast<SolveCtl*>(p);
  ML_HIP(ctx, hipMalloc(&p, sizeof(unsigned int)));
  m_arrive = static_cast<unsigned int*>(p);
  ML_HIP(ctx, hipMemsetAsync(m_arrive, 0, sizeof(unsigned int), ctx->stream));
  ML_HIP(ctx, hipHostMalloc(&p, sizeof(SolveCtl), hipHostMallocDefault));
  m_hctl = static_cast<SolveCtl*>(p);
  return ML_OK;
}

int HCEngine::read_norm(int slot, double* out) {
  ml_ctx ctx = m_op->ctx;
  ML_HIP(ctx, hipMemcpyAsync(m_hslots + slot, m_slots + slot, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memcpy(out, m_hslots + slot, sizeof(double));
  return ML_OK;
}

int HCEngine::read_ctl(SolveCtl* out) {
  ml_ctx ctx = m_op->ctx;
  ML_HIP(ctx, hipMemcpyAsync(m_hctl, m_ctl, sizeof(SolveCtl), hipMemcpyDeviceToHost, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = *m_hctl;
  return ML_OK;
}

int HCEngine::write_ctl(const SolveCtl& c) {
  ml_ctx ctx = m_op->ctx;
  *m_hctl = c;
  ML_HIP(ctx, hipMemcpyAsync(m_ctl, m_hctl, sizeof(SolveCtl), hipMemcpyHostToDevice, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ML_OK;
}

// ------------------------------------------------------------------------
// FP64 FAS.
int HCEngine::upload_desc(double* phi0, const double* rhs0, const ml_solve_opts& o) {
  ml_ctx ctx = m_op->ctx;
  CoarseDesc<double> D;
  std::memset(&D, 0, sizeof(D));
  D.l0 = m_lc;
  D.nlev = m_L;
  D.nu1 = o.nu1;
  D.nu2 = o.nu2;
  D.bottom = o.bottom;
  D.cp = m_op->cp;
  D.bc = m_op->bc;
  D.alpha = m_op->alpha;
  D.beta = m_op->beta;
  D.skip = &m_ctl->done;
  for (int l = 0; l < m_L; ++l) {
    LevelDesc<double>& L = D.lev[l];
    L.g = m_op->lev[l].g;
    L.phi = l == 0 ? phi0 : m_phi[l];
    L.rhs = l == 0 ? const_cast<double*>(rhs0) : m_rhs[l];
    L.phi0 = m_phi0[l];
    L.tmp = nullptr;
    L.a = m_op->lev[l].a;
    for (int d = 0; d < 3; ++d) L.b[d] = m_op->lev[l].b[d];
  }
  ML_HIP(ctx, m_stage_fas.put(m_desc, m_hdesc, &D, sizeof(D), ctx->stream));
  return ML_OK;
}

int HCEngine::vcycle(double* phi, const double* rhs, const ml_solve_opts& o, const int* skip) {
  ml_ctx ctx = m_op->ctx;
  hipStream_t s = ctx->stream;
  auto P = [&](int l) { return cur(l, phi); };
  auto R = [&](int l) { return l == 0 ? const_cast<double*>(rhs) : m_rhs[l]; };
  auto args = [&](int l) {
    HCArgs<double> A = m_op->args(l);
    A.skip = skip;
    return A;
  };
  double* B[2];
  const int ltiles = m_lc < m_L ? m_lc : m_L - 1;   // tiled levels above the coarse kernel
  if (m_lc >= m_L && m_L == 1) {
    // a single level too large for one workgroup: relax bottom with tiles
    bufs(0, phi, B);
    ML_HIP(ctx, hc_relax<double>(args(0), B, &m_cur[0], rhs, 2 * o.bottom, 0, s));
    return ML_OK;
  }
  for (int l = 0; l < ltiles; ++l) {
    const HCArgs<double> A = args(l);
    bufs(l, phi, B);
    ML_HIP(ctx, hc_relax<double>(A, B, &m_cur[l], R(l), 2 * o.nu1, 0, s));
    ML_HIP(ctx, hc_restrict_all<double>(A, m_op->lev[l + 1].g, P(l), R(l), R(l + 1), P(l + 1), m_phi0[l + 1], 0,
                                        s));
    if (l + 1 < m_lc) ML_HIP(ctx, hc_apply_op<double>(args(l + 1), P(l + 1), R(l + 1), R(l + 1), 0, s));
  }
  if (m_lc < m_L) {
    ML_HIP(ctx, hc_coarse_vcycle<double>(m_desc, m_op->dim, m_lc > 0 ? 1 : 0, s));
  } else {
    // no level small enough for one workgroup: the coarsest tiled level is
    // the bottom
    bufs(m_L - 1, phi, B);
    ML_HIP(ctx, hc_relax<double>(args(m_L - 1), B, &m_cur[m_L - 1], R(m_L - 1), 2 * o.bottom, 0, s));
  }
  for (int l = ltiles - 1; l >= 0; --l) {
    const HCArgs<double> A = args(l);
    const Geom gc = m_op->lev[l + 1].g;
    bufs(l, phi, B);
    ML_HIP(ctx, hc_relax_tiles<double>(A, B, &m_cur[l], R(l), 2 * o.nu2, 0, &gc, P(l + 1), m_phi0[l + 1], s));
  }
  return ML_OK;
}

int HCEngine::residual_norm(double* phi, const double* rhs, int slot, const int* skip, bool check) {
  ml_ctx ctx = m_op->ctx;
  HCArgs<double> A = m_op->args(0);
  A.skip = skip;
  const CtlFuse f = check ? CtlFuse{m_ctl, m_slots, m_arrive, 0} : CtlFuse{};
  ML_HIP(ctx, hc_residual<double>(A, phi, rhs, nullptr, m_slots + slot, 0, ctx->stream, f));
  return ML_OK;
}

int HCEngine::solve(double* phi, const double* rhs, const ml_solve_opts& o, ml_solve_stats* st) {
  ml_ctx ctx = m_op->ctx;
  hipStream_t s = ctx->stream;
  if (!m_ok) return fail(ctx, ML_ENOMEM, "FAS engine initialisation failed");
  if (o.precision == ML_PRECISION_FP32_REFINED) return solve_refined(phi, rhs, o, st);
  if (o.max_iter + 1 > m_nslots) return fail(ctx, ML_EINVAL, "max_iter too large");
  if (o.nu1 < 0 || o.nu2 < 0 || o.bottom < 0) return fail(ctx, ML_EINVAL, "negative smoothing count");
  int rc = upload_desc(phi, rhs, o);
  if (rc) return rc;
  // the level-0 partner starts as a full copy (ghost cells included) so
  // that a state ending in it can be copied back whole
  const size_t bytes0 = sizeof(double) * 2 * m_op->lev[0].g.ncell;
  ML_HIP(ctx, hipMemcpyAsync(m_alt[0], phi, bytes0, hipMemcpyDeviceToDevice, s));
  std::fill(m_cur.begin(), m_cur.end(), 0);
  ML_HIP(ctx, zero_u64(m_slots, o.max_iter + 1, s));
  if ((rc = residual_norm(phi, rhs, 0, nullptr, false))) return rc;
  ML_HIP(ctx, ctl_init(m_ctl, m_slots, o.max_iter, o.eps, o.hang, o.norm_thresh, s));
  const int* skip = &m_ctl->done;
  int flip0 = -1;   // level-0 buffer flips per V-cycle (fixed by the launch sequence)
  int launched = 0;
  const bool lookahead = !(o.flags & ML_SOLVE_NO_LOOKAHEAD);
  int batch = lookahead ? std::max(1, std::min(m_batch_fas, o.max_iter)) : 1;
  SolveCtl c{};
  for (;;) {
    for (int b = 0; b < batch && launched < o.max_iter; ++b) {
      const int before = m_cur[0];
      if ((rc = vcycle(phi, rhs, o, skip))) return rc;
      if (flip0 < 0) flip0 = m_cur[0] != before ? 1 : 0;
      ++launched;
      // the residual kernel's last workgroup applies the stopping rule
      if ((rc = residual_norm(cur(0, phi), rhs, launched, skip, true))) return rc;
    }
    if ((rc = read_ctl(&c))) return rc;
    if (c.done || launched >= o.max_iter) break;
    batch = lookahead ? 2 : 1;
  }
  if (lookahead) m_batch_fas = std::max(1, c.iter);
  // the state lies in the buffer reached after c.iter V-cycles
  const int cur0 = flip0 > 0 ? (c.iter & 1) : 0;
  if (cur0) ML_HIP(ctx, hipMemcpyAsync(phi, m_alt[0], bytes0, hipMemcpyDeviceToDevice, s));
  std::fill(m_cur.begin(), m_cur.end(), 0);
  if (st) {
    st->exit_status = c.status != 0 ? c.status : 2;
    st->iterations = c.iter;
    st->initial_residual = c.r0;
    st->final_residual = c.norm;
  }
  return ML_OK;
}

int HCEngine::solve_be(double* phi_new, const double* phi_old, const double* src, double dt,
                       const ml_solve_opts& o, ml_solve_stats* st) {
  ml_ctx ctx = m_op->ctx;
  if (!m_ok) return fail(ctx, ML_ENOMEM, "FAS engine initialisation failed");
  m_op->alpha = 1.0;
  m_op->beta = -dt;
  ML_HIP(ctx, be_rhs<double>(m_op->lev[0].g, 2, m_rhs_be, phi_old, m_op->lev[0].a, dt, src, ctx->stream));
  return solve(phi_new, m_rhs_be, o, st);
}

// ------------------------------------------------------------------------
// FP32-refined form.  Each outer iteration (FP64): phi += e; r = rhs - L(phi)
// with the nonlinear operator; M = closure Jacobian.  Then one FP32 linear
// V-cycle on J e = r (tiled levels above m_lc, one workgroup below, FP64
// bottom solve).  Stopping rule and status bits as for the FP64 FAS solve,
// with iterations counting outer iterations.  A step phi + e whose residual
// does not decrease is handed back to the host, which backtracks on the
// step length.
int HCEngine::init_refined() {
  ml_ctx ctx = m_op->ctx;
  if (m_lc >= m_L) return fail(ctx, ML_ENOTSUP, "FP32-refined form needs an MG level of at most 1024 cells");
  if (m_op->lev[m_L - 1].g.nvalid() > 512)
    return fail(ctx, ML_ENOTSUP, "FP32-refined form needs a coarsest level of at most 512 cells");
  if (m_L - m_lc > lin_coarse_levels_max(m_op->dim))
    return fail(ctx, ML_ENOTSUP, "FP32-refined form: too many levels below the single-workgroup level");
  {
    std::vector<Geom> gs(m_L);
    for (int l = 0; l < m_L; ++l) gs[l] = m_op->lev[l].g;
    m_lin_lds = lin_coarse_lds_bytes(gs.data(), m_lc, m_L);
    size_t maxlds = 0;
    int rc = max_lds(ctx, &maxlds);
    if (rc) return rc;
    if (m_lin_lds + lin_coarse_static_lds() > maxlds)
      return fail(ctx, ML_ENOTSUP, "FP32-refined form: the coarse levels exceed the workgroup's LDS");
  }
  m_lin.assign(m_L, LinLevel<float>{});
  auto alloc = [&](size_t n, float** p) -> int {
    void* v = nullptr;
    ML_HIP(ctx, hipMalloc(&v, n * sizeof(float)));
    *p = static_cast<float*>(v);
    m_lin_owned.push_back(*p);
    ML_HIP(ctx, hipMemsetAsync(v, 0, n * sizeof(float), ctx->stream));
    return ML_OK;
  };
  for (int l = 0; l < m_L; ++l) {
    LinLevel<float>& L = m_lin[l];
    const Geom& g = m_op->lev[l].g;
    L.g = g;
    float* p = nullptr;
    int rc;
    if ((rc = alloc(2 * g.ncell, &p))) return rc;
    L.e = p;
    if ((rc = alloc(2 * g.ncell, &p))) return rc;
    L.r = p;
    if ((rc = alloc(4 * g.ncell, &p))) return rc;
    L.M = p;
    for (int d = 0; d < 3; ++d) {
      L.b[d] = nullptr;
      if (d < g.dim) {
        if ((rc = alloc(2 * g.nface[d], &p))) return rc;
        L.b[d] = p;
      }
    }
    L.a = nullptr;
    L.skip = &m_ctl->done;
  }
  m_ealt.assign(m_L, nullptr);
  m_ecur.assign(m_L, 0);
  for (int l = 0; l < m_lc; ++l) {
    float* p = nullptr;
    int rc;
    if ((rc = alloc(2 * m_op->lev[l].g.ncell, &p))) return rc;
    m_ealt[l] = p;
  }
  void* v = nullptr;
  ML_HIP(ctx, hipMalloc(&v, sizeof(LinDesc)));
  m_lindesc = static_cast<LinDesc*>(v);
  ML_HIP(ctx, hipHostMalloc(&v, sizeof(LinDesc), hipHostMallocDefault));
  m_hlindesc = static_cast<LinDesc*>(v);
  m_refined_ok = true;
  return ML_OK;
}

int HCEngine::lin_vcycle(const ml_solve_opts& o) {
  ml_ctx ctx = m_op->ctx;
  hipStream_t s = ctx->stream;
  const float alpha = float(m_op->alpha), beta = float(m_op->beta);
  auto ecur = [&](int l) { return m_ecur[l] ? m_ealt[l] : m_lin[l].e; };
  for (int l = 0; l < m_lc; ++l) {
    float* const eb[2] = {m_lin[l].e, m_ealt[l]};
    ML_HIP(ctx, lin_relax_restrict(m_lin[l], eb, &m_ecur[l], m_lin[l + 1], m_op->bc, alpha, beta, 2 * o.nu1, true,
                                   s));
  }
  ML_HIP(ctx, lin_coarse_vcycle(m_lindesc, m_op->dim, m_lin_lds, s));
  for (int l = m_lc - 1; l >= 0; --l) {
    const Geom gc = m_lin[l + 1].g;
    float* const eb[2] = {m_lin[l].e, m_ealt[l]};
    ML_HIP(ctx, lin_relax(m_lin[l], eb, &m_ecur[l], m_op->bc, alpha, beta, 2 * o.nu2, false, &gc, ecur(l + 1), s));
  }
  return ML_OK;
}

int HCEngine::solve_refined(double* phi, const double* rhs, const ml_solve_opts& o, ml_solve_stats* st) {
  ml_ctx ctx = m_op->ctx;
  hipStream_t s = ctx->stream;
  if (!m_refined_ok) {
    int rc = init_refined();
    if (rc) return rc;
  }
  if (o.max_iter + 1 > m_nslots) return fail(ctx, ML_EINVAL, "max_iter too large");
  if (o.nu1 < 1 || o.nu2 < 1 || o.bottom < 1)
    return fail(ctx, ML_EINVAL, "FP32-refined form needs nu1, nu2, bottom >= 1");
  // FP32 copies of the face coefficients of every level, in one launch
  CastJobs J{};
  long long total = 0;
  for (int l = 0; l < m_L; ++l) {
    const Geom& g = m_op->lev[l].g;
    if (m_op->lev[l].a) return fail(ctx, ML_ENOTSUP, "FP32-refined form supports aCoef == 1 only");
    for (int d = 0; d < g.dim; ++d) {
      J.begin[J.n] = total;
      J.src[J.n] = m_op->lev[l].b[d];
      J.dst[J.n] = const_cast<float*>(m_lin[l].b[d]);
      total += 2 * g.nface[d];
      ++J.n;
    }
  }
  J.begin[J.n] = total;
  ML_HIP(ctx, cast_many(J, s));
  LinDesc D;
  std::memset(&D, 0, sizeof(D));
  D.l0 = m_lc;
  D.nlev = m_L;
  D.nu1 = o.nu1;
  D.nu2 = o.nu2;
  D.bottom = o.bottom;
  D.bc = m_op->bc;
  D.alpha = m_op->alpha;
  D.beta = m_op->beta;
  D.skip = &m_ctl->done;
  for (int l = 0; l < m_L; ++l) D.lev[l] = m_lin[l];
  for (int d = 0; d < 3; ++d) D.b64[d] = m_op->lev[m_L - 1].b[d];
  D.a64 = nullptr;
  ML_HIP(ctx, m_stage_lin.put(m_lindesc, m_hlindesc, &D, sizeof(D), s));
  ML_HIP(ctx, zero_u64(m_slots, o.max_iter + 1, s));

  HCArgs<double> A = m_op->args(0);
  const size_t bytes0 = sizeof(double) * 2 * m_op->lev[0].g.ncell;
  ML_HIP(ctx, hipMemcpyAsync(m_alt[0], phi, bytes0, hipMemcpyDeviceToDevice, s));
  double* const pb[2] = {phi, m_alt[0]};
  ML_HIP(ctx, hc_residual_jacobian(A, phi, nullptr, rhs, nullptr, 1.0, m_lin[0].r, m_lin[0].M, m_slots, s));
  ML_HIP(ctx, ctl_init(m_ctl, m_slots, o.max_iter, o.eps, o.hang, o.norm_thresh, s));
  A.skip = &m_ctl->done;
  std::fill(m_ecur.begin(), m_ecur.end(), 0);
  std::vector<int> eflip;   // per-level correction-buffer flips of one linear V-cycle
  int executed = 0;         // linear V-cycles that ran
  int rc;
  SolveCtl c{};
  const bool lookahead = !(o.flags & ML_SOLVE_NO_LOOKAHEAD);
  int batch = lookahead ? std::max(1, std::min(m_batch_refined, o.max_iter)) : 1;
  for (;;) {
    const int accepted0 = c.iter;   // accepted iterations before this batch
    int launched = 0;
    for (int b = 0; b < batch && accepted0 + launched < o.max_iter; ++b) {
      const std::vector<int> before = m_ecur;
      if ((rc = lin_vcycle(o))) return rc;
      if (eflip.empty()) {
        eflip.resize(m_L);
        for (int l = 0; l < m_L; ++l) eflip[l] = m_ecur[l] != before[l] ? 1 : 0;
      }
      const int it = accepted0 + launched + 1;
      const int pc = (it - 1) & 1;   // each earlier iteration accepted one step
      const float* e0 = m_ecur[0] ? m_ealt[0] : m_lin[0].e;
      // the residual kernel's last workgroup applies the stopping rule
      ML_HIP(ctx, hc_residual_jacobian(A, pb[pc], pb[1 - pc], rhs, e0, 1.0, m_lin[0].r, m_lin[0].M,
                                       m_slots + it, s, CtlFuse{m_ctl, m_slots, m_arrive, 1}));
      ++launched;
    }
    if ((rc = read_ctl(&c))) return rc;
    // V-cycles that ran in this batch: the accepted ones, plus a rejected one
    executed += (c.iter - accepted0) + (c.need_host ? 1 : 0);
    for (int l = 0; l < m_L; ++l) m_ecur[l] = (eflip[l] * executed) & 1;
    if (c.need_host) {
      // backtracking on the step length: phi + omega*e is accepted once its
      // FP64 residual max-norm falls below the previous one (or after
      // kMaxHalvings halvings); a trial overwrites only the partner array
      const int it = c.iter + 1;
      const int pc = c.iter & 1;
      const float* e0 = m_ecur[0] ? m_ealt[0] : m_lin[0].e;
      const HCArgs<double> An = m_op->args(0);
      double norm = c.norm, omega = 1.0;
      for (int k = 1; k <= kMaxHalvings; ++k) {
        omega *= 0.5;
        ML_HIP(ctx, zero_u64(m_slots + it, 1, s));
        ML_HIP(ctx, hc_residual_jacobian(An, pb[pc], pb[1 - pc], rhs, e0, omega, m_lin[0].r, m_lin[0].M,
                                         m_slots + it, s));
        if ((rc = read_norm(it, &norm))) return rc;
        if (norm < c.prev) break;
      }
      c.iter = it;
      int status = 0;
      if (norm <= c.eps * c.r0) status |= 1;
      if (norm <= c.thresh) status |= 8;
      if (status == 0 && norm > (1.0 - c.hang) * c.prev) status |= 4;
      c.prev = norm;
      c.norm = norm;
      c.need_host = 0;
      c.done = 0;
      if (status != 0 || c.iter >= c.max_iter) {
        c.status = status != 0 ? status : 2;
        c.done = 1;
      }
      if ((rc = write_ctl(c))) return rc;
    }
    if (c.done || c.iter >= o.max_iter) break;
    batch = lookahead ? 2 : 1;
  }
  if (lookahead) m_batch_refined = std::max(1, c.iter);
  if (c.iter & 1) ML_HIP(ctx, hipMemcpyAsync(phi, m_alt[0], bytes0, hipMemcpyDeviceToDevice, s));
  std::fill(m_ecur.begin(), m_ecur.end(), 0);
  if (st) {
    st->exit_status = c.status != 0 ? c.status : 2;
    st->iterations = c.iter;
    st->initial_residual = c.r0;
    st->final_residual = c.norm;
  }
  return ML_OK;
}

}  // namespace ml
