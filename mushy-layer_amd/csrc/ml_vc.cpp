// ml_vc.cpp — C ABI of the linear variable-coefficient operator (the
// DarcyBrinkmanOp / AMRProjectionOp family) and its multigrid solve.
//
// The solve is the residual-correction multigrid the reference drives
// through Chombo's AMRMultiGrid on one AMR level (build-owned specification,
// restated in the CPU oracle):
//   res = rhs - L(phi) (inhomogeneous BCs), r0 = |res|_inf
//   per iteration: e = 0; V-cycle on L e = res with homogeneous BCs;
//                  phi += e; res = rhs - L(phi); stop as the FAS solve does
//   V-cycle on level l: relax(nu1); restrict the residual; V-cycle(l+1) from
//   a zero correction; piecewise-constant prolongation (fused into the first
//   post-smoothing launch); relax(nu2).  The coarsest level relaxes
//   `bottom` times.
// Tiled kernels serve the levels above `lc`; one single-workgroup kernel
// runs the V-cycle from `lc` down.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>

#include "../host/HipVCOps.H"
#include "../host/LinearMultiGrid.H"
#include "ml_internal.h"

namespace ml {

// ML_SOLVE_HOST_DRIVER: one reference-API operator object per MG depth
// (host/HipVCOps.H) and the linear MG driver (host/LinearMultiGrid.H), every
// operator method one C-ABI call; bit-identical to the fused engine
struct VCHostDriver {
  std::vector<std::unique_ptr<mlhost::HipVCLevelOp>> owned;
  std::vector<mlhost::HipVCLevelOp*> ops;
  mlhost::LinearMultiGrid<mlhost::HipVCLevelOp> mg;
};

}  // namespace ml

namespace {

using ml::fail;

void delete_vc_host_driver(ml::VCHostDriver* p) { delete p; }

ml::VCHostDriver& vc_host_driver_of(ml_vc_op op) {
  if (!op->host_driver) {
    std::unique_ptr<ml::VCHostDriver> S(new ml::VCHostDriver());
    const ml::Geom& g = op->lev[0].g;
    for (int l = 0; l < int(op->lev.size()); ++l) {
      S->owned.emplace_back(new mlhost::HipDarcyBrinkmanOp(op, op->ctx, l, g.dim, g.n));
      S->ops.push_back(S->owned.back().get());
    }
    mlhost::LevelData proto(op->ctx, g.dim, g.n, 1);
    S->mg.define(S->ops, proto);
    op->host_driver = std::unique_ptr<ml::VCHostDriver, void (*)(ml::VCHostDriver*)>(S.release(), delete_vc_host_driver);
  }
  return *op->host_driver;
}

// the solves of ml_vc_solve_many through the host driver, one after another
// on the context stream
int vc_solve_host(ml_vc_op op, int n, double* const phi[], const double* const rhs[], const ml_solve_opts& o,
                  ml_solve_stats* st) {
  try {
    ml::VCHostDriver& S = vc_host_driver_of(op);
    mlhost::FASParams& p = S.mg.params();
    p.nu1 = o.nu1;
    p.nu2 = o.nu2;
    p.bottom = o.bottom;
    p.max_iter = o.max_iter;
    p.eps = o.eps;
    p.hang = o.hang;
    p.norm_thresh = o.norm_thresh;
    p.bicgstab_bottom = o.bottom_solver == ML_BOTTOM_BICGSTAB;
    p.bottom_imax = o.bottom_max_iter;
    const ml::Geom& g = op->lev[0].g;
    for (int i = 0; i < n; ++i) {
      mlhost::LevelData vphi, vrhs;
      vphi.alias(op->ctx, g.dim, g.n, 1, phi[i]);
      vrhs.alias(op->ctx, g.dim, g.n, 1, const_cast<double*>(rhs[i]));
      mlhost::SolveStats s;
      S.mg.solve(vphi, vrhs, &s);
      if (st) {
        st[i].exit_status = s.exit_status;
        st[i].iterations = s.iterations;
        st[i].initial_residual = s.initial_residual;
        st[i].final_residual = s.final_residual;
      }
    }
  } catch (const std::exception& e) {
    return fail(op->ctx, ML_EHIP, e.what());
  }
  return ML_OK;
}

int vc_alloc(ml_ctx ctx, size_t n, double** p) {
  void* v = nullptr;
  ML_HIP(ctx, hipMalloc(&v, n * sizeof(double)));
  *p = static_cast<double*>(v);
  ML_HIP(ctx, ml::vec_set((long long)n, *p, 0.0, ctx->stream));
  return ML_OK;
}

void vc_work_free(ml::VCWork& W) {
  auto f = [](double* q) {
    if (q) hipFree(q);
  };
  for (double* q : W.e0) f(q);
  for (double* q : W.e1) f(q);
  for (double* q : W.r) f(q);
  f(W.r0alt);
  f(W.alt0);
  f(W.rhs_be);
  if (W.slots) hipFree(W.slots);
  if (W.desc) hipFree(W.desc);
  if (W.hdesc) hipHostFree(W.hdesc);
  W.stage.release();
  if (W.ctl) hipFree(W.ctl);
  if (W.hctl) hipHostFree(W.hctl);
  if (W.arrive) hipFree(W.arrive);
  if (W.joined) hipEventDestroy(W.joined);
  if (W.gexec) hipGraphExecDestroy(W.gexec);
  if (W.own_stream && W.stream) hipStreamDestroy(W.stream);
  W = ml::VCWork{};
}

// allocate workspace W (own_stream: a new stream of default priority)
int vc_work_init(ml_vc_op op, ml::VCWork& W, bool own_stream) {
  ml_ctx ctx = op->ctx;
  const int L = int(op->lev.size());
  W.e0.assign(L, nullptr);
  W.e1.assign(L, nullptr);
  W.r.assign(L, nullptr);
  int rc;
  for (int l = 0; l < L; ++l) {
    const ml::Geom& g = op->lev[l].g;
    if ((rc = vc_alloc(ctx, g.ncell, &W.e0[l]))) return rc;
    if ((rc = vc_alloc(ctx, g.ncell, &W.r[l]))) return rc;
    // tiled levels ping-pong their correction
    if (l < op->lc || op->lc >= L)
      if ((rc = vc_alloc(ctx, g.ncell, &W.e1[l]))) return rc;
  }
  const ml::Geom& g0 = op->lev[0].g;
  if ((rc = vc_alloc(ctx, g0.ncell, &W.r0alt))) return rc;
  if ((rc = vc_alloc(ctx, g0.ncell, &W.alt0))) return rc;
  if ((rc = vc_alloc(ctx, g0.ncell, &W.rhs_be))) return rc;
  void* p = nullptr;
  ML_HIP(ctx, hipMalloc(&p, sizeof(unsigned long long) * op->nslots));
  W.slots = static_cast<unsigned long long*>(p);
  ML_HIP(ctx, hipMalloc(&p, sizeof(ml::VCDesc<double>)));
  W.desc = static_cast<ml::VCDesc<double>*>(p);
  ML_HIP(ctx, hipHostMalloc(&p, sizeof(ml::VCDesc<double>), hipHostMallocDefault));
  W.hdesc = static_cast<ml::VCDesc<double>*>(p);
  ML_HIP(ctx, hipMalloc(&p, sizeof(ml::SolveCtl)));
  W.ctl = static_cast<ml::SolveCtl*>(p);
  ML_HIP(ctx, hipHostMalloc(&p, sizeof(ml::SolveCtl), hipHostMallocDefault));
  W.hctl = static_cast<ml::SolveCtl*>(p);
  ML_HIP(ctx, hipMalloc(&p, sizeof(unsigned int)));
  W.arrive = static_cast<unsigned int*>(p);
  ML_HIP(ctx, hipMemsetAsync(W.arrive, 0, sizeof(unsigned int), ctx->stream));
  ML_HIP(ctx, hipEventCreateWithFlags(&W.joined, hipEventDisableTiming));
  if (own_stream) {
    ML_HIP(ctx, hipStreamCreateWithFlags(&W.stream, hipStreamNonBlocking));
    W.own_stream = true;
  } else {
    W.stream = ctx->stream;
  }
  // the allocations' zero fills run on the context stream
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ML_OK;
}

// workspaces 0..n-1, creating the missing ones
int vc_work_ensure(ml_vc_op op, int n) {
  while (int(op->work.size()) < n) {
    op->work.emplace_back();
    const int rc = vc_work_init(op, op->work.back(), op->work.size() > 1);
    if (rc) {
      vc_work_free(op->work.back());
      op->work.pop_back();
      return rc;
    }
  }
  return ML_OK;
}

int vc_upload_desc(ml_vc_op op, ml::VCWork& W, const ml_solve_opts& o, hipStream_t stream = nullptr) {
  ml_ctx ctx = op->ctx;
  ml::VCDesc<double> D;
  std::memset(&D, 0, sizeof(D));
  D.l0 = op->lc;
  D.nlev = int(op->lev.size());
  D.nu1 = o.nu1;
  D.nu2 = o.nu2;
  D.bottom = o.bottom;
  D.bottom_solver = o.bottom_solver;
  D.bottom_imax = o.bottom_max_iter;
  D.bc = op->bc;
  D.alpha = op->alpha;
  D.beta = op->beta;
  D.skip = &W.ctl->done;
  for (int l = 0; l < D.nlev; ++l) {
    const ml::VCArgs<double> A = op->args(l);
    ml::VCLevel<double>& L = D.lev[l];
    L.g = A.g;
    L.a = A.a;
    L.c = A.c;
    for (int d = 0; d < 3; ++d) L.b[d] = A.b[d];
    L.e = W.e0[l];
    L.r = W.r[l];
  }
  ML_HIP(ctx, W.stage.put(W.desc, W.hdesc, &D, sizeof(D), stream ? stream : W.stream));
  return ML_OK;
}

// one V-cycle of the correction equation with level-0 right-hand side r0;
// on return the level-0 correction is W's level-0 buffer ecur[0].  With fin,
// the last launch on a tiled level 0 also forms phi + e and the full
// residual (*fused = true).
int vc_vcycle(ml_vc_op op, ml::VCWork& W, const ml_solve_opts& o, int* ecur, const int* skip, double* r0,
              const ml::VCFinal<double>* fin = nullptr, bool* fused = nullptr) {
  ml_ctx ctx = op->ctx;
  hipStream_t s = W.stream;
  const int L = int(op->lev.size());
  const int lc = op->lc;
  auto buf = [&](int l, double** b) {
    b[0] = W.e0[l];
    b[1] = W.e1[l];
  };
  auto E = [&](int l) { return ecur[l] ? W.e1[l] : W.e0[l]; };
  auto R = [&](int l) { return l == 0 ? r0 : W.r[l]; };
  auto args = [&](int l) {
    ml::VCArgs<double> A = op->args(l);
    A.skip = skip;
    return A;
  };
  const int ltiles = lc < L ? lc : L - 1;
  const ml::LaunchTimer T = ctx->timer(ML_KT_VC_RELAX);
  const ml::LaunchTimer* tm = T.start ? &T : nullptr;
  double* B[2];
  if (lc >= L && L == 1) {
    buf(0, B);
    ML_HIP(ctx, ml::vc_relax<double>(args(0), B, &ecur[0], R(0), 2 * o.bottom, true, nullptr, nullptr, s, tm));
    return ML_OK;
  }
  for (int l = 0; l < ltiles; ++l) {
    buf(l, B);
    ML_HIP(ctx, ml::vc_relax_restrict<double>(args(l), B, &ecur[l], R(l), 2 * o.nu1, true, op->lev[l + 1].g,
                                               W.r[l + 1], s, tm));
  }
  if (lc < L) {
    ML_HIP(ctx, ml::vc_coarse_vcycle<double>(W.desc, op->dim,
                                             o.bottom_solver == ML_BOTTOM_BICGSTAB ? op->coarse_lds_bicg
                                                                                   : op->coarse_lds,
                                             s));
  } else {
    buf(L - 1, B);
    ML_HIP(ctx, ml::vc_relax<double>(args(L - 1), B, &ecur[L - 1], R(L - 1), 2 * o.bottom, true, nullptr, nullptr,
                                      s, tm));
  }
  if (fused) *fused = false;
  for (int l = ltiles - 1; l >= 0; --l) {
    const ml::Geom gc = op->lev[l + 1].g;
    buf(l, B);
    if (l == 0 && fin) {
      ML_HIP(ctx, ml::vc_relax_final<double>(args(l), B, &ecur[l], R(l), 2 * o.nu2, &gc, E(l + 1), *fin, fused, s,
                                              tm));
      if (*fused) continue;
    }
    ML_HIP(ctx, ml::vc_relax<double>(args(l), B, &ecur[l], R(l), 2 * o.nu2, false, &gc, E(l + 1), s, tm));
  }
  return ML_OK;
}

int vc_read_ctl(ml_vc_op op, ml::VCWork& W, ml::SolveCtl* out) {
  ml_ctx ctx = op->ctx;
  ML_HIP(ctx, hipMemcpyAsync(W.hctl, W.ctl, sizeof(ml::SolveCtl), hipMemcpyDeviceToHost, W.stream));
  ML_HIP(ctx, hipStreamSynchronize(W.stream));
  *out = *W.hctl;
  return ML_OK;
}

// One solve on a workspace, split so that several can be interleaved:
// begin (initial residual, control block), batches of iterations, end.
struct VCSolve {
  ml::VCWork* W = nullptr;
  double* phi = nullptr;
  const double* rhs = nullptr;
  std::vector<int> ecur;
  double* r = nullptr;   // level-0 residual the next iteration corrects
  int launched = 0, batch = 1;
  bool lookahead = true, finished = false, graph = false;
  ml::SolveCtl c{};
};

// one iteration: V-cycle on the correction equation with rhs *r (the
// previous residual), then phi(out) = phi(in) + e and the new residual (its
// norm into the slot of the iteration it completes, the stopping rule in the
// kernel's last workgroup).  The new residual goes to the other level-0
// buffer when it is formed inside the last relaxation launch (which still
// reads *r in other workgroups); *r then points to it.
int vc_iteration(ml_vc_op op, ml::VCWork& W, const ml_solve_opts& o, int* ecur, double* phi_in, double* phi_out,
                 const double* rhs, double** r) {
  ml_ctx ctx = op->ctx;
  const int* skip = &W.ctl->done;
  const ml::CtlFuse fuse{W.ctl, W.slots, W.arrive, 0};
  // the residual, phi_out = phi_in + e and the stopping rule ride on the last
  // relaxation launch when it fits, else run as their own kernel
  double* rnext = *r == W.r[0] ? W.r0alt : W.r[0];
  const ml::VCFinal<double> fin{phi_in, phi_out, rhs, rnext, fuse};
  bool fused = false;
  int rc = vc_vcycle(op, W, o, ecur, skip, *r, (o.flags & ML_SOLVE_NO_GRAPH) ? nullptr : &fin, &fused);
  if (rc) return rc;
  if (fused) {
    *r = rnext;
    return ML_OK;
  }
  ml::VCArgs<double> A0 = op->args(0);
  A0.skip = skip;
  double* e0 = ecur[0] ? W.e1[0] : W.e0[0];
  ML_HIP(ctx, ml::vc_residual<double>(A0, phi_in, e0, phi_out, rhs, *r, nullptr, 0, W.stream, fuse));
  return ML_OK;
}

// Capture two iterations (state phi -> partner -> phi) as a HIP graph on W's
// stream; every pointer and value they use is fixed for the solve (the norm
// slot comes from the control block).  Returns false when the capture is
// not usable (the V-cycle's ping-pong state does not return to its start).
bool vc_capture(ml_vc_op op, ml::VCWork& W, VCSolve& S, const ml_solve_opts& o) {
  if (W.gexec && W.g_phi == S.phi && W.g_rhs == S.rhs && W.g_nu1 == o.nu1 && W.g_nu2 == o.nu2 &&
      W.g_bottom == o.bottom && W.g_bsolver == o.bottom_solver && W.g_bimax == o.bottom_max_iter &&
      W.g_alpha == op->alpha && W.g_beta == op->beta)
    return true;
  if (W.gexec) hipGraphExecDestroy(W.gexec);
  W.gexec = nullptr;
  if (hipStreamBeginCapture(W.stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
  std::vector<int> ec(op->lev.size(), 0);
  double* r = W.r[0];
  int rc = vc_iteration(op, W, o, ec.data(), S.phi, W.alt0, S.rhs, &r);
  if (!rc) rc = vc_iteration(op, W, o, ec.data(), W.alt0, S.phi, S.rhs, &r);
  hipGraph_t graph = nullptr;
  const hipError_t ee = hipStreamEndCapture(W.stream, &graph);
  bool ok = !rc && ee == hipSuccess && graph && r == W.r[0];
  for (int v : ec) ok = ok && v == 0;
  if (ok) ok = hipGraphInstantiate(&W.gexec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (graph) hipGraphDestroy(graph);
  if (!ok) {
    W.gexec = nullptr;
    return false;
  }
  W.g_phi = S.phi;
  W.g_rhs = S.rhs;
  W.g_nu1 = o.nu1;
  W.g_nu2 = o.nu2;
  W.g_bottom = o.bottom;
  W.g_bsolver = o.bottom_solver;
  W.g_bimax = o.bottom_max_iter;
  W.g_alpha = op->alpha;
  W.g_beta = op->beta;
  return true;
}

int vc_solve_check(ml_vc_op op, const ml_solve_opts& o) {
  ml_ctx ctx = op->ctx;
  if (o.max_iter + 1 > op->nslots) return fail(ctx, ML_EINVAL, "max_iter too large");
  if (o.nu1 < 0 || o.nu2 < 0 || o.bottom < 0) return fail(ctx, ML_EINVAL, "negative smoothing count");
  if (o.precision != ML_PRECISION_FP64) return fail(ctx, ML_ENOTSUP, "the linear operator solves in FP64 only");
  if (o.bottom_solver != ML_BOTTOM_RELAX && o.bottom_solver != ML_BOTTOM_BICGSTAB)
    return fail(ctx, ML_EINVAL, "unknown bottom solver");
  if (o.bottom_solver == ML_BOTTOM_BICGSTAB) {
    if (o.bottom_max_iter < 0) return fail(ctx, ML_EINVAL, "negative bottom_max_iter");
    // the BiCGStab bottom runs inside the single-workgroup V-cycle
    if (op->lc >= int(op->lev.size()))
      return fail(ctx, ML_ENOTSUP, "BiCGStab bottom: the coarsest MG level exceeds one workgroup's cells");
    if (op->coarse_lds_bicg == 0) return fail(ctx, ML_ENOTSUP, "BiCGStab bottom: the coarse levels exceed the LDS");
  }
  return ML_OK;
}

int vc_solve_begin(ml_vc_op op, VCSolve& S, const ml_solve_opts& o) {
  ml_ctx ctx = op->ctx;
  ml::VCWork& W = *S.W;
  hipStream_t s = W.stream;
  int rc = vc_upload_desc(op, W, o);
  if (rc) return rc;
  const ml::Geom& g0 = op->lev[0].g;
  ML_HIP(ctx, hipMemcpyAsync(W.alt0, S.phi, sizeof(double) * g0.ncell, hipMemcpyDeviceToDevice, s));
  S.ecur.assign(op->lev.size(), 0);
  S.r = W.r[0];
  ML_HIP(ctx, ml::zero_u64(W.slots, o.max_iter + 1, s));
  ML_HIP(ctx, ml::vc_residual<double>(op->args(0), S.phi, nullptr, nullptr, S.rhs, W.r[0], W.slots, 0, s));
  ML_HIP(ctx, ml::ctl_init(W.ctl, W.slots, o.max_iter, o.eps, o.hang, o.norm_thresh, s));
  S.lookahead = !(o.flags & ML_SOLVE_NO_LOOKAHEAD);
  S.batch = S.lookahead ? std::max(1, std::min(W.batch, o.max_iter)) : 1;
  S.launched = 0;
  S.finished = false;
  // graphs replay two iterations at a time; not with one-at-a-time
  // iterations or while launches are being timed
  S.graph = !(o.flags & ML_SOLVE_NO_GRAPH) && S.lookahead && ctx->timing_kernel == 0 && vc_capture(op, W, S, o);
  return ML_OK;
}

// iterations are enqueued in batches; the device stops them (SolveCtl)
int vc_solve_enqueue(ml_vc_op op, VCSolve& S, const ml_solve_opts& o) {
  ml_ctx ctx = op->ctx;
  ml::VCWork& W = *S.W;
  if (S.graph) {
    // iterations past the stopping point return at once (done flag)
    for (int b = 0; b < S.batch && S.launched < o.max_iter; b += 2) {
      ML_HIP(ctx, hipGraphLaunch(W.gexec, W.stream));
      S.launched += 2;
    }
    return ML_OK;
  }
  double* const pb[2] = {S.phi, W.alt0};
  for (int b = 0; b < S.batch && S.launched < o.max_iter; ++b) {
    ++S.launched;
    const int pcur = (S.launched - 1) & 1;   // phi += e alternates the state buffer
    int rc = vc_iteration(op, W, o, S.ecur.data(), pb[pcur], pb[1 - pcur], S.rhs, &S.r);
    if (rc) return rc;
  }
  return ML_OK;
}

int vc_solve_poll(ml_vc_op op, VCSolve& S, const ml_solve_opts& o) {
  int rc = vc_read_ctl(op, *S.W, &S.c);
  if (rc) return rc;
  if (S.c.done || S.launched >= o.max_iter) S.finished = true;
  S.batch = S.lookahead ? 2 : 1;
  return ML_OK;
}

int vc_solve_end(ml_vc_op op, VCSolve& S, ml_solve_stats* st) {
  ml_ctx ctx = op->ctx;
  ml::VCWork& W = *S.W;
  if (S.lookahead) W.batch = std::max(1, S.c.iter);
  const size_t bytes0 = sizeof(double) * op->lev[0].g.ncell;
  if (S.c.iter & 1) ML_HIP(ctx, hipMemcpyAsync(S.phi, W.alt0, bytes0, hipMemcpyDeviceToDevice, W.stream));
  if (st) {
    st->exit_status = S.c.status != 0 ? S.c.status : 2;
    st->iterations = S.c.iter;
    st->initial_residual = S.c.r0;
    st->final_residual = S.c.norm;
  }
  return ML_OK;
}

// n solves on workspaces 0..n-1 (phi[i] updated in place, rhs[i]): the
// workspaces' streams start after the work already on the context stream,
// and the context stream continues after all of them
int vc_solve_many(ml_vc_op op, int n, double* const phi[], const double* const rhs[], const ml_solve_opts& o,
                  ml_solve_stats* st) {
  ml_ctx ctx = op->ctx;
  int rc = vc_solve_check(op, o);
  if (rc) return rc;
  if (o.flags & ML_SOLVE_HOST_DRIVER) return vc_solve_host(op, n, phi, rhs, o, st);
  if ((rc = vc_work_ensure(op, n))) return rc;
  std::vector<VCSolve> S(n);
  for (int i = 0; i < n; ++i) {
    S[i].W = &op->work[i];
    S[i].phi = phi[i];
    S[i].rhs = rhs[i];
  }
  if (n > 1) {
    ML_HIP(ctx, hipEventRecord(op->work[0].joined, ctx->stream));
    for (int i = 1; i < n; ++i) ML_HIP(ctx, hipStreamWaitEvent(op->work[i].stream, op->work[0].joined, 0));
  }
  for (int i = 0; i < n; ++i)
    if ((rc = vc_solve_begin(op, S[i], o))) return rc;
  for (;;) {
    bool any = false;
    for (int i = 0; i < n; ++i)
      if (!S[i].finished && (rc = vc_solve_enqueue(op, S[i], o))) return rc;
    for (int i = 0; i < n; ++i) {
      if (S[i].finished) continue;
      if ((rc = vc_solve_poll(op, S[i], o))) return rc;
      any = any || !S[i].finished;
    }
    if (!any) break;
  }
  for (int i = 0; i < n; ++i)
    if ((rc = vc_solve_end(op, S[i], st ? st + i : nullptr))) return rc;
  for (int i = 1; i < n; ++i) {
    ML_HIP(ctx, hipEventRecord(op->work[i].joined, op->work[i].stream));
    ML_HIP(ctx, hipStreamWaitEvent(ctx->stream, op->work[i].joined, 0));
  }
  return ML_OK;
}

}  // namespace

extern "C" {

int ml_vc_op_create(ml_ctx ctx, int dim, const int lo[], const int hi[], double dx, const int periodic[],
                    const ml_bc* bc, ml_vc_op* out) {
  if (!ctx || !lo || !hi || !periodic || !bc || !out) return ML_EINVAL;
  *out = nullptr;
  int n[3] = {1, 1, 1};
  for (int d = 0; d < dim && d < 3; ++d) {
    if (lo[d] != 0) return fail(ctx, ML_ENOTSUP, "ml_vc_op_create: boxes must start at the origin (lo == 0)");
    n[d] = hi[d] - lo[d] + 1;
  }
  if (!ml::valid_grid(dim, n, periodic, dx)) return fail(ctx, ML_EINVAL, "ml_vc_op_create: bad grid");
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      if (bc->type[d][s][0] != ML_BC_DIRICHLET && bc->type[d][s][0] != ML_BC_NEUMANN)
        return fail(ctx, ML_EINVAL, "ml_vc_op_create: unknown BC type");
  std::unique_ptr<ml_vc_op_s> op(new (std::nothrow) ml_vc_op_s());
  if (!op) return ML_ENOMEM;
  op->ctx = ctx;
  op->dim = dim;
  std::memcpy(op->bc.type, bc->type, sizeof(bc->type));
  std::memcpy(op->bc.val, bc->val, sizeof(bc->val));
  op->bc.apply = 1;
  const int depth = ml_mg_depth(dim, n);
  if (depth > ml::kMaxLevels) return fail(ctx, ML_ENOTSUP, "too many MG levels");
  if (ml::odd_periodic_small_level(dim, n, periodic))
    return fail(ctx, ML_ENOTSUP, "odd periodic length on a level of at most 1024 cells that does not coarsen");
  const ml::Geom g0 = ml::make_geom(dim, n, periodic, dx);
  op->lev.resize(depth);
  op->lc = depth;
  for (int l = 0; l < depth; ++l)
    if (op->lev.size() && ml::vc_small_level(l == 0 ? g0 : ml::coarsen_geom(g0, 1 << l))) {
      op->lc = l;
      break;
    }
  if (op->lc < depth && depth - op->lc > ml::vc_coarse_levels_max(dim))
    return fail(ctx, ML_ENOTSUP, "too many levels below the single-workgroup level");
  ml_vc_op raw = op.get();
  auto bail = [&](int rc) {
    ml_vc_op_destroy(op.release());
    return rc;
  };
  for (int l = 0; l < depth; ++l) {
    ml::VCLevelStore& L = raw->lev[l];
    L.g = l == 0 ? g0 : ml::coarsen_geom(g0, 1 << l);
    int rc;
    for (int d = 0; d < dim; ++d)
      if ((rc = vc_alloc(ctx, L.g.nface[d], &L.b[d]))) return bail(rc);
  }
  int rc;
  if (raw->lc < depth) {
    std::vector<ml::Geom> gs(depth);
    for (int l = 0; l < depth; ++l) gs[l] = raw->lev[l].g;
    raw->coarse_lds = ml::vc_coarse_lds_bytes<double>(gs.data(), raw->lc, depth);
    size_t mx = 0;
    if ((rc = ml::max_lds(ctx, &mx))) return bail(rc);
    const size_t lb = ml::vc_coarse_lds_bytes<double>(gs.data(), raw->lc, depth, true);
    raw->coarse_lds_bicg = lb <= mx ? lb : 0;
    if (raw->coarse_lds > mx) {
      ml::fail(ctx, ML_ENOTSUP, "the coarse MG levels exceed the workgroup's LDS");
      return bail(ML_ENOTSUP);
    }
  }
  if ((rc = vc_work_ensure(raw, 1))) return bail(rc);
  *out = op.release();
  return ML_OK;
}

int ml_vc_op_destroy(ml_vc_op op) {
  if (!op) return ML_EINVAL;
  if (op->ctx) hipStreamSynchronize(op->ctx->stream);
  for (auto& W : op->work) {
    if (W.stream) hipStreamSynchronize(W.stream);
    vc_work_free(W);
  }
  op->host_driver.reset();
  for (auto& L : op->lev) {
    for (int d = 0; d < 3; ++d)
      if (L.b[d]) hipFree(L.b[d]);
    if (L.a) hipFree(L.a);
    if (L.c) hipFree(L.c);
    if (L.tmp) hipFree(L.tmp);
  }
  delete op;
  return ML_OK;
}

int ml_vc_num_levels(ml_vc_op op) { return op ? int(op->lev.size()) : 0; }

int ml_vc_set_coefs(ml_vc_op op, const double* aCoef, const double* const bCoef[], const double* cCoef,
                    double alpha, double beta) {
  if (!op || !bCoef) return ML_EINVAL;
  ml_ctx ctx = op->ctx;
  for (int d = 0; d < op->dim; ++d)
    if (!bCoef[d]) return fail(ctx, ML_EINVAL, "ml_vc_set_coefs: null bCoef");
  op->alpha = alpha;
  op->beta = beta;
  op->has_a = aCoef != nullptr;
  op->has_c = cCoef != nullptr;
  const ml::Geom& g0 = op->lev[0].g;
  hipStream_t s = ctx->stream;
  for (auto& L : op->lev) {
    int rc;
    if (aCoef && !L.a && (rc = vc_alloc(ctx, L.g.ncell, &L.a))) return rc;
    if (cCoef && !L.c && (rc = vc_alloc(ctx, L.g.ncell, &L.c))) return rc;
  }
  for (int d = 0; d < op->dim; ++d)
    ML_HIP(ctx, hipMemcpyAsync(op->lev[0].b[d], bCoef[d], sizeof(double) * g0.nface[d], hipMemcpyDeviceToDevice, s));
  if (aCoef) ML_HIP(ctx, hipMemcpyAsync(op->lev[0].a, aCoef, sizeof(double) * g0.ncell, hipMemcpyDeviceToDevice, s));
  if (cCoef) ML_HIP(ctx, hipMemcpyAsync(op->lev[0].c, cCoef, sizeof(double) * g0.ncell, hipMemcpyDeviceToDevice, s));
  // arithmetic coarsening straight from the finest level
  // (DarcyBrinkmanOpFactory::MGnewOp, src/DarcyBrinkmanOp.cpp:977-1008)
  const int L = int(op->lev.size());
  const double* bf[3] = {op->lev[0].b[0], op->lev[0].b[1], op->lev[0].b[2]};
  std::vector<double* const*> bc(L);
  std::vector<double*> ac(L, nullptr), cc(L, nullptr);
  for (int l = 0; l < L; ++l) {
    bc[l] = op->lev[l].b;
    ac[l] = op->lev[l].a;
    cc[l] = op->lev[l].c;
  }
  ML_HIP(ctx, ml::coarsen_hierarchy<double>(g0, 1, L, bf, bc.data(), aCoef ? op->lev[0].a : nullptr, ac.data(), s));
  if (cCoef) {
    const double* none[3] = {nullptr, nullptr, nullptr};
    std::vector<double* const*> nob(L, nullptr);
    ML_HIP(ctx, ml::coarsen_hierarchy<double>(g0, 1, L, none, nob.data(), op->lev[0].c, cc.data(), s, false));
  }
  return ML_OK;
}

int ml_vc_set_alpha_beta(ml_vc_op op, double alpha, double beta) {
  if (!op) return ML_EINVAL;
  op->alpha = alpha;
  op->beta = beta;
  return ML_OK;
}

#define VC_DEPTH_CHECK(op, depth) \
  if (!(op) || (depth) < 0 || (depth) >= int((op)->lev.size())) return ML_EINVAL

int ml_vc_relax(ml_vc_op op, int depth, double* phi, const double* rhs, int iters) {
  VC_DEPTH_CHECK(op, depth);
  if (!phi || !rhs || iters < 0) return ML_EINVAL;
  if (iters == 0) return ML_OK;
  ml_ctx ctx = op->ctx;
  ml::VCLevelStore& L = op->lev[depth];
  const size_t bytes = sizeof(double) * L.g.ncell;
  int rc;
  if (!L.tmp && (rc = vc_alloc(ctx, L.g.ncell, &L.tmp))) return rc;
  // the tiled smoother ping-pongs between phi and the level's partner array,
  // which starts as a full copy so a state ending there is copied back whole
  ML_HIP(ctx, hipMemcpyAsync(L.tmp, phi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  double* const buf[2] = {phi, L.tmp};
  int cur = 0;
  const ml::LaunchTimer T = ctx->timer(ML_KT_VC_RELAX);
  ML_HIP(ctx, ml::vc_relax<double>(op->args(depth), buf, &cur, rhs, 2 * iters, false, nullptr, nullptr,
                                   ctx->stream, T.start ? &T : nullptr));
  if (cur) ML_HIP(ctx, hipMemcpyAsync(phi, L.tmp, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return ML_OK;
}

int ml_vc_residual(ml_vc_op op, int depth, double* res, double* phi, const double* rhs, int homogeneous) {
  VC_DEPTH_CHECK(op, depth);
  if (!res || !phi || !rhs) return ML_EINVAL;
  ML_HIP(op->ctx, ml::vc_residual<double>(op->args(depth), phi, nullptr, nullptr, rhs, res, nullptr, homogeneous,
                                          op->ctx->stream));
  return ML_OK;
}

int ml_vc_apply_op(ml_vc_op op, int depth, double* lhs, double* phi, int homogeneous) {
  VC_DEPTH_CHECK(op, depth);
  if (!lhs || !phi) return ML_EINVAL;
  ML_HIP(op->ctx, ml::vc_apply_op<double>(op->args(depth), phi, lhs, homogeneous, op->ctx->stream));
  return ML_OK;
}

int ml_vc_precond(ml_vc_op op, int depth, double* phi, const double* rhs) {
  VC_DEPTH_CHECK(op, depth);
  if (!phi || !rhs) return ML_EINVAL;
  ML_HIP(op->ctx, ml::vc_lambda_mult<double>(op->args(depth), rhs, phi, op->ctx->stream));
  return ml_vc_relax(op, depth, phi, rhs, 2);
}

int ml_vc_dot(ml_vc_op op, int depth, const double* x, const double* y, double* out) {
  VC_DEPTH_CHECK(op, depth);
  if (!x || !y || !out) return ML_EINVAL;
  ml_ctx ctx = op->ctx;
  ml::VCLevelStore& L = op->lev[depth];
  int rc;
  if (!L.tmp && (rc = vc_alloc(ctx, L.g.ncell, &L.tmp))) return rc;
  // the level's partner array holds the result (a relaxation overwrites it)
  ML_HIP(ctx, ml::vc_dot<double>(L.g, x, y, L.tmp, ctx->stream));
  ML_HIP(ctx, hipMemcpyAsync(out, L.tmp, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ML_OK;
}

int ml_vc_restrict_residual(ml_vc_op op, int depth, double* resCoarse, double* phiFine, const double* rhsFine) {
  VC_DEPTH_CHECK(op, depth);
  if (depth + 1 >= int(op->lev.size())) return fail(op->ctx, ML_EINVAL, "no coarser level");
  if (!resCoarse || !phiFine || !rhsFine) return ML_EINVAL;
  ML_HIP(op->ctx, ml::vc_restrict_residual<double>(op->args(depth), op->lev[depth + 1].g, phiFine, rhsFine,
                                                   resCoarse, op->ctx->stream));
  return ML_OK;
}

int ml_vc_prolong_increment(ml_vc_op op, int depth, double* phiFine, const double* corrCoarse) {
  VC_DEPTH_CHECK(op, depth);
  if (depth + 1 >= int(op->lev.size())) return fail(op->ctx, ML_EINVAL, "no coarser level");
  if (!phiFine || !corrCoarse) return ML_EINVAL;
  ML_HIP(op->ctx, ml::prolong_constant<double>(op->lev[depth].g, 1, phiFine, corrCoarse, op->ctx->stream));
  return ML_OK;
}

int ml_vc_solve(ml_vc_op op, double* phi, const double* rhs, const ml_solve_opts* opts, ml_solve_stats* stats) {
  if (!op || !phi || !rhs) return ML_EINVAL;
  ml_solve_opts o;
  if (opts)
    o = *opts;
  else
    ml_default_opts(&o);
  double* const p[1] = {phi};
  const double* const r[1] = {rhs};
  return vc_solve_many(op, 1, p, r, o, stats);
}

int ml_vc_solve_be(ml_vc_op op, double* phi_new, const double* phi_old, const double* src, double dt,
                   const ml_solve_opts* opts, ml_solve_stats* stats) {
  double* const pn[1] = {phi_new};
  const double* const po[1] = {phi_old};
  const double* const ps[1] = {src};
  return ml_vc_solve_be_many(op, 1, pn, po, src ? ps : nullptr, dt, opts, stats);
}

int ml_vc_solve_be_many(ml_vc_op op, int n, double* const phi_new[], const double* const phi_old[],
                        const double* const src[], double dt, const ml_solve_opts* opts, ml_solve_stats* stats) {
  if (!op || n < 1 || n > ML_VC_MAX_SOLVES || !phi_new || !phi_old) return ML_EINVAL;
  for (int i = 0; i < n; ++i)
    if (!phi_new[i] || !phi_old[i]) return ML_EINVAL;
  ml_solve_opts o;
  if (opts)
    o = *opts;
  else
    ml_default_opts(&o);
  int rc = vc_solve_check(op, o);
  if (rc) return rc;
  if ((rc = vc_work_ensure(op, n))) return rc;
  op->alpha = 1.0;
  op->beta = -dt;
  std::vector<const double*> rhs(n);
  for (int i = 0; i < n; ++i) {
    ML_HIP(op->ctx, ml::be_rhs<double>(op->lev[0].g, 1, op->work[i].rhs_be, phi_old[i],
                                       op->has_a ? op->lev[0].a : nullptr, dt, src ? src[i] : nullptr,
                                       op->ctx->stream));
    rhs[i] = op->work[i].rhs_be;
  }
  return vc_solve_many(op, n, phi_new, rhs.data(), o, stats);
}

int ml_vc_darcy_ccoef(ml_vc_op op, const double* chi, double darcy_coeff, double* cCoef) {
  if (!op || !chi || !cCoef) return ML_EINVAL;
  ML_HIP(op->ctx, ml::darcy_ccoef(op->lev[0].g, darcy_coeff, chi, cCoef, op->ctx->stream));
  return ML_OK;
}

int ml_vc_porosity_faces(ml_vc_op op, const double* chi, double scale, double* const bCoef[]) {
  if (!op || !chi || !bCoef) return ML_EINVAL;
  for (int d = 0; d < op->dim; ++d)
    if (!bCoef[d]) return ML_EINVAL;
  ML_HIP(op->ctx, ml::porosity_faces(op->lev[0].g, scale, chi, bCoef, op->ctx->stream));
  return ML_OK;
}

}  // extern "C"

// ---- batched linear solves -----------------------------------------------
// Many solve units (an operator and one of its workspaces each) iterate
// together: per iteration the units' tiled levels are aligned by their
// distance from the single-workgroup level, so down-sweep launch d relaxes
// and restricts level d - (K - t_u) of every unit with t_u tiled levels
// (K = max t_u); one launch runs every unit's coarse V-cycle (one workgroup
// each); the up-sweep mirrors the down-sweep, its last launch forming every
// unit's phi + e, residual, norm and stopping rule.  2K + 1 launches per
// iteration for the whole batch.  A unit's launches are exactly those of its
// single-unit solve (same kernels' bodies, same buffers, same stopping
// rule), so results and statistics are bit-identical to ml_vc_solve_many.

namespace ml {
namespace {

struct VCBatchUnit {
  ml_vc_op op = nullptr;
  int w = 0;          // workspace index
  int tiled = 0;      // tiled levels (lc)
};

struct VCBatchLaunch {
  size_t job_off = 0, map_off = 0;
  int blocks = 0;
  size_t lds = 0;
};

}  // namespace
}  // namespace ml

// level-0 cells of a batch from which its relaxation uses 32x32 tiles
constexpr long long kVCBatchWideCells = 1LL << 19;

struct ml_vc_batch_s {
  ml_ctx ctx = nullptr;
  int dim = 2;
  std::vector<ml::VCBatchUnit> units;
  int K = 0;
  size_t coarse_lds = 0;
  // device tables: [jobs, maps of both iteration parities | coarse
  // descriptor pointers | control pointers | gathered controls], one pinned mirror
  unsigned char* arena = nullptr;
  unsigned char* harena = nullptr;
  size_t arena_cap = 0;
  std::vector<ml::VCBatchLaunch> launch[2];   // per parity: K down, K up
  size_t desc_off = 0, ctlp_off = 0, ctl_off = 0;
  // set-up tables of a solve: phi copies and slot zero fills, residuals,
  // slot pointers; the copy-back table written after the iterations
  size_t bw_job_off = 0, bw_map_off = 0, br_job_off = 0, br_map_off = 0, slotp_off = 0;
  size_t ew_job_off = 0, ew_map_off = 0;
  int bw_blocks = 0, br_blocks = 0;
  std::vector<const void*> key;   // what the uploaded tables were built for
  int batch_iters = 4;
  int shape = 0;   // relaxation tile shape (vc_batch_tiles)

  ~ml_vc_batch_s() {
    if (arena) hipFree(arena);
    if (harena) hipHostFree(harena);
  }
};

namespace {

using ml::fail;

size_t vcb_align(size_t x) { return (x + 255) & ~size_t(255); }

// the job and workgroup tables of both iteration parities, the coarse
// descriptor and control pointers, built for phi/rhs and uploaded
int vcb_tables(ml_vc_batch b, double* const phi[], const double* const rhs[], const ml_solve_opts& o) {
  ml_ctx ctx = b->ctx;
  const int n = int(b->units.size());
  // the jobs depend on the state and rhs buffers, alpha and beta (through
  // the operator arguments) and the smoothing counts
  std::vector<const void*> key;
  auto bits_of = [](double v) {
    std::uint64_t x;
    std::memcpy(&x, &v, sizeof(x));
    return reinterpret_cast<const void*>(std::uintptr_t(x));
  };
  for (int u = 0; u < n; ++u) {
    key.push_back(phi[u]);
    key.push_back(rhs[u]);
    key.push_back(bits_of(b->units[u].op->alpha));
    key.push_back(bits_of(b->units[u].op->beta));
  }
  key.push_back(reinterpret_cast<const void*>(std::uintptr_t(o.nu1 * 1000 + o.nu2)));
  key.push_back(reinterpret_cast<const void*>(std::uintptr_t(o.max_iter)));
  if (key == b->key) return ML_OK;
  b->key.clear();

  const int K = b->K;
  const int hmax = ml::vc_batch_hmax(b->dim);
  std::vector<ml::VCRelaxJob<double>> jobs[2];
  std::vector<unsigned> maps[2];
  for (int par = 0; par < 2; ++par) {
    b->launch[par].assign(2 * K, ml::VCBatchLaunch{});
    for (int k = 0; k < 2 * K; ++k) {
      const bool down = k < K;
      const int d = down ? k : 2 * K - 1 - k;   // aligned depth of the launch
      ml::VCBatchLaunch& Lh = b->launch[par][k];
      Lh.job_off = jobs[par].size();
      Lh.map_off = maps[par].size();
      for (int u = 0; u < n; ++u) {
        const ml::VCBatchUnit& U = b->units[u];
        const int l = d - (K - U.tiled);
        if (l < 0) continue;
        ml_vc_op op = U.op;
        ml::VCWork& W = op->work[U.w];
        ml::VCRelaxJob<double> J{};
        J.A = op->args(l);
        J.A.skip = &W.ctl->done;
        double* const r_in = par == 0 ? W.r[0] : W.r0alt;
        double* const rnext = par == 0 ? W.r0alt : W.r[0];
        double* const phi_in = par == 0 ? phi[u] : W.alt0;
        double* const phi_out = par == 0 ? W.alt0 : phi[u];
        J.rhs = l == 0 ? r_in : W.r[l];
        J.first = 0;
        J.gc = J.A.g;
        J.gr = J.A.g;
        if (down) {
          // relax from zero, restrict r - L e into the next level's rhs
          J.in = W.e0[l];
          J.out = W.e1[l];
          J.npass = 2 * o.nu1;
          J.zero_init = 1;
          J.gr = op->lev[l + 1].g;
          J.resc = W.r[l + 1];
        } else {
          // prolong the next level's correction, relax; level 0 also forms
          // phi + e, the residual and the stopping rule
          J.in = W.e1[l];
          J.out = W.e0[l];
          J.npass = 2 * o.nu2;
          J.zero_init = 0;
          J.gc = op->lev[l + 1].g;
          J.pc = W.e0[l + 1];
        }
        int ntx = 0, nty = 0;
        const int nt = ml::vc_batch_tiles(J.A.g, b->shape, &ntx, &nty);
        J.ntx = ntx;
        J.nty = nty;
        if (!down && l == 0) {
          J.fin.phi = phi_in;
          J.fin.phi_out = phi_out;
          J.fin.rhs = rhs[u];
          J.fin.res = rnext;
          J.fin.fuse = ml::CtlFuse{W.ctl, W.slots, W.arrive, 0, unsigned(nt)};
        }
        if (J.npass + 1 > hmax) return fail(ctx, ML_ENOTSUP, "batched solve: too many smoothing passes per launch");
        const size_t lds = ml::vc_batch_lds(b->dim, b->shape, J.npass, J.resc != nullptr, J.fin.phi != nullptr);
        Lh.lds = std::max(Lh.lds, lds);
        const unsigned jl = unsigned(jobs[par].size() - Lh.job_off);
        for (int t = 0; t < nt; ++t) maps[par].push_back((jl << ml::kBatchTileBits) | unsigned(t));
        jobs[par].push_back(J);
      }
      Lh.blocks = int(maps[par].size() - Lh.map_off);
    }
  }
  // set-up of a solve: every unit's phi copied to its ping-pong partner and
  // its norm slots zeroed, then its initial residual, then its control block
  std::vector<ml::WordsJob> bw, ew;
  std::vector<unsigned> bwm, ewm;
  std::vector<ml::VCResidJob<double>> br;
  std::vector<unsigned> brm;
  auto add_words = [](std::vector<ml::WordsJob>& jv, std::vector<unsigned>& mv, const void* src, void* dst,
                      long long nw) {
    const unsigned j = unsigned(jv.size());
    jv.push_back(ml::WordsJob{static_cast<const unsigned long long*>(src), static_cast<unsigned long long*>(dst), nw});
    for (long long c = 0; c * ml::kWordsChunk < nw; ++c) mv.push_back((j << ml::kBatchTileBits) | unsigned(c));
  };
  for (int u = 0; u < n; ++u) {
    ml_vc_op op = b->units[u].op;
    ml::VCWork& W = op->work[b->units[u].w];
    const ml::Geom& g0 = op->lev[0].g;
    add_words(bw, bwm, phi[u], W.alt0, g0.ncell);
    add_words(bw, bwm, nullptr, W.slots, o.max_iter + 1);
    add_words(ew, ewm, W.alt0, phi[u], g0.ncell);   // capacity of the copy-back table
    ml::VCResidJob<double> R{};
    R.A = op->args(0);
    R.io.phi = phi[u];
    R.io.rhs = rhs[u];
    R.io.out = W.r[0];
    R.io.norm = W.slots;
    R.io.homog = 0;
    int ntx = 0, nty = 0;
    const int nt = ml::vc_resid_tiles(g0, &ntx, &nty);
    R.io.ntx = ntx;
    R.io.nty = nty;
    const unsigned j = unsigned(br.size());
    for (int t = 0; t < nt; ++t) brm.push_back((j << ml::kBatchTileBits) | unsigned(t));
    br.push_back(R);
  }
  if (bw.size() > (size_t(1) << (32 - ml::kBatchTileBits)))
    return fail(ctx, ML_ENOTSUP, "batched solve: more than 2048 units");
  // layout: jobs0 | maps0 | jobs1 | maps1 | descs | ctl pointers | set-up
  // tables | slot pointers | copy-back tables | gathered ctls
  size_t off = 0;
  size_t joff[2], moff[2];
  for (int par = 0; par < 2; ++par) {
    joff[par] = off;
    off = vcb_align(off + jobs[par].size() * sizeof(ml::VCRelaxJob<double>));
    moff[par] = off;
    off = vcb_align(off + maps[par].size() * sizeof(unsigned));
  }
  b->desc_off = off;
  off = vcb_align(off + n * sizeof(void*));
  b->ctlp_off = off;
  off = vcb_align(off + n * sizeof(void*));
  b->bw_job_off = off;
  off = vcb_align(off + bw.size() * sizeof(ml::WordsJob));
  b->bw_map_off = off;
  off = vcb_align(off + bwm.size() * sizeof(unsigned));
  b->br_job_off = off;
  off = vcb_align(off + br.size() * sizeof(ml::VCResidJob<double>));
  b->br_map_off = off;
  off = vcb_align(off + brm.size() * sizeof(unsigned));
  b->slotp_off = off;
  off = vcb_align(off + n * sizeof(void*));
  b->ew_job_off = off;
  off = vcb_align(off + ew.size() * sizeof(ml::WordsJob));
  b->ew_map_off = off;
  off = vcb_align(off + ewm.size() * sizeof(unsigned));
  b->bw_blocks = int(bwm.size());
  b->br_blocks = int(brm.size());
  b->ctl_off = off;
  off = vcb_align(off + n * sizeof(ml::SolveCtl));
  if (off > b->arena_cap) {
    ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (b->arena) hipFree(b->arena);
    if (b->harena) hipHostFree(b->harena);
    b->arena = nullptr;
    b->harena = nullptr;
    b->arena_cap = 0;
    void* p = nullptr;
    ML_HIP(ctx, hipMalloc(&p, off));
    b->arena = static_cast<unsigned char*>(p);
    ML_HIP(ctx, hipHostMalloc(&p, off, hipHostMallocDefault));
    b->harena = static_cast<unsigned char*>(p);
    b->arena_cap = off;
  } else {
    // the previous upload may still be in flight from the pinned mirror
    ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  for (int par = 0; par < 2; ++par) {
    std::memcpy(b->harena + joff[par], jobs[par].data(), jobs[par].size() * sizeof(ml::VCRelaxJob<double>));
    std::memcpy(b->harena + moff[par], maps[par].data(), maps[par].size() * sizeof(unsigned));
    for (auto& Lh : b->launch[par]) {
      Lh.job_off = joff[par] + Lh.job_off * sizeof(ml::VCRelaxJob<double>);
      Lh.map_off = moff[par] + Lh.map_off * sizeof(unsigned);
    }
  }
  void** dp = reinterpret_cast<void**>(b->harena + b->desc_off);
  void** cp = reinterpret_cast<void**>(b->harena + b->ctlp_off);
  void** sp = reinterpret_cast<void**>(b->harena + b->slotp_off);
  for (int u = 0; u < n; ++u) {
    ml::VCWork& W = b->units[u].op->work[b->units[u].w];
    dp[u] = W.desc;
    cp[u] = W.ctl;
    sp[u] = W.slots;
  }
  std::memcpy(b->harena + b->bw_job_off, bw.data(), bw.size() * sizeof(ml::WordsJob));
  std::memcpy(b->harena + b->bw_map_off, bwm.data(), bwm.size() * sizeof(unsigned));
  std::memcpy(b->harena + b->br_job_off, br.data(), br.size() * sizeof(ml::VCResidJob<double>));
  std::memcpy(b->harena + b->br_map_off, brm.data(), brm.size() * sizeof(unsigned));
  ML_HIP(ctx, hipMemcpyAsync(b->arena, b->harena, b->ctl_off, hipMemcpyHostToDevice, ctx->stream));
  b->key = key;
  return ML_OK;
}

// one iteration of every unit (parity par = (iteration - 1) & 1)
int vcb_iteration(ml_vc_batch b, int par) {
  ml_ctx ctx = b->ctx;
  hipStream_t s = ctx->stream;
  const int K = b->K;
  auto launch = [&](const ml::VCBatchLaunch& Lh) -> int {
    ML_HIP(ctx, ml::vc_relax_batch<double>(b->dim, b->shape, reinterpret_cast<const ml::VCRelaxJob<double>*>(b->arena + Lh.job_off),
                                           reinterpret_cast<const unsigned*>(b->arena + Lh.map_off), Lh.blocks, Lh.lds,
                                           s));
    return ML_OK;
  };
  int rc;
  for (int k = 0; k < K; ++k)
    if ((rc = launch(b->launch[par][k]))) return rc;
  ML_HIP(ctx, ml::vc_coarse_batch<double>(b->dim, reinterpret_cast<const ml::VCDesc<double>* const*>(b->arena + b->desc_off),
                                          int(b->units.size()), b->coarse_lds, s));
  for (int k = K; k < 2 * K; ++k)
    if ((rc = launch(b->launch[par][k]))) return rc;
  return ML_OK;
}

int vcb_solve(ml_vc_batch b, double* const phi[], const double* const rhs[], const ml_solve_opts& o,
              ml_solve_stats* st) {
  ml_ctx ctx = b->ctx;
  hipStream_t s = ctx->stream;
  const int n = int(b->units.size());
  if (o.flags & ML_SOLVE_HOST_DRIVER) return fail(ctx, ML_ENOTSUP, "batched solve: no host driver");
  if (o.nu2 < 1) return fail(ctx, ML_ENOTSUP, "batched solve: needs post-smoothing (nu2 >= 1)");
  int rc;
  for (int u = 0; u < n; ++u)
    if ((rc = vc_solve_check(b->units[u].op, o))) return rc;
  // coarse LDS of the bottom solver in use
  b->coarse_lds = 0;
  for (auto& U : b->units)
    b->coarse_lds = std::max(b->coarse_lds, o.bottom_solver == ML_BOTTOM_BICGSTAB ? U.op->coarse_lds_bicg : U.op->coarse_lds);
  // begin: descriptors, then for every unit in three launches its phi copy
  // and zeroed norm slots, its initial residual, its control block
  for (int u = 0; u < n; ++u) {
    ml_vc_op op = b->units[u].op;
    if ((rc = vc_upload_desc(op, op->work[b->units[u].w], o, s))) return rc;
  }
  if ((rc = vcb_tables(b, phi, rhs, o))) return rc;
  ML_HIP(ctx, ml::words_batch(reinterpret_cast<const ml::WordsJob*>(b->arena + b->bw_job_off),
                              reinterpret_cast<const unsigned*>(b->arena + b->bw_map_off), b->bw_blocks, s));
  ML_HIP(ctx, ml::vc_resid_batch<double>(b->dim, reinterpret_cast<const ml::VCResidJob<double>*>(b->arena + b->br_job_off),
                                         reinterpret_cast<const unsigned*>(b->arena + b->br_map_off), b->br_blocks, s));
  ML_HIP(ctx, ml::ctl_init_many(reinterpret_cast<ml::SolveCtl* const*>(b->arena + b->ctlp_off),
                                reinterpret_cast<const unsigned long long* const*>(b->arena + b->slotp_off), n,
                                o.max_iter, o.eps, o.hang, o.norm_thresh, s));
  std::vector<ml::SolveCtl> c(n);
  int launched = 0;
  int nb = std::max(1, std::min(b->batch_iters, o.max_iter));
  for (;;) {
    for (int i = 0; i < nb && launched < o.max_iter; ++i) {
      ++launched;
      if ((rc = vcb_iteration(b, (launched - 1) & 1))) return rc;
    }
    ML_HIP(ctx, ml::ctl_gather(reinterpret_cast<ml::SolveCtl* const*>(b->arena + b->ctlp_off),
                               reinterpret_cast<ml::SolveCtl*>(b->arena + b->ctl_off), n, s));
    ML_HIP(ctx, hipMemcpyAsync(c.data(), b->arena + b->ctl_off, n * sizeof(ml::SolveCtl), hipMemcpyDeviceToHost, s));
    ML_HIP(ctx, hipStreamSynchronize(s));
    bool all = true;
    for (int u = 0; u < n; ++u) all = all && c[u].done;
    if (all || launched >= o.max_iter) break;
    nb = 2;
  }
  // units that stopped after an odd number of iterations hold their state
  // in the partner buffer: copy it back, all in one launch
  int most = 1;
  std::vector<ml::WordsJob> ew;
  std::vector<unsigned> ewm;
  for (int u = 0; u < n; ++u) {
    ml_vc_op op = b->units[u].op;
    ml::VCWork& W = op->work[b->units[u].w];
    most = std::max(most, c[u].iter);
    if (c[u].iter & 1) {
      const long long nw = op->lev[0].g.ncell;
      const unsigned j = unsigned(ew.size());
      ew.push_back(ml::WordsJob{reinterpret_cast<const unsigned long long*>(W.alt0),
                                reinterpret_cast<unsigned long long*>(phi[u]), nw});
      for (long long ch = 0; ch * ml::kWordsChunk < nw; ++ch) ewm.push_back((j << ml::kBatchTileBits) | unsigned(ch));
    }
  }
  if (!ew.empty()) {
    // the host mirror is idle here (the stream was synchronised above)
    std::memcpy(b->harena + b->ew_job_off, ew.data(), ew.size() * sizeof(ml::WordsJob));
    std::memcpy(b->harena + b->ew_map_off, ewm.data(), ewm.size() * sizeof(unsigned));
    ML_HIP(ctx, hipMemcpyAsync(b->arena + b->ew_job_off, b->harena + b->ew_job_off, b->ctl_off - b->ew_job_off,
                               hipMemcpyHostToDevice, s));
    ML_HIP(ctx, ml::words_batch(reinterpret_cast<const ml::WordsJob*>(b->arena + b->ew_job_off),
                                reinterpret_cast<const unsigned*>(b->arena + b->ew_map_off), int(ewm.size()), s));
  }
  for (int u = 0; u < n; ++u) {
    if (st) {
      st[u].exit_status = c[u].status != 0 ? c[u].status : 2;
      st[u].iterations = c[u].iter;
      st[u].initial_residual = c[u].r0;
      st[u].final_residual = c[u].norm;
    }
  }
  b->batch_iters = most;
  return ML_OK;
}

}  // namespace

extern "C" {

int ml_vc_batch_create(ml_ctx ctx, int nops, const ml_vc_op ops[], int per_op, ml_vc_batch* out) {
  if (!ctx || nops < 1 || !ops || per_op < 1 || per_op > ML_VC_MAX_SOLVES || !out) return ML_EINVAL;
  *out = nullptr;
  std::unique_ptr<ml_vc_batch_s> b(new (std::nothrow) ml_vc_batch_s());
  if (!b) return ML_ENOMEM;
  b->ctx = ctx;
  b->dim = ops[0] ? ops[0]->dim : 2;
  for (int i = 0; i < nops; ++i) {
    ml_vc_op op = ops[i];
    if (!op || op->ctx != ctx) return fail(ctx, ML_EINVAL, "ml_vc_batch_create: operator of another context");
    for (int j = 0; j < i; ++j)
      if (ops[j] == op) return fail(ctx, ML_EINVAL, "ml_vc_batch_create: an operator appears twice");
    if (op->dim != 2) return fail(ctx, ML_ENOTSUP, "ml_vc_batch_create: 2D operators only");
    const int L = int(op->lev.size());
    if (op->lc >= L || op->lc < 1)
      return fail(ctx, ML_ENOTSUP, "ml_vc_batch_create: needs tiled levels above a single-workgroup coarse level");
    for (int l = 0; l < L; ++l)
      if (!ml::vc_batch_level_ok(op->lev[l].g))
        return fail(ctx, ML_ENOTSUP, "ml_vc_batch_create: levels above 512^2 cells are not batched");
    int rc = vc_work_ensure(op, per_op);
    if (rc) return rc;
    for (int w = 0; w < per_op; ++w) b->units.push_back(ml::VCBatchUnit{op, w, op->lc});
    b->K = std::max(b->K, op->lc);
  }
  // large batches are bandwidth-bound: wider tiles restage less halo
  long long cells = 0;
  for (const auto& U : b->units) cells += U.op->lev[0].g.nvalid();
  b->shape = (b->dim == 2 && cells >= kVCBatchWideCells) ? 1 : 0;
  *out = b.release();
  return ML_OK;
}

int ml_vc_batch_destroy(ml_vc_batch b) {
  if (!b) return ML_EINVAL;
  hipStreamSynchronize(b->ctx->stream);
  delete b;
  return ML_OK;
}

int ml_vc_batch_solve(ml_vc_batch b, double* const phi[], const double* const rhs[], const ml_solve_opts* opts,
                      ml_solve_stats stats[]) {
  if (!b || !phi || !rhs) return ML_EINVAL;
  const int n = int(b->units.size());
  for (int u = 0; u < n; ++u)
    if (!phi[u] || !rhs[u]) return ML_EINVAL;
  ml_solve_opts o;
  if (opts)
    o = *opts;
  else
    ml_default_opts(&o);
  return vcb_solve(b, phi, rhs, o, stats);
}

int ml_vc_batch_solve_be(ml_vc_batch b, double* const phi_new[], const double* const phi_old[],
                         const double* const src[], double dt, const ml_solve_opts* opts, ml_solve_stats stats[]) {
  if (!b || !phi_new || !phi_old) return ML_EINVAL;
  const int n = int(b->units.size());
  ml_solve_opts o;
  if (opts)
    o = *opts;
  else
    ml_default_opts(&o);
  std::vector<const double*> rhs(n);
  for (int u = 0; u < n; ++u) {
    if (!phi_new[u] || !phi_old[u]) return ML_EINVAL;
    ml_vc_op op = b->units[u].op;
    op->alpha = 1.0;
    op->beta = -dt;
    ml::VCWork& W = op->work[b->units[u].w];
    ML_HIP(b->ctx, ml::be_rhs<double>(op->lev[0].g, 1, W.rhs_be, phi_old[u], op->has_a ? op->lev[0].a : nullptr, dt,
                                      src ? src[u] : nullptr, b->ctx->stream));
    rhs[u] = W.rhs_be;
  }
  return vcb_solve(b, phi_new, rhs.data(), o, stats);
}

}  // extern "C"
