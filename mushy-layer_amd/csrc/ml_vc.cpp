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
#include <cstring>
#include <memory>

#include "ml_internal.h"

namespace {

using ml::fail;

int vc_alloc(ml_ctx ctx, size_t n, double** p) {
  void* v = nullptr;
  ML_HIP(ctx, hipMalloc(&v, n * sizeof(double)));
  *p = static_cast<double*>(v);
  ML_HIP(ctx, ml::vec_set((long long)n, *p, 0.0, ctx->stream));
  return ML_OK;
}

int vc_upload_desc(ml_vc_op op, const ml_solve_opts& o) {
  ml_ctx ctx = op->ctx;
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));   // the staging buffer may still be in flight
  ml::VCDesc<double>& D = *op->hdesc;
  std::memset(&D, 0, sizeof(D));
  D.l0 = op->lc;
  D.nlev = int(op->lev.size());
  D.nu1 = o.nu1;
  D.nu2 = o.nu2;
  D.bottom = o.bottom;
  D.bc = op->bc;
  D.alpha = op->alpha;
  D.beta = op->beta;
  D.skip = &op->ctl->done;
  for (int l = 0; l < D.nlev; ++l) {
    const ml::VCArgs<double> A = op->args(l);
    ml::VCLevel<double>& L = D.lev[l];
    L.g = A.g;
    L.a = A.a;
    L.c = A.c;
    for (int d = 0; d < 3; ++d) L.b[d] = A.b[d];
    L.e = op->lev[l].e[0];
    L.r = op->lev[l].r;
  }
  ML_HIP(ctx, hipMemcpyAsync(op->desc, op->hdesc, sizeof(D), hipMemcpyHostToDevice, ctx->stream));
  return ML_OK;
}

// one V-cycle of the correction equation; on return the level-0 correction
// is op->lev[0].e[ecur[0]]
int vc_vcycle(ml_vc_op op, const ml_solve_opts& o, int* ecur, const int* skip) {
  ml_ctx ctx = op->ctx;
  hipStream_t s = ctx->stream;
  const int L = int(op->lev.size());
  const int lc = op->lc;
  auto E = [&](int l) { return op->lev[l].e[ecur[l]]; };
  auto args = [&](int l) {
    ml::VCArgs<double> A = op->args(l);
    A.skip = skip;
    return A;
  };
  const int ltiles = lc < L ? lc : L - 1;
  const ml::LaunchTimer T = ctx->timer(ML_KT_VC_RELAX);
  const ml::LaunchTimer* tm = T.start ? &T : nullptr;
  if (lc >= L && L == 1) {
    ML_HIP(ctx, ml::vc_relax<double>(args(0), op->lev[0].e, &ecur[0], op->lev[0].r, 2 * o.bottom, true, nullptr,
                                      nullptr, s, tm));
    return ML_OK;
  }
  for (int l = 0; l < ltiles; ++l) {
    ML_HIP(ctx, ml::vc_relax_restrict<double>(args(l), op->lev[l].e, &ecur[l], op->lev[l].r, 2 * o.nu1, true,
                                               op->lev[l + 1].g, op->lev[l + 1].r, s, tm));
  }
  if (lc < L) {
    ML_HIP(ctx, ml::vc_coarse_vcycle<double>(op->desc, op->dim, op->coarse_lds, s));
  } else {
    ML_HIP(ctx, ml::vc_relax<double>(args(L - 1), op->lev[L - 1].e, &ecur[L - 1], op->lev[L - 1].r, 2 * o.bottom,
                                      true, nullptr, nullptr, s, tm));
  }
  for (int l = ltiles - 1; l >= 0; --l) {
    const ml::Geom gc = op->lev[l + 1].g;
    ML_HIP(ctx, ml::vc_relax<double>(args(l), op->lev[l].e, &ecur[l], op->lev[l].r, 2 * o.nu2, false, &gc,
                                      E(l + 1), s, tm));
  }
  return ML_OK;
}

int vc_read_ctl(ml_vc_op op, ml::SolveCtl* out) {
  ml_ctx ctx = op->ctx;
  ML_HIP(ctx, hipMemcpyAsync(op->hctl, op->ctl, sizeof(ml::SolveCtl), hipMemcpyDeviceToHost, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = *op->hctl;
  return ML_OK;
}

int vc_solve(ml_vc_op op, double* phi, const double* rhs, const ml_solve_opts& o, ml_solve_stats* st) {
  ml_ctx ctx = op->ctx;
  hipStream_t s = ctx->stream;
  if (o.max_iter + 1 > op->nslots) return fail(ctx, ML_EINVAL, "max_iter too large");
  if (o.nu1 < 0 || o.nu2 < 0 || o.bottom < 0) return fail(ctx, ML_EINVAL, "negative smoothing count");
  if (o.precision != ML_PRECISION_FP64) return fail(ctx, ML_ENOTSUP, "the linear operator solves in FP64 only");
  int rc = vc_upload_desc(op, o);
  if (rc) return rc;
  const int L = int(op->lev.size());
  const ml::Geom& g0 = op->lev[0].g;
  const size_t bytes0 = sizeof(double) * g0.ncell;
  ML_HIP(ctx, hipMemcpyAsync(op->alt0, phi, bytes0, hipMemcpyDeviceToDevice, s));
  double* const pb[2] = {phi, op->alt0};
  std::vector<int> ecur(L, 0);
  ML_HIP(ctx, ml::zero_u64(op->slots, o.max_iter + 1, s));
  ml::VCArgs<double> A0 = op->args(0);
  ML_HIP(ctx, ml::vc_residual<double>(A0, phi, nullptr, nullptr, rhs, op->lev[0].r, op->slots, 0, s));
  ML_HIP(ctx, ml::ctl_init(op->ctl, op->slots, o.max_iter, o.eps, o.hang, o.norm_thresh, s));
  const int* skip = &op->ctl->done;
  A0.skip = skip;
  // iterations are enqueued in batches; the device stops them (SolveCtl)
  int launched = 0;
  const bool lookahead = !(o.flags & ML_SOLVE_NO_LOOKAHEAD);
  int batch = lookahead ? std::max(1, std::min(op->batch, o.max_iter)) : 1;
  ml::SolveCtl c{};
  for (;;) {
    for (int b = 0; b < batch && launched < o.max_iter; ++b) {
      if ((rc = vc_vcycle(op, o, ecur.data(), skip))) return rc;
      ++launched;
      const int pcur = (launched - 1) & 1;   // phi += e alternates the state buffer
      // the residual kernel's last workgroup applies the stopping rule
      ML_HIP(ctx, ml::vc_residual<double>(A0, pb[pcur], op->lev[0].e[ecur[0]], pb[1 - pcur], rhs, op->lev[0].r,
                                           op->slots + launched, 0, s, ml::CtlFuse{op->ctl, op->slots, op->arrive, 0}));
    }
    if ((rc = vc_read_ctl(op, &c))) return rc;
    if (c.done || launched >= o.max_iter) break;
    batch = lookahead ? 2 : 1;
  }
  if (lookahead) op->batch = std::max(1, c.iter);
  if (c.iter & 1) ML_HIP(ctx, hipMemcpyAsync(phi, op->alt0, bytes0, hipMemcpyDeviceToDevice, s));
  if (st) {
    st->exit_status = c.status != 0 ? c.status : 2;
    st->iterations = c.iter;
    st->initial_residual = c.r0;
    st->final_residual = c.norm;
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
    if ((rc = vc_alloc(ctx, L.g.ncell, &L.e[0]))) return bail(rc);
    if ((rc = vc_alloc(ctx, L.g.ncell, &L.r))) return bail(rc);
    // tiled levels ping-pong their correction
    if (l < raw->lc || raw->lc >= depth)
      if ((rc = vc_alloc(ctx, L.g.ncell, &L.e[1]))) return bail(rc);
  }
  int rc;
  if (raw->lc < depth) {
    std::vector<ml::Geom> gs(depth);
    for (int l = 0; l < depth; ++l) gs[l] = raw->lev[l].g;
    raw->coarse_lds = ml::vc_coarse_lds_bytes<double>(gs.data(), raw->lc, depth);
    size_t mx = 0;
    if ((rc = ml::max_lds(ctx, &mx))) return bail(rc);
    if (raw->coarse_lds > mx) {
      ml::fail(ctx, ML_ENOTSUP, "the coarse MG levels exceed the workgroup's LDS");
      return bail(ML_ENOTSUP);
    }
  }
  if ((rc = vc_alloc(ctx, g0.ncell, &raw->alt0))) return bail(rc);
  if ((rc = vc_alloc(ctx, g0.ncell, &raw->rhs_be))) return bail(rc);
  raw->nslots = 1024;
  void* p = nullptr;
  if (hipMalloc(&p, sizeof(unsigned long long) * raw->nslots) != hipSuccess) return bail(ML_ENOMEM);
  raw->slots = static_cast<unsigned long long*>(p);
  if (hipHostMalloc(&p, sizeof(unsigned long long) * raw->nslots, hipHostMallocDefault) != hipSuccess)
    return bail(ML_ENOMEM);
  raw->hslots = static_cast<unsigned long long*>(p);
  if (hipMalloc(&p, sizeof(ml::VCDesc<double>)) != hipSuccess) return bail(ML_ENOMEM);
  raw->desc = static_cast<ml::VCDesc<double>*>(p);
  if (hipHostMalloc(&p, sizeof(ml::VCDesc<double>), hipHostMallocDefault) != hipSuccess) return bail(ML_ENOMEM);
  raw->hdesc = static_cast<ml::VCDesc<double>*>(p);
  if (hipMalloc(&p, sizeof(ml::SolveCtl)) != hipSuccess) return bail(ML_ENOMEM);
  raw->ctl = static_cast<ml::SolveCtl*>(p);
  if (hipMalloc(&p, sizeof(unsigned int)) != hipSuccess) return bail(ML_ENOMEM);
  raw->arrive = static_cast<unsigned int*>(p);
  if (hipMemsetAsync(raw->arrive, 0, sizeof(unsigned int), ctx->stream) != hipSuccess) return bail(ML_EHIP);
  if (hipHostMalloc(&p, sizeof(ml::SolveCtl), hipHostMallocDefault) != hipSuccess) return bail(ML_ENOMEM);
  raw->hctl = static_cast<ml::SolveCtl*>(p);
  *out = op.release();
  return ML_OK;
}

int ml_vc_op_destroy(ml_vc_op op) {
  if (!op) return ML_EINVAL;
  if (op->ctx) hipStreamSynchronize(op->ctx->stream);
  for (auto& L : op->lev) {
    for (int d = 0; d < 3; ++d)
      if (L.b[d]) hipFree(L.b[d]);
    if (L.a) hipFree(L.a);
    if (L.c) hipFree(L.c);
    if (L.e[0]) hipFree(L.e[0]);
    if (L.e[1]) hipFree(L.e[1]);
    if (L.r) hipFree(L.r);
    if (L.tmp) hipFree(L.tmp);
  }
  if (op->alt0) hipFree(op->alt0);
  if (op->rhs_be) hipFree(op->rhs_be);
  if (op->slots) hipFree(op->slots);
  if (op->hslots) hipHostFree(op->hslots);
  if (op->desc) hipFree(op->desc);
  if (op->hdesc) hipHostFree(op->hdesc);
  if (op->ctl) hipFree(op->ctl);
  if (op->arrive) hipFree(op->arrive);
  if (op->hctl) hipHostFree(op->hctl);
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
  ML_HIP(ctx, ml::vc_relax<double>(op->args(depth), buf, &cur, rhs, 2 * iters, false, nullptr, nullptr,
                                   ctx->stream));
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
  if (o.flags & ML_SOLVE_HOST_DRIVER) return fail(op->ctx, ML_ENOTSUP, "the linear solve runs on the fused engine");
  return vc_solve(op, phi, rhs, o, stats);
}

int ml_vc_solve_be(ml_vc_op op, double* phi_new, const double* phi_old, const double* src, double dt,
                   const ml_solve_opts* opts, ml_solve_stats* stats) {
  if (!op || !phi_new || !phi_old) return ML_EINVAL;
  op->alpha = 1.0;
  op->beta = -dt;
  ML_HIP(op->ctx, ml::be_rhs<double>(op->lev[0].g, 1, op->rhs_be, phi_old, op->has_a ? op->lev[0].a : nullptr, dt,
                                     src, op->ctx->stream));
  return ml_vc_solve(op, phi_new, op->rhs_be, opts, stats);
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
