// ml_capi.cpp — implementation of the C ABI declared in include/mlhip.h:
// contexts, device buffers, the nonlinear (H, Theta) operator with its MG
// hierarchy, and the whole-solve entry points.  Whole solves run the fused
// FAS engine (ml_hc_engine.cpp); with ML_SOLVE_HOST_DRIVER they instead run
// the reference-API driver (host/FASMultiGrid.H) over the per-operator ABI
// calls, as a host C++ caller of this library would.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../host/FASMultiGrid.H"
#include "../host/HipNonLinearMultiCompOp.H"
#include "ml_internal.h"

namespace ml {

int fail(ml_ctx ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int max_lds(ml_ctx ctx, size_t* out) {
  int v = 0;
  ML_HIP(ctx, hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  *out = size_t(v);
  return ML_OK;
}

// A level that does not coarsen and holds at most 1024 cells is relaxed by a
// single-workgroup kernel that reads periodic images in place; with an odd
// periodic length the two cells across the seam have the same colour, so
// their red/black order is not the reference's (which reads the seam from
// ghost cells filled before each colour).  Coarser levels are always even.
bool odd_periodic_small_level(int dim, const int n[], const int periodic[]) {
  if (ml_mg_depth(dim, n) != 1) return false;
  long long cells = 1;
  bool odd = false;
  for (int d = 0; d < dim; ++d) {
    cells *= n[d];
    if (periodic && periodic[d] && (n[d] & 1)) odd = true;
  }
  return odd && cells <= 1024;
}

bool valid_grid(int dim, const int n[], const int periodic[], double dx) {
  if (dim != 2 && dim != 3) return false;
  if (!(dx > 0)) return false;
  for (int d = 0; d < dim; ++d) {
    if (n[d] <= 0) return false;
    if (periodic && periodic[d] != 0 && periodic[d] != 1) return false;
  }
  return true;
}

struct HostDriver {
  std::vector<std::unique_ptr<mlhost::HipNonLinearMultiCompOp>> owned;
  std::vector<mlhost::HipNonLinearMultiCompOp*> ops;
  mlhost::FASMultiGrid<mlhost::HipNonLinearMultiCompOp> fas;
  mlhost::LevelData rhs;   // backward-Euler right-hand side
};

}  // namespace ml

namespace {

using ml::fail;
using ml::valid_grid;

void delete_host_driver(ml::HostDriver* p) { delete p; }
void delete_engine(ml::HCEngine* p) { delete p; }

}  // namespace

extern "C" {

// ---------------------------------------------------------------- context
int ml_ctx_create(int device, unsigned flags, ml_ctx* out) {
  (void)flags;
  if (!out) return ML_EINVAL;
  *out = nullptr;
  std::unique_ptr<ml_ctx_s> c(new (std::nothrow) ml_ctx_s());
  if (!c) return ML_ENOMEM;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ML_EHIP;
  if (device < 0 || device >= ndev) return ML_EINVAL;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) return ML_EHIP;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return ML_EHIP;
  if (hipMalloc(&c->d_norm, 64 * sizeof(unsigned long long)) != hipSuccess) return ML_ENOMEM;
  if (hipHostMalloc(&c->h_norm, 64 * sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess)
    return ML_ENOMEM;
  // the no-allocation entry point's static scratch: its address on this device
  if (ml::scratch_init(c.get(), device) != ML_OK) return ML_EHIP;
  *out = c.release();
  return ML_OK;
}

int ml_ctx_destroy(ml_ctx ctx) {
  if (!ctx) return ML_EINVAL;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->d_norm) hipFree(ctx->d_norm);
  if (ctx->h_norm) hipHostFree(ctx->h_norm);
  for (auto& e : ctx->events)
    if (e) hipEventDestroy(e);
  for (auto& e : ctx->timing_start) hipEventDestroy(e);
  for (auto& e : ctx->timing_stop) hipEventDestroy(e);
  ctx->tables.release();
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return ML_OK;
}

const char* ml_last_error(ml_ctx ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int ml_ctx_synchronize(ml_ctx ctx) {
  if (!ctx) return ML_EINVAL;
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ML_OK;
}

void* ml_ctx_stream(ml_ctx ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int ml_ctx_event_record(ml_ctx ctx, int slot) {
  if (!ctx || slot < 0 || slot >= 16) return ML_EINVAL;
  if (!ctx->events[slot]) ML_HIP(ctx, hipEventCreate(&ctx->events[slot]));
  ML_HIP(ctx, hipEventRecord(ctx->events[slot], ctx->stream));
  return ML_OK;
}

int ml_ctx_event_elapsed(ml_ctx ctx, int a, int b, float* ms) {
  if (!ctx || !ms || a < 0 || a >= 16 || b < 0 || b >= 16 || !ctx->events[a] || !ctx->events[b])
    return ML_EINVAL;
  ML_HIP(ctx, hipEventSynchronize(ctx->events[b]));
  ML_HIP(ctx, hipEventElapsedTime(ms, ctx->events[a], ctx->events[b]));
  return ML_OK;
}

int ml_ctx_timing_start(ml_ctx ctx, int kernel, int max_launches) {
  if (!ctx || (kernel != ML_KT_VC_RELAX && kernel != ML_KT_HC_RELAX && kernel != ML_KT_BATCH_RELAX) ||
      max_launches <= 0)
    return ML_EINVAL;
  while (int(ctx->timing_start.size()) < max_launches) {
    hipEvent_t a, b;
    ML_HIP(ctx, hipEventCreate(&a));
    ML_HIP(ctx, hipEventCreate(&b));
    ctx->timing_start.push_back(a);
    ctx->timing_stop.push_back(b);
  }
  ctx->timing_cap = max_launches;   // the pool keeps its events (destroyed with the context)
  ctx->timing_kernel = kernel;
  ctx->timing_n = 0;
  ctx->timing_missed = 0;
  ctx->timing_bytes = 0;
  return ML_OK;
}

int ml_ctx_timing_stop(ml_ctx ctx, long long* launches, double* total_ms, double* bytes, long long* missed) {
  if (!ctx || !launches || !total_ms || !bytes || !missed) return ML_EINVAL;
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double ms = 0;
  for (int i = 0; i < ctx->timing_n; ++i) {
    float t = 0;
    ML_HIP(ctx, hipEventElapsedTime(&t, ctx->timing_start[i], ctx->timing_stop[i]));
    ms += t;
  }
  *launches = ctx->timing_n;
  *total_ms = ms;
  *bytes = ctx->timing_bytes;
  *missed = ctx->timing_missed;
  ctx->timing_kernel = 0;
  ctx->timing_n = 0;
  return ML_OK;
}

int ml_buf_alloc(ml_ctx ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return ML_EINVAL;
  *dptr = nullptr;
  if (hipMalloc(dptr, bytes ? bytes : 8) != hipSuccess) return fail(ctx, ML_ENOMEM, "hipMalloc failed");
  return ML_OK;
}

int ml_buf_free(ml_ctx ctx, void* dptr) {
  if (!ctx) return ML_EINVAL;
  if (dptr) {
    hipStreamSynchronize(ctx->stream);
    ML_HIP(ctx, hipFree(dptr));
  }
  return ML_OK;
}

int ml_buf_upload(ml_ctx ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || !dst || !src) return ML_EINVAL;
  ML_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ML_OK;
}

int ml_buf_download(ml_ctx ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || !dst || !src) return ML_EINVAL;
  ML_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return ML_OK;
}

int ml_buf_copy(ml_ctx ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || !dst || !src) return ML_EINVAL;
  ML_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return ML_OK;
}

size_t ml_cell_count(int dim, const int n[]) {
  if ((dim != 2 && dim != 3) || !n) return 0;
  size_t r = size_t(n[0] + 2) * size_t(n[1] + 2);
  if (dim == 3) r *= size_t(n[2] + 2);
  return r;
}

size_t ml_face_count(int dim, const int n[], int dir) {
  if ((dim != 2 && dim != 3) || !n || dir < 0 || dir >= dim) return 0;
  size_t r = 1;
  for (int d = 0; d < dim; ++d) r *= size_t(n[d] + (d == dir));
  return r;
}

int ml_mg_depth(int dim, const int n[]) {
  if ((dim != 2 && dim != 3) || !n) return 0;
  int depth = 1;
  long long coarsening = 2;
  for (;;) {
    bool ok = true;
    for (int d = 0; d < dim; ++d)
      if (n[d] % (coarsening * 2) != 0) ok = false;
    if (!ok) break;
    ++depth;
    coarsening *= 2;
  }
  return depth;
}

void ml_default_opts(ml_solve_opts* o) {
  if (!o) return;
  o->nu1 = 1;
  o->nu2 = 4;
  o->bottom = 40;
  o->max_iter = 10;
  o->eps = 1e-10;
  o->hang = 1e-10;
  o->norm_thresh = 1e-10;
  o->precision = ML_PRECISION_FP64;
  o->flags = 0;
  o->bottom_solver = ML_BOTTOM_RELAX;
  o->bottom_max_iter = 80;
}

int ml_norm_max(ml_ctx ctx, int dim, const int n[], int ncomp, const double* x, double* out) {
  if (!ctx || !n || !x || !out || ncomp <= 0) return ML_EINVAL;
  int per[3] = {0, 0, 0};
  if (!valid_grid(dim, n, per, 1.0)) return fail(ctx, ML_EINVAL, "ml_norm_max: bad grid");
  ml::Geom g = ml::make_geom(dim, n, per, 1.0);
  ML_HIP(ctx, ml::zero_u64(ctx->d_norm, 1, ctx->stream));
  ML_HIP(ctx, ml::norm_max<double>(g, ncomp, x, ctx->d_norm, ctx->stream));
  ML_HIP(ctx, hipMemcpyAsync(ctx->h_norm, ctx->d_norm, 8, hipMemcpyDeviceToHost, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  unsigned long long bits = ctx->h_norm[0];
  std::memcpy(out, &bits, 8);
  return ML_OK;
}

int ml_vec_axby(ml_ctx ctx, size_t n, double* lhs, const double* x, const double* y, double a, double b) {
  if (!ctx || !lhs || !x) return ML_EINVAL;
  ML_HIP(ctx, ml::vec_axby((long long)n, lhs, x, y, a, b, ctx->stream));
  return ML_OK;
}
int ml_vec_incr(ml_ctx ctx, size_t n, double* lhs, const double* x, double scale) {
  if (!ctx || !lhs || !x) return ML_EINVAL;
  ML_HIP(ctx, ml::vec_incr((long long)n, lhs, x, scale, ctx->stream));
  return ML_OK;
}
int ml_vec_mult(ml_ctx ctx, size_t n, double* lhs, const double* x) {
  if (!ctx || !lhs || !x) return ML_EINVAL;
  ML_HIP(ctx, ml::vec_mult((long long)n, lhs, x, ctx->stream));
  return ML_OK;
}
int ml_vec_set(ml_ctx ctx, size_t n, double* lhs, double v) {
  if (!ctx || !lhs) return ML_EINVAL;
  ML_HIP(ctx, ml::vec_set((long long)n, lhs, v, ctx->stream));
  return ML_OK;
}

// ------------------------------------------------------------ HC operator
int ml_hc_op_create(ml_ctx ctx, int dim, const int lo[], const int hi[], double dx, const int periodic[],
                    const ml_params* params, const ml_bc* bc, ml_hc_op* out) {
  if (!ctx || !lo || !hi || !periodic || !params || !bc || !out) return ML_EINVAL;
  *out = nullptr;
  int n[3] = {1, 1, 1};
  for (int d = 0; d < dim && d < 3; ++d) {
    if (lo[d] != 0) return fail(ctx, ML_ENOTSUP, "ml_hc_op_create: boxes must start at the origin (lo == 0)");
    n[d] = hi[d] - lo[d] + 1;
  }
  if (!valid_grid(dim, n, periodic, dx)) return fail(ctx, ML_EINVAL, "ml_hc_op_create: bad grid");
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      for (int c = 0; c < 2; ++c)
        if (bc->type[d][s][c] != ML_BC_DIRICHLET && bc->type[d][s][c] != ML_BC_NEUMANN)
          return fail(ctx, ML_EINVAL, "ml_hc_op_create: unknown BC type");
  if (params->waterDistributionCoeff == 0.0 || params->specificHeatRatio == 0.0)
    return fail(ctx, ML_EINVAL, "ml_hc_op_create: pc and c must be non-zero");
  std::unique_ptr<ml_hc_op_s> op(new (std::nothrow) ml_hc_op_s());
  if (!op) return ML_ENOMEM;
  op->ctx = ctx;
  op->dim = dim;
  op->params = *params;
  op->cp = {params->stefan, params->compositionRatio, params->waterDistributionCoeff, params->specificHeatRatio,
            params->thetaEutectic, params->ThetaEutectic};
  std::memcpy(op->bc.type, bc->type, sizeof(bc->type));
  std::memcpy(op->bc.val, bc->val, sizeof(bc->val));
  op->bc.apply = bc->apply;
  const int depth = ml_mg_depth(dim, n);
  if (depth > ml::kMaxLevels) return fail(ctx, ML_ENOTSUP, "too many MG levels");
  if (ml::odd_periodic_small_level(dim, n, periodic))
    return fail(ctx, ML_ENOTSUP, "odd periodic length on a level of at most 1024 cells that does not coarsen");
  ml::Geom g0 = ml::make_geom(dim, n, periodic, dx);
  op->lev.resize(depth);
  for (int l = 0; l < depth; ++l) {
    ml::HCLevel& L = op->lev[l];
    L.g = l == 0 ? g0 : ml::coarsen_geom(g0, 1 << l);
    for (int d = 0; d < dim; ++d) {
      void* p = nullptr;
      if (hipMalloc(&p, sizeof(double) * 2 * L.g.nface[d]) != hipSuccess) {
        ml_hc_op_destroy(op.release());
        return fail(ctx, ML_ENOMEM, "hipMalloc bCoef");
      }
      L.b[d] = static_cast<double*>(p);
      ml::vec_set(2 * L.g.nface[d], L.b[d], 0.0, ctx->stream);
    }
  }
  *out = op.release();
  return ML_OK;
}

int ml_hc_op_destroy(ml_hc_op op) {
  if (!op) return ML_EINVAL;
  if (op->ctx) hipStreamSynchronize(op->ctx->stream);
  for (auto& L : op->lev) {
    for (int d = 0; d < 3; ++d)
      if (L.b[d]) hipFree(L.b[d]);
    if (L.a) hipFree(L.a);
    if (L.tmp) hipFree(L.tmp);
  }
  if (op->adv) hipFree(op->adv);
  delete op;
  return ML_OK;
}

int ml_hc_num_levels(ml_hc_op op) { return op ? int(op->lev.size()) : 0; }

int ml_hc_level_info(ml_hc_op op, int depth, int n_out[3], double* dx_out) {
  if (!op || depth < 0 || depth >= int(op->lev.size())) return ML_EINVAL;
  const ml::Geom& g = op->lev[depth].g;
  if (n_out)
    for (int d = 0; d < 3; ++d) n_out[d] = g.n[d];
  if (dx_out) *dx_out = g.dx;
  return ML_OK;
}

int ml_hc_set_coefs(ml_hc_op op, const double* aCoef, const double* const bCoef[], double alpha, double beta) {
  if (!op || !bCoef) return ML_EINVAL;
  ml_ctx ctx = op->ctx;
  for (int d = 0; d < op->dim; ++d)
    if (!bCoef[d]) return fail(ctx, ML_EINVAL, "ml_hc_set_coefs: null bCoef");
  op->alpha = alpha;
  op->beta = beta;
  const ml::Geom& g0 = op->lev[0].g;
  for (int d = 0; d < op->dim; ++d)
    ML_HIP(ctx, hipMemcpyAsync(op->lev[0].b[d], bCoef[d], sizeof(double) * 2 * g0.nface[d], hipMemcpyDeviceToDevice,
                               ctx->stream));
  for (size_t l = 0; l < op->lev.size(); ++l) {
    ml::HCLevel& L = op->lev[l];
    if (aCoef && !L.a) {
      void* p = nullptr;
      if (hipMalloc(&p, sizeof(double) * 2 * L.g.ncell) != hipSuccess) return fail(ctx, ML_ENOMEM, "hipMalloc aCoef");
      L.a = static_cast<double*>(p);
      ML_HIP(ctx, ml::vec_set(2 * L.g.ncell, L.a, 0.0, ctx->stream));
    }
    if (!aCoef && L.a) {
      hipStreamSynchronize(ctx->stream);
      hipFree(L.a);
      L.a = nullptr;
    }
  }
  if (aCoef)
    ML_HIP(ctx, hipMemcpyAsync(op->lev[0].a, aCoef, sizeof(double) * 2 * g0.ncell, hipMemcpyDeviceToDevice,
                               ctx->stream));
  // MG coarsening straight from the finest level (MGnewOp,
  // src/AMRNonLinearMultiCompOp.cpp:1423-1494)
  const int L = int(op->lev.size());
  const double* bf[3] = {op->lev[0].b[0], op->lev[0].b[1], op->lev[0].b[2]};
  std::vector<double* const*> bc(L);
  std::vector<double*> ac(L, nullptr);
  for (int l = 0; l < L; ++l) {
    bc[l] = op->lev[l].b;
    ac[l] = op->lev[l].a;
  }
  ML_HIP(ctx, ml::coarsen_hierarchy<double>(g0, 2, L, bf, bc.data(), aCoef ? op->lev[0].a : nullptr, ac.data(),
                                            ctx->stream));
  return ML_OK;
}

int ml_hc_set_alpha_beta(ml_hc_op op, double alpha, double beta) {
  if (!op) return ML_EINVAL;
  op->alpha = alpha;
  op->beta = beta;
  return ML_OK;
}

int ml_hc_diagonal_scale(ml_hc_op op, int depth, double* rhs) {
  if (!op || !rhs || depth < 0 || depth >= int(op->lev.size())) return ML_EINVAL;
  if (!op->lev[depth].a) return ML_OK;
  ML_HIP(op->ctx, ml::vec_mult(2 * op->lev[depth].g.ncell, rhs, op->lev[depth].a, op->ctx->stream));
  return ML_OK;
}

#define HC_DEPTH_CHECK(op, depth)                                              \
  if (!(op) || (depth) < 0 || (depth) >= int((op)->lev.size())) return ML_EINVAL

int ml_hc_compute_diffused_var(ml_hc_op op, int depth, const double* phi, double* out, int homogeneous) {
  HC_DEPTH_CHECK(op, depth);
  if (!phi || !out) return ML_EINVAL;
  ML_HIP(op->ctx, ml::hc_diffused_var<double>(op->args(depth), phi, out, homogeneous, op->ctx->stream));
  return ML_OK;
}

int ml_hc_relax(ml_hc_op op, int depth, double* phi, const double* rhs, int iters) {
  HC_DEPTH_CHECK(op, depth);
  if (!phi || !rhs || iters < 0) return ML_EINVAL;
  if (iters == 0) return ML_OK;
  ml::HCLevel& L = op->lev[depth];
  const size_t bytes = sizeof(double) * 2 * L.g.ncell;
  if (!L.tmp) {
    void* p = nullptr;
    ML_HIP(op->ctx, hipMalloc(&p, bytes));
    L.tmp = static_cast<double*>(p);
  }
  // the tiled smoother ping-pongs between phi and the level's partner
  // array; the partner starts as a full copy so a state ending there can be
  // copied back whole
  ML_HIP(op->ctx, hipMemcpyAsync(L.tmp, phi, bytes, hipMemcpyDeviceToDevice, op->ctx->stream));
  double* const buf[2] = {phi, L.tmp};
  int cur = 0;
  const ml::LaunchTimer T = op->ctx->timer(ML_KT_HC_RELAX);
  ML_HIP(op->ctx, ml::hc_relax<double>(op->args(depth), buf, &cur, rhs, 2 * iters, 0, op->ctx->stream,
                                       T.start ? &T : nullptr));
  if (cur) ML_HIP(op->ctx, hipMemcpyAsync(phi, L.tmp, bytes, hipMemcpyDeviceToDevice, op->ctx->stream));
  return ML_OK;
}

int ml_hc_residual(ml_hc_op op, int depth, double* res, double* phi, const double* rhs, int homogeneous) {
  HC_DEPTH_CHECK(op, depth);
  if (!res || !phi || !rhs) return ML_EINVAL;
  ML_HIP(op->ctx, ml::hc_residual<double>(op->args(depth), phi, rhs, res, nullptr, homogeneous, op->ctx->stream));
  return ML_OK;
}

int ml_hc_apply_op(ml_hc_op op, int depth, double* lhs, double* phi, int homogeneous) {
  HC_DEPTH_CHECK(op, depth);
  if (!lhs || !phi) return ML_EINVAL;
  ML_HIP(op->ctx, ml::hc_apply_op<double>(op->args(depth), phi, lhs, nullptr, homogeneous, op->ctx->stream));
  return ML_OK;
}

int ml_hc_restrict_residual(ml_hc_op op, int depth, double* resCoarse, double* phiFine, const double* rhsFine) {
  HC_DEPTH_CHECK(op, depth);
  if (depth + 1 >= int(op->lev.size())) return fail(op->ctx, ML_EINVAL, "no coarser level");
  if (!resCoarse || !phiFine || !rhsFine) return ML_EINVAL;
  ML_HIP(op->ctx, ml::hc_restrict_residual<double>(op->args(depth), op->lev[depth + 1].g, phiFine, rhsFine,
                                                   resCoarse, 0, op->ctx->stream));
  return ML_OK;
}

int ml_hc_restrict_r(ml_hc_op op, int depth, double* phiCoarse, const double* phiFine) {
  HC_DEPTH_CHECK(op, depth);
  if (depth + 1 >= int(op->lev.size())) return fail(op->ctx, ML_EINVAL, "no coarser level");
  if (!phiCoarse || !phiFine) return ML_EINVAL;
  ML_HIP(op->ctx, ml::restrict_r<double>(op->lev[depth].g, 2, phiFine, phiCoarse, nullptr, op->ctx->stream));
  return ML_OK;
}

int ml_hc_prolong_increment(ml_hc_op op, int depth, double* phiFine, const double* corrCoarse) {
  HC_DEPTH_CHECK(op, depth);
  if (depth + 1 >= int(op->lev.size())) return fail(op->ctx, ML_EINVAL, "no coarser level");
  if (!phiFine || !corrCoarse) return ML_EINVAL;
  ML_HIP(op->ctx, ml::prolong_linear<double>(op->lev[depth].g, 2, phiFine, corrCoarse, nullptr, op->ctx->stream));
  return ML_OK;
}

int ml_hc_porosity(ml_hc_op op, const double* phi, double* chi) {
  if (!op || !phi || !chi) return ML_EINVAL;
  ML_HIP(op->ctx, ml::porosity_field(op->lev[0].g, op->cp, phi, chi, op->ctx->stream));
  return ML_OK;
}

int ml_hc_assemble_bcoef(ml_hc_op op, const double* chi, double* const bCoef[]) {
  if (!op || !chi || !bCoef) return ML_EINVAL;
  for (int d = 0; d < op->dim; ++d)
    if (!bCoef[d]) return ML_EINVAL;
  ML_HIP(op->ctx, ml::hc_bcoef(op->lev[0].g, op->params.heatConductivityRatio, op->params.heatDiffusionCoeff,
                               op->params.saltDiffusionCoeff, chi, bCoef, op->ctx->stream));
  return ML_OK;
}

// ------------------------------------------------------------- whole solves
namespace {

void set_params(mlhost::FASParams& p, const ml_solve_opts& o) {
  p.nu1 = o.nu1;
  p.nu2 = o.nu2;
  p.bottom = o.bottom;
  p.max_iter = o.max_iter;
  p.eps = o.eps;
  p.hang = o.hang;
  p.norm_thresh = o.norm_thresh;
}

ml::HostDriver& host_driver_of(ml_hc_op op) {
  if (!op->host_driver) {
    std::unique_ptr<ml::HostDriver> S(new ml::HostDriver());
    for (int l = 0; l < int(op->lev.size()); ++l) {
      S->owned.emplace_back(new mlhost::HipNonLinearMultiCompOp(op, op->ctx, l));
      S->ops.push_back(S->owned.back().get());
    }
    const ml::Geom& g = op->lev[0].g;
    S->rhs.define(op->ctx, g.dim, g.n, 2);
    S->fas.define(S->ops, S->rhs);
    op->host_driver = std::unique_ptr<ml::HostDriver, void (*)(ml::HostDriver*)>(S.release(), delete_host_driver);
  }
  return *op->host_driver;
}

ml::HCEngine& engine_of(ml_hc_op op) {
  if (!op->engine)
    op->engine = std::unique_ptr<ml::HCEngine, void (*)(ml::HCEngine*)>(new ml::HCEngine(op), delete_engine);
  return *op->engine;
}

void fill_stats(ml_solve_stats* st, const mlhost::SolveStats& s) {
  if (!st) return;
  st->exit_status = s.exit_status;
  st->iterations = s.iterations;
  st->initial_residual = s.initial_residual;
  st->final_residual = s.final_residual;
}

int check_opts(ml_hc_op op, const ml_solve_opts* opts, ml_solve_opts& o) {
  ml_default_opts(&o);
  if (opts) o = *opts;
  if (o.precision != ML_PRECISION_FP64 && o.precision != ML_PRECISION_FP32_REFINED)
    return fail(op->ctx, ML_EINVAL, "unknown precision form");
  if (o.precision != ML_PRECISION_FP64 && (o.flags & ML_SOLVE_HOST_DRIVER))
    return fail(op->ctx, ML_ENOTSUP, "the host driver runs the FP64 form only");
  if (o.nu1 < 0 || o.nu2 < 0 || o.bottom < 0 || o.max_iter < 0)
    return fail(op->ctx, ML_EINVAL, "negative smoothing or iteration count");
  return ML_OK;
}

}  // namespace

int ml_hc_solve(ml_hc_op op, double* phi, const double* rhs, const ml_solve_opts* opts, ml_solve_stats* stats) {
  if (!op || !phi || !rhs) return ML_EINVAL;
  ml_solve_opts o;
  if (int rc = check_opts(op, opts, o)) return rc;
  try {
    if (!(o.flags & ML_SOLVE_HOST_DRIVER)) return engine_of(op).solve(phi, rhs, o, stats);
    const ml::Geom& g = op->lev[0].g;
    ml::HostDriver& S = host_driver_of(op);
    mlhost::LevelData vphi, vrhs;
    vphi.alias(op->ctx, g.dim, g.n, 2, phi);
    vrhs.alias(op->ctx, g.dim, g.n, 2, const_cast<double*>(rhs));
    set_params(S.fas.params(), o);
    mlhost::SolveStats st;
    S.fas.solve(vphi, vrhs, &st);
    fill_stats(stats, st);
  } catch (const std::exception& e) {
    return fail(op->ctx, ML_EHIP, e.what());
  }
  return ML_OK;
}

int ml_hc_solve_be(ml_hc_op op, double* phi_new, const double* phi_old, const double* src, double dt,
                   const ml_solve_opts* opts, ml_solve_stats* stats) {
  if (!op || !phi_new || !phi_old) return ML_EINVAL;
  ml_solve_opts o;
  if (int rc = check_opts(op, opts, o)) return rc;
  try {
    if (!(o.flags & ML_SOLVE_HOST_DRIVER)) return engine_of(op).solve_be(phi_new, phi_old, src, dt, o, stats);
    const ml::Geom& g = op->lev[0].g;
    ml::HostDriver& S = host_driver_of(op);
    mlhost::LevelData vnew, vold, vsrc;
    vnew.alias(op->ctx, g.dim, g.n, 2, phi_new);
    vold.alias(op->ctx, g.dim, g.n, 2, const_cast<double*>(phi_old));
    if (src) vsrc.alias(op->ctx, g.dim, g.n, 2, const_cast<double*>(src));
    set_params(S.fas.params(), o);
    mlhost::SolveStats st;
    mlhost::backwardEulerStep(S.ops, S.fas, vnew, vold, src ? &vsrc : nullptr, dt, S.rhs, &st);
    fill_stats(stats, st);
  } catch (const std::exception& e) {
    return fail(op->ctx, ML_EHIP, e.what());
  }
  return ML_OK;
}

}  // extern "C"
