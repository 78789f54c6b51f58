// ml_noalloc.cpp — the entry point that allocates nothing: ml_hc_relax_noalloc.
//
// ml_hc_relax runs the tiled red/black smoother between the caller's state
// and a partner array of the same size (every launch reads one and writes
// the other).  ml_hc_relax allocates that partner per MG level on its first
// call; this variant takes it from a file-scope device array instead, so a
// caller that cannot allocate device memory (HIP graph capture, serving
// loops) can relax a level.  The kernels and the pass sequence are those of
// ml_hc_relax, so both give the same bits.
//
// The array is per GPU (each device holds its own copy for the life of the
// process).  Host code never uses its name as an address: its device address
// is looked up once per device with hipGetSymbolAddress while the context is
// created (ml_scratch_init, called from ml_ctx_create with that device
// current) and kept in g_scratch[].  Calls that overlap on one GPU race on
// the array.
#include <mutex>

#include "ml_internal.h"

namespace {

constexpr size_t kScratchBytes = size_t(ML_NOALLOC_SCRATCH_BYTES);
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

namespace ml {

// the device address of `scratch` on `device` (the current device), looked
// up once per device
int scratch_init(ml_ctx ctx, int device) {
  if (device < 0 || device >= kMaxDevices) return fail(ctx, ML_ENOTSUP, "no-allocation scratch: device index too large");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return ML_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(ctx, ML_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch[device] = p;
  return ML_OK;
}

}  // namespace ml

extern "C" {

int ml_hc_relax_noalloc(ml_hc_op op, int depth, double* phi, const double* rhs, int iters) {
  if (!op || depth < 0 || depth >= int(op->lev.size())) return ML_EINVAL;
  if (!phi || !rhs || iters < 0) return ML_EINVAL;
  ml_ctx ctx = op->ctx;
  const ml::HCLevel& L = op->lev[depth];
  const size_t bytes = sizeof(double) * 2 * L.g.ncell;
  if (bytes > kScratchBytes)
    return ml::fail(ctx, ML_ENOTSUP,
                    "ml_hc_relax_noalloc: the level's state (" + std::to_string(bytes) +
                        " bytes) exceeds the static scratch of " + std::to_string(kScratchBytes) + " bytes");
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    part = static_cast<double*>(g_scratch[ctx->device]);
  }
  if (!part) return ml::fail(ctx, ML_EHIP, "ml_hc_relax_noalloc: scratch not set up on this device");
  if (iters == 0) return ML_OK;
  // as ml_hc_relax: the partner starts as a full copy of phi, so a state
  // that ends in it can be copied back whole
  ML_HIP(ctx, hipMemcpyAsync(part, phi, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  double* const buf[2] = {phi, part};
  int cur = 0;
  const ml::LaunchTimer T = ctx->timer(ML_KT_HC_RELAX);
  ML_HIP(ctx, ml::hc_relax<double>(op->args(depth), buf, &cur, rhs, 2 * iters, 0, ctx->stream, T.start ? &T : nullptr));
  if (cur) ML_HIP(ctx, hipMemcpyAsync(phi, part, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return ML_OK;
}

}  // extern "C"
