// ml_batch.hip — step set-up and wrap-up kernels of the batched form: many
// independent (H, Theta) instances of different grid sizes and parameters
// share every launch.  A launch's workgroups are mapped to (instance, chunk)
// pairs by a table built once per batch (map[w] = instance << 20 | chunk).
//
// Per instance the set-up reproduces, value for value, what a single
// instance's step computes before its FP32-refined solve:
//   porosity (CALCULATEPOROSITY_MULTICOMP, src/EnthalpyVariablesF.ChF:56-94)
//   -> bCoef on the faces (srcSubcycle/AMRLevelMushyLayerInit.cpp:914-953,
//   geometric face porosity src/CellToEdge2F.ChF:20-46)
//   -> the MG hierarchy of bCoef, each coarse face the average of its fine
//   sub-faces (MGnewOp, src/AMRNonLinearMultiCompOp.cpp:1423-1494)
//   -> FP32 copies of every level's faces
//   -> backward-Euler rhs a*phi_old with a == 1, plus dt*src with a source.
#include <hip/hip_runtime.h>

#include "ml_kernels.h"

namespace ml {
namespace {

constexpr int kChunkThreads = 256;
constexpr int kChunkItems = 4 * kChunkThreads;

__device__ inline void unflat(long long c, int nx, int ny, int& i, int& j, int& k) {
  i = int(c % nx);
  const long long r = c / nx;
  j = int(r % ny);
  k = int(r / ny);
}

// items of an instance: [state copies (2 comps x ncell) | faces d = 0 ..
// dim-1 | norm slots]
__global__ __launch_bounds__(kChunkThreads) void batch_setup_kernel(const BatchSetupJob* __restrict__ jobs,
                                                                   const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const BatchSetupJob& J = jobs[w >> kBatchTileBits];
  const Geom& g = J.g;
  const long long ncopy = 2 * g.ncell;
  long long fbeg[4];
  fbeg[0] = ncopy;
  for (int d = 0; d < 3; ++d) fbeg[d + 1] = fbeg[d] + (d < g.dim ? g.nface[d] : 0);
  const long long total = fbeg[3] + J.nslots;
  const long long first = (long long)(w & kBatchTileMask) * kChunkItems;
  // the initial residual runs with the instance's done flag cleared
  if ((w & kBatchTileMask) == 0 && threadIdx.x == 0) J.ctl->done = 0;
  for (int u = threadIdx.x; u < kChunkItems; u += kChunkThreads) {
    const long long t = first + u;
    if (t >= total) break;
    if (t < ncopy) {
      const double v = J.phi[t];
      J.alt[t] = v;
      double r = v * 1.0;
      if (J.src) r = r + J.dt * J.src[t];
      J.rhs[t] = r;
      continue;
    }
    if (t >= fbeg[3]) {
      J.slots[t - fbeg[3]] = 0ull;
      continue;
    }
    int d = 0;
    while (t >= fbeg[d + 1]) ++d;
    const long long c = t - fbeg[d];
    int i, j, k;
    unflat(c, int(g.fn[d][0]), int(g.fn[d][1]), i, j, k);
    const int li = wrap_clamp(i - (d == 0), g.n[0], g.periodic[0]);
    const int lj = wrap_clamp(j - (d == 1), g.n[1], g.periodic[1]);
    const int lk = g.dim == 3 ? wrap_clamp(k - (d == 2), g.n[2], g.periodic[2]) : 0;
    const int hi = wrap_clamp(i, g.n[0], g.periodic[0]);
    const int hj = wrap_clamp(j, g.n[1], g.periodic[1]);
    const int hk = g.dim == 3 ? wrap_clamp(k, g.n[2], g.periodic[2]) : 0;
    const long long ql = g.ci(li, lj, lk), qh = g.ci(hi, hj, hk);
    const double lo = porosity<double>(J.phi[ql], J.phi[ql + g.ncell], J.cp);
    const double up = porosity<double>(J.phi[qh], J.phi[qh + g.ncell], J.cp);
    const double cf = sqrt(fabs(lo * up));
    double bC = 1.0 * cf;
    double bH = 1.0 - cf;
    bH = bH * J.kcond;
    bH = bH + cf;
    bH = bH * (-J.DH);
    bC = bC * (-J.DC);
    const long long N = g.nface[d];
    J.b64[d][c] = bH;
    J.b64[d][c + N] = bC;
    J.b32[d][c] = float(bH);
    J.b32[d][c + N] = float(bC);
  }
}

// coarse faces of levels 1.. of an instance, straight from level 0: the
// ratio^(d-1) fine sub-faces summed lowest transverse index fastest, times
// 1/ratio^(d-1); FP64 and FP32 copies
__global__ __launch_bounds__(kChunkThreads) void batch_coarsen_kernel(const BatchCoarsenJob* __restrict__ jobs,
                                                                     const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const BatchCoarsenJob& J = jobs[w >> kBatchTileBits];
  const Geom& gf = J.g0;
  const long long total = J.begin[J.njobs];
  const long long first = (long long)(w & kBatchTileMask) * kChunkItems;
  for (int u = threadIdx.x; u < kChunkItems; u += kChunkThreads) {
    const long long t = first + u;
    if (t >= total) break;
    int jb = 0;
    while (t >= J.begin[jb + 1]) ++jb;
    const long long item = t - J.begin[jb];
    const int l = J.level[jb], d = J.dir[jb];
    const int ratio = 1 << l;
    const Geom gc = coarsen_geom(gf, ratio);
    int r[3] = {ratio, ratio, gf.dim == 3 ? ratio : 1};
    r[d] = 1;
    const double scale = 1.0 / double(r[0] * r[1] * r[2]);
    const long long N = gc.nface[d];
    const int n = int(item / N);
    int I, Jj, K;
    unflat(item % N, int(gc.fn[d][0]), int(gc.fn[d][1]), I, Jj, K);
    const int base[3] = {I * (d == 0 ? ratio : r[0]), Jj * (d == 1 ? ratio : r[1]), K * (d == 2 ? ratio : r[2])};
    const double* bf = J.b0[d];
    double sum = 0.0;
    for (int kk = 0; kk < r[2]; ++kk)
      for (int jj = 0; jj < r[1]; ++jj)
        for (int ii = 0; ii < r[0]; ++ii)
          sum = sum + bf[gf.fi(d, base[0] + ii, base[1] + jj, base[2] + kk) + n * gf.nface[d]];
    const double v = sum * scale;
    const long long q = gc.fi(d, I, Jj, K) + n * N;
    J.b64[jb][q] = v;
    J.b32[jb][q] = float(v);
  }
}

// the iteration control of every instance from its initial residual norm
__global__ void batch_ctl_init_kernel(SolveCtl* ctl, const unsigned long long* slots, int nslots, int n, int max_iter,
                                      double eps, double hang, double thresh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SolveCtl* c = ctl + i;
  const double r0 = __longlong_as_double((long long)slots[(long long)i * nslots]);
  c->iter = 0;
  c->need_host = 0;
  c->max_iter = max_iter;
  c->eps = eps;
  c->hang = hang;
  c->thresh = thresh;
  c->r0 = r0;
  c->prev = r0;
  c->norm = r0;
  c->status = 0;
  c->done = 0;
  if (r0 <= thresh) {
    c->status = 8;
    c->done = 1;
  } else if (max_iter <= 0) {
    c->status = 2;
    c->done = 1;
  }
}

// an instance whose accepted iterations are odd ends in its partner array:
// copy it back (ghost cells included)
__global__ __launch_bounds__(kChunkThreads) void batch_finish_kernel(const BatchSetupJob* __restrict__ jobs,
                                                                    const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const BatchSetupJob& J = jobs[w >> kBatchTileBits];
  if (!(J.ctl->iter & 1)) return;
  const long long total = 2 * J.g.ncell;
  const long long first = (long long)(w & kBatchTileMask) * kChunkItems;
  for (int u = threadIdx.x; u < kChunkItems; u += kChunkThreads) {
    const long long t = first + u;
    if (t >= total) break;
    J.phi[t] = J.alt[t];
  }
}

}  // namespace

long long batch_chunk_items() { return kChunkItems; }

long long batch_setup_items(const BatchSetupJob& J) {
  long long n = 2 * J.g.ncell + J.nslots;
  for (int d = 0; d < J.g.dim; ++d) n += J.g.nface[d];
  return n;
}

hipError_t batch_setup(const BatchSetupJob* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_setup_kernel, dim3(nblocks), dim3(kChunkThreads), 0, s, jobs, map);
  return hipGetLastError();
}

hipError_t batch_coarsen(const BatchCoarsenJob* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_coarsen_kernel, dim3(nblocks), dim3(kChunkThreads), 0, s, jobs, map);
  return hipGetLastError();
}

hipError_t batch_ctl_init(SolveCtl* ctl, const unsigned long long* slots, int nslots, int n, int max_iter, double eps,
                          double hang, double thresh, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_ctl_init_kernel, dim3((n + 255) / 256), dim3(256), 0, s, ctl, slots, nslots, n, max_iter,
                     eps, hang, thresh);
  return hipGetLastError();
}

hipError_t batch_finish(const BatchSetupJob* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_finish_kernel, dim3(nblocks), dim3(kChunkThreads), 0, s, jobs, map);
  return hipGetLastError();
}

}  // namespace ml
