// ml_transfer_kernels.hip — MG transfers (restriction of the state,
// prolongation of corrections), max-norms, coefficient averaging for the MG
// hierarchy and the HC coefficient assembly, for gfx950.
#include "ml_common.h"
#include "ml_kernels.h"

namespace ml {

namespace {

__device__ inline void unflatten(long long c, int nx, int ny, int& i, int& j, int& k) {
  i = int(c % nx);
  const long long r = c / nx;
  j = int(r % ny);
  k = int(r / ny);
}

inline int grid_for(long long n, int threads = 256) {
  long long b = (n + threads - 1) / threads;
  if (b > 65535 * 4) b = 65535 * 4;
  if (b < 1) b = 1;
  return int(b);
}

// RESTRICT (src/AMRNonLinearMultiCompOpF.ChF:555-585): phi_c = 0 + sum of
// children/2^d in Fortran order; optionally also stored in phic0 (the FAS
// initial coarse state).
template <typename T, int DIM>
__global__ void restrict_r_kernel(Geom gf, Geom gc, int ncomp, const T* __restrict__ phif,
                                  T* __restrict__ phic, T* __restrict__ phic0) {
  const long long N = gc.nvalid();
  const T denom = DIM == 3 ? T(8) : T(4);
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N;
       c += (long long)gridDim.x * blockDim.x) {
    int I, J, K;
    unflatten(c, gc.n[0], gc.n[1], I, J, K);
    const long long qc = gc.ci(I, J, K);
    for (int n = 0; n < ncomp; ++n) {
      T s = T(0);
      for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
        for (int jj = 0; jj < 2; ++jj)
          for (int ii = 0; ii < 2; ++ii)
            s = s + phif[gf.ci(2 * I + ii, 2 * J + jj, DIM == 3 ? 2 * K + kk : 0) + n * gf.ncell] / denom;
      phic[qc + n * gc.ncell] = s;
      if (phic0) phic0[qc + n * gc.ncell] = s;
    }
  }
}

// Linear prolongation of e_c = phic - phic0 (PROLONG_JP weights,
// src/AMRProjectionOpF.ChF:15-77); coarse ghosts by periodic image or the
// no-flux rule (nearest interior index).
template <typename T, int DIM>
__global__ void prolong_linear_kernel(Geom gf, Geom gc, int ncomp, T* __restrict__ phif,
                                      const T* __restrict__ phic, const T* __restrict__ phic0) {
  const long long N = gf.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N;
       c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    unflatten(c, gf.n[0], gf.n[1], i, j, k);
    const long long qf = gf.ci(i, j, k);
    for (int n = 0; n < ncomp; ++n) {
      const T* pc = phic + n * gc.ncell;
      const T* p0 = phic0 ? phic0 + n * gc.ncell : nullptr;
      phif[qf + n * gf.ncell] = prolong_value<T, DIM>(gc, i, j, k, pc, p0, phif[qf + n * gf.ncell]);
    }
  }
}

// phi_f += e_c(i/2) (piecewise constant, Chombo FORT_PROLONG semantics).
template <typename T, int DIM>
__global__ void prolong_constant_kernel(Geom gf, Geom gc, int ncomp, T* __restrict__ phif,
                                        const T* __restrict__ corr) {
  const long long N = gf.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N;
       c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    unflatten(c, gf.n[0], gf.n[1], i, j, k);
    const long long qf = gf.ci(i, j, k), qc = gc.ci(i / 2, j / 2, DIM == 3 ? k / 2 : 0);
    for (int n = 0; n < ncomp; ++n) phif[qf + n * gf.ncell] = phif[qf + n * gf.ncell] + corr[qc + n * gc.ncell];
  }
}

template <typename T>
__global__ void norm_max_kernel(Geom g, int ncomp, const T* __restrict__ x, unsigned long long* out) {
  const long long N = g.nvalid();
  T m = T(0);
  int nan = 0;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N * ncomp;
       c += (long long)gridDim.x * blockDim.x) {
    const int n = int(c / N);
    int i, j, k;
    unflatten(c % N, g.n[0], g.n[1], i, j, k);
    const T v = fabs(x[g.ci(i, j, k) + n * g.ncell]);
    if (v != v) nan = 1;
    else if (v > m) m = v;
  }
  block_max_atomic<T>(m, nan, out);
}

__global__ void zero_u64_kernel(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// stopping rule of the multigrid solves (status bits: 1 |r| <= eps*r0,
// 2 iteration maximum, 4 hang, 8 |r| <= thresh), evaluated by one thread

__device__ void ctl_init_one(SolveCtl* c, const unsigned long long* slots, int max_iter, double eps, double hang,
                             double thresh) {
  const double r0 = __longlong_as_double((long long)slots[0]);
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

__global__ void ctl_init_kernel(SolveCtl* c, const unsigned long long* slots, int max_iter, double eps, double hang,
                                double thresh) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctl_init_one(c, slots, max_iter, eps, hang, thresh);
}

__global__ void ctl_init_many_kernel(SolveCtl* const* ctls, const unsigned long long* const* slots, int n,
                                     int max_iter, double eps, double hang, double thresh) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ctl_init_one(ctls[i], slots[i], max_iter, eps, hang, thresh);
}

__global__ void ctl_check_kernel(SolveCtl* c, const unsigned long long* slots, int refined) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ctl_check_step(c, slots, refined);
}

// MG coefficient hierarchy (CoarseAverage / CoarseAverageFace from the
// finest level, MGnewOp src/AMRNonLinearMultiCompOp.cpp:1423-1494,
// src/DarcyBrinkmanOp.cpp:977-1008).  Faces: the ratio^(d-1) fine sub-faces
// summed lowest transverse index fastest, times 1/ratio^(d-1).  Cells: each
// fine row (fixed j, k) summed i fastest, the row sums added in (j, k)
// order, times 1/ratio^d.
// One coarse item (face or cell, all components flattened) of the average
// of level `level` from the finest level gf: src the fine faces of
// direction kind (kind < 3) or the fine cells (kind 3).
template <typename T, int DIM>
__device__ __forceinline__ void coarsen_one(const Geom& gf, int level, int kind, const T* __restrict__ src,
                                            T* __restrict__ dst, long long item) {
  const int ratio = 1 << level;
  const Geom gc = coarsen_geom(gf, ratio);
  if (kind < 3) {
    const int d = kind;
    const T* bf = src;
    int r[3] = {ratio, ratio, DIM == 3 ? ratio : 1};
    r[d] = 1;
    const T scale = T(1) / T(r[0] * r[1] * r[2]);
    const long long N = gc.nface[d];
    const int n = int(item / N);
    int I, Jj, K;
    unflatten(item % N, int(gc.fn[d][0]), int(gc.fn[d][1]), I, Jj, K);
    const int base[3] = {I * (d == 0 ? ratio : r[0]), Jj * (d == 1 ? ratio : r[1]), K * (d == 2 ? ratio : r[2])};
    T sum = T(0);
    for (int kk = 0; kk < r[2]; ++kk)
      for (int jj = 0; jj < r[1]; ++jj)
#pragma unroll 8
        for (int ii = 0; ii < r[0]; ++ii)
          sum = sum + bf[gf.fi(d, base[0] + ii, base[1] + jj, base[2] + kk) + n * gf.nface[d]];
    dst[gc.fi(d, I, Jj, K) + n * gc.nface[d]] = sum * scale;
  } else {
    const T* af = src;
    const int rz = DIM == 3 ? ratio : 1;
    const T scale = T(1) / T(ratio * ratio * rz);
    const long long N = gc.nvalid();
    const int n = int(item / N);
    int I, Jj, K;
    unflatten(item % N, gc.n[0], gc.n[1], I, Jj, K);
    T sum = T(0);
    for (int kk = 0; kk < rz; ++kk)
      for (int jj = 0; jj < ratio; ++jj) {
        const T* row = af + gf.ci(I * ratio, Jj * ratio + jj, K * rz + kk) + n * gf.ncell;
        T rs = T(0);
#pragma unroll 8
        for (int ii = 0; ii < ratio; ++ii) rs = rs + row[ii];
        sum = sum + rs;
      }
    dst[gc.ci(I, Jj, K) + n * gc.ncell] = sum * scale;
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void coarsen_hierarchy_kernel(Geom gf, int ncomp, CoarsenJobs<T> J, const T* bf0,
                                                                const T* bf1, const T* bf2, const T* af) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= J.begin[J.njobs]) return;
  int jb = 0;
  while (t >= J.begin[jb + 1]) ++jb;
  const int kind = J.kind[jb];
  const T* src = kind == 0 ? bf0 : (kind == 1 ? bf1 : (kind == 2 ? bf2 : af));
  coarsen_one<T, DIM>(gf, J.level[jb], kind, src, J.dst[jb], t - J.begin[jb]);
}

// batched hierarchies: workgroup w averages 256 items of record map[w] >> 32
template <typename T, int DIM>
__global__ __launch_bounds__(256) void coarsen_rec_kernel(const CoarsenRec<T>* __restrict__ recs,
                                                          const unsigned long long* __restrict__ map) {
  const unsigned long long w = map[blockIdx.x];
  const CoarsenRec<T>& R = recs[w >> 32];
  const long long item = (long long)(w & 0xffffffffull) * 256 + threadIdx.x;
  if (item >= R.count) return;
  coarsen_one<T, DIM>(R.gf, R.level, R.kind, R.src, R.dst, item);
}

// Heavy items of the hierarchy (large ratios): one workgroup per coarse
// item.  Faces: the fine sub-faces loaded in parallel, summed in the light
// kernel's order (lowest transverse index fastest) by one thread.  Cells:
// each thread sums fine rows (i fastest) into LDS, then one thread adds the
// row sums in (j, k) order.
template <typename T, int DIM>
__device__ __forceinline__ void coarsen_block_one(const Geom& gf, int level, int kind, const T* __restrict__ src,
                                                  T* __restrict__ dst, long long item) {
  constexpr int kRows = 2048;
  __shared__ T rows[kRows];
  const int ratio = 1 << level;
  const Geom gc = coarsen_geom(gf, ratio);
  if (kind < 3) {
    const int d = kind;
    const T* bf = src;
    int r[3] = {ratio, ratio, DIM == 3 ? ratio : 1};
    r[d] = 1;
    const T scale = T(1) / T(r[0] * r[1] * r[2]);
    const long long NF = gc.nface[d];
    const int n = int(item / NF);
    int I, Jj, K;
    unflatten(item % NF, int(gc.fn[d][0]), int(gc.fn[d][1]), I, Jj, K);
    const int base[3] = {I * (d == 0 ? ratio : r[0]), Jj * (d == 1 ? ratio : r[1]), K * (d == 2 ? ratio : r[2])};
    const int count = r[0] * r[1] * r[2];
    T sum = T(0);
    for (int u0 = 0; u0 < count; u0 += kRows) {
      const int nu = (count - u0) < kRows ? (count - u0) : kRows;
      for (int u = threadIdx.x; u < nu; u += blockDim.x) {
        const int w = u0 + u, ii = w % r[0], rest = w / r[0], jj = rest % r[1], kk = rest / r[1];
        rows[u] = bf[gf.fi(d, base[0] + ii, base[1] + jj, base[2] + kk) + n * gf.nface[d]];
      }
      __syncthreads();
      if (threadIdx.x == 0)
        for (int u = 0; u < nu; ++u) sum = sum + rows[u];
      __syncthreads();
    }
    if (threadIdx.x == 0) dst[gc.fi(d, I, Jj, K) + n * gc.nface[d]] = sum * scale;
    return;
  }
  const T* af = src;
  const int rz = DIM == 3 ? ratio : 1;
  const T scale = T(1) / T(ratio * ratio * rz);
  const long long N = gc.nvalid();
  const int n = int(item / N);
  int I, Jj, K;
  unflatten(item % N, gc.n[0], gc.n[1], I, Jj, K);
  const int nrows = ratio * rz;   // rows (jj, kk), jj fastest
  T sum = T(0);
  for (int r0 = 0; r0 < nrows; r0 += kRows) {
    const int nr = (nrows - r0) < kRows ? (nrows - r0) : kRows;
    for (int u = threadIdx.x; u < nr; u += blockDim.x) {
      const int row = r0 + u, jj = row % ratio, kk = row / ratio;
      const T* srow = af + gf.ci(I * ratio, Jj * ratio + jj, K * rz + kk) + n * gf.ncell;
      T rs = T(0);
      for (int ii = 0; ii < ratio; ++ii) rs = rs + srow[ii];
      rows[u] = rs;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int u = 0; u < nr; ++u) sum = sum + rows[u];
    __syncthreads();
  }
  if (threadIdx.x == 0) dst[gc.ci(I, Jj, K) + n * gc.ncell] = sum * scale;
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void coarsen_cells_block_kernel(Geom gf, int ncomp, CoarsenJobs<T> J,
                                                                  const T* __restrict__ af, const T* bf0,
                                                                  const T* bf1, const T* bf2) {
  const long long t = blockIdx.x;
  if (t >= J.begin[J.njobs]) return;
  int jb = 0;
  while (t >= J.begin[jb + 1]) ++jb;
  const int kind = J.kind[jb];
  const T* src = kind == 0 ? bf0 : (kind == 1 ? bf1 : (kind == 2 ? bf2 : af));
  coarsen_block_one<T, DIM>(gf, J.level[jb], kind, src, J.dst[jb], t - J.begin[jb]);
}

// batched heavy items: workgroup w averages item map[w] & 0xffffffff of record map[w] >> 32
template <typename T, int DIM>
__global__ __launch_bounds__(256) void coarsen_rec_block_kernel(const CoarsenRec<T>* __restrict__ recs,
                                                                const unsigned long long* __restrict__ map) {
  const unsigned long long w = map[blockIdx.x];
  const CoarsenRec<T>& R = recs[w >> 32];
  coarsen_block_one<T, DIM>(R.gf, R.level, R.kind, R.src, R.dst, (long long)(w & 0xffffffffull));
}

__global__ void porosity_kernel(Geom g, ClosureP<double> cp, const double* __restrict__ phi,
                                double* __restrict__ chi) {
  const long long N = g.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N;
       c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    unflatten(c, g.n[0], g.n[1], i, j, k);
    const long long q = g.ci(i, j, k);
    chi[q] = porosity<double>(phi[q], phi[q + g.ncell], cp);
  }
}

// HC bCoef assembly (srcSubcycle/AMRLevelMushyLayerInit.cpp:914-953) with
// geometric face porosity (src/CellToEdge2F.ChF:20-46).
template <int DIM, class IX>
__device__ __forceinline__ void hc_bcoef_body(Geom g, int d, double kcond, double DH, double DC,
                                const double* __restrict__ chi, double* __restrict__ b, IX ix) {
  const long long N = g.nface[d];
  for (long long c = ix.start(); c < N;
       c += ix.stride()) {
    int i, j, k;
    unflatten(c, int(g.fn[d][0]), int(g.fn[d][1]), i, j, k);
    const int li = wrap_clamp(i - (d == 0), g.n[0], g.periodic[0]);
    const int lj = wrap_clamp(j - (d == 1), g.n[1], g.periodic[1]);
    const int lk = DIM == 3 ? wrap_clamp(k - (d == 2), g.n[2], g.periodic[2]) : 0;
    const int hi = wrap_clamp(i, g.n[0], g.periodic[0]);
    const int hj = wrap_clamp(j, g.n[1], g.periodic[1]);
    const int hk = DIM == 3 ? wrap_clamp(k, g.n[2], g.periodic[2]) : 0;
    const double lo = chi[g.ci(li, lj, lk)], up = chi[g.ci(hi, hj, hk)];
    const double cf = sqrt(fabs(lo * up));
    double bC = 1.0 * cf;
    double bH = 1.0 - cf;
    bH = bH * kcond;
    bH = bH + cf;
    bH = bH * (-DH);
    bC = bC * (-DC);
    b[c] = bH;
    b[c + N] = bC;
  }
}

template <int DIM>
__global__ void hc_bcoef_kernel(Geom g, int d, double kcond, double DH, double DC,
                                const double* __restrict__ chi, double* __restrict__ b) {
  hc_bcoef_body<DIM>(g, d, kcond, DH, DC, chi, b, GridIx{});
}

template <int DIM>
__global__ __launch_bounds__(256) void hc_bcoef_batch_kernel(int d, const FlowJob* __restrict__ jobs,
                                                             const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const FlowJob& J = jobs[w >> kBatchTileBits];
  hc_bcoef_body<DIM>(J.g, d, J.kcond, J.DH, J.DC, J.chi, J.b[d], JobIx{w & kBatchTileMask, J.nbf[d]});
}

// Darcy coefficient of the U* solve: c = (chi / Pi(chi)) * coeff with the
// Kozeny-Carman permeability Pi = chi^3/(1-chi)^2
// (srcSubcycle/AMRLevelMushyLayerInit.cpp:665-689: copy porosity, divide by
// permeability, multiply by m_darcyCoeff; src/MushyLayerParams.cpp:1046-1049),
// pow(x, 3.0) and pow(x, 2.0) as the once-rounded cube and square.
__global__ void darcy_ccoef_kernel(Geom g, double coeff, const double* __restrict__ chi, double* __restrict__ c) {
  const long long N = g.nvalid();
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < N;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    unflatten(t, g.n[0], g.n[1], i, j, k);
    const long long q = g.ci(i, j, k);
    const double liquid = chi[q], solid = 1.0 - liquid;
    const double perm = cube_rn(liquid) / (solid * solid);
    double v = liquid;
    v = v / perm;
    c[q] = v * coeff;
  }
}

// scale * sqrt(|chi_lo chi_hi|) on the faces of direction d
// (CELLTOEDGEGEOMETRIC, src/CellToEdge2F.ChF:20-46); wall faces use the
// interior cell.
template <int DIM>
__global__ void porosity_faces_kernel(Geom g, int d, double scale, const double* __restrict__ chi,
                                      double* __restrict__ b) {
  const long long N = g.nface[d];
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N;
       c += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    unflatten(c, int(g.fn[d][0]), int(g.fn[d][1]), i, j, k);
    const int li = wrap_clamp(i - (d == 0), g.n[0], g.periodic[0]);
    const int lj = wrap_clamp(j - (d == 1), g.n[1], g.periodic[1]);
    const int lk = DIM == 3 ? wrap_clamp(k - (d == 2), g.n[2], g.periodic[2]) : 0;
    const int hi = wrap_clamp(i, g.n[0], g.periodic[0]);
    const int hj = wrap_clamp(j, g.n[1], g.periodic[1]);
    const int hk = DIM == 3 ? wrap_clamp(k, g.n[2], g.periodic[2]) : 0;
    const double lo = chi[g.ci(li, lj, lk)], up = chi[g.ci(hi, hj, hk)];
    b[c] = scale * sqrt(fabs(lo * up));
  }
}

// y = x*a + b*z over valid cells (z may be null): backward-Euler rhs
// a*phi_old + dt*src with a == 1 when the aCoef array is absent.
template <typename T>
__global__ void be_rhs_kernel(Geom g, int ncomp, T* __restrict__ y, const T* __restrict__ x,
                              const T* __restrict__ acoef, T dt, const T* __restrict__ src) {
  const long long N = g.ncell * ncomp;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N;
       c += (long long)gridDim.x * blockDim.x) {
    T r = x[c] * (acoef ? acoef[c] : T(1));
    if (src) r = r + dt * src[c];
    y[c] = r;
  }
}

}  // namespace

template <typename T>
hipError_t restrict_r(const Geom& gf, int ncomp, const T* phif, T* phic, T* phic0, hipStream_t s) {
  Geom gc = coarsen_geom(gf, 2);
  const int b = grid_for(gc.nvalid());
  if (gf.dim == 2)
    hipLaunchKernelGGL((restrict_r_kernel<T, 2>), dim3(b), dim3(256), 0, s, gf, gc, ncomp, phif, phic, phic0);
  else
    hipLaunchKernelGGL((restrict_r_kernel<T, 3>), dim3(b), dim3(256), 0, s, gf, gc, ncomp, phif, phic, phic0);
  return hipGetLastError();
}

template <typename T>
hipError_t prolong_linear(const Geom& gf, int ncomp, T* phif, const T* phic, const T* phic0, hipStream_t s) {
  Geom gc = coarsen_geom(gf, 2);
  const int b = grid_for(gf.nvalid());
  if (gf.dim == 2)
    hipLaunchKernelGGL((prolong_linear_kernel<T, 2>), dim3(b), dim3(256), 0, s, gf, gc, ncomp, phif, phic, phic0);
  else
    hipLaunchKernelGGL((prolong_linear_kernel<T, 3>), dim3(b), dim3(256), 0, s, gf, gc, ncomp, phif, phic, phic0);
  return hipGetLastError();
}

template <typename T>
hipError_t prolong_constant(const Geom& gf, int ncomp, T* phif, const T* corr, hipStream_t s) {
  Geom gc = coarsen_geom(gf, 2);
  const int b = grid_for(gf.nvalid());
  if (gf.dim == 2)
    hipLaunchKernelGGL((prolong_constant_kernel<T, 2>), dim3(b), dim3(256), 0, s, gf, gc, ncomp, phif, corr);
  else
    hipLaunchKernelGGL((prolong_constant_kernel<T, 3>), dim3(b), dim3(256), 0, s, gf, gc, ncomp, phif, corr);
  return hipGetLastError();
}

template <typename T>
hipError_t norm_max(const Geom& g, int ncomp, const T* x, unsigned long long* out, hipStream_t s) {
  int b = grid_for(g.nvalid() * ncomp);
  if (b > 1024) b = 1024;
  hipLaunchKernelGGL((norm_max_kernel<T>), dim3(b), dim3(256), 0, s, g, ncomp, x, out);
  return hipGetLastError();
}

hipError_t zero_u64(unsigned long long* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

hipError_t ctl_init(SolveCtl* ctl, const unsigned long long* slots, int max_iter, double eps, double hang,
                    double thresh, hipStream_t s) {
  hipLaunchKernelGGL(ctl_init_kernel, dim3(1), dim3(64), 0, s, ctl, slots, max_iter, eps, hang, thresh);
  return hipGetLastError();
}

hipError_t ctl_init_many(SolveCtl* const* ctls, const unsigned long long* const* slots, int n, int max_iter, double eps,
                         double hang, double thresh, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(ctl_init_many_kernel, dim3((n + 63) / 64), dim3(64), 0, s, ctls, slots, n, max_iter, eps, hang,
                     thresh);
  return hipGetLastError();
}

hipError_t ctl_check(SolveCtl* ctl, const unsigned long long* slots, int refined, hipStream_t s) {
  hipLaunchKernelGGL(ctl_check_kernel, dim3(1), dim3(64), 0, s, ctl, slots, refined);
  return hipGetLastError();
}

template <typename T>
hipError_t coarsen_hierarchy(const Geom& gf, int ncomp, int nlev, const T* const* bf, T* const* const* bc, const T* af,
                             T* const* ac, hipStream_t s, bool faces) {
  CoarsenJobs<T> J{}, H{};
  long long total = 0, htotal = 0;
  // averages over at least kHeavy fine cells or faces go to the
  // workgroup-per-item kernel
  constexpr long long kHeavy = 64;
  for (int l = nlev - 1; l >= 1; --l) {
    const Geom gc = coarsen_geom(gf, 1 << l);
    const long long block = (1LL << l) * (1LL << l) * (gf.dim == 3 ? (1LL << l) : 1);
    const long long fblock = block >> l;   // fine faces per coarse face
    if (faces && fblock >= kHeavy)
      for (int d = 0; d < gf.dim; ++d) {
        if (H.njobs >= kMaxCoarsenJobs) return hipErrorInvalidValue;
        H.begin[H.njobs] = htotal;
        H.level[H.njobs] = short(l);
        H.kind[H.njobs] = short(d);
        H.dst[H.njobs] = bc[l][d];
        htotal += gc.nface[d] * ncomp;
        ++H.njobs;
      }
    if (af && block >= kHeavy) {
      if (H.njobs >= kMaxCoarsenJobs) return hipErrorInvalidValue;
      H.begin[H.njobs] = htotal;
      H.level[H.njobs] = short(l);
      H.kind[H.njobs] = 3;
      H.dst[H.njobs] = ac[l];
      htotal += gc.nvalid() * ncomp;
      ++H.njobs;
    }
    for (int kind = 0; kind < 4; ++kind) {
      if (kind < 3 && (!faces || kind >= gf.dim || fblock >= kHeavy)) continue;
      if (kind == 3 && (!af || block >= kHeavy)) continue;
      if (J.njobs >= kMaxCoarsenJobs) return hipErrorInvalidValue;
      J.begin[J.njobs] = total;
      J.level[J.njobs] = short(l);
      J.kind[J.njobs] = short(kind);
      J.dst[J.njobs] = kind < 3 ? bc[l][kind] : ac[l];
      total += (kind < 3 ? gc.nface[kind] : gc.nvalid()) * ncomp;
      ++J.njobs;
    }
  }
  J.begin[J.njobs] = total;
  H.begin[H.njobs] = htotal;
  if (htotal > 0) {
    if (gf.dim == 2)
      hipLaunchKernelGGL((coarsen_cells_block_kernel<T, 2>), dim3(unsigned(htotal)), dim3(256), 0, s, gf, ncomp, H, af,
                         bf[0], bf[1], bf[2]);
    else
      hipLaunchKernelGGL((coarsen_cells_block_kernel<T, 3>), dim3(unsigned(htotal)), dim3(256), 0, s, gf, ncomp, H, af,
                         bf[0], bf[1], bf[2]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (total == 0) return hipSuccess;
  const unsigned blocks = unsigned((total + 255) / 256);
  if (gf.dim == 2)
    hipLaunchKernelGGL((coarsen_hierarchy_kernel<T, 2>), dim3(blocks), dim3(256), 0, s, gf, ncomp, J, bf[0], bf[1],
                       bf[2], af);
  else
    hipLaunchKernelGGL((coarsen_hierarchy_kernel<T, 3>), dim3(blocks), dim3(256), 0, s, gf, ncomp, J, bf[0], bf[1],
                       bf[2], af);
  return hipGetLastError();
}

template <typename T>
void coarsen_records(const Geom& gf, int ncomp, int nlev, const T* const* bf, T* const* const* bc, const T* af,
                     T* const* ac, bool faces, std::vector<CoarsenRec<T>>& light, std::vector<CoarsenRec<T>>& heavy) {
  // averages over at least 1024 fine cells or faces are heavy (one
  // workgroup each); below that one thread sums an item, which in a batch of
  // many operators keeps every lane busy.  Light and heavy items add in the
  // same order, so the split does not change a bit of the result.
  constexpr long long kHeavy = 1024;
  for (int l = nlev - 1; l >= 1; --l) {
    const Geom gc = coarsen_geom(gf, 1 << l);
    const long long block = (1LL << l) * (1LL << l) * (gf.dim == 3 ? (1LL << l) : 1);
    const long long fblock = block >> l;
    if (faces)
      for (int d = 0; d < gf.dim; ++d)
        (fblock >= kHeavy ? heavy : light).push_back(CoarsenRec<T>{gf, l, d, bf[d], bc[l][d], gc.nface[d] * ncomp});
    if (af) (block >= kHeavy ? heavy : light).push_back(CoarsenRec<T>{gf, l, 3, af, ac[l], gc.nvalid() * ncomp});
  }
}

template <typename T>
hipError_t coarsen_recs(int dim, const CoarsenRec<T>* light, const unsigned long long* lmap, int lblocks,
                        const CoarsenRec<T>* heavy, const unsigned long long* hmap, int hblocks, hipStream_t s) {
  if (hblocks > 0) {
    if (dim == 2)
      hipLaunchKernelGGL((coarsen_rec_block_kernel<T, 2>), dim3(hblocks), dim3(256), 0, s, heavy, hmap);
    else
      hipLaunchKernelGGL((coarsen_rec_block_kernel<T, 3>), dim3(hblocks), dim3(256), 0, s, heavy, hmap);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (lblocks > 0) {
    if (dim == 2)
      hipLaunchKernelGGL((coarsen_rec_kernel<T, 2>), dim3(lblocks), dim3(256), 0, s, light, lmap);
    else
      hipLaunchKernelGGL((coarsen_rec_kernel<T, 3>), dim3(lblocks), dim3(256), 0, s, light, lmap);
  }
  return hipGetLastError();
}

template void coarsen_records<double>(const Geom&, int, int, const double* const*, double* const* const*,
                                      const double*, double* const*, bool, std::vector<CoarsenRec<double>>&,
                                      std::vector<CoarsenRec<double>>&);
template hipError_t coarsen_recs<double>(int, const CoarsenRec<double>*, const unsigned long long*, int,
                                         const CoarsenRec<double>*, const unsigned long long*, int, hipStream_t);

hipError_t porosity_field(const Geom& g, const ClosureP<double>& cp, const double* phi, double* chi,
                          hipStream_t s) {
  hipLaunchKernelGGL(porosity_kernel, dim3(grid_for(g.nvalid())), dim3(256), 0, s, g, cp, phi, chi);
  return hipGetLastError();
}

hipError_t hc_bcoef(const Geom& g, double k, double DH, double DC, const double* chi, double* const* b,
                    hipStream_t s) {
  for (int d = 0; d < g.dim; ++d) {
    const int blocks = grid_for(g.nface[d]);
    if (g.dim == 2)
      hipLaunchKernelGGL((hc_bcoef_kernel<2>), dim3(blocks), dim3(256), 0, s, g, d, k, DH, DC, chi, b[d]);
    else
      hipLaunchKernelGGL((hc_bcoef_kernel<3>), dim3(blocks), dim3(256), 0, s, g, d, k, DH, DC, chi, b[d]);
  }
  return hipGetLastError();
}

hipError_t hc_bcoef_batch(int dim, int d, const FlowJob* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  if (dim == 2)
    hipLaunchKernelGGL((hc_bcoef_batch_kernel<2>), dim3(nblocks), dim3(256), 0, s, d, jobs, map);
  else
    hipLaunchKernelGGL((hc_bcoef_batch_kernel<3>), dim3(nblocks), dim3(256), 0, s, d, jobs, map);
  return hipGetLastError();
}

hipError_t darcy_ccoef(const Geom& g, double coeff, const double* chi, double* c, hipStream_t s) {
  hipLaunchKernelGGL(darcy_ccoef_kernel, dim3(grid_for(g.nvalid())), dim3(256), 0, s, g, coeff, chi, c);
  return hipGetLastError();
}

hipError_t porosity_faces(const Geom& g, double scale, const double* chi, double* const* b, hipStream_t s) {
  for (int d = 0; d < g.dim; ++d) {
    const int blocks = grid_for(g.nface[d]);
    if (g.dim == 2)
      hipLaunchKernelGGL((porosity_faces_kernel<2>), dim3(blocks), dim3(256), 0, s, g, d, scale, chi, b[d]);
    else
      hipLaunchKernelGGL((porosity_faces_kernel<3>), dim3(blocks), dim3(256), 0, s, g, d, scale, chi, b[d]);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t be_rhs(const Geom& g, int ncomp, T* y, const T* x, const T* acoef, T dt, const T* src, hipStream_t s) {
  hipLaunchKernelGGL((be_rhs_kernel<T>), dim3(grid_for(g.ncell * ncomp)), dim3(256), 0, s, g, ncomp, y, x, acoef,
                     dt, src);
  return hipGetLastError();
}

template hipError_t restrict_r<double>(const Geom&, int, const double*, double*, double*, hipStream_t);
template hipError_t prolong_linear<double>(const Geom&, int, double*, const double*, const double*, hipStream_t);
template hipError_t prolong_constant<double>(const Geom&, int, double*, const double*, hipStream_t);
template hipError_t norm_max<double>(const Geom&, int, const double*, unsigned long long*, hipStream_t);
template hipError_t coarsen_hierarchy<double>(const Geom&, int, int, const double* const*, double* const* const*,
                                              const double*, double* const*, hipStream_t, bool);
template hipError_t be_rhs<double>(const Geom&, int, double*, const double*, const double*, double, const double*,
                                   hipStream_t);

}  // namespace ml

// ---- elementwise LevelData arithmetic ------------------------------------
namespace ml {
namespace {
__global__ void vec_axby_kernel(long long n, double* __restrict__ lhs, const double* __restrict__ x,
                                const double* __restrict__ y, double a, double b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    lhs[i] = y ? a * x[i] + b * y[i] : a * x[i];
}
__global__ void vec_incr_kernel(long long n, double* __restrict__ lhs, const double* __restrict__ x, double s) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    lhs[i] = lhs[i] + s * x[i];
}
__global__ void vec_mult_kernel(long long n, double* __restrict__ lhs, const double* __restrict__ x) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    lhs[i] = lhs[i] * x[i];
}
__global__ void vec_set_kernel(long long n, double* __restrict__ lhs, double v) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    lhs[i] = v;
}
inline int vgrid(long long n) {
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  return b < 1 ? 1 : int(b);
}
}  // namespace

hipError_t vec_axby(long long n, double* lhs, const double* x, const double* y, double a, double b, hipStream_t s) {
  hipLaunchKernelGGL(vec_axby_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, lhs, x, y, a, b);
  return hipGetLastError();
}
hipError_t vec_incr(long long n, double* lhs, const double* x, double sc, hipStream_t s) {
  hipLaunchKernelGGL(vec_incr_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, lhs, x, sc);
  return hipGetLastError();
}
hipError_t vec_mult(long long n, double* lhs, const double* x, hipStream_t s) {
  hipLaunchKernelGGL(vec_mult_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, lhs, x);
  return hipGetLastError();
}
hipError_t vec_set(long long n, double* lhs, double v, hipStream_t s) {
  hipLaunchKernelGGL(vec_set_kernel, dim3(vgrid(n)), dim3(256), 0, s, n, lhs, v);
  return hipGetLastError();
}
}  // namespace ml
