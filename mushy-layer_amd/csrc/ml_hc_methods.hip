// ml_hc_methods.hip — the (H, Theta) operator's methods off the default
// single-level relaxation path: levelJacobi (relax mode 4), preCond (the
// preconditioner of a Krylov bottom solver) and getFlux (face fluxes for flux
// registers), src/AMRNonLinearMultiCompOp.cpp:1077-1162, :172-206.
// Each is one or two grid-stride passes over the level; the arithmetic
// follows the oracle (oracle/mlref_hc.c) term by term.
#include <hip/hip_runtime.h>

#include "ml_internal.h"
#include "ml_tile.h"

namespace ml {
namespace {

__device__ inline void cell_of(const Geom& g, long long t, int& i, int& j, int& k) {
  i = int(t % g.n[0]);
  const long long r = t / g.n[0];
  j = int(r % g.n[1]);
  k = int(r / g.n[1]);
}

inline int grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65535 * 4) b = 65535 * 4;
  return b < 1 ? 1 : int(b);
}

// levelJacobi's update: phi += 0.5 * (res * lambda)
// mode 1: preCond's start phi = rhs * lambda (res = rhs)
template <int DIM>
__global__ void hc_diag_kernel(HCArgs<double> A, double* __restrict__ phi, const double* __restrict__ res, int mode) {
  const Geom& g = A.g;
  const double scale = 1.0 / (g.dx * g.dx);
  const long long N = g.nvalid();
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < N; t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const long long q = g.ci(i, j, k);
    CellFaces<double, DIM> F;
    load_faces<double, DIM>(A, i, j, k, F);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const long long qn = q + n * g.ncell;
      const double aq = A.a ? A.a[qn] : 1.0;
      const double lam = lambda_of<double, DIM>(F, n, aq, A.alpha, A.beta, scale);
      if (mode == 0) {
        const double r = res[qn] * lam;
        phi[qn] = phi[qn] + 0.5 * r;
      } else {
        phi[qn] = res[qn] * lam;
      }
    }
  }
}

// getFlux: flux = -b * ((phil_hi - phil_lo) * (beta * ref / dx)), ref = 1,
// phil with its ghosts (the diffused variable of the level)
__global__ void hc_flux_kernel(Geom g, int d, double beta, const double* __restrict__ phil,
                               const double* __restrict__ b, double* __restrict__ flux) {
  const long long N = g.nface[d];
  const double scale = beta * 1 / g.dx;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < N; t += (long long)gridDim.x * blockDim.x) {
    const int i = int(t % g.fn[d][0]);
    const long long r = t / g.fn[d][0];
    const int j = int(r % g.fn[d][1]);
    const int k = int(r / g.fn[d][1]);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const double phihi = phil[g.ci(i, j, k) + n * g.ncell];
      const double philo = phil[g.ci(i - (d == 0), j - (d == 1), k - (d == 2)) + n * g.ncell];
      const double gradphi = (phihi - philo) * scale;
      const long long f = t + n * N;
      flux[f] = -b[f] * gradphi;
    }
  }
}

hipError_t launch_diag(const HCArgs<double>& A, double* phi, const double* res, int mode, hipStream_t s) {
  if (A.g.dim == 2)
    hipLaunchKernelGGL((hc_diag_kernel<2>), dim3(grid_for(A.g.nvalid())), dim3(256), 0, s, A, phi, res, mode);
  else
    hipLaunchKernelGGL((hc_diag_kernel<3>), dim3(grid_for(A.g.nvalid())), dim3(256), 0, s, A, phi, res, mode);
  return hipGetLastError();
}

// the level's scratch array (2 components), allocated at its first use
int level_tmp(ml_hc_op op, int depth, double** out) {
  ml::HCLevel& L = op->lev[depth];
  if (!L.tmp) {
    void* p = nullptr;
    ML_HIP(op->ctx, hipMalloc(&p, sizeof(double) * 2 * L.g.ncell));
    L.tmp = static_cast<double*>(p);
  }
  *out = L.tmp;
  return ML_OK;
}

}  // namespace
}  // namespace ml

extern "C" {

int ml_hc_jacobi(ml_hc_op op, int depth, double* phi, const double* rhs, int iters) {
  if (!op || depth < 0 || depth >= int(op->lev.size()) || !phi || !rhs || iters < 0) return ML_EINVAL;
  double* res = nullptr;
  if (int rc = ml::level_tmp(op, depth, &res)) return rc;
  for (int it = 0; it < iters; ++it) {
    ML_HIP(op->ctx, ml::hc_residual<double>(op->args(depth), phi, rhs, res, nullptr, 0, op->ctx->stream));
    ML_HIP(op->ctx, ml::launch_diag(op->args(depth), phi, res, 0, op->ctx->stream));
  }
  return ML_OK;
}

int ml_hc_precond(ml_hc_op op, int depth, double* phi, const double* rhs) {
  if (!op || depth < 0 || depth >= int(op->lev.size()) || !phi || !rhs) return ML_EINVAL;
  ML_HIP(op->ctx, ml::launch_diag(op->args(depth), phi, rhs, 1, op->ctx->stream));
  return ml_hc_relax(op, depth, phi, rhs, 2);
}

int ml_hc_get_flux(ml_hc_op op, int depth, const double* phi, double* const flux[]) {
  if (!op || depth < 0 || depth >= int(op->lev.size()) || !phi || !flux) return ML_EINVAL;
  for (int d = 0; d < op->dim; ++d)
    if (!flux[d]) return ML_EINVAL;
  double* phil = nullptr;
  if (int rc = ml::level_tmp(op, depth, &phil)) return rc;
  const ml::HCArgs<double> A = op->args(depth);
  ML_HIP(op->ctx, ml::hc_diffused_var<double>(A, phi, phil, 0, op->ctx->stream));
  for (int d = 0; d < op->dim; ++d) {
    hipLaunchKernelGGL(ml::hc_flux_kernel, dim3(ml::grid_for(A.g.nface[d])), dim3(256), 0, op->ctx->stream, A.g, d,
                       A.beta, phil, A.b[d], flux[d]);
    ML_HIP(op->ctx, hipGetLastError());
  }
  return ML_OK;
}

}  // extern "C"
