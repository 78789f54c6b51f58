// ml_advect.hip — the advective source of the (H, Theta) equations, the step
// before the implicit solve (multiCompAdvectDiffuse ->
// computeScalarAdvectiveSrcHC, srcSubcycle/AMRLevelMushyLayer.cpp:1645-1672,
// 2074-2104), and the diffusive source the predictor carries
// (computeScalDiffusion, :1280-1340).
//
// The face values of W = (theta, Theta_l) at t^n + dt/2 come from the
// unsplit PLM Godunov predictor the reference runs through Chombo's
// PatchGodunov with AdvectionPhysics (srcSubcycle/AdvectionPhysics.cpp):
// van Leer limited slopes (one-sided next to a wall), the normal predictor
// with the cell velocity as every wave speed, the source S*(dt/2), then the
// corner-transport-upwind corrections (2D: one transverse update with
// -dt/(2dx); 3D: intermediate face values with -dt/(3dx) and the final
// update from them with -dt/(2dx)) and the upwind Riemann problem; faces on
// a wall follow AdvectIBC::primBC (BCutil/AdvectIBC.cpp:130-260).  Fluxes
// are W_half*u_face*advectionCoeff as computeScalarAdvectiveFlux forms them
// (:1165-1198) and src -= div(flux) (util/DivergenceF.ChF:22-47).
//
// One pass per stage, each a grid-stride loop over cells or faces:
//   adv_predict_kernel   cells: cell velocity (EdgeToCell), slopes, W-/W+
//                        in every direction
//   adv_face_kernel<1>   faces of every direction: WHalf1
//   adv_face_kernel<2>   3D only, faces of d1 corrected by d2: WHalf2[d1][d2]
//   adv_face_kernel<3>   faces: final corrected states, Riemann, flux
//   adv_div_kernel       cells: src = src - div(flux)
// Intermediate states live in a workspace the operator keeps (compact
// valid-cell arrays and face arrays).  A face reads the corrected states of
// its two cells directly (each corrected state is formed by the two faces
// that need it rather than stored).  Periodic directions wrap indices.
// Operation order follows the oracle (oracle/mlref_advect.c) term by term,
// so FP64 results are bit-identical.
#include <hip/hip_runtime.h>

#include <string>

#include "ml_internal.h"
#include "ml_kernels.h"
#include "mlhip.h"

namespace ml {
namespace {

constexpr int kExtrap = ML_ADV_EXTRAP, kDirichlet = ML_ADV_DIRICHLET, kNeumann = ML_ADV_NEUMANN;

struct AdvBC {
  int type[3][2][2];
  double val[3][2][2];
};

struct AdvArgs {
  Geom g;
  long long nv;
  double dt, dtbydx, hdt, s2, s3, coeff, one_on_dx;
  AdvBC bc;
  const double* W;     // 2 comps, cell layout
  const double* S;     // 2 comps or null
  const double* uf[3];
  double* vc[3];       // compact valid-cell arrays
  double* pred[3][2];  // [d][0 = W-, 1 = W+], 2 comps (c*nv)
  double* h1[3];       // WHalf1[d], 2 comps, face layout
  double* h2[3][3];    // WHalf2[d1][d2] (3D)
  double* flux[3];     // 2 comps, face layout
};

__device__ inline int wrapc(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ inline long long cidx(const Geom& g, int i, int j, int k) {
  i = wrapc(i, g.n[0]);
  j = wrapc(j, g.n[1]);
  if (g.dim == 3) k = wrapc(k, g.n[2]);
  return (long long)i + (long long)g.n[0] * ((long long)j + (long long)g.n[1] * (long long)k);
}

__device__ inline long long wcell(const Geom& g, int i, int j, int k) {
  i = wrapc(i, g.n[0]);
  j = wrapc(j, g.n[1]);
  if (g.dim == 3) k = wrapc(k, g.n[2]);
  return g.ci(i, j, k);
}

// face of direction d; transverse coordinates wrap, the normal one wraps
// only in a periodic direction
__device__ inline long long fidx(const Geom& g, int d, int i, int j, int k) {
  int c[3] = {i, j, k};
  for (int e = 0; e < g.dim; ++e) {
    if (e == d) {
      if (g.periodic[e]) {
        if (c[e] < 0) c[e] += g.n[e];
        if (c[e] > g.n[e]) c[e] -= g.n[e];
      }
    } else {
      c[e] = wrapc(c[e], g.n[e]);
    }
  }
  return g.fi(d, c[0], c[1], c[2]);
}

__device__ inline double slope(const AdvArgs& A, int d, int c, int i, int j, int k) {
  const Geom& g = A.g;
  const double* W = A.W + (long long)c * g.ncell;
  int p[3] = {i, j, k}, m[3] = {i, j, k};
  p[d] += 1;
  m[d] -= 1;
  const double w0 = W[wcell(g, i, j, k)];
  const int at = d == 0 ? i : d == 1 ? j : k;
  if (!g.periodic[d] && at == 0) return W[wcell(g, p[0], p[1], p[2])] - w0;
  if (!g.periodic[d] && at == g.n[d] - 1) return w0 - W[wcell(g, m[0], m[1], m[2])];
  const double dWr = W[wcell(g, p[0], p[1], p[2])] - w0;
  const double dWl = w0 - W[wcell(g, m[0], m[1], m[2])];
  const double dWc = 0.5 * (dWr + dWl);
  double lim = fmin(2.0 * fabs(dWl), 2.0 * fabs(dWr));
  lim = fmin(lim, fabs(dWc));
  if (dWl * dWr < 0.0) lim = 0.0;
  return copysign(lim, dWc);
}

template <int DIM>
__global__ void adv_predict_kernel(AdvArgs A) {
  const Geom& g = A.g;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < A.nv;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = int(t % g.n[0]);
    const long long r = t / g.n[0];
    const int j = int(r % g.n[1]);
    const int k = int(r / g.n[1]);
    const long long w = g.ci(i, j, k);
    for (int d = 0; d < DIM; ++d) {
      int h[3] = {i, j, k};
      h[d] += 1;
      const double lam = 0.5 * (A.uf[d][g.fi(d, i, j, k)] + A.uf[d][g.fi(d, h[0], h[1], h[2])]);
      A.vc[d][t] = lam;
      const double lmin = fmin(lam, 0.0), lmax = fmax(lam, 0.0);
      for (int c = 0; c < 2; ++c) {
        const double dW = slope(A, d, c, i, j, k);
        double dm, dp;
        if (lam > 0.0) {
          dm = -0.5 * (1.0 + A.dtbydx * lmin) * dW;
          dp = 0.5 * (1.0 - A.dtbydx * lam) * dW;
        } else if (lam < 0.0) {
          dm = -0.5 * (1.0 + A.dtbydx * lam) * dW;
          dp = 0.5 * (1.0 - A.dtbydx * lmax) * dW;
        } else {
          dm = -0.5 * (1.0 + A.dtbydx * lmin) * dW;
          dp = 0.5 * (1.0 - A.dtbydx * lmax) * dW;
        }
        const double wc = A.W[w + (long long)c * g.ncell];
        double wm = dm + wc, wp = dp + wc;
        if (A.S) {
          const double sl = A.S[w + (long long)c * g.ncell] * A.hdt;
          wm = wm + sl;
          wp = wp + sl;
        }
        A.pred[d][0][t + (long long)c * A.nv] = wm;
        A.pred[d][1][t + (long long)c * A.nv] = wp;
      }
    }
  }
}

// quasilinear update (scale*u_cell[e])*(WHalf[e][hi] - WHalf[e][lo]) at a cell
__device__ inline double qlu(const AdvArgs& A, const double* wh, int e, double scale, int c, int i, int j, int k) {
  const Geom& g = A.g;
  int h[3] = {i, j, k};
  h[e] += 1;
  const double diff =
      wh[fidx(g, e, h[0], h[1], h[2]) + (long long)c * g.nface[e]] - wh[fidx(g, e, i, j, k) + (long long)c * g.nface[e]];
  return (scale * A.vc[e][cidx(g, i, j, k)]) * diff;
}

// state side s (0: W-, 1: W+) of a cell for the faces of direction d at
// stage STAGE (1 predicted, 2 corrected by d2 for WHalf2, 3 final)
template <int DIM, int STAGE>
__device__ inline double state(const AdvArgs& A, int d, int d2, int s, int c, int i, int j, int k) {
  double w = A.pred[d][s][cidx(A.g, i, j, k) + (long long)c * A.nv];
  if (STAGE == 1) return w;
  if (STAGE == 2) return w + qlu(A, A.h1[d2], d2, A.s3, c, i, j, k);
  if (DIM == 2) {
    const int e = 1 - d;
    return w + qlu(A, A.h1[e], e, A.s2, c, i, j, k);
  }
  for (int e = 0; e < 3; ++e) {
    if (e == d) continue;
    w = w + qlu(A, A.h2[e][3 - d - e], e, A.s2, c, i, j, k);
  }
  return w;
}

__device__ inline double riemann(double wl, double wr, double u) {
  if (u > 0.0) return wl;
  if (u < 0.0) return wr;
  return 0.5 * (wl + wr);
}

template <int DIM, int STAGE>
__device__ inline double face_interior(const AdvArgs& A, int d, int d2, int c, int i, int j, int k) {
  int l[3] = {i, j, k};
  l[d] -= 1;
  const double wl = state<DIM, STAGE>(A, d, d2, 1, c, l[0], l[1], l[2]);
  const double wr = state<DIM, STAGE>(A, d, d2, 0, c, i, j, k);
  return riemann(wl, wr, A.uf[d][fidx(A.g, d, i, j, k)]);
}

template <int DIM, int STAGE>
__device__ inline double face_value(const AdvArgs& A, int d, int d2, int c, int i, int j, int k) {
  const Geom& g = A.g;
  const int at = d == 0 ? i : d == 1 ? j : k;
  if (g.periodic[d] || (at > 0 && at < g.n[d])) return face_interior<DIM, STAGE>(A, d, d2, c, i, j, k);
  const int side = at == 0 ? 0 : 1;
  const int lohisign = side == 0 ? -1 : 1;
  const int type = A.bc.type[d][side][c];
  const double val = A.bc.val[d][side][c];
  if (type == kDirichlet) return val;
  int cell[3] = {i, j, k};
  if (side == 1) cell[d] -= 1;
  const double wext = state<DIM, STAGE>(A, d, d2, side, c, cell[0], cell[1], cell[2]);
  if (type == kExtrap) return wext;
  if (type == kNeumann) {
    int in[3] = {i, j, k};
    in[d] -= lohisign;
    const double phi1 = face_interior<DIM, STAGE>(A, d, d2, c, in[0], in[1], in[2]);
    return phi1 + ((double)lohisign * g.dx) * val;
  }
  const double inOut = A.uf[d][fidx(g, d, i, j, k)] * (double)lohisign;   // inflow/outflow
  return inOut > 0.0 ? wext : val;
}

// the face groups of one stage: group q covers the faces of direction
// dir[q] (stage 2: corrected by dir2[q]); first[q] its first item
struct FaceGroups {
  int n;
  int dir[6], dir2[6];
  long long first[7];
};

template <int DIM, int STAGE>
__global__ void adv_face_kernel(AdvArgs A, FaceGroups G) {
  const Geom& g = A.g;
  const long long total = G.first[G.n];
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    int q = 0;
    while (q + 1 < G.n && t >= G.first[q + 1]) ++q;
    const int d = G.dir[q], d2 = G.dir2[q];
    const long long f = t - G.first[q];
    const int i = int(f % g.fn[d][0]);
    const long long r = f / g.fn[d][0];
    const int j = int(r % g.fn[d][1]);
    const int k = int(r / g.fn[d][1]);
    const long long nf = g.nface[d];
    const double v0 = face_value<DIM, STAGE>(A, d, d2, 0, i, j, k);
    const double v1 = face_value<DIM, STAGE>(A, d, d2, 1, i, j, k);
    if (STAGE == 1) {
      A.h1[d][f] = v0;
      A.h1[d][f + nf] = v1;
    } else if (STAGE == 2) {
      A.h2[d][d2][f] = v0;
      A.h2[d][d2][f + nf] = v1;
    } else {
      // computeScalarAdvectiveFlux: the coefficient multiplies every
      // component once per component of the loop
      const double u = A.uf[d][f];
      A.flux[d][f] = ((v0 * u) * A.coeff) * A.coeff;
      A.flux[d][f + nf] = ((v1 * A.coeff) * u) * A.coeff;
    }
  }
}

template <int DIM>
__global__ void adv_div_kernel(AdvArgs A, double* __restrict__ src) {
  const Geom& g = A.g;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < A.nv;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = int(t % g.n[0]);
    const long long r = t / g.n[0];
    const int j = int(r % g.n[1]);
    const int k = int(r / g.n[1]);
    const long long q = g.ci(i, j, k);
    for (int c = 0; c < 2; ++c) {
      double div = 0.0;
      for (int d = 0; d < DIM; ++d) {
        int h[3] = {i, j, k};
        h[d] += 1;
        const long long nf = g.nface[d];
        div = div + A.one_on_dx * (A.flux[d][g.fi(d, h[0], h[1], h[2]) + c * nf] - A.flux[d][g.fi(d, i, j, k) + c * nf]);
      }
      src[q + c * g.ncell] = src[q + c * g.ncell] + (-1.0 * div);
    }
  }
}

// CellToEdge of a no-slip cell-centred velocity: interior (and periodic)
// faces 0.5*(u_lo + u_hi), wall faces 0 (the ghost 2*0 - u1 averages to 0)
template <int DIM>
__global__ void face_velocity_kernel(Geom g, int d, const double* __restrict__ u, double* __restrict__ uf) {
  for (long long f = blockIdx.x * (long long)blockDim.x + threadIdx.x; f < g.nface[d];
       f += (long long)gridDim.x * blockDim.x) {
    int c[3];
    c[0] = int(f % g.fn[d][0]);
    const long long r = f / g.fn[d][0];
    c[1] = int(r % g.fn[d][1]);
    c[2] = int(r / g.fn[d][1]);
    const int at = c[d];
    int lo[3] = {c[0], c[1], c[2]};
    lo[d] -= 1;
    if (g.periodic[d]) {
      lo[d] = wrapc(lo[d], g.n[d]);
      c[d] = wrapc(c[d], g.n[d]);
      uf[f] = 0.5 * (u[g.ci(lo[0], lo[1], lo[2])] + u[g.ci(c[0], c[1], c[2])]);
    } else if (at == 0 || at == g.n[d]) {
      uf[f] = 0.0;
    } else {
      uf[f] = 0.5 * (u[g.ci(lo[0], lo[1], lo[2])] + u[g.ci(c[0], c[1], c[2])]);
    }
  }
}

inline int grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65535 * 4) b = 65535 * 4;
  if (b < 1) b = 1;
  return int(b);
}

// workspace of the advective source, kept by the operator
int adv_workspace(ml_hc_op op, size_t need, double** out) {
  if (op->adv_bytes < need) {
    if (op->adv) hipFree(op->adv);
    op->adv = nullptr;
    op->adv_bytes = 0;
    void* p = nullptr;
    ML_HIP(op->ctx, hipMalloc(&p, need));
    op->adv = static_cast<double*>(p);
    op->adv_bytes = need;
  }
  *out = op->adv;
  return ML_OK;
}

template <int DIM>
hipError_t launch_advect(AdvArgs A, double* src, hipStream_t s) {
  const Geom& g = A.g;
  hipLaunchKernelGGL((adv_predict_kernel<DIM>), dim3(grid_for(A.nv)), dim3(256), 0, s, A);
  FaceGroups G{};
  G.n = DIM;
  G.first[0] = 0;
  for (int d = 0; d < DIM; ++d) {
    G.dir[d] = d;
    G.dir2[d] = -1;
    G.first[d + 1] = G.first[d] + g.nface[d];
  }
  hipLaunchKernelGGL((adv_face_kernel<DIM, 1>), dim3(grid_for(G.first[DIM])), dim3(256), 0, s, A, G);
  if (DIM == 3) {
    FaceGroups H{};
    H.first[0] = 0;
    for (int d = 0; d < 3; ++d)
      for (int e = 0; e < 3; ++e) {
        if (e == d) continue;
        H.dir[H.n] = d;
        H.dir2[H.n] = e;
        H.first[H.n + 1] = H.first[H.n] + g.nface[d];
        ++H.n;
      }
    hipLaunchKernelGGL((adv_face_kernel<DIM, 2>), dim3(grid_for(H.first[H.n])), dim3(256), 0, s, A, H);
  }
  hipLaunchKernelGGL((adv_face_kernel<DIM, 3>), dim3(grid_for(G.first[DIM])), dim3(256), 0, s, A, G);
  hipLaunchKernelGGL((adv_div_kernel<DIM>), dim3(grid_for(A.nv)), dim3(256), 0, s, A, src);
  return hipGetLastError();
}

bool fields_ok(int dim, const void* const* f) {
  if (!f) return false;
  for (int d = 0; d < dim; ++d)
    if (!f[d]) return false;
  return true;
}

}  // namespace
}  // namespace ml

extern "C" {

int ml_hc_diffusive_src(ml_hc_op op, double* phi, double* out) {
  if (!op || !phi || !out) return ML_EINVAL;
  ml::HCArgs<double> A = op->args(0);
  A.alpha = 0.0;
  A.beta = 1.0;
  ML_HIP(op->ctx, ml::hc_apply_op<double>(A, phi, out, nullptr, 0, op->ctx->stream));
  return ML_OK;
}

int ml_hc_advect_src(ml_hc_op op, const ml_advect_params* ap, double dt, const double* W,
                     const double* const u_face[], const double* S, double* src, double* const flux[]) {
  if (!op || !ap || !W || !src || !ml::fields_ok(op->dim, reinterpret_cast<const void* const*>(u_face)) ||
      !(dt > 0))
    return ML_EINVAL;
  const ml::Geom& g = op->lev[0].g;
  for (int d = 0; d < g.dim; ++d) {
    if (g.periodic[d]) continue;
    if (g.n[d] < 3) return ml::fail(op->ctx, ML_ENOTSUP, "ml_hc_advect_src: a walled direction needs at least 3 cells");
    for (int s = 0; s < 2; ++s)
      for (int c = 0; c < 2; ++c) {
        const int t = ap->type[d][s][c];
        if (t != ML_ADV_DIRICHLET && t != ML_ADV_EXTRAP && t != ML_ADV_INFLOW_OUTFLOW && t != ML_ADV_NEUMANN)
          return ml::fail(op->ctx, ML_ENOTSUP, "ml_hc_advect_src: unsupported advection boundary condition");
      }
  }
  const long long nv = g.nvalid();
  long long nface = 0;
  for (int d = 0; d < g.dim; ++d) nface += g.nface[d];
  const bool own_flux = !ml::fields_ok(g.dim, reinterpret_cast<const void* const*>(flux));
  // vc (dim*nv) + pred (dim*2*2*nv) + h1 (2*sum nface) + h2 (3D: 2*2*sum
  // nface) + flux (2*sum nface, when the caller passes none)
  const long long ndbl = g.dim * nv + g.dim * 4 * nv + 2 * nface + (g.dim == 3 ? 4 * nface : 0) +
                         (own_flux ? 2 * nface : 0);
  double* ws = nullptr;
  if (int rc = ml::adv_workspace(op, sizeof(double) * size_t(ndbl), &ws)) return rc;
  ml::AdvArgs A{};
  A.g = g;
  A.nv = nv;
  A.dt = dt;
  A.dtbydx = dt / g.dx;
  A.hdt = 0.5 * dt;
  A.s2 = -(1.0 / 2.0) * dt / g.dx;
  A.s3 = -(1.0 / 3.0) * dt / g.dx;
  A.coeff = ap->advectionCoeff;
  A.one_on_dx = 1.0 / g.dx;
  for (int d = 0; d < 3; ++d)
    for (int s = 0; s < 2; ++s)
      for (int c = 0; c < 2; ++c) {
        A.bc.type[d][s][c] = ap->type[d][s][c];
        A.bc.val[d][s][c] = ap->val[d][s][c];
      }
  A.W = W;
  A.S = S;
  double* p = ws;
  for (int d = 0; d < g.dim; ++d) {
    A.uf[d] = u_face[d];
    A.vc[d] = p;
    p += nv;
  }
  for (int d = 0; d < g.dim; ++d)
    for (int s = 0; s < 2; ++s) {
      A.pred[d][s] = p;
      p += 2 * nv;
    }
  for (int d = 0; d < g.dim; ++d) {
    A.h1[d] = p;
    p += 2 * g.nface[d];
  }
  if (g.dim == 3)
    for (int d = 0; d < 3; ++d)
      for (int e = 0; e < 3; ++e) {
        if (e == d) continue;
        A.h2[d][e] = p;
        p += 2 * g.nface[d];
      }
  for (int d = 0; d < g.dim; ++d) {
    if (own_flux) {
      A.flux[d] = p;
      p += 2 * g.nface[d];
    } else {
      A.flux[d] = flux[d];
    }
  }
  const hipError_t e = g.dim == 2 ? ml::launch_advect<2>(A, src, op->ctx->stream)
                                  : ml::launch_advect<3>(A, src, op->ctx->stream);
  if (e != hipSuccess) return ml::fail(op->ctx, ML_EHIP, std::string("ml_hc_advect_src: ") + hipGetErrorString(e));
  return ML_OK;
}

int ml_vc_face_velocity(ml_vc_op op, const double* const u[], double* const u_face[]) {
  if (!op || !ml::fields_ok(op->dim, reinterpret_cast<const void* const*>(u)) ||
      !ml::fields_ok(op->dim, reinterpret_cast<const void* const*>(u_face)))
    return ML_EINVAL;
  const ml::Geom& g = op->lev[0].g;
  for (int d = 0; d < g.dim; ++d) {
    if (g.dim == 2)
      hipLaunchKernelGGL((ml::face_velocity_kernel<2>), dim3(ml::grid_for(g.nface[d])), dim3(256), 0,
                         op->ctx->stream, g, d, u[d], u_face[d]);
    else
      hipLaunchKernelGGL((ml::face_velocity_kernel<3>), dim3(ml::grid_for(g.nface[d])), dim3(256), 0,
                         op->ctx->stream, g, d, u[d], u_face[d]);
  }
  ML_HIP(op->ctx, hipGetLastError());
  return ML_OK;
}

}  // extern "C"
