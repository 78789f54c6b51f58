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

template <int DIM, class IX>
__device__ __forceinline__ void adv_predict_body(const AdvArgs& A, IX ix) {
  const Geom& g = A.g;
  for (long long t = ix.start(); t < A.nv; t += ix.stride()) {
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

template <int DIM>
__global__ void adv_predict_kernel(AdvArgs A) {
  adv_predict_body<DIM>(A, GridIx{});
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

template <int DIM, int STAGE, class IX>
__device__ __forceinline__ void adv_face_body(const AdvArgs& A, const FaceGroups& G, IX ix) {
  const Geom& g = A.g;
  const long long total = G.first[G.n];
  for (long long t = ix.start(); t < total; t += ix.stride()) {
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

template <int DIM, int STAGE>
__global__ void adv_face_kernel(AdvArgs A, FaceGroups G) {
  adv_face_body<DIM, STAGE>(A, G, GridIx{});
}

template <int DIM, class IX>
__device__ __forceinline__ void adv_div_body(const AdvArgs& A, double* __restrict__ src, IX ix) {
  const Geom& g = A.g;
  for (long long t = ix.start(); t < A.nv; t += ix.stride()) {
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

template <int DIM>
__global__ void adv_div_kernel(AdvArgs A, double* __restrict__ src) {
  adv_div_body<DIM>(A, src, GridIx{});
}

// CellToEdge of a no-slip cell-centred velocity: interior (and periodic)
// faces 0.5*(u_lo + u_hi), wall faces 0 (the ghost 2*0 - u1 averages to 0)
template <int DIM, class IX>
__device__ __forceinline__ void face_velocity_body(const Geom& g, int d, const double* __restrict__ u,
                                                   double* __restrict__ uf, IX ix) {
  for (long long f = ix.start(); f < g.nface[d]; f += ix.stride()) {
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

template <int DIM>
__global__ void face_velocity_kernel(Geom g, int d, const double* __restrict__ u, double* __restrict__ uf) {
  face_velocity_body<DIM>(g, d, u, uf, GridIx{});
}

// ---- batched form: one launch per stage for many operators --------------
// job of one operator: its arguments, face groups, the cells' and faces'
// workgroup counts (nb[0] cells, nb[1] faces of every direction, nb[2] the
// 3D corrected faces) and its face velocities
struct AdvJob {
  AdvArgs A;
  FaceGroups G1, G2;
  double* src;
  long long nb[3];
  const double* u[3];
  long long nbf[3];   // face-velocity workgroups per direction
};

// STAGE 0 cells (predictor), 1 faces (WHalf1), 2 faces (3D WHalf2), 3 faces
// (flux), 4 cells (divergence), 5 + d face velocities of direction d
template <int DIM, int STAGE>
__global__ __launch_bounds__(256) void adv_batch_kernel(const AdvJob* __restrict__ jobs,
                                                        const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const AdvJob& J = jobs[w >> kBatchTileBits];
  const long long b = w & kBatchTileMask;
  if (STAGE == 0) adv_predict_body<DIM>(J.A, JobIx{b, J.nb[0]});
  if (STAGE == 1) adv_face_body<DIM, 1>(J.A, J.G1, JobIx{b, J.nb[1]});
  if (STAGE == 2) adv_face_body<DIM, 2>(J.A, J.G2, JobIx{b, J.nb[2]});
  if (STAGE == 3) adv_face_body<DIM, 3>(J.A, J.G1, JobIx{b, J.nb[1]});
  if (STAGE == 4) adv_div_body<DIM>(J.A, J.src, JobIx{b, J.nb[0]});
  if (STAGE >= 5) face_velocity_body<DIM>(J.A.g, STAGE - 5, J.u[STAGE - 5], const_cast<double*>(J.A.uf[STAGE - 5]),
                                          JobIx{b, J.nbf[STAGE - 5]});
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

// checks and arguments of one operator's advective source (the stages'
// workspace kept by the operator)
int adv_prepare(ml_hc_op op, const ml_advect_params* ap, double dt, const double* W, const double* const u_face[],
                const double* S, double* src, double* const flux[], AdvArgs* out) {
  if (!op || !ap || !W || !src || !fields_ok(op->dim, reinterpret_cast<const void* const*>(u_face)) ||
      !(dt > 0))
    return ML_EINVAL;
  const ml::Geom& g = op->lev[0].g;
  for (int d = 0; d < g.dim; ++d) {
    if (g.periodic[d]) continue;
    if (g.n[d] < 3) return fail(op->ctx, ML_ENOTSUP, "ml_hc_advect_src: a walled direction needs at least 3 cells");
    for (int s = 0; s < 2; ++s)
      for (int c = 0; c < 2; ++c) {
        const int t = ap->type[d][s][c];
        if (t != ML_ADV_DIRICHLET && t != ML_ADV_EXTRAP && t != ML_ADV_INFLOW_OUTFLOW && t != ML_ADV_NEUMANN)
          return fail(op->ctx, ML_ENOTSUP, "ml_hc_advect_src: unsupported advection boundary condition");
      }
  }
  const long long nv = g.nvalid();
  long long nface = 0;
  for (int d = 0; d < g.dim; ++d) nface += g.nface[d];
  const bool own_flux = !fields_ok(g.dim, reinterpret_cast<const void* const*>(flux));
  // vc (dim*nv) + pred (dim*2*2*nv) + h1 (2*sum nface) + h2 (3D: 2*2*sum
  // nface) + flux (2*sum nface, when the caller passes none)
  const long long ndbl = g.dim * nv + g.dim * 4 * nv + 2 * nface + (g.dim == 3 ? 4 * nface : 0) +
                         (own_flux ? 2 * nface : 0);
  double* ws = nullptr;
  if (int rc = adv_workspace(op, sizeof(double) * size_t(ndbl), &ws)) return rc;
  AdvArgs A{};
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
  *out = A;
  return ML_OK;
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
  ml::AdvArgs A{};
  if (int rc = ml::adv_prepare(op, ap, dt, W, u_face, S, src, flux, &A)) return rc;
  const hipError_t e = A.g.dim == 2 ? ml::launch_advect<2>(A, src, op->ctx->stream)
                                    : ml::launch_advect<3>(A, src, op->ctx->stream);
  if (e != hipSuccess) return ml::fail(op->ctx, ML_EHIP, std::string("ml_hc_advect_src: ") + hipGetErrorString(e));
  return ML_OK;
}

int ml_hc_db_advect_src_many(ml_ctx ctx, int n, const ml_hc_op ops[], const ml_advect_params* const ap[], double dt,
                             double* const phi[], const double* const W[], const double* const u[],
                             double* const dsrc[], double* const u_face[], double* const src[]) {
  if (!ctx || n < 0 || (n > 0 && (!ops || !ap || !phi || !W || !u || !dsrc || !u_face || !src))) return ML_EINVAL;
  if (n == 0) return ML_OK;
  if (n > (1 << (32 - ml::kBatchTileBits)))
    return ml::fail(ctx, ML_ENOTSUP, "ml_hc_db_advect_src_many: more than 4096 operators");
  const int dim = ops[0] ? ops[0]->dim : 0;
  std::vector<ml::AdvJob> jobs(n);
  std::vector<ml::HCApplyJob<double>> aj(n);
  std::vector<ml::WordsJob> zj(n);
  std::vector<unsigned> amap, zmap, mcell, mface1, mface2, mfv[3];
  auto blocks = [](long long items) { return (items + 255) / 256; };
  for (int i = 0; i < n; ++i) {
    ml_hc_op op = ops[i];
    if (!op || op->ctx != ctx || op->dim != dim)
      return ml::fail(ctx, ML_EINVAL, "ml_hc_db_advect_src_many: operators of one context and dimension");
    for (int j = 0; j < i; ++j)
      if (ops[j] == op) return ml::fail(ctx, ML_EINVAL, "ml_hc_db_advect_src_many: an operator appears twice");
    const double* ufi[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < dim; ++d) {
      if (!u[i * dim + d] || !u_face[i * dim + d]) return ML_EINVAL;
      ufi[d] = u_face[i * dim + d];
    }
    if (!phi[i] || !dsrc[i]) return ML_EINVAL;
    ml::AdvJob& J = jobs[i];
    if (int rc = ml::adv_prepare(op, ap[i], dt, W[i], ufi, dsrc[i], src[i], nullptr, &J.A)) return rc;
    const ml::Geom& g = J.A.g;
    J.src = src[i];
    J.G1 = ml::FaceGroups{};
    J.G1.n = dim;
    for (int d = 0; d < dim; ++d) {
      J.G1.dir[d] = d;
      J.G1.dir2[d] = -1;
      J.G1.first[d + 1] = J.G1.first[d] + g.nface[d];
    }
    J.G2 = ml::FaceGroups{};
    if (dim == 3)
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < 3; ++e) {
          if (e == d) continue;
          J.G2.dir[J.G2.n] = d;
          J.G2.dir2[J.G2.n] = e;
          J.G2.first[J.G2.n + 1] = J.G2.first[J.G2.n] + g.nface[d];
          ++J.G2.n;
        }
    J.nb[0] = blocks(J.A.nv);
    J.nb[1] = blocks(J.G1.first[dim]);
    J.nb[2] = dim == 3 ? blocks(J.G2.first[J.G2.n]) : 0;
    for (int d = 0; d < 3; ++d) {
      J.u[d] = d < dim ? u[i * dim + d] : nullptr;
      J.nbf[d] = d < dim ? blocks(g.nface[d]) : 0;
    }
    const unsigned hi = unsigned(i) << ml::kBatchTileBits;
    for (long long b = 0; b < J.nb[0]; ++b) mcell.push_back(hi | unsigned(b));
    for (long long b = 0; b < J.nb[1]; ++b) mface1.push_back(hi | unsigned(b));
    for (long long b = 0; b < J.nb[2]; ++b) mface2.push_back(hi | unsigned(b));
    for (int d = 0; d < dim; ++d)
      for (long long b = 0; b < J.nbf[d]; ++b) mfv[d].push_back(hi | unsigned(b));
    // the diffusive source (ml_hc_diffusive_src): the operator with alpha 0, beta 1
    ml::HCApplyJob<double>& H = aj[i];
    H.A = op->args(0);
    H.A.alpha = 0.0;
    H.A.beta = 1.0;
    H.io = ml::TileIO<double>{};
    H.io.phi = phi[i];
    H.io.out = dsrc[i];
    H.io.homog = 0;
    int ntx = 0, nty = 0;
    const int nt = ml::hc_apply_tiles(g, &ntx, &nty);
    H.io.ntx = ntx;
    H.io.nty = nty;
    for (int t = 0; t < nt; ++t) amap.push_back(hi | unsigned(t));
    zj[i] = ml::WordsJob{nullptr, reinterpret_cast<unsigned long long*>(src[i]), 2 * g.ncell};
    for (long long c = 0; c * ml::kWordsChunk < 2 * g.ncell; ++c) zmap.push_back(hi | unsigned(c));
  }
  // one staged table: jobs, then the maps of every launch
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  std::vector<std::pair<const void*, size_t>> parts = {
      {jobs.data(), jobs.size() * sizeof(ml::AdvJob)}, {aj.data(), aj.size() * sizeof(ml::HCApplyJob<double>)},
      {zj.data(), zj.size() * sizeof(ml::WordsJob)},   {amap.data(), amap.size() * sizeof(unsigned)},
      {zmap.data(), zmap.size() * sizeof(unsigned)},   {mcell.data(), mcell.size() * sizeof(unsigned)},
      {mface1.data(), mface1.size() * sizeof(unsigned)}, {mface2.data(), mface2.size() * sizeof(unsigned)},
      {mfv[0].data(), mfv[0].size() * sizeof(unsigned)}, {mfv[1].data(), mfv[1].size() * sizeof(unsigned)},
      {mfv[2].data(), mfv[2].size() * sizeof(unsigned)}};
  std::vector<size_t> off(parts.size() + 1, 0);
  for (size_t k = 0; k < parts.size(); ++k) off[k + 1] = al(off[k] + parts[k].second);
  std::vector<unsigned char> blob(off.back());
  for (size_t k = 0; k < parts.size(); ++k)
    if (parts[k].second) std::memcpy(blob.data() + off[k], parts[k].first, parts[k].second);
  void* dv = nullptr;
  ML_HIP(ctx, ctx->tables.put(blob.data(), blob.size(), ctx->stream, &dv));
  unsigned char* d = static_cast<unsigned char*>(dv);
  const ml::AdvJob* dj = reinterpret_cast<const ml::AdvJob*>(d + off[0]);
  auto map_at = [&](int k) { return reinterpret_cast<const unsigned*>(d + off[k]); };
  hipStream_t s = ctx->stream;
  ML_HIP(ctx, ml::hc_apply_batch<double>(dim, reinterpret_cast<const ml::HCApplyJob<double>*>(d + off[1]), map_at(3),
                                         int(amap.size()), s));
  ML_HIP(ctx, ml::words_batch(reinterpret_cast<const ml::WordsJob*>(d + off[2]), map_at(4), int(zmap.size()), s));
  auto stage = [&](auto kern, int k, size_t nblk) -> int {
    if (nblk == 0) return ML_OK;
    hipLaunchKernelGGL(kern, dim3(unsigned(nblk)), dim3(256), 0, s, dj, map_at(k));
    ML_HIP(ctx, hipGetLastError());
    return ML_OK;
  };
  int rc;
  if (dim == 2) {
    if ((rc = stage(ml::adv_batch_kernel<2, 5>, 8, mfv[0].size())) || (rc = stage(ml::adv_batch_kernel<2, 6>, 9, mfv[1].size())) ||
        (rc = stage(ml::adv_batch_kernel<2, 0>, 5, mcell.size())) || (rc = stage(ml::adv_batch_kernel<2, 1>, 6, mface1.size())) ||
        (rc = stage(ml::adv_batch_kernel<2, 3>, 6, mface1.size())) || (rc = stage(ml::adv_batch_kernel<2, 4>, 5, mcell.size())))
      return rc;
  } else {
    if ((rc = stage(ml::adv_batch_kernel<3, 5>, 8, mfv[0].size())) || (rc = stage(ml::adv_batch_kernel<3, 6>, 9, mfv[1].size())) ||
        (rc = stage(ml::adv_batch_kernel<3, 7>, 10, mfv[2].size())) || (rc = stage(ml::adv_batch_kernel<3, 0>, 5, mcell.size())) ||
        (rc = stage(ml::adv_batch_kernel<3, 1>, 6, mface1.size())) || (rc = stage(ml::adv_batch_kernel<3, 2>, 7, mface2.size())) ||
        (rc = stage(ml::adv_batch_kernel<3, 3>, 6, mface1.size())) || (rc = stage(ml::adv_batch_kernel<3, 4>, 5, mcell.size())))
      return rc;
  }
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
