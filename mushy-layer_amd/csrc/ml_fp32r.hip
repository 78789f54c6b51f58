// ml_fp32r.hip — kernels of the FP32-refined form of the (H, Theta) solve:
// an FP64 defect-correction loop around FP32 multigrid cycles.
//
//   outer (FP64): phi += e;  r = rhs - L(phi) with the full nonlinear
//                 operator (closure regime decided in FP64);  M = the 2x2
//                 Jacobian of the closure (theta, Theta_l)(H, Theta), FP64
//                 evaluated, stored FP32
//   inner (FP32): one V-cycle on J e = r, J e = alpha a e - beta div(b grad(M e)),
//                 red/black smoothing with the 2x2 point-block diagonal,
//                 residual restriction, linear prolongation; the coarsest
//                 level is solved in FP64.
// The nonlinear operator is the reference's (AMRNonLinearMultiCompOp,
// src/AMRNonLinearMultiCompOpF.ChF:125-160); the Jacobian differentiates
// its closure branch by branch (CALCULATE_T_CL / CALCULATEPOROSITYMUSH,
// src/EnthalpyVariablesF.ChF:419-467, :103-150).
#include <hip/hip_ext.h>

#include "ml_common.h"
#include "ml_kernels.h"
#include "ml_tile.h"

namespace ml {

namespace {

constexpr int kLinThreads = 256;
constexpr int kWGT = 512;      // single-workgroup kernel threads
constexpr int kWGC = 1024;     // cells of the largest level it holds
constexpr int kWGM = kWGC / kWGT;

// --------------------------------------------------------------------------
// FP64 residual of the nonlinear operator, fused with phi += e and the
// Jacobian, over a tile with a one-cell halo.
// Tile `blk` of the level.
template <int DIM, int TX, int TY, int TZ>
__device__ __forceinline__ void hc_residual_jacobian_body(const HCArgs<double>& A, const double* __restrict__ phi,
                                                          double* __restrict__ phi_out,
                                                          const double* __restrict__ rhs,
                                                          const float* __restrict__ ecorr, double omega,
                                                          float* __restrict__ r32, float* __restrict__ M32,
                                                          unsigned long long* norm, int ntx, int nty,
                                                          const CtlFuse& fuse, int blk, const int* skip) {
  if (skip && *skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2) * (TY + 2) * (DIM == 3 ? TZ + 2 : 1);
  constexpr int KMAX = (EMAX + kLinThreads - 1) / kLinThreads;
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init_at(1, ntx, nty, blk);
  const int E = t.E;
  double* sphi = reinterpret_cast<double*>(smem_raw);
  double* sphl = sphi + 2 * E;
  const double dxinv = 1.0 / (g.dx * g.dx);
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * kLinThreads, R[kk]);
  // 2D: the owners' face coefficients and rhs are loaded before anything
  // waits (3D keeps them for later: too many registers)
  constexpr bool kPre = DIM == 2;
  CellFaces<double, DIM> F[KMAX];
  double rq[KMAX][2];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (!kPre || !r.owner) continue;
    load_faces<double, DIM>(A, r.i, r.j, r.k, F[kk]);
    rq[kk][0] = rhs[r.q];
    rq[kk][1] = rhs[r.q + g.ncell];
  }
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    double h = phi[r.q], c = phi[r.q + g.ncell];
    if (ecorr) {
      h = h + omega * double(ecorr[r.q]);
      c = c + omega * double(ecorr[r.q + g.ncell]);
    }
    sphi[r.e] = h;
    sphi[E + r.e] = c;
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind == 0 || r.dist > 1) continue;
    const int es = r.kind == 1 ? r.e : r.eint;
    double th, cl;
    closure<double>(sphi[es], sphi[E + es], A.cp, th, cl);
    if (r.kind == 2 && A.bc.apply) {
      const int isign = r.gs == 0 ? -1 : 1;
      th = bc_ghost<double>(A.bc.type[r.gd][r.gs][0], A.bc.val[r.gd][r.gs][0], th, isign, g.dx);
      cl = bc_ghost<double>(A.bc.type[r.gd][r.gs][1], A.bc.val[r.gd][r.gs][1], cl, isign, g.dx);
    }
    sphl[r.e] = th;
    sphl[E + r.e] = cl;
  }
  __syncthreads();
  double mx = 0.0;
  int nan = 0;
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (!r.owner) continue;
    if (ecorr) {
      phi_out[r.q] = sphi[r.e];
      phi_out[r.q + g.ncell] = sphi[E + r.e];
    }
    if (!kPre) {
      load_faces<double, DIM>(A, r.i, r.j, r.k, F[kk]);
      rq[kk][0] = rhs[r.q];
      rq[kk][1] = rhs[r.q + g.ncell];
    }
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const long long q = r.q + n * g.ncell;
      double st[7];
      gather_lds<double, DIM>(sphl + n * E, r.e, t.EX, t.EXY, st);
      const double aq = A.a ? A.a[q] : 1.0;
      const double lof = lof_of<double, DIM>(F[kk], n, st, aq, sphi[n * E + r.e], A.alpha, A.beta, dxinv);
      const double res = rq[kk][n] - lof;
      r32[q] = float(res);
      const double ar = fabs(res);
      if (ar != ar) nan = 1;
      else if (ar > mx) mx = ar;
    }
    double M[4];
    closure_jacobian(sphi[r.e], sphi[E + r.e], A.cp, M);
#pragma unroll
    for (int m = 0; m < 4; ++m) M32[r.q + m * g.ncell] = float(M[m]);
  }
  block_max_atomic<double>(mx, nan, norm, fuse);
}

template <int DIM, int TX, int TY, int TZ>
__global__ __launch_bounds__(kLinThreads) void hc_residual_jacobian_kernel(HCArgs<double> A,
                                                                           const double* __restrict__ phi,
                                                                           double* __restrict__ phi_out,
                                                                           const double* __restrict__ rhs,
                                                                           const float* __restrict__ ecorr,
                                                                           double omega,
                                                                           float* __restrict__ r32,
                                                                           float* __restrict__ M32,
                                                                           unsigned long long* norm, int ntx, int nty,
                                                                           CtlFuse fuse) {
  hc_residual_jacobian_body<DIM, TX, TY, TZ>(A, phi, phi_out, rhs, ecorr, omega, r32, M32, norm, ntx, nty, fuse,
                                             blockIdx.x, A.skip);
}

// --------------------------------------------------------------------------
// Linear operator pieces (any precision T).

// u = M e at a staged position
template <typename T>
__device__ inline void mvec(const T* sM, const T* se, int E, int e, T& u0, T& u1) {
  u0 = sM[e] * se[e] + sM[E + e] * se[E + e];
  u1 = sM[2 * E + e] * se[e] + sM[3 * E + e] * se[E + e];
}

// homogeneous wall rule for a correction of the derived variable
__device__ inline float wall_sign(const BCDesc& bc, int d, int s, int n) {
  return (bc.apply && bc.type[d][s][n] == 0) ? -1.0f : 1.0f;
}

// Inverse of the 2x2 diagonal block of J at a cell:
// D = alpha*a + beta*sum(b_lo + b_hi)*dxinv * M (per component row).
template <typename T, int DIM>
__device__ __forceinline__ void block_inverse(const CellFaces<T, DIM>& F, const T* aq, const T* Mc, T alpha, T beta, T dxinv,
                                     T* inv) {
  T sig[2];
#pragma unroll
  for (int n = 0; n < 2; ++n) {
    T sb = T(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) sb = sb + (F.v[n][2 * d + 1] + F.v[n][2 * d]);
    sig[n] = beta * sb * dxinv;
  }
  const T d00 = alpha * aq[0] + sig[0] * Mc[0], d01 = sig[0] * Mc[1];
  const T d10 = sig[1] * Mc[2], d11 = alpha * aq[1] + sig[1] * Mc[3];
  const T rdet = T(1) / (d00 * d11 - d01 * d10);
  inv[0] = d11 * rdet;
  inv[1] = -d01 * rdet;
  inv[2] = -d10 * rdet;
  inv[3] = d00 * rdet;
}

// (J e)_n from the staged (M e) stencils, then one block Gauss-Seidel update.
template <typename T, int DIM>
__device__ __forceinline__ void block_step(const CellFaces<T, DIM>& F, const T* aq, const T* rr, const T* inv, T alpha, T beta,
                                  T dxinv, const T* u0s, const T* u1s, T& e0, T& e1) {
  const T r0 = rr[0] - lof_of<T, DIM>(F, 0, u0s, aq[0], e0, alpha, beta, dxinv);
  const T r1 = rr[1] - lof_of<T, DIM>(F, 1, u1s, aq[1], e1, alpha, beta, dxinv);
  e0 = e0 + (inv[0] * r0 + inv[1] * r1);
  e1 = e1 + (inv[2] * r0 + inv[3] * r1);
}

// Red/black block relaxation of J e = r over tile `blk` of the level with
// halo npass.
template <int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__device__ __forceinline__ void lin_relax_tile_body(const LinLevel<float>& L, const float* __restrict__ ein,
                                                    float* __restrict__ eout, const BCDesc& bc, float alpha,
                                                    float beta, int npass, int zero_init, const Geom& gc,
                                                    const float* __restrict__ pc, int ntx, int nty, const Geom& gr,
                                                    float* __restrict__ rc, float* __restrict__ Mc_out, int blk) {
  if (L.skip && *L.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2 * HMAX) * (TY + 2 * HMAX) * (DIM == 3 ? TZ + 2 * HMAX : 1);
  constexpr int KMAX = (EMAX + NT - 1) / NT;
  const Geom& g = L.g;
  Tile<DIM, TX, TY, TZ> t;
  // with a restriction to produce, one more halo cell keeps the core's
  // neighbours exact after the last pass
  t.init_at(npass + (rc ? 1 : 0), ntx, nty, blk);
  const int E = t.E;
  float* se = reinterpret_cast<float*>(smem_raw);
  float* su = se + 2 * E;
  float* sM = su + 2 * E;
  const float dxinv = float(1.0 / (g.dx * g.dx));
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * NT, R[kk]);
  // every global load of the tile is issued before the first use
  float ev[KMAX][2], Mv[KMAX][4];
  ProlongCorners<float, DIM> P0[KMAX], P1[KMAX];
  CellFaces<float, DIM> F[KMAX];
  float rr[KMAX][2], aq[KMAX][2], inv[KMAX][4];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    ev[kk][0] = zero_init ? 0.0f : ein[r.q];
    ev[kk][1] = zero_init ? 0.0f : ein[r.q + g.ncell];
    if (pc) {
      prolong_fetch<float, DIM>(gc, r.i, r.j, r.k, pc, P0[kk]);
      prolong_fetch<float, DIM>(gc, r.i, r.j, r.k, pc + gc.ncell, P1[kk]);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) Mv[kk][m] = L.M[r.q + m * g.ncell];
  }
  // operator data of every position this thread may update, in registers
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1 || r.dist > t.H - 1) continue;
    load_faces<float, DIM>(g, L.b, r.i, r.j, r.k, F[kk]);
    rr[kk][0] = L.r[r.q];
    rr[kk][1] = L.r[r.q + g.ncell];
    aq[kk][0] = L.a ? L.a[r.q] : 1.0f;
    aq[kk][1] = L.a ? L.a[r.q + g.ncell] : 1.0f;
  }
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    float e0 = ev[kk][0], e1 = ev[kk][1];
    if (pc) {
      e0 = prolong_combine<float, DIM>(P0[kk], e0);
      e1 = prolong_combine<float, DIM>(P1[kk], e1);
    }
    se[r.e] = e0;
    se[E + r.e] = e1;
#pragma unroll
    for (int m = 0; m < 4; ++m) sM[m * E + r.e] = Mv[kk][m];
    if (r.dist > t.H - 1) continue;
    block_inverse<float, DIM>(F[kk], aq[kk], Mv[kk], alpha, beta, dxinv, inv[kk]);
  }
  __syncthreads();
  // u = M e at every staged position (wall ghosts by the homogeneous rule)
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind == 0) continue;
    float u0, u1;
    mvec<float>(sM, se, E, r.kind == 1 ? r.e : r.eint, u0, u1);
    if (r.kind == 2) {
      u0 *= wall_sign(bc, r.gd, r.gs, 0);
      u1 *= wall_sign(bc, r.gd, r.gs, 1);
    }
    su[r.e] = u0;
    su[E + r.e] = u1;
  }
  __syncthreads();
  // each pass updates one colour and refreshes M e at the updated cells and
  // their wall ghosts: one barrier per pass
  for (int p = 0; p < npass; ++p) {
    const int gext = t.H - p;
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      const PosRec& r = R[kk];
      if (!(r.kind == 1 && r.dist <= gext - 1 && !((r.i + r.j + r.k + p) & 1))) continue;
      const int e = r.e;
      const float Mc[4] = {sM[e], sM[E + e], sM[2 * E + e], sM[3 * E + e]};
      float u0s[7], u1s[7];
      gather_lds<float, DIM>(su, e, t.EX, t.EXY, u0s);
      gather_lds<float, DIM>(su + E, e, t.EX, t.EXY, u1s);
      float e0 = se[e], e1 = se[E + e];
      block_step<float, DIM>(F[kk], aq[kk], rr[kk], inv[kk], alpha, beta, dxinv, u0s, u1s, e0, e1);
      se[e] = e0;
      se[E + e] = e1;
      const float u0 = Mc[0] * e0 + Mc[1] * e1, u1 = Mc[2] * e0 + Mc[3] * e1;
      su[e] = u0;
      su[E + e] = u1;
      const int iv[3] = {r.i, r.j, r.k};
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        if (g.periodic[d]) continue;
        const int stride = d == 0 ? 1 : (d == 1 ? t.EX : t.EXY);
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
          if (iv[d] != (sd ? g.n[d] - 1 : 0)) continue;
          const int ge = sd ? e + stride : e - stride;
          su[ge] = u0 * wall_sign(bc, d, sd, 0);
          su[E + ge] = u1 * wall_sign(bc, d, sd, 1);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (!r.owner) continue;
    eout[r.q] = se[r.e];
    eout[r.q + g.ncell] = se[E + r.e];
  }
  if (!rc) return;
  // r_c = R(r - J e), M_c = R(M) of the relaxed correction (the restriction
  // of lin_restrict_kernel on the tile's own cells)
  constexpr int NC = TX * TY * (DIM == 3 ? TZ : 1);
  float* sres = sM + 4 * E;
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (!r.owner) continue;
    int ex, ey, ez;
    t.coords(r.e, ex, ey, ez);
    const int c = (ex - t.H) + TX * ((ey - t.H) + TY * (DIM == 3 ? ez - t.H : 0));
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      float us[7];
      gather_lds<float, DIM>(su + n * E, r.e, t.EX, t.EXY, us);
      const float je = lof_of<float, DIM>(F[kk], n, us, aq[kk][n], se[n * E + r.e], alpha, beta, dxinv);
      sres[n * NC + c] = rr[kk][n] - je;
    }
  }
  __syncthreads();
  constexpr int CX = TX / 2, CY = TY / 2, CZ = DIM == 3 ? TZ / 2 : 1;
  const float w = DIM == 3 ? 0.125f : 0.25f;
  for (int c = threadIdx.x; c < CX * CY * CZ; c += NT) {
    const int cx = c % CX, cr = c / CX, cy = cr % CY, cz = cr / CY;
    const int I = t.x0 / 2 + cx, J = t.y0 / 2 + cy, K = DIM == 3 ? t.z0 / 2 + cz : 0;
    if (I >= gr.n[0] || J >= gr.n[1] || (DIM == 3 && K >= gr.n[2])) continue;
    const long long qc = gr.ci(I, J, K);
    float s0 = 0.f, s1 = 0.f, m[4] = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
      for (int jj = 0; jj < 2; ++jj)
        for (int ii = 0; ii < 2; ++ii) {
          const int f = (2 * cx + ii) + TX * ((2 * cy + jj) + TY * (DIM == 3 ? 2 * cz + kk : 0));
          s0 += sres[f];
          s1 += sres[NC + f];
          const int e = (2 * cx + ii + t.H) + t.EX * (2 * cy + jj + t.H) + t.EXY * (DIM == 3 ? 2 * cz + kk + t.H : 0);
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) m[mm] += sM[mm * E + e];
        }
    rc[qc] = s0 * w;
    rc[qc + gr.ncell] = s1 * w;
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) Mc_out[qc + mm * gr.ncell] = m[mm] * w;
  }
}

template <int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__global__ __launch_bounds__(NT) void lin_relax_tile_kernel(LinLevel<float> L, const float* __restrict__ ein,
                                                            float* __restrict__ eout, BCDesc bc, float alpha,
                                                            float beta, int npass, int zero_init, Geom gc,
                                                            const float* __restrict__ pc, int ntx, int nty, Geom gr,
                                                            float* __restrict__ rc, float* __restrict__ Mc_out) {
  lin_relax_tile_body<DIM, TX, TY, TZ, HMAX, NT>(L, ein, eout, bc, alpha, beta, npass, zero_init, gc, pc, ntx, nty,
                                                 gr, rc, Mc_out, blockIdx.x);
}

// r_c = R(r - J e), M_c = R(M) over a tile with a one-cell halo.
template <int DIM, int TX, int TY, int TZ>
__global__ __launch_bounds__(kLinThreads) void lin_restrict_kernel(LinLevel<float> L, LinLevel<float> C, BCDesc bc,
                                                                   float alpha, float beta, int ntx, int nty) {
  if (L.skip && *L.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2) * (TY + 2) * (DIM == 3 ? TZ + 2 : 1);
  constexpr int KMAX = (EMAX + kLinThreads - 1) / kLinThreads;
  constexpr int NC = TX * TY * (DIM == 3 ? TZ : 1);
  const Geom& g = L.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init(1, ntx, nty);
  const int E = t.E;
  float* se = reinterpret_cast<float*>(smem_raw);
  float* su = se + 2 * E;
  float* sM = su + 2 * E;
  float* sres = sM + 4 * E;
  const float dxinv = float(1.0 / (g.dx * g.dx));
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * kLinThreads, R[kk]);
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    se[r.e] = L.e[r.q];
    se[E + r.e] = L.e[r.q + g.ncell];
#pragma unroll
    for (int m = 0; m < 4; ++m) sM[m * E + r.e] = L.M[r.q + m * g.ncell];
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind == 0 || r.dist > 1) continue;
    float u0, u1;
    mvec<float>(sM, se, E, r.kind == 1 ? r.e : r.eint, u0, u1);
    if (r.kind == 2) {
      u0 *= wall_sign(bc, r.gd, r.gs, 0);
      u1 *= wall_sign(bc, r.gd, r.gs, 1);
    }
    su[r.e] = u0;
    su[E + r.e] = u1;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < NC; c += blockDim.x) {
    const int cx = c % TX, cr = c / TX, cy = cr % TY, cz = cr / TY;
    const int gi = t.x0 + cx, gj = t.y0 + cy, gk = DIM == 3 ? t.z0 + cz : 0;
    if (gi >= g.n[0] || gj >= g.n[1] || (DIM == 3 && gk >= g.n[2])) continue;
    const int e = (cx + 1) + t.EX * (cy + 1) + t.EXY * (DIM == 3 ? cz + 1 : 0);
    const long long q = g.ci(gi, gj, gk);
    CellFaces<float, DIM> F;
    load_faces<float, DIM>(g, L.b, gi, gj, gk, F);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      float us[7];
      gather_lds<float, DIM>(su + n * E, e, t.EX, t.EXY, us);
      const float aq = L.a ? L.a[q + n * g.ncell] : 1.0f;
      const float je = lof_of<float, DIM>(F, n, us, aq, se[n * E + e], alpha, beta, dxinv);
      sres[n * NC + c] = L.r[q + n * g.ncell] - je;
    }
  }
  __syncthreads();
  constexpr int CX = TX / 2, CY = TY / 2, CZ = DIM == 3 ? TZ / 2 : 1;
  constexpr int NCC = CX * CY * CZ;
  const Geom& gcc = C.g;
  const float w = DIM == 3 ? 0.125f : 0.25f;
  for (int c = threadIdx.x; c < NCC; c += blockDim.x) {
    const int cx = c % CX, cr = c / CX, cy = cr % CY, cz = cr / CY;
    const int I = t.x0 / 2 + cx, J = t.y0 / 2 + cy, K = DIM == 3 ? t.z0 / 2 + cz : 0;
    if (I >= gcc.n[0] || J >= gcc.n[1] || (DIM == 3 && K >= gcc.n[2])) continue;
    const long long qc = gcc.ci(I, J, K);
    float s0 = 0.f, s1 = 0.f, m[4] = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
      for (int jj = 0; jj < 2; ++jj)
        for (int ii = 0; ii < 2; ++ii) {
          const int f = (2 * cx + ii) + TX * ((2 * cy + jj) + TY * (DIM == 3 ? 2 * cz + kk : 0));
          s0 += sres[f];
          s1 += sres[NC + f];
          const int e = (2 * cx + ii + 1) + t.EX * (2 * cy + jj + 1) + t.EXY * (DIM == 3 ? 2 * cz + kk + 1 : 0);
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) m[mm] += sM[mm * E + e];
        }
    C.r[qc] = s0 * w;
    C.r[qc + gcc.ncell] = s1 * w;
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) C.M[qc + mm * gcc.ncell] = m[mm] * w;
  }
}

// --------------------------------------------------------------------------
// Single-workgroup linear V-cycle below the tiled levels.
// unknowns of the largest coarsest level solved exactly
constexpr int kDirect = 16;

// Linear V-cycle for levels l0.. in one workgroup.  e, M e, M, r, the faces,
// a and the block inverses of every level stay in LDS for the whole cycle
// (compact, x fastest, levels one after another; the largest level's faces
// and inverses of this thread's cells in registers).  Each thread owns one
// cell of each colour of a level (even row length) or the cells t and
// t + 512; levels whose owners fit one wavefront are relaxed by wave 0
// alone.  The coarsest level is solved in FP64: exactly (Gauss-Jordan with
// partial pivoting by one wavefront) when it has at most kDirect unknowns,
// otherwise by `bottom` block Gauss-Seidel iterations.
constexpr int kLL = 6;   // levels one launch handles: 1024 cells coarsen to one cell in 6 (2D), 4 (3D)
template <int DIM>
struct LinLevels {
  static constexpr int n = DIM == 3 ? 4 : kLL;
};

template <int DIM>
struct LinLds {
  float *se, *su, *sM, *sr;   // level at 2*off (e, M e, r: 2 comps) and 4*off (M)
  float *sF, *sA, *sI;        // faces [2][2*DIM][tot], a [2][tot], inverses [4][tot]
  int tot;
  __device__ float* e(const CLev& L) const { return se + 2 * L.off; }
  __device__ float* u(const CLev& L) const { return su + 2 * L.off; }
  __device__ float* m(const CLev& L) const { return sM + 4 * L.off; }
  __device__ float* r(const CLev& L) const { return sr + 2 * L.off; }
};

// wall sign of component n's M e (bit 6n + 2d + s of bcw: -1)
__device__ __forceinline__ float lc_wsign(int bcw, int d, int s, int n) {
  return ((bcw >> (6 * n + 2 * d + s)) & 1) ? -1.0f : 1.0f;
}

// A thread's cells of one level for the duration of a level visit:
// coordinates, colour, neighbour indices, faces, a, M, r and the block
// inverses in registers.
template <int DIM>
struct LCCtx {
  int pk[2], c[2], col[2], wall[2];
  int nb[2][2 * DIM];
  CellFaces<float, DIM> F[2];
  float aq[2][2], inv[2][4], M[2][4], rr[2][2];
};

// inverses: compute the block inverses (and keep them in LDS for the
// level's next visit) or read them back
template <int DIM>
__device__ __forceinline__ void lcx_build(LCCtx<DIM>& X, const CLev& L, const int* per, const LinLds<DIM>& S,
                                          float alpha, float beta, bool inverses) {
  const float dxinv = float(1.0 / (L.dx * L.dx));
  const int N = L.N;
  const float* sM = S.m(L);
  const float* sr = S.r(L);
  COwn o;
  cown_set<kWGT>(L, o);
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    X.pk[m] = o.p[m];
    if (o.p[m] < 0) continue;
    int i, j, k;
    cunpack(o.p[m], i, j, k);
    const int c = i + L.nx * j + L.nxy * k;
    X.c[m] = c;
    X.col[m] = (i + j + k) & 1;
    clev_nbrs<DIM>(L, per, i, j, k, X.nb[m], X.wall[m]);
    const int q = L.off + c;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
#pragma unroll
      for (int f = 0; f < 2 * DIM; ++f) X.F[m].v[n][f] = S.sF[(n * 2 * DIM + f) * S.tot + q];
      X.aq[m][n] = S.sA[n * S.tot + q];
      X.rr[m][n] = sr[n * N + c];
    }
#pragma unroll
    for (int z = 0; z < 4; ++z) X.M[m][z] = sM[z * N + c];
    if (inverses) {
      block_inverse<float, DIM>(X.F[m], X.aq[m], X.M[m], alpha, beta, dxinv, X.inv[m]);
#pragma unroll
      for (int z = 0; z < 4; ++z) S.sI[z * S.tot + q] = X.inv[m][z];
    } else {
#pragma unroll
      for (int z = 0; z < 4; ++z) X.inv[m][z] = S.sI[z * S.tot + q];
    }
  }
}

// (M e) stencil of component n of cell m from the level's su
template <int DIM>
__device__ __forceinline__ void lcx_ustencil(const LCCtx<DIM>& X, int m, int n, int bcw, const float* su, float* p) {
  p[0] = su[X.c[m]];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) {
    const int d = f >> 1, s = f & 1;
    float v = su[X.nb[m][f]];
    if ((X.wall[m] >> f) & 1) v = v * lc_wsign(bcw, d, s, n);
    p[1 + 2 * d + (1 - s)] = v;
  }
}

// u = M e at the context's cells
template <int DIM>
__device__ __forceinline__ void lcx_mvec(const LCCtx<DIM>& X, const CLev& L, const LinLds<DIM>& S) {
  const int N = L.N;
  const float* se = S.e(L);
  float* su = S.u(L);
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    if (X.pk[m] < 0) continue;
    const int c = X.c[m];
    const float e0 = se[c], e1 = se[N + c];
    su[c] = X.M[m][0] * e0 + X.M[m][1] * e1;
    su[N + c] = X.M[m][2] * e0 + X.M[m][3] * e1;
  }
  __syncthreads();
}

template <int DIM>
__device__ __forceinline__ void lcx_relax(const LCCtx<DIM>& X, const CLev& L, int bcw, const LinLds<DIM>& S,
                                          float alpha, float beta, int npass) {
  const float dxinv = float(1.0 / (L.dx * L.dx));
  const int N = L.N;
  float* se = S.e(L);
  float* su = S.u(L);
  auto pass = [&](int p) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      if (X.pk[m] < 0 || X.col[m] != (p & 1)) continue;
      const int c = X.c[m];
      float u0s[7], u1s[7];
      lcx_ustencil<DIM>(X, m, 0, bcw, su, u0s);
      lcx_ustencil<DIM>(X, m, 1, bcw, su + N, u1s);
      float e0 = se[c], e1 = se[N + c];
      block_step<float, DIM>(X.F[m], X.aq[m], X.rr[m], X.inv[m], alpha, beta, dxinv, u0s, u1s, e0, e1);
      se[c] = e0;
      se[N + c] = e1;
      su[c] = X.M[m][0] * e0 + X.M[m][1] * e1;
      su[N + c] = X.M[m][2] * e0 + X.M[m][3] * e1;
    }
  };
  // levels whose owners fit one wavefront: wave 0 alone with wavefront-scope
  // syncs (small is uniform, so the workgroup barrier is reached by all
  // waves or by none)
  const bool small = cown_active<kWGT>(L) <= 64;
  if (!small || threadIdx.x < 64) {
#pragma unroll 1
    for (int p = 0; p < npass; ++p) {
      pass(p);
      if (small)
        wave_sync();
      else
        __syncthreads();
    }
  }
  if (small) __syncthreads();
}

// r_c = R(r - J e), M_c = R(M); the coarse correction starts at zero.  The
// fine level's M e array holds the residual between the two phases (the
// next visit of the level recomputes it).
template <int DIM>
__device__ __forceinline__ void lcx_restrict(const LCCtx<DIM>& X, const CLev& Lf, const CLev& Lc, int bcw,
                                             const LinLds<DIM>& S, float alpha, float beta) {
  const float dxinv = float(1.0 / (Lf.dx * Lf.dx));
  const int N = Lf.N;
  const float* se = S.e(Lf);
  const float* sM = S.m(Lf);
  float* su = S.u(Lf);
  float res[2][2];
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    if (X.pk[m] < 0) continue;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      float us[7];
      lcx_ustencil<DIM>(X, m, n, bcw, su + n * N, us);
      res[m][n] = X.rr[m][n] - lof_of<float, DIM>(X.F[m], n, us, X.aq[m][n], se[n * N + X.c[m]], alpha, beta, dxinv);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    if (X.pk[m] < 0) continue;
    su[X.c[m]] = res[m][0];
    su[N + X.c[m]] = res[m][1];
  }
  __syncthreads();
  const int NC = Lc.N;
  float* rc = S.r(Lc);
  float* Mc = S.m(Lc);
  float* ec = S.e(Lc);
  const float w = DIM == 3 ? 0.125f : 0.25f;
  for (int cc = threadIdx.x; cc < NC; cc += kWGT) {
    int I, J, K;
    clev_coords(Lc, cc, I, J, K);
    float s0 = 0.f, s1 = 0.f, mm4[4] = {0.f, 0.f, 0.f, 0.f};
    for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
      for (int jj = 0; jj < 2; ++jj)
        for (int ii = 0; ii < 2; ++ii) {
          const int fc = (2 * I + ii) + Lf.nx * (2 * J + jj) + Lf.nxy * (DIM == 3 ? 2 * K + kk : 0);
          s0 += su[fc];
          s1 += su[N + fc];
#pragma unroll
          for (int mm = 0; mm < 4; ++mm) mm4[mm] += sM[mm * N + fc];
        }
    rc[cc] = s0 * w;
    rc[NC + cc] = s1 * w;
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) Mc[mm * NC + cc] = mm4[mm] * w;
    ec[cc] = 0.f;
    ec[NC + cc] = 0.f;
  }
  __syncthreads();
}

// e_f += P(e_c) (linear prolongation from the coarse level in LDS)
template <int DIM>
__device__ __forceinline__ void lcx_prolong(const LCCtx<DIM>& X, const CLev& Lf, const CLev& Lc, const int* per,
                                            const LinLds<DIM>& S) {
  float* ef = S.e(Lf);
  const float* ec = S.e(Lc);
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    if (X.pk[m] < 0) continue;
    int i, j, k;
    cunpack(X.pk[m], i, j, k);
    const int c = X.c[m];
    ef[c] = prolong_value_lds<float, DIM>(Lc.n, per, i, j, k, ec, ef[c]);
    ef[Lf.N + c] = prolong_value_lds<float, DIM>(Lc.n, per, i, j, k, ec + Lc.N, ef[Lf.N + c]);
  }
  __syncthreads();
}

// Exact FP64 solve of J e = r on a tiny level (U = 2N <= kDirect unknowns)
// by wave 0: lane R builds row R of J (unknown R = comp*N + cell) from the
// unit vectors, then Gauss-Jordan elimination with partial pivoting, the
// rows staying in the lanes' registers.  Returns false (e untouched) on a
// zero pivot.  bF/bA: the level's FP64 faces and a in LDS; buf and prow:
// 64 + kDirect + 1 doubles of LDS scratch.
template <int DIM>
__device__ __forceinline__ bool lc_direct(const CLev& L, const int* per, int bcw, double alpha, double beta, const double* bF,
                          const double* bA, const double* dM, const double* dr, double* de, double* buf,
                          double* prow) {
  const int N = L.N, U = 2 * N;
  const int lane = threadIdx.x;
  const double dxinv = 1.0 / (L.dx * L.dx);
  double a[kDirect];
  double b = 0.0;
#pragma unroll
  for (int t = 0; t < kDirect; ++t) a[t] = 0.0;
  if (lane < U) {
    const int rc = lane % N, kc = lane / N;
    int i, j, k;
    clev_coords(L, rc, i, j, k);
    const int iv[3] = {i, j, k};
    // faces of component kc (held as component 0 of F)
    CellFaces<double, DIM> F;
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) {
      F.v[0][f] = bF[(rc * 2 + kc) * 2 * DIM + f];
      F.v[1][f] = 0.0;
    }
    const double aq = bA[rc * 2 + kc];
#pragma unroll
    for (int t = 0; t < kDirect; ++t) {
      if (t >= U) continue;
      const int ct = t % N, nt = t / N;
      // u = M e of the unit vector of unknown t: nonzero only at cell ct
      auto u_at = [&](int cell) -> double { return cell == ct ? dM[(2 * kc + nt) * N + ct] : 0.0; };
      double p[7];
      p[0] = u_at(rc);
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int nd = L.n[d];
        const int st = d == 0 ? 1 : (d == 1 ? L.nx : L.nxy);
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
          const int v = iv[d] + (sd ? 1 : -1);
          double val;
          if (v >= 0 && v < nd)
            val = u_at(rc + (sd ? st : -st));
          else if (per[d])
            val = u_at(rc + (sd ? -(nd - 1) * st : (nd - 1) * st));
          else
            val = p[0] * double(lc_wsign(bcw, d, sd, kc));
          p[1 + 2 * d + (sd ? 0 : 1)] = val;
        }
      }
      const double ev = (rc == ct && kc == nt) ? 1.0 : 0.0;
      a[t] = lof_of<double, DIM>(F, 0, p, aq, ev, alpha, beta, dxinv);
    }
    b = dr[lane];
  }
  bool used = false;
  int mycol = -1;
#pragma unroll
  for (int c = 0; c < kDirect; ++c) {
    if (c >= U) break;
    buf[lane] = (lane < U && !used) ? fabs(a[c]) : -1.0;
    wave_sync();
    int piv = 0;
    double best = -1.0;
    for (int r = 0; r < U; ++r) {
      const double v = buf[r];
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    if (!(best > 0.0)) return false;
    if (lane == piv) {
#pragma unroll
      for (int t = 0; t < kDirect; ++t) prow[t] = a[t];
      prow[kDirect] = b;
      used = true;
      mycol = c;
    }
    wave_sync();
    if (lane < U && lane != piv) {
      const double f = a[c] / prow[c];
#pragma unroll
      for (int t = 0; t < kDirect; ++t) a[t] = a[t] - f * prow[t];
      b = b - f * prow[kDirect];
    }
    wave_sync();
  }
  if (lane < U) {
    double dcol = 1.0;
#pragma unroll
    for (int t = 0; t < kDirect; ++t)
      if (t == mycol) dcol = a[t];
    de[mycol] = b / dcol;
  }
  wave_sync();
  return true;
}

// one cell's operator data (level l0 also: state and Jacobian; the
// coarsest level of an exact solve also: FP64 faces and a) read from global
// memory
template <int DIM>
struct LinPre {
  CellFaces<float, DIM> F;
  float aq[2], r[2], M[4];
  CellFaces<double, DIM> Fd;
  double ad[2];
};

template <int DIM>
__device__ __forceinline__ void linpre_load(LinPre<DIM>& P, const LinLevel<float>& lv, const CLev& L, int c,
                                            bool state, bool fp64, const LinDesc& D) {
  if (c >= L.N) return;
  int i, j, k;
  clev_coords(L, c, i, j, k);
  load_faces<float, DIM>(lv.g, lv.b, i, j, k, P.F);
  const long long g = lv.g.ci(i, j, k);
#pragma unroll
  for (int n = 0; n < 2; ++n) P.aq[n] = lv.a ? lv.a[g + n * lv.g.ncell] : 1.0f;
  if (state) {
#pragma unroll
    for (int n = 0; n < 2; ++n) P.r[n] = lv.r[g + n * lv.g.ncell];
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) P.M[mm] = lv.M[g + mm * lv.g.ncell];
  }
  if (fp64) {
    load_faces<double, DIM>(lv.g, D.b64, i, j, k, P.Fd);
#pragma unroll
    for (int n = 0; n < 2; ++n) P.ad[n] = D.a64 ? D.a64[g + n * lv.g.ncell] : 1.0;
  }
}

template <int DIM>
__device__ __forceinline__ void linpre_store(const LinPre<DIM>& P, const LinLds<DIM>& S, const CLev& L, int c,
                                             bool state, bool fp64, double* bF, double* bA) {
  if (c >= L.N) return;
  const int q = L.off + c;
  const int tot = S.tot;
#pragma unroll
  for (int n = 0; n < 2; ++n) {
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) S.sF[(n * 2 * DIM + f) * tot + q] = P.F.v[n][f];
    S.sA[n * tot + q] = P.aq[n];
  }
  if (state) {
    const int N = L.N;
    S.se[c] = 0.f;
    S.se[N + c] = 0.f;
    S.sr[c] = P.r[0];
    S.sr[N + c] = P.r[1];
#pragma unroll
    for (int mm = 0; mm < 4; ++mm) S.sM[mm * N + c] = P.M[mm];
  }
  if (fp64) {
#pragma unroll
    for (int n = 0; n < 2; ++n) {
#pragma unroll
      for (int f = 0; f < 2 * DIM; ++f) bF[(c * 2 + n) * 2 * DIM + f] = P.Fd.v[n][f];
      bA[c * 2 + n] = P.ad[n];
    }
  }
}

template <int DIM>
__global__ __launch_bounds__(kWGT) void lin_coarse_vcycle_kernel(const LinDesc* __restrict__ Dp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int LL = LinLevels<DIM>::n;
  // one workgroup per descriptor: the instances of the batched form share
  // the launch (a single solve launches one workgroup)
  const LinDesc& D = Dp[blockIdx.x];
  if (D.skip && *D.skip) return;
  const int l0 = D.l0, nl = D.nlev - D.l0;
  const float alpha = float(D.alpha), beta = float(D.beta);
  int per[3];
  for (int d = 0; d < 3; ++d) per[d] = D.lev[l0].g.periodic[d];
  int bcw = 0;
  for (int n = 0; n < 2; ++n)
    for (int d = 0; d < DIM; ++d)
      for (int sd = 0; sd < 2; ++sd)
        bcw |= ((D.bc.apply && D.bc.type[d][sd][n] == 0) ? 1 : 0) << (6 * n + 2 * d + sd);
  // level geometry in LDS (uniform reads) rather than registers
  __shared__ CLev G[LL];
  int tot = 0;
  for (int l = 0; l < nl && l < LL; ++l) tot += D.lev[l0 + l].g.n[0] * D.lev[l0 + l].g.n[1] * D.lev[l0 + l].g.n[2];
  if (threadIdx.x == 0) {
    int off = 0;
    for (int l = 0; l < nl && l < LL; ++l) {
      clev_set(G[l], D.lev[l0 + l].g, off);
      off += G[l].N;
    }
  }
  __syncthreads();
  LinLds<DIM> S;
  S.tot = tot;
  S.se = reinterpret_cast<float*>(smem_raw);
  S.su = S.se + 2 * tot;
  S.sM = S.su + 2 * tot;
  S.sr = S.sM + 4 * tot;
  S.sF = S.sr + 2 * tot;
  S.sA = S.sF + 4 * DIM * tot;
  S.sI = S.sA + 2 * tot;
  // FP64 arrays of the coarsest level after the FP32 ones
  const int lb = nl - 1;
  const CLev& Lb = G[lb];
  const int Nb = Lb.N;
  const bool direct = 2 * Nb <= kDirect;
  // exact solve only: FP64 e, M, r, faces [cell][2][2*DIM], a [cell][2] and
  // 64 + kDirect + 1 scratch
  double* de = reinterpret_cast<double*>(smem_raw + ((size_t(16 + 4 * DIM) * tot * sizeof(float) + 15) & ~size_t(15)));
  double* dM = de + 2 * Nb;
  double* dr = dM + 4 * Nb;
  double* bF = dr + 2 * Nb;
  double* bA = bF + 4 * DIM * Nb;
  double* dbuf = bA + 2 * Nb;
  // operator data of every level (and the coarsest level's FP64 copy for
  // the exact solve), level l0's state: all global loads first (level l0:
  // cells t and t + 512, coarser levels: cell t), then the LDS stores
  LinPre<DIM> P0[2], P1[LL - 1];
#pragma unroll
  for (int m = 0; m < 2; ++m)
    linpre_load<DIM>(P0[m], D.lev[l0], G[0], threadIdx.x + m * kWGT, true, direct && lb == 0, D);
#pragma unroll
  for (int l = 1; l < LL; ++l)
    if (l < nl) linpre_load<DIM>(P1[l - 1], D.lev[l0 + l], G[l], threadIdx.x, false, direct && lb == l, D);
#pragma unroll
  for (int m = 0; m < 2; ++m)
    linpre_store<DIM>(P0[m], S, G[0], threadIdx.x + m * kWGT, true, direct && lb == 0, bF, bA);
#pragma unroll
  for (int l = 1; l < LL; ++l)
    if (l < nl) linpre_store<DIM>(P1[l - 1], S, G[l], threadIdx.x, false, direct && lb == l, bF, bA);
  __syncthreads();
  // The level visits are runtime loops over one copy of their code (level
  // geometry from LDS): unrolled per level, the kernel's code is several
  // times the instruction cache and a launch spends most of its time
  // fetching it.
  // down: pre-smoothing and restriction to the coarsest level
#pragma unroll 1
  for (int l = 0; l + 1 < nl; ++l) {
    LCCtx<DIM> X;
    lcx_build<DIM>(X, G[l], per, S, alpha, beta, true);
    lcx_mvec<DIM>(X, G[l], S);
    lcx_relax<DIM>(X, G[l], bcw, S, alpha, beta, 2 * D.nu1);
    lcx_restrict<DIM>(X, G[l], G[l + 1], bcw, S, alpha, beta);
  }
  // bottom: exact FP64 solve of a tiny coarsest level; otherwise (or on a
  // zero pivot) `bottom` block Gauss-Seidel iterations
  {
    __shared__ int direct_ok;
    bool solved = false;
    if (direct) {
      const float* fr = S.r(Lb);
      const float* fM = S.m(Lb);
      for (int c = threadIdx.x; c < Nb; c += kWGT) {
        de[c] = 0.0;
        de[Nb + c] = 0.0;
        dr[c] = double(fr[c]);
        dr[Nb + c] = double(fr[Nb + c]);
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) dM[mm * Nb + c] = double(fM[mm * Nb + c]);
      }
      __syncthreads();
      if (threadIdx.x < 64) {
        const bool ok = lc_direct<DIM>(Lb, per, bcw, D.alpha, D.beta, bF, bA, dM, dr, de, dbuf, dbuf + 64);
        if (threadIdx.x == 0) direct_ok = ok ? 1 : 0;
      }
      __syncthreads();
      solved = direct_ok != 0;
      if (solved) {
        float* fe = S.e(Lb);
        for (int c = threadIdx.x; c < Nb; c += kWGT) {
          fe[c] = float(de[c]);
          fe[Nb + c] = float(de[Nb + c]);
        }
        __syncthreads();
      }
    }
    if (!solved) {
      LCCtx<DIM> X;
      lcx_build<DIM>(X, Lb, per, S, alpha, beta, true);
      lcx_mvec<DIM>(X, Lb, S);
      lcx_relax<DIM>(X, Lb, bcw, S, alpha, beta, 2 * D.bottom);
    }
  }
  // up: prolongation and post-smoothing
#pragma unroll 1
  for (int l = nl - 2; l >= 0; --l) {
    LCCtx<DIM> X;
    lcx_build<DIM>(X, G[l], per, S, alpha, beta, false);
    lcx_prolong<DIM>(X, G[l], G[l + 1], per, S);
    lcx_mvec<DIM>(X, G[l], S);
    lcx_relax<DIM>(X, G[l], bcw, S, alpha, beta, 2 * D.nu2);
  }
  const LinLevel<float>& lv = D.lev[l0];
  const float* se = S.e(G[0]);
  for (int c = threadIdx.x; c < G[0].N; c += kWGT) {
    int i, j, k;
    clev_coords(G[0], c, i, j, k);
    const long long q = lv.g.ci(i, j, k);
    lv.e[q] = se[c];
    lv.e[q + lv.g.ncell] = se[G[0].N + c];
  }
}

__global__ void cast_many_kernel(CastJobs J) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= J.begin[J.n]) return;
  int j = 0;
  while (t >= J.begin[j + 1]) ++j;
  J.dst[j][t - J.begin[j]] = float(J.src[j][t - J.begin[j]]);
}

__global__ void cast_kernel(long long n, const double* __restrict__ x, float* __restrict__ y) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    y[i] = float(x[i]);
}

template <int DIM>
struct LinTile;
template <>
struct LinTile<2> {
  static constexpr int TX = 32, TY = 8, TZ = 1;
};
template <>
struct LinTile<3> {
  static constexpr int TX = 16, TY = 8, TZ = 4;
};

// tiles of the deep (up to 8 passes) 2D relaxation launch
struct LinRelaxTile2 {
  static constexpr int TX = 16, TY = 16, TZ = 1, NT = 1024;   // measured best of 7 shapes at 256^2
};

template <int TX, int TY>
int tiles2(const Geom& g, int& ntx, int& nty) {
  ntx = (g.n[0] + TX - 1) / TX;
  nty = (g.n[1] + TY - 1) / TY;
  return ntx * nty;
}

template <int DIM>
int lin_tiles(const Geom& g, int& ntx, int& nty) {
  using D = LinTile<DIM>;
  ntx = (g.n[0] + D::TX - 1) / D::TX;
  nty = (g.n[1] + D::TY - 1) / D::TY;
  const int ntz = DIM == 3 ? (g.n[2] + D::TZ - 1) / D::TZ : 1;
  return ntx * nty * ntz;
}

template <int DIM>
size_t lin_extent(int halo) {
  using D = LinTile<DIM>;
  return size_t(D::TX + 2 * halo) * (D::TY + 2 * halo) * (DIM == 3 ? D::TZ + 2 * halo : 1);
}

}  // namespace

hipError_t hc_residual_jacobian(const HCArgs<double>& A, const double* phi, double* phi_out, const double* rhs,
                                const float* e, double omega, float* r32, float* M32, unsigned long long* norm,
                                hipStream_t s, const CtlFuse& fuse) {
  if (e && (!phi_out || phi_out == phi)) return hipErrorInvalidValue;
  int ntx, nty;
  if (A.g.dim == 2) {
    using D = LinTile<2>;
    const int nt = lin_tiles<2>(A.g, ntx, nty);
    const size_t lds = 4 * lin_extent<2>(1) * sizeof(double);
    hipLaunchKernelGGL((hc_residual_jacobian_kernel<2, D::TX, D::TY, D::TZ>), dim3(nt), dim3(kLinThreads), lds, s, A,
                       phi, phi_out, rhs, e, omega, r32, M32, norm, ntx, nty, fuse);
  } else {
    using D = LinTile<3>;
    const int nt = lin_tiles<3>(A.g, ntx, nty);
    const size_t lds = 4 * lin_extent<3>(1) * sizeof(double);
    hipLaunchKernelGGL((hc_residual_jacobian_kernel<3, D::TX, D::TY, D::TZ>), dim3(nt), dim3(kLinThreads), lds, s, A,
                       phi, phi_out, rhs, e, omega, r32, M32, norm, ntx, nty, fuse);
  }
  return hipGetLastError();
}

namespace {

// fused: the restriction into C (rc != nullptr) needs npass + 1 <= the
// kernel's halo bound, so that one launch does all passes
hipError_t lin_relax_impl(const LinLevel<float>& L, float* const ebuf[2], int* cur, const BCDesc& bc, float alpha,
                          float beta, int npass, bool zero_init, const Geom* gc, const float* pc,
                          const LinLevel<float>* C, hipStream_t s) {
  int ntx, nty;
  const Geom gcc = gc ? *gc : L.g;
  const Geom grr = C ? C->g : L.g;
  float* rc = C ? C->r : nullptr;
  float* Mc = C ? C->M : nullptr;
  const int hx = C ? 1 : 0;
  int done = 0;
  // npass == 0 still runs one launch when there is a zero start, a
  // prolongation or a restriction to apply
  bool first = true;
  while (done < npass || (first && npass == 0 && (zero_init || pc || C))) {
    first = false;
    const float* ein = ebuf[*cur];
    float* eout = ebuf[1 - *cur];
    const int zi = (done == 0 && zero_init) ? 1 : 0;
    const float* p0 = done == 0 ? pc : nullptr;
    if (L.g.dim == 2) {
      using D = LinTile<2>;
      using DR = LinRelaxTile2;
      const int nt = lin_tiles<2>(L.g, ntx, nty);
      const int pmax = (nt <= 1024 || (C && npass + hx > 2)) ? 8 : 2;
      const int p = (npass - done) < pmax ? (npass - done) : pmax;
      if (C && p + hx > pmax) return hipErrorInvalidValue;
      if (pmax == 8) {
        const int ntr = tiles2<DR::TX, DR::TY>(L.g, ntx, nty);
        const size_t lds = (8 * size_t(DR::TX + 2 * (p + hx)) * (DR::TY + 2 * (p + hx)) +
                            (C ? 2 * DR::TX * DR::TY : 0)) * sizeof(float);
        hipLaunchKernelGGL((lin_relax_tile_kernel<2, DR::TX, DR::TY, DR::TZ, 8, DR::NT>), dim3(ntr), dim3(DR::NT), lds,
                           s, L, ein, eout, bc, alpha, beta, p, zi, gcc, p0, ntx, nty, grr, rc, Mc);
      } else {
        const size_t lds = (8 * lin_extent<2>(p + hx) + (C ? 2 * D::TX * D::TY : 0)) * sizeof(float);
        hipLaunchKernelGGL((lin_relax_tile_kernel<2, D::TX, D::TY, D::TZ, 2, 256>), dim3(nt), dim3(256), lds, s, L,
                           ein, eout, bc, alpha, beta, p, zi, gcc, p0, ntx, nty, grr, rc, Mc);
      }
      done += p;
    } else {
      using D = LinTile<3>;
      const int nt = lin_tiles<3>(L.g, ntx, nty);
      const int p = (npass - done) < 2 ? (npass - done) : 2;
      if (C && p + hx > 2) return hipErrorInvalidValue;
      const size_t lds = (8 * lin_extent<3>(p + hx) + (C ? 2 * D::TX * D::TY * D::TZ : 0)) * sizeof(float);
      hipLaunchKernelGGL((lin_relax_tile_kernel<3, D::TX, D::TY, D::TZ, 2, 1024>), dim3(nt), dim3(1024), lds, s, L, ein,
                         eout, bc, alpha, beta, p, zi, gcc, p0, ntx, nty, grr, rc, Mc);
      done += p;
    }
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    *cur = 1 - *cur;
  }
  return hipSuccess;
}

}  // namespace

hipError_t lin_relax(const LinLevel<float>& L, float* const ebuf[2], int* cur, const BCDesc& bc, float alpha,
                     float beta, int npass, bool zero_init, const Geom* gc, const float* pc, hipStream_t s) {
  return lin_relax_impl(L, ebuf, cur, bc, alpha, beta, npass, zero_init, gc, pc, nullptr, s);
}

hipError_t lin_relax_restrict(const LinLevel<float>& L, float* const ebuf[2], int* cur, const LinLevel<float>& C,
                              const BCDesc& bc, float alpha, float beta, int npass, bool zero_init, hipStream_t s) {
  if (npass + 1 <= (L.g.dim == 2 ? 8 : 2))
    return lin_relax_impl(L, ebuf, cur, bc, alpha, beta, npass, zero_init, nullptr, nullptr, &C, s);
  hipError_t e = lin_relax_impl(L, ebuf, cur, bc, alpha, beta, npass, zero_init, nullptr, nullptr, nullptr, s);
  if (e != hipSuccess) return e;
  LinLevel<float> Lf = L;
  Lf.e = ebuf[*cur];
  return lin_restrict(Lf, C, bc, alpha, beta, s);
}

hipError_t lin_restrict(const LinLevel<float>& L, const LinLevel<float>& C, const BCDesc& bc, float alpha, float beta,
                        hipStream_t s) {
  int ntx, nty;
  if (L.g.dim == 2) {
    using D = LinTile<2>;
    const int nt = lin_tiles<2>(L.g, ntx, nty);
    const size_t lds = (8 * lin_extent<2>(1) + 2 * D::TX * D::TY) * sizeof(float);
    hipLaunchKernelGGL((lin_restrict_kernel<2, D::TX, D::TY, D::TZ>), dim3(nt), dim3(kLinThreads), lds, s, L, C, bc,
                       alpha, beta, ntx, nty);
  } else {
    using D = LinTile<3>;
    const int nt = lin_tiles<3>(L.g, ntx, nty);
    const size_t lds = (8 * lin_extent<3>(1) + 2 * D::TX * D::TY * D::TZ) * sizeof(float);
    hipLaunchKernelGGL((lin_restrict_kernel<3, D::TX, D::TY, D::TZ>), dim3(nt), dim3(kLinThreads), lds, s, L, C, bc,
                       alpha, beta, ntx, nty);
  }
  return hipGetLastError();
}

size_t lin_coarse_lds_bytes(const Geom* g, int l0, int nlev) {
  long long tot = 0;
  for (int l = l0; l < nlev; ++l) tot += g[l].nvalid();
  const long long nb = g[nlev - 1].nvalid();
  const int dim = g[l0].dim;
  const bool direct = 2 * nb <= kDirect;
  const size_t dbl = direct ? size_t((8 + 4 * dim + 2) * nb + 64 + kDirect + 1) : 0;
  return ((size_t(16 + 4 * dim) * tot * sizeof(float) + 15) & ~size_t(15)) + dbl * sizeof(double);
}

size_t lin_coarse_static_lds() {
  return 1024;   // bound on the kernel's static LDS (level geometry, flags)
}

int lin_coarse_levels_max(int dim) { return dim == 3 ? LinLevels<3>::n : LinLevels<2>::n; }

hipError_t lin_coarse_vcycle(const LinDesc* dev_desc, int dim, size_t lds, hipStream_t s) {
  if (dim == 2)
    hipLaunchKernelGGL((lin_coarse_vcycle_kernel<2>), dim3(1), dim3(kWGT), lds, s, dev_desc);
  else
    hipLaunchKernelGGL((lin_coarse_vcycle_kernel<3>), dim3(1), dim3(kWGT), lds, s, dev_desc);
  return hipGetLastError();
}

hipError_t cast_many(const CastJobs& J, hipStream_t s) {
  const long long n = J.begin[J.n];
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(cast_many_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, J);
  return hipGetLastError();
}

hipError_t cast_f64_f32(long long n, const double* x, float* y, hipStream_t s) {
  long long b = (n + 255) / 256;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(cast_kernel, dim3(int(b < 1 ? 1 : b)), dim3(256), 0, s, n, x, y);
  return hipGetLastError();
}


// ---- batched form -----------------------------------------------------------
namespace {

template <int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__global__ __launch_bounds__(NT) void lin_relax_batch_kernel(const BatchLinJob* __restrict__ jobs,
                                                             const unsigned* __restrict__ map, int npass,
                                                             int zero_init) {
  const unsigned w = map[blockIdx.x];
  const BatchLinJob& J = jobs[w >> kBatchTileBits];
  lin_relax_tile_body<DIM, TX, TY, TZ, HMAX, NT>(J.L, J.ein, J.eout, J.bc, J.alpha, J.beta, npass, zero_init, J.gc,
                                                 J.pc, J.ntx, J.nty, J.gr, J.rc, J.Mc, int(w & kBatchTileMask));
}

// mode 0: iteration ctl->iter + 1 with the stopping rule; 1: initial
// residual of pb[0] into slots[0]; 2: a backtracking trial of step length
// omega for an instance whose step was rejected (its done flag is set; the
// norm goes to slots[ctl->iter + 1], which the caller has zeroed)
template <int DIM, int TX, int TY, int TZ>
__global__ __launch_bounds__(kLinThreads) void hc_residual_jacobian_batch_kernel(const BatchResJob* __restrict__ jobs,
                                                                                 const unsigned* __restrict__ map,
                                                                                 int mode, double omega) {
  const unsigned w = map[blockIdx.x];
  const BatchResJob& J = jobs[w >> kBatchTileBits];
  const int blk = int(w & kBatchTileMask);
  if (mode == 1) {
    hc_residual_jacobian_body<DIM, TX, TY, TZ>(J.A, J.pb[0], nullptr, J.rhs, nullptr, 1.0, J.r32, J.M32, J.slots,
                                               J.ntx, J.nty, CtlFuse{}, blk, J.A.skip);
    return;
  }
  if (mode == 2) {
    const int it = J.ctl->iter + 1;
    const int pc = (it - 1) & 1;
    hc_residual_jacobian_body<DIM, TX, TY, TZ>(J.A, J.pb[pc], J.pb[1 - pc], J.rhs, J.e, omega, J.r32, J.M32,
                                               J.slots + it, J.ntx, J.nty, CtlFuse{}, blk, nullptr);
    return;
  }
  if (*J.A.skip) return;
  // the instance's own iteration: the last workgroup of this launch moves
  // ctl->iter on only after every workgroup has read it
  const int it = J.ctl->iter + 1;
  const int pc = (it - 1) & 1;
  const CtlFuse f{J.ctl, J.slots, J.arrive, 1, J.ntiles};
  hc_residual_jacobian_body<DIM, TX, TY, TZ>(J.A, J.pb[pc], J.pb[1 - pc], J.rhs, J.e, 1.0, J.r32, J.M32,
                                             J.slots + it, J.ntx, J.nty, f, blk, J.A.skip);
}

// slots[i][ctl_i->iter + 1] of the listed instances: zeroed (gather == 0)
// or copied to out[j] (gather == 1)
__global__ void batch_slots_kernel(const BatchResJob* __restrict__ jobs, const int* __restrict__ list, int n,
                                   int gather, unsigned long long* out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const BatchResJob& J = jobs[list[j]];
  unsigned long long* slot = J.slots + J.ctl->iter + 1;
  if (gather)
    out[j] = *slot;
  else
    *slot = 0ull;
}

}  // namespace

int lin_relax_tiles2(const Geom& g, int& ntx, int& nty) { return tiles2<LinRelaxTile2::TX, LinRelaxTile2::TY>(g, ntx, nty); }

// Tile shapes of the batched relaxation (TX, TY, threads); every shape
// stages a halo of up to 8 cells, so a launch takes up to 8 passes.  The
// results do not depend on the shape.
namespace {
struct BatchTile {
  int TX, TY, NT;
};
constexpr BatchTile kBatchTiles[] = {{16, 16, 1024}, {32, 16, 512}, {32, 32, 1024}, {64, 16, 1024},
                                     {32, 8, 256},   {64, 8, 512},  {32, 16, 256},  {16, 32, 512}};
constexpr int kNumBatchTiles = int(sizeof(kBatchTiles) / sizeof(kBatchTiles[0]));

template <int V>
hipError_t launch_batch_variant(const BatchLinJob* jobs, const unsigned* map, int nblocks, int npass, bool zero_init,
                                size_t lds, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  constexpr BatchTile T = kBatchTiles[V];
  hipExtLaunchKernelGGL((lin_relax_batch_kernel<2, T.TX, T.TY, 1, 8, T.NT>), dim3(nblocks), dim3(T.NT), lds, s, e0, e1,
                        0, jobs, map, npass, zero_init ? 1 : 0);
  return hipGetLastError();
}
}  // namespace

int lin_relax_batch_tiles() { return kNumBatchTiles; }

int lin_relax_batch_tile_shape(int v, int* tx, int* ty, int* nt) {
  if (v < 0 || v >= kNumBatchTiles) return -1;
  *tx = kBatchTiles[v].TX;
  *ty = kBatchTiles[v].TY;
  *nt = kBatchTiles[v].NT;
  return 0;
}

int lin_relax_batch_tiles_of(int v, const Geom& g, int& ntx, int& nty) {
  const BatchTile& T = kBatchTiles[v];
  ntx = (g.n[0] + T.TX - 1) / T.TX;
  nty = (g.n[1] + T.TY - 1) / T.TY;
  return ntx * nty;
}
int lin_residual_tiles2(const Geom& g, int& ntx, int& nty) { return lin_tiles<2>(g, ntx, nty); }
long long lin_deep_tiles_max2() { return 1024; }

hipError_t lin_relax_batch(const BatchLinJob* jobs, const unsigned* map, int nblocks, int npass, bool zero_init,
                           bool restrict_, int variant, hipStream_t s, const LaunchTimer* tm, double bytes) {
  const int h = npass + (restrict_ ? 1 : 0);
  if (npass < 1 || h > 8 || variant < 0 || variant >= kNumBatchTiles) return hipErrorInvalidValue;
  if (nblocks == 0) return hipSuccess;
  const BatchTile& T = kBatchTiles[variant];
  const size_t lds =
      (8 * size_t(T.TX + 2 * h) * (T.TY + 2 * h) + (restrict_ ? 2 * T.TX * T.TY : 0)) * sizeof(float);
  const bool timed = tm && tm->start && *tm->n < tm->cap;
  if (tm && tm->start && !timed) ++*tm->missed;
  hipEvent_t e0 = timed ? tm->start[*tm->n] : nullptr, e1 = timed ? tm->stop[*tm->n] : nullptr;
  hipError_t e;
  switch (variant) {
    case 0: e = launch_batch_variant<0>(jobs, map, nblocks, npass, zero_init, lds, s, e0, e1); break;
    case 1: e = launch_batch_variant<1>(jobs, map, nblocks, npass, zero_init, lds, s, e0, e1); break;
    case 2: e = launch_batch_variant<2>(jobs, map, nblocks, npass, zero_init, lds, s, e0, e1); break;
    case 3: e = launch_batch_variant<3>(jobs, map, nblocks, npass, zero_init, lds, s, e0, e1); break;
    case 4: e = launch_batch_variant<4>(jobs, map, nblocks, npass, zero_init, lds, s, e0, e1); break;
    case 5: e = launch_batch_variant<5>(jobs, map, nblocks, npass, zero_init, lds, s, e0, e1); break;
    case 6: e = launch_batch_variant<6>(jobs, map, nblocks, npass, zero_init, lds, s, e0, e1); break;
    default: e = launch_batch_variant<7>(jobs, map, nblocks, npass, zero_init, lds, s, e0, e1); break;
  }
  if (e == hipSuccess && timed) {
    *tm->bytes += bytes;
    ++*tm->n;
  }
  return e;
}

hipError_t hc_residual_jacobian_batch(const BatchResJob* jobs, const unsigned* map, int nblocks, bool initial,
                                      hipStream_t s) {
  using D = LinTile<2>;
  if (nblocks == 0) return hipSuccess;
  const size_t lds = 4 * lin_extent<2>(1) * sizeof(double);
  hipLaunchKernelGGL((hc_residual_jacobian_batch_kernel<2, D::TX, D::TY, D::TZ>), dim3(nblocks), dim3(kLinThreads), lds,
                     s, jobs, map, initial ? 1 : 0, 1.0);
  return hipGetLastError();
}

hipError_t hc_residual_jacobian_batch_trial(const BatchResJob* jobs, const unsigned* map, int nblocks, double omega,
                                            hipStream_t s) {
  using D = LinTile<2>;
  if (nblocks == 0) return hipSuccess;
  const size_t lds = 4 * lin_extent<2>(1) * sizeof(double);
  hipLaunchKernelGGL((hc_residual_jacobian_batch_kernel<2, D::TX, D::TY, D::TZ>), dim3(nblocks), dim3(kLinThreads), lds,
                     s, jobs, map, 2, omega);
  return hipGetLastError();
}

hipError_t batch_trial_slots(const BatchResJob* jobs, const int* list, int n, bool gather, unsigned long long* out,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_slots_kernel, dim3((n + 255) / 256), dim3(256), 0, s, jobs, list, n, gather ? 1 : 0, out);
  return hipGetLastError();
}

hipError_t lin_coarse_vcycle_batch(const LinDesc* descs, int n, size_t lds, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((lin_coarse_vcycle_kernel<2>), dim3(n), dim3(kWGT), lds, s, descs);
  return hipGetLastError();
}

}  // namespace ml
