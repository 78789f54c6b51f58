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
#include "ml_common.h"
#include "ml_kernels.h"
#include "ml_tile.h"

namespace ml {

namespace {

constexpr int kLinThreads = 256;
constexpr int kWGT = 1024;     // single-workgroup kernel threads (one per cell)
constexpr int kWGC = 1024;     // cells of the largest level it holds
constexpr int kWGM = kWGC / kWGT;

// Jacobian of the closure at (H, C), regime decided exactly as the closure
// decides it.  M[0] = dtheta/dH, M[1] = dtheta/dTheta, M[2] = dTheta_l/dH,
// M[3] = dTheta_l/dTheta.
__device__ inline void closure_jacobian(double H, double C, const ClosureP<double>& p, double* M) {
  double hs;
  if (C < -p.CR)
    hs = 0.0;
  else
    hs = p.c * ((-C - p.CR) / p.pc);
  if (hs < 0.0) hs = 0.0;
  hs = hs + p.te * p.c;
  const double chie = (p.CR + C) / (p.CR + p.Te);
  const double he = chie * (p.St + p.te * (1.0 - p.c)) + p.c * p.te;
  const double hl = p.St - C + p.te + p.Te;
  M[0] = M[1] = M[2] = M[3] = 0.0;
  if (H <= hs) {
    M[0] = 1.0 / p.c;
  } else if (H > hs && H <= he) {
    // eutectic plateau: theta and Theta_l constant
  } else if (H > he && H < hl) {
    const double chi = porosity_mush<double>(C, H, p);
    double chiH = 0.0, chiC = 0.0;
    if (p.St != 0.0) {
      const double Cv = (C >= -p.CR) ? C : -p.CR;
      const double a = p.CR * (p.c - 1.0) + p.St * (p.pc - 1.0);
      const double b = p.CR * (1.0 - 2.0 * p.c) + H * (1.0 - p.pc) - Cv * (p.c - 1.0) - p.pc * p.St;
      const double Fchi = 2.0 * a * chi + b;
      chiH = -(chi * (1.0 - p.pc) + p.pc) / Fchi;
      chiC = (C >= -p.CR) ? -(p.c - chi * (p.c - 1.0)) / Fchi : 0.0;
    }
    const double D = chi + p.pc * (1.0 - chi);
    const double Tl = (C + p.CR * (1.0 - chi)) / D;
    const double dTl = -(p.CR + Tl * (1.0 - p.pc)) / D;
    M[2] = dTl * chiH;
    M[3] = 1.0 / D + dTl * chiC;
    M[0] = -M[2];
    M[1] = -M[3];
  } else if (H >= hl) {
    M[0] = 1.0;
    M[3] = 1.0;
  }
}

// --------------------------------------------------------------------------
// FP64 residual of the nonlinear operator, fused with phi += e and the
// Jacobian, over a tile with a one-cell halo.
template <int DIM, int TX, int TY, int TZ>
__global__ __launch_bounds__(kLinThreads) void hc_residual_jacobian_kernel(HCArgs<double> A,
                                                                           const double* __restrict__ phi,
                                                                           double* __restrict__ phi_out,
                                                                           const double* __restrict__ rhs,
                                                                           const float* __restrict__ ecorr,
                                                                           double omega,
                                                                           float* __restrict__ r32,
                                                                           float* __restrict__ M32,
                                                                           unsigned long long* norm, int ntx, int nty) {
  if (A.skip && *A.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2) * (TY + 2) * (DIM == 3 ? TZ + 2 : 1);
  constexpr int KMAX = (EMAX + kLinThreads - 1) / kLinThreads;
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init(1, ntx, nty);
  const int E = t.E;
  double* sphi = reinterpret_cast<double*>(smem_raw);
  double* sphl = sphi + 2 * E;
  const double dxinv = 1.0 / (g.dx * g.dx);
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * kLinThreads, R[kk]);
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
    CellFaces<double, DIM> F;
    load_faces<double, DIM>(A, r.i, r.j, r.k, F);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const long long q = r.q + n * g.ncell;
      double st[7];
      gather_lds<double, DIM>(sphl + n * E, r.e, t.EX, t.EXY, st);
      const double aq = A.a ? A.a[q] : 1.0;
      const double lof = lof_of<double, DIM>(F, n, st, aq, sphi[n * E + r.e], A.alpha, A.beta, dxinv);
      const double res = rhs[q] - lof;
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
  block_max_atomic<double>(mx, nan, norm);
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
__device__ inline void block_inverse(const CellFaces<T, DIM>& F, const T* aq, const T* Mc, T alpha, T beta, T dxinv,
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
__device__ inline void block_step(const CellFaces<T, DIM>& F, const T* aq, const T* rr, const T* inv, T alpha, T beta,
                                  T dxinv, const T* u0s, const T* u1s, T& e0, T& e1) {
  const T r0 = rr[0] - lof_of<T, DIM>(F, 0, u0s, aq[0], e0, alpha, beta, dxinv);
  const T r1 = rr[1] - lof_of<T, DIM>(F, 1, u1s, aq[1], e1, alpha, beta, dxinv);
  e0 = e0 + (inv[0] * r0 + inv[1] * r1);
  e1 = e1 + (inv[2] * r0 + inv[3] * r1);
}

// Red/black block relaxation of J e = r over a tile with halo npass.
template <int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__global__ __launch_bounds__(NT) void lin_relax_tile_kernel(LinLevel<float> L, const float* __restrict__ ein,
                                                            float* __restrict__ eout, BCDesc bc, float alpha,
                                                            float beta, int npass, int zero_init, Geom gc,
                                                            const float* __restrict__ pc, int ntx, int nty, Geom gr,
                                                            float* __restrict__ rc, float* __restrict__ Mc_out) {
  if (L.skip && *L.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2 * HMAX) * (TY + 2 * HMAX) * (DIM == 3 ? TZ + 2 * HMAX : 1);
  constexpr int KMAX = (EMAX + NT - 1) / NT;
  const Geom& g = L.g;
  Tile<DIM, TX, TY, TZ> t;
  // with a restriction to produce, one more halo cell keeps the core's
  // neighbours exact after the last pass
  t.init(npass + (rc ? 1 : 0), ntx, nty);
  const int E = t.E;
  float* se = reinterpret_cast<float*>(smem_raw);
  float* su = se + 2 * E;
  float* sM = su + 2 * E;
  const float dxinv = float(1.0 / (g.dx * g.dx));
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * NT, R[kk]);
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    float e0 = zero_init ? 0.0f : ein[r.q], e1 = zero_init ? 0.0f : ein[r.q + g.ncell];
    if (pc) {
      e0 = prolong_value<float, DIM>(gc, r.i, r.j, r.k, pc, nullptr, e0);
      e1 = prolong_value<float, DIM>(gc, r.i, r.j, r.k, pc + gc.ncell, nullptr, e1);
    }
    se[r.e] = e0;
    se[E + r.e] = e1;
#pragma unroll
    for (int m = 0; m < 4; ++m) sM[m * E + r.e] = L.M[r.q + m * g.ncell];
  }
  // operator data of every position this thread may update, in registers
  CellFaces<float, DIM> F[KMAX];
  float rr[KMAX][2], aq[KMAX][2], inv[KMAX][4];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1 || r.dist > t.H - 1) continue;
    load_faces<float, DIM>(g, L.b, r.i, r.j, r.k, F[kk]);
    rr[kk][0] = L.r[r.q];
    rr[kk][1] = L.r[r.q + g.ncell];
    aq[kk][0] = L.a ? L.a[r.q] : 1.0f;
    aq[kk][1] = L.a ? L.a[r.q + g.ncell] : 1.0f;
    const float Mc[4] = {L.M[r.q], L.M[r.q + g.ncell], L.M[r.q + 2 * g.ncell], L.M[r.q + 3 * g.ncell]};
    block_inverse<float, DIM>(F[kk], aq[kk], Mc, alpha, beta, dxinv, inv[kk]);
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
template <typename T, int DIM>
struct WGLin {
  Geom g;
  int N, nx, nxy;
  int c[kWGM], i[kWGM], j[kWGM], k[kWGM];
  long long q[kWGM];
  bool own[kWGM];
  __device__ inline void init(const Geom& gg) {
    g = gg;
    nx = g.n[0];
    nxy = g.n[0] * g.n[1];
    N = nxy * (DIM == 3 ? g.n[2] : 1);
#pragma unroll
    for (int m = 0; m < kWGM; ++m) {
      c[m] = threadIdx.x + m * kWGT;
      own[m] = c[m] < N;
      i[m] = c[m] % nx;
      const int r = c[m] / nx;
      j[m] = r % g.n[1];
      k[m] = DIM == 3 ? r / g.n[1] : 0;
      q[m] = own[m] ? g.ci(i[m], j[m], k[m]) : 0;
    }
  }
};

// gather the (M e) stencil of cell (i,j,k) from LDS with wrap / wall rule
template <typename T, int DIM>
__device__ inline void wg_ustencil(const Geom& g, const BCDesc& bc, const T* su, int c, int i, int j, int k, int n,
                                   int nx, int nxy, T* p) {
  const int iv[3] = {i, j, k};
  p[0] = su[c];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int nd = g.n[d];
    const int st = d == 0 ? 1 : (d == 1 ? nx : nxy);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int v = iv[d] + (s ? 1 : -1);
      T val;
      if (v >= 0 && v < nd)
        val = su[c + (s ? st : -st)];
      else if (g.periodic[d])
        val = su[c + (s ? -(nd - 1) * st : (nd - 1) * st)];
      else
        val = su[c] * T(wall_sign(bc, d, s, n));
      p[1 + 2 * d + (s ? 0 : 1)] = val;
    }
  }
}

// npass red/black block passes on a whole small level held in LDS
// (se: e, su: M e, sM: Jacobian; all of this level).  M e is refreshed at
// the updated cells only, so each pass needs one barrier.
template <typename T, int DIM>
__device__ void wg_lin_relax(const WGLin<T, DIM>& W, const BCDesc& bc, T alpha, T beta, const T* const* b,
                             const T* a, const T* rg, T* se, T* su, const T* sM, int npass) {
  const Geom& g = W.g;
  const T dxinv = T(1.0 / (g.dx * g.dx));
  const int N = W.N;
  CellFaces<T, DIM> F[kWGM];
  T rr[kWGM][2], aq[kWGM][2], inv[kWGM][4];
#pragma unroll
  for (int m = 0; m < kWGM; ++m) {
    if (!W.own[m]) continue;
    const int c = W.c[m];
    load_faces<T, DIM>(g, b, W.i[m], W.j[m], W.k[m], F[m]);
    rr[m][0] = rg[c];
    rr[m][1] = rg[N + c];
    aq[m][0] = a ? a[W.q[m]] : T(1);
    aq[m][1] = a ? a[W.q[m] + g.ncell] : T(1);
    const T Mc[4] = {sM[c], sM[N + c], sM[2 * N + c], sM[3 * N + c]};
    block_inverse<T, DIM>(F[m], aq[m], Mc, alpha, beta, dxinv, inv[m]);
    su[c] = Mc[0] * se[c] + Mc[1] * se[N + c];
    su[N + c] = Mc[2] * se[c] + Mc[3] * se[N + c];
  }
  __syncthreads();
  for (int p = 0; p < npass; ++p) {
#pragma unroll
    for (int m = 0; m < kWGM; ++m) {
      if (!W.own[m] || ((W.i[m] + W.j[m] + W.k[m] + p) & 1)) continue;
      const int c = W.c[m];
      T u0s[7], u1s[7];
      wg_ustencil<T, DIM>(g, bc, su, c, W.i[m], W.j[m], W.k[m], 0, W.nx, W.nxy, u0s);
      wg_ustencil<T, DIM>(g, bc, su + N, c, W.i[m], W.j[m], W.k[m], 1, W.nx, W.nxy, u1s);
      T e0 = se[c], e1 = se[N + c];
      block_step<T, DIM>(F[m], aq[m], rr[m], inv[m], alpha, beta, dxinv, u0s, u1s, e0, e1);
      se[c] = e0;
      se[N + c] = e1;
      su[c] = sM[c] * e0 + sM[N + c] * e1;
      su[N + c] = sM[2 * N + c] * e0 + sM[3 * N + c] * e1;
    }
    __syncthreads();
  }
}

// Direct FP64 solve of J e = r on a whole tiny level (2N <= kDirect
// unknowns): each thread t < 2N builds column t of J by applying the
// operator to the unit vector of unknown t; thread 0 then eliminates with
// partial pivoting.  Returns false (leaving e untouched) on a zero pivot.
constexpr int kDirect = 16;
template <int DIM>
__device__ bool wg_direct_solve(const Geom& g, const BCDesc& bc, double alpha, double beta, const double* const* b,
                                const double* a, const double* sMd, const double* rd, double* ed, double* A,
                                int* flag) {
  const int N = g.n[0] * g.n[1] * (DIM == 3 ? g.n[2] : 1);
  const int U = 2 * N;
  const double dxinv = 1.0 / (g.dx * g.dx);
  const int nx = g.n[0], nxy = g.n[0] * g.n[1];
  const int t = threadIdx.x;
  if (t < U) {
    const int ct = t % N, nt = t / N;
    // u = M e for the unit vector: nonzero only at cell ct
    auto u_at = [&](int cell, int k) -> double {
      return cell == ct ? sMd[(2 * k + nt) * N + ct] : 0.0;
    };
    for (int r = 0; r < N; ++r) {
      const int i = r % nx, j = (r / nx) % g.n[1], k = DIM == 3 ? r / nxy : 0;
      CellFaces<double, DIM> F;
      load_faces<double, DIM>(g, b, i, j, k, F);
      const int iv[3] = {i, j, k};
#pragma unroll
      for (int kc = 0; kc < 2; ++kc) {
        double p[7];
        p[0] = u_at(r, kc);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const int nd = g.n[d];
          const int st = d == 0 ? 1 : (d == 1 ? nx : nxy);
#pragma unroll
          for (int sd = 0; sd < 2; ++sd) {
            const int v = iv[d] + (sd ? 1 : -1);
            double val;
            if (v >= 0 && v < nd)
              val = u_at(r + (sd ? st : -st), kc);
            else if (g.periodic[d])
              val = u_at(r + (sd ? -(nd - 1) * st : (nd - 1) * st), kc);
            else
              val = p[0] * double(wall_sign(bc, d, sd, kc));
            p[1 + 2 * d + (sd ? 0 : 1)] = val;
          }
        }
        const long long q = g.ci(i, j, k) + kc * g.ncell;
        const double aq = a ? a[q] : 1.0;
        const double ev = (r == ct && kc == nt) ? 1.0 : 0.0;
        A[(kc * N + r) * kDirect + t] = lof_of<double, DIM>(F, kc, p, aq, ev, alpha, beta, dxinv);
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    double x[kDirect];
    for (int r = 0; r < U; ++r) x[r] = rd[r];
    int ok = 1;
    for (int c = 0; c < U && ok; ++c) {
      int piv = c;
      double best = fabs(A[c * kDirect + c]);
      for (int r = c + 1; r < U; ++r)
        if (fabs(A[r * kDirect + c]) > best) {
          best = fabs(A[r * kDirect + c]);
          piv = r;
        }
      if (!(best > 0.0)) {
        ok = 0;
        break;
      }
      if (piv != c) {
        for (int k = 0; k < U; ++k) {
          const double tmp = A[c * kDirect + k];
          A[c * kDirect + k] = A[piv * kDirect + k];
          A[piv * kDirect + k] = tmp;
        }
        const double tmp = x[c];
        x[c] = x[piv];
        x[piv] = tmp;
      }
      const double rp = 1.0 / A[c * kDirect + c];
      for (int r = c + 1; r < U; ++r) {
        const double f = A[r * kDirect + c] * rp;
        if (f == 0.0) continue;
        for (int k = c; k < U; ++k) A[r * kDirect + k] -= f * A[c * kDirect + k];
        x[r] -= f * x[c];
      }
    }
    if (ok)
      for (int c = U - 1; c >= 0; --c) {
        double v = x[c];
        for (int k = c + 1; k < U; ++k) v -= A[c * kDirect + k] * x[k];
        x[c] = v / A[c * kDirect + c];
      }
    if (ok)
      for (int r = 0; r < U; ++r) ed[r] = x[r];
    *flag = ok;
  }
  __syncthreads();
  return *flag != 0;
}

// Linear V-cycle for levels l0.. in one workgroup: per level the state
// (e, M, r) is staged in LDS; the coarsest level is solved in FP64 with
// `bottom` block Gauss-Seidel iterations.
template <int DIM>
__global__ __launch_bounds__(kWGT) void lin_coarse_vcycle_kernel(const LinDesc* __restrict__ Dp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const LinDesc& D = *Dp;
  if (D.skip && *D.skip) return;
  float* se = reinterpret_cast<float*>(smem_raw);   // 2*kWGC
  float* su = se + 2 * kWGC;                         // 2*kWGC
  float* sM = su + 2 * kWGC;                         // 4*kWGC
  float* sr = sM + 4 * kWGC;                         // 2*kWGC
  const int l0 = D.l0, L = D.nlev;
  const float alpha = float(D.alpha), beta = float(D.beta);

  auto stage = [&](const LinLevel<float>& lv, const WGLin<float, DIM>& W, bool zero_e) {
#pragma unroll
    for (int m = 0; m < kWGM; ++m) {
      if (!W.own[m]) continue;
      const int c = W.c[m];
      const long long q = W.q[m];
      se[c] = zero_e ? 0.0f : lv.e[q];
      se[W.N + c] = zero_e ? 0.0f : lv.e[q + lv.g.ncell];
      sr[c] = lv.r[q];
      sr[W.N + c] = lv.r[q + lv.g.ncell];
#pragma unroll
      for (int mm = 0; mm < 4; ++mm) sM[mm * W.N + c] = lv.M[q + mm * lv.g.ncell];
    }
    __syncthreads();
  };
  auto store_e = [&](const LinLevel<float>& lv, const WGLin<float, DIM>& W) {
#pragma unroll
    for (int m = 0; m < kWGM; ++m) {
      if (!W.own[m]) continue;
      lv.e[W.q[m]] = se[W.c[m]];
      lv.e[W.q[m] + lv.g.ncell] = se[W.N + W.c[m]];
    }
    __syncthreads();
  };

  for (int l = l0; l < L - 1; ++l) {
    const LinLevel<float>& lv = D.lev[l];
    const LinLevel<float>& cv = D.lev[l + 1];
    WGLin<float, DIM> W;
    W.init(lv.g);
    stage(lv, W, true);
    wg_lin_relax<float, DIM>(W, D.bc, alpha, beta, lv.b, lv.a, sr, se, su, sM, 2 * D.nu1);
    store_e(lv, W);
    // r_c = R(r - J e), M_c = R(M)
#pragma unroll
    for (int m = 0; m < kWGM; ++m) {
      if (!W.own[m]) continue;
      const int c = W.c[m];
      su[c] = sM[c] * se[c] + sM[W.N + c] * se[W.N + c];
      su[W.N + c] = sM[2 * W.N + c] * se[c] + sM[3 * W.N + c] * se[W.N + c];
    }
    __syncthreads();
    {
      const Geom& g = lv.g;
      const Geom& gcc = cv.g;
      const int NC = gcc.n[0] * gcc.n[1] * (DIM == 3 ? gcc.n[2] : 1);
      const float dxinv = float(1.0 / (g.dx * g.dx));
      const float w = DIM == 3 ? 0.125f : 0.25f;
      for (int cc = threadIdx.x; cc < NC; cc += kWGT) {
        const int I = cc % gcc.n[0], rq = cc / gcc.n[0], J = rq % gcc.n[1], K = DIM == 3 ? rq / gcc.n[1] : 0;
        float s0 = 0.f, s1 = 0.f, mm4[4] = {0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
          for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) {
              const int fi = 2 * I + ii, fj = 2 * J + jj, fk = DIM == 3 ? 2 * K + kk : 0;
              const int fc = fi + W.nx * fj + W.nxy * fk;
              CellFaces<float, DIM> F;
              load_faces<float, DIM>(g, lv.b, fi, fj, fk, F);
              const long long q = g.ci(fi, fj, fk);
              float u0s[7], u1s[7];
              wg_ustencil<float, DIM>(g, D.bc, su, fc, fi, fj, fk, 0, W.nx, W.nxy, u0s);
              wg_ustencil<float, DIM>(g, D.bc, su + W.N, fc, fi, fj, fk, 1, W.nx, W.nxy, u1s);
              const float a0 = lv.a ? lv.a[q] : 1.0f, a1 = lv.a ? lv.a[q + g.ncell] : 1.0f;
              s0 += sr[fc] - lof_of<float, DIM>(F, 0, u0s, a0, se[fc], alpha, beta, dxinv);
              s1 += sr[W.N + fc] - lof_of<float, DIM>(F, 1, u1s, a1, se[W.N + fc], alpha, beta, dxinv);
#pragma unroll
              for (int mm = 0; mm < 4; ++mm) mm4[mm] += sM[mm * W.N + fc];
            }
        const long long qc = gcc.ci(I, J, K);
        cv.r[qc] = s0 * w;
        cv.r[qc + gcc.ncell] = s1 * w;
#pragma unroll
        for (int mm = 0; mm < 4; ++mm) cv.M[qc + mm * gcc.ncell] = mm4[mm] * w;
      }
    }
    __syncthreads();
  }

  // bottom: FP64 solve on the coarsest level
  {
    const LinLevel<float>& lv = D.lev[L - 1];
    WGLin<double, DIM> W;
    W.init(lv.g);
    double* de = reinterpret_cast<double*>(smem_raw);   // reuse LDS: 2N + 2N + 4N + 2N doubles
    double* du = de + 2 * W.N;
    double* dM = du + 2 * W.N;
    double* dr = dM + 4 * W.N;
#pragma unroll
    for (int m = 0; m < kWGM; ++m) {
      if (!W.own[m]) continue;
      const int c = W.c[m];
      const long long q = W.q[m];
      de[c] = 0.0;
      de[W.N + c] = 0.0;
      dr[c] = double(lv.r[q]);
      dr[W.N + c] = double(lv.r[q + lv.g.ncell]);
#pragma unroll
      for (int mm = 0; mm < 4; ++mm) dM[mm * W.N + c] = double(lv.M[q + mm * lv.g.ncell]);
    }
    __syncthreads();
    // a tiny coarsest level is solved exactly; otherwise (or on a zero
    // pivot) `bottom` block Gauss-Seidel iterations
    __shared__ int direct_ok;
    bool solved = false;
    if (2 * W.N <= kDirect)
      solved = wg_direct_solve<DIM>(lv.g, D.bc, D.alpha, D.beta, D.b64, D.a64, dM, dr, de, dr + 2 * W.N,
                                    &direct_ok);
    if (!solved) wg_lin_relax<double, DIM>(W, D.bc, D.alpha, D.beta, D.b64, D.a64, dr, de, du, dM, 2 * D.bottom);
#pragma unroll
    for (int m = 0; m < kWGM; ++m) {
      if (!W.own[m]) continue;
      lv.e[W.q[m]] = float(de[W.c[m]]);
      lv.e[W.q[m] + lv.g.ncell] = float(de[W.N + W.c[m]]);
    }
    __syncthreads();
  }

  for (int l = L - 2; l >= l0; --l) {
    const LinLevel<float>& lv = D.lev[l];
    const LinLevel<float>& cv = D.lev[l + 1];
    WGLin<float, DIM> W;
    W.init(lv.g);
    stage(lv, W, false);
#pragma unroll
    for (int m = 0; m < kWGM; ++m) {
      if (!W.own[m]) continue;
      const int c = W.c[m];
      se[c] = prolong_value<float, DIM>(cv.g, W.i[m], W.j[m], W.k[m], cv.e, nullptr, se[c]);
      se[W.N + c] = prolong_value<float, DIM>(cv.g, W.i[m], W.j[m], W.k[m], cv.e + cv.g.ncell, nullptr, se[W.N + c]);
    }
    __syncthreads();
    wg_lin_relax<float, DIM>(W, D.bc, alpha, beta, lv.b, lv.a, sr, se, su, sM, 2 * D.nu2);
    store_e(lv, W);
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
                                hipStream_t s) {
  if (e && (!phi_out || phi_out == phi)) return hipErrorInvalidValue;
  int ntx, nty;
  if (A.g.dim == 2) {
    using D = LinTile<2>;
    const int nt = lin_tiles<2>(A.g, ntx, nty);
    const size_t lds = 4 * lin_extent<2>(1) * sizeof(double);
    hipLaunchKernelGGL((hc_residual_jacobian_kernel<2, D::TX, D::TY, D::TZ>), dim3(nt), dim3(kLinThreads), lds, s, A,
                       phi, phi_out, rhs, e, omega, r32, M32, norm, ntx, nty);
  } else {
    using D = LinTile<3>;
    const int nt = lin_tiles<3>(A.g, ntx, nty);
    const size_t lds = 4 * lin_extent<3>(1) * sizeof(double);
    hipLaunchKernelGGL((hc_residual_jacobian_kernel<3, D::TX, D::TY, D::TZ>), dim3(nt), dim3(kLinThreads), lds, s, A,
                       phi, phi_out, rhs, e, omega, r32, M32, norm, ntx, nty);
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
      const int nt = lin_tiles<2>(L.g, ntx, nty);
      const int pmax = (nt <= 1024 || (C && npass + hx > 2)) ? 8 : 2;
      const int p = (npass - done) < pmax ? (npass - done) : pmax;
      if (C && p + hx > pmax) return hipErrorInvalidValue;
      const size_t lds = (8 * lin_extent<2>(p + hx) + (C ? 2 * D::TX * D::TY : 0)) * sizeof(float);
      if (pmax == 8)
        hipLaunchKernelGGL((lin_relax_tile_kernel<2, D::TX, D::TY, D::TZ, 8, 512>), dim3(nt), dim3(512), lds, s, L,
                           ein, eout, bc, alpha, beta, p, zi, gcc, p0, ntx, nty, grr, rc, Mc);
      else
        hipLaunchKernelGGL((lin_relax_tile_kernel<2, D::TX, D::TY, D::TZ, 2, 256>), dim3(nt), dim3(256), lds, s, L,
                           ein, eout, bc, alpha, beta, p, zi, gcc, p0, ntx, nty, grr, rc, Mc);
      done += p;
    } else {
      using D = LinTile<3>;
      const int nt = lin_tiles<3>(L.g, ntx, nty);
      const int p = (npass - done) < 2 ? (npass - done) : 2;
      if (C && p + hx > 2) return hipErrorInvalidValue;
      const size_t lds = (8 * lin_extent<3>(p + hx) + (C ? 2 * D::TX * D::TY * D::TZ : 0)) * sizeof(float);
      hipLaunchKernelGGL((lin_relax_tile_kernel<3, D::TX, D::TY, D::TZ, 2, 512>), dim3(nt), dim3(512), lds, s, L, ein,
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

hipError_t lin_coarse_vcycle(const LinDesc* dev_desc, int dim, hipStream_t s) {
  const size_t lds = 10 * size_t(kWGC) * sizeof(float);
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

}  // namespace ml
