// ml_vc_kernels.hip — kernels of the linear variable-coefficient operator
//
//   L u = alpha*a*u - beta*c*u - beta*div(b grad u)      (one component)
//
// behind DarcyBrinkmanOp (src/DarcyBrinkmanOp.cpp, src/DarcyBrinkmanOpF.ChF;
// a = 1, b = -Pr, c = (Pr/Da) chi/Pi) and AMRProjectionOp (c absent,
// alpha = 0, beta = 1, b = -S_face).  The arithmetic of each cell follows
// GSRBHELMHOLTZDB / DBCOMPUTERES / DBCOMPUTEOP / RESTRICTRESDB term by term
// (compiled with -ffp-contract=off), so the FP64 results are those of the
// CPU restatement bit for bit.
//
//   vc_relax_tile_kernel   red/black Gauss-Seidel, homogeneous BCs (the
//                          residual-correction smoother), npass colour
//                          passes per launch over a tile grown by npass
//                          cells, input and output arrays distinct;
//                          optional zero start and piecewise-constant
//                          prolongation of a coarse correction on load
//   vc_tile_kernel         residual (+ max-norm, + fused phi += e),
//                          operator application, restricted residual
//   vc_coarse_vcycle_kernel  the V-cycle below the tiled levels in one
//                          workgroup (levels of at most 1024 cells)
#include "ml_common.h"
#include "ml_kernels.h"
#include "ml_tile.h"

namespace ml {

namespace {

constexpr int kVCThreads = 256;   // halo-1 tile kernels
constexpr int kVCWG = 512;        // single-workgroup kernel threads
constexpr int kVCWGCells = 1024;  // largest level it holds
constexpr int kVCM = kVCWGCells / kVCWG;

enum { kVCResid = 0, kVCApply = 1, kVCRestrict = 2 };

// Face coefficients of one component around a cell: v[2d] lo, v[2d+1] hi.
template <typename T, int DIM>
struct Faces1 {
  T v[2 * DIM];
};

template <typename T, int DIM>
__device__ inline void load_faces1(const VCArgs<T>& A, int i, int j, int k, Faces1<T, DIM>& F) {
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    F.v[2 * d] = A.b[d][A.g.fi(d, i, j, k)];
    F.v[2 * d + 1] = A.b[d][A.g.fi(d, i + (d == 0), j + (d == 1), k + (d == 2))];
  }
}

// lambda = 1/(alpha*a - beta*c + sum_d scale*beta*(b_hi + b_lo))
// (DarcyBrinkmanOp::resetLambda, src/DarcyBrinkmanOp.cpp:265-296, SUMFACES3
// src/DarcyBrinkmanOpF.ChF:511-545).
template <typename T, int DIM>
__device__ inline T vc_lambda(const Faces1<T, DIM>& F, T aq, bool hasc, T cq, T alpha, T beta, T scale) {
  T lam = aq * alpha;
  if (hasc) lam = lam + (-beta) * cq;
#pragma unroll
  for (int d = 0; d < DIM; ++d) lam = lam + scale * beta * (F.v[2 * d + 1] + F.v[2 * d]);
  return T(1) / lam;
}

// lofphi (src/DarcyBrinkmanOpF.ChF:125-147): alpha*a*u - beta*c*u - beta*S*dxinv
// from p[0] centre, p[2d+1] = +e_d, p[2d+2] = -e_d.
template <typename T, int DIM>
__device__ inline T vc_lof(const Faces1<T, DIM>& F, const T* p, T aq, bool hasc, T cq, T alpha, T beta, T dxinv) {
  const T p0 = p[0];
  T s = F.v[1] * (p[1] - p0) - F.v[0] * (p0 - p[2]) + F.v[3] * (p[3] - p0) - F.v[2] * (p0 - p[4]);
  if (DIM == 3) s = s + F.v[5] * (p[5] - p0) - F.v[4] * (p0 - p[6]);
  T lof = alpha * aq * p0;
  if (hasc) lof = lof - beta * cq * p0;
  return lof - beta * s * dxinv;
}

// Relaxation of L e = r on one tile with a halo of npass cells.  The
// coefficients, lambda and rhs of every position a thread may update are
// loaded once into registers; a pass updates one colour and refreshes the
// homogeneous wall ghosts of the cells it updated, so one barrier per pass.
template <typename T, int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__global__ __launch_bounds__(NT) void vc_relax_tile_kernel(VCArgs<T> A, const T* __restrict__ in,
                                                           T* __restrict__ out, const T* __restrict__ rhs, int npass,
                                                           int first, int zero_init, Geom gc,
                                                           const T* __restrict__ pc, int ntx, int nty, Geom gr,
                                                           T* __restrict__ resc) {
  if (A.skip && *A.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2 * HMAX) * (TY + 2 * HMAX) * (DIM == 3 ? TZ + 2 * HMAX : 1);
  constexpr int KMAX = (EMAX + NT - 1) / NT;
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  // with a restriction to produce, one more halo cell keeps the core's
  // neighbours exact after the last pass
  t.init(npass + (resc ? 1 : 0), ntx, nty);
  T* su = reinterpret_cast<T*>(smem_raw);
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * NT, R[kk]);
  Faces1<T, DIM> F[KMAX];
  T aq[KMAX], cv[KMAX], lam[KMAX], rr[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    T v = zero_init ? T(0) : in[r.q];
    if (pc) v = v + pc[gc.ci(r.i / 2, r.j / 2, DIM == 3 ? r.k / 2 : 0)];
    su[r.e] = v;
    if (r.dist > t.H - 1) continue;
    load_faces1<T, DIM>(A, r.i, r.j, r.k, F[kk]);
    aq[kk] = A.a ? A.a[r.q] : T(1);
    cv[kk] = A.c ? A.c[r.q] : T(0);
    rr[kk] = rhs[r.q];
    lam[kk] = vc_lambda<T, DIM>(F[kk], aq[kk], A.c != nullptr, cv[kk], A.alpha, A.beta, dxinv);
  }
  __syncthreads();
  // homogeneous wall ghosts of the staged state
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 2) continue;
    su[r.e] = bc_ghost<T>(A.bc.type[r.gd][r.gs][0], T(0), su[r.eint], r.gs == 0 ? -1 : 1, T(g.dx));
  }
  __syncthreads();
  for (int p = 0; p < npass; ++p) {
    const int gext = t.H - p;
    const int colour = (first + p) & 1;
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      const PosRec& r = R[kk];
      if (!(r.kind == 1 && r.dist <= gext - 1 && !((r.i + r.j + r.k + colour) & 1))) continue;
      T st[7];
      gather_lds<T, DIM>(su, r.e, t.EX, t.EXY, st);
      const T lof = vc_lof<T, DIM>(F[kk], st, aq[kk], A.c != nullptr, cv[kk], A.alpha, A.beta, dxinv);
      const T nv = st[0] - lam[kk] * (lof - rr[kk]);
      su[r.e] = nv;
      // wall ghosts of this cell (it is their interior neighbour)
      const int iv[3] = {r.i, r.j, r.k};
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        if (g.periodic[d]) continue;
        const int stride = d == 0 ? 1 : (d == 1 ? t.EX : t.EXY);
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
          if (iv[d] != (sd ? g.n[d] - 1 : 0)) continue;
          su[sd ? r.e + stride : r.e - stride] = bc_ghost<T>(A.bc.type[d][sd][0], T(0), nv, sd ? 1 : -1, T(g.dx));
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.owner) out[r.q] = su[r.e];
  }
  if (!resc) return;
  // restricted residual of the relaxed correction (RESTRICTRESDB,
  // src/DarcyBrinkmanOpF.ChF:402-509, homogeneous BCs): (r - L e)/2^d per
  // fine cell, child sums in Fortran order
  T* sres = su + t.E;
  const T denom = DIM == 3 ? T(8) : T(4);
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (!r.owner) continue;
    int ex, ey, ez;
    t.coords(r.e, ex, ey, ez);
    const int c = (ex - t.H) + TX * ((ey - t.H) + TY * (DIM == 3 ? ez - t.H : 0));
    T st[7];
    gather_lds<T, DIM>(su, r.e, t.EX, t.EXY, st);
    const T lof = vc_lof<T, DIM>(F[kk], st, aq[kk], A.c != nullptr, cv[kk], A.alpha, A.beta, dxinv);
    sres[c] = (rr[kk] - lof) / denom;
  }
  __syncthreads();
  constexpr int CX = TX / 2, CY = TY / 2, CZ = DIM == 3 ? TZ / 2 : 1;
  for (int c = threadIdx.x; c < CX * CY * CZ; c += NT) {
    const int cx = c % CX, cr = c / CX, cy = cr % CY, cz = cr / CY;
    const int I = t.x0 / 2 + cx, J = t.y0 / 2 + cy, K = DIM == 3 ? t.z0 / 2 + cz : 0;
    if (I >= gr.n[0] || J >= gr.n[1] || (DIM == 3 && K >= gr.n[2])) continue;
    T sum = T(0);
    for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
      for (int jj = 0; jj < 2; ++jj)
        for (int ii = 0; ii < 2; ++ii)
          sum = sum + sres[(2 * cx + ii) + TX * ((2 * cy + jj) + TY * (DIM == 3 ? 2 * cz + kk : 0))];
    resc[gr.ci(I, J, K)] = sum;
  }
}

// Residual / operator / restricted residual over a tile with a one-cell halo.
// kVCResid: u = phi (+ e when e != nullptr, written to phi_out), res = rhs - L u
//           (res optional), max-norm into *norm (optional);
// kVCApply: out = L u;  kVCRestrict: out (coarse) = child sums of (rhs - L u)/2^d.
template <typename T, int DIM, int TX, int TY, int TZ, int MODE>
__global__ __launch_bounds__(kVCThreads) void vc_tile_kernel(VCArgs<T> A, VCTileIO<T> io) {
  if (A.skip && *A.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2) * (TY + 2) * (DIM == 3 ? TZ + 2 : 1);
  constexpr int KMAX = (EMAX + kVCThreads - 1) / kVCThreads;
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init(1, io.ntx, io.nty);
  T* su = reinterpret_cast<T*>(smem_raw);
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * kVCThreads, R[kk]);
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    T v = io.phi[r.q];
    if (io.e) v = v + io.e[r.q];
    su[r.e] = v;
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 2) continue;
    const T v = io.homog ? T(0) : T(A.bc.val[r.gd][r.gs][0]);
    su[r.e] = bc_ghost<T>(A.bc.type[r.gd][r.gs][0], v, su[r.eint], r.gs == 0 ? -1 : 1, T(g.dx));
  }
  __syncthreads();
  if (MODE == kVCResid || MODE == kVCApply) {
    T mx = T(0);
    int nan = 0;
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      const PosRec& r = R[kk];
      if (!r.owner) continue;
      if (MODE == kVCResid && io.e) io.phi_out[r.q] = su[r.e];
      Faces1<T, DIM> F;
      load_faces1<T, DIM>(A, r.i, r.j, r.k, F);
      const T aq = A.a ? A.a[r.q] : T(1);
      const T cv = A.c ? A.c[r.q] : T(0);
      T st[7];
      gather_lds<T, DIM>(su, r.e, t.EX, t.EXY, st);
      const T lof = vc_lof<T, DIM>(F, st, aq, A.c != nullptr, cv, A.alpha, A.beta, dxinv);
      if (MODE == kVCResid) {
        const T res = io.rhs[r.q] - lof;
        if (io.out) io.out[r.q] = res;
        const T ar = fabs(res);
        if (ar != ar) nan = 1;
        else if (ar > mx) mx = ar;
      } else {
        io.out[r.q] = lof;
      }
    }
    if (MODE == kVCResid && io.norm) block_max_atomic<T>(mx, nan, io.norm);
    return;
  }
  // kVCRestrict (RESTRICTRESDB, src/DarcyBrinkmanOpF.ChF:402-509)
  constexpr int NC = TX * TY * (DIM == 3 ? TZ : 1);
  T* sres = su + t.E;
  const T denom = DIM == 3 ? T(8) : T(4);
  for (int c = threadIdx.x; c < NC; c += blockDim.x) {
    const int cx = c % TX, cr = c / TX, cy = cr % TY, cz = cr / TY;
    const int gi = t.x0 + cx, gj = t.y0 + cy, gk = DIM == 3 ? t.z0 + cz : 0;
    if (gi >= g.n[0] || gj >= g.n[1] || (DIM == 3 && gk >= g.n[2])) continue;
    const int e = (cx + 1) + t.EX * (cy + 1) + t.EXY * (DIM == 3 ? cz + 1 : 0);
    const long long q = g.ci(gi, gj, gk);
    Faces1<T, DIM> F;
    load_faces1<T, DIM>(A, gi, gj, gk, F);
    const T aq = A.a ? A.a[q] : T(1);
    const T cv = A.c ? A.c[q] : T(0);
    T st[7];
    gather_lds<T, DIM>(su, e, t.EX, t.EXY, st);
    const T lof = vc_lof<T, DIM>(F, st, aq, A.c != nullptr, cv, A.alpha, A.beta, dxinv);
    sres[c] = (io.rhs[q] - lof) / denom;
  }
  __syncthreads();
  constexpr int CX = TX / 2, CY = TY / 2, CZ = DIM == 3 ? TZ / 2 : 1;
  constexpr int NCC = CX * CY * CZ;
  const Geom gcc = io.gc;
  for (int c = threadIdx.x; c < NCC; c += blockDim.x) {
    const int cx = c % CX, cr = c / CX, cy = cr % CY, cz = cr / CY;
    const int I = t.x0 / 2 + cx, J = t.y0 / 2 + cy, K = DIM == 3 ? t.z0 / 2 + cz : 0;
    if (I >= gcc.n[0] || J >= gcc.n[1] || (DIM == 3 && K >= gcc.n[2])) continue;
    T s = T(0);
    for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
      for (int jj = 0; jj < 2; ++jj)
        for (int ii = 0; ii < 2; ++ii)
          s = s + sres[(2 * cx + ii) + TX * ((2 * cy + jj) + TY * (DIM == 3 ? 2 * cz + kk : 0))];
    io.out[gcc.ci(I, J, K)] = s;
  }
}

// ---------------------------------------------------------------------------
// Single-workgroup V-cycle over levels l0.. of a device descriptor.
template <typename T, int DIM>
struct VCWG {
  Geom g;
  int N, nx, nxy;
  int c[kVCM], i[kVCM], j[kVCM], k[kVCM];
  long long q[kVCM];
  bool own[kVCM];
  __device__ inline void init(const Geom& gg) {
    g = gg;
    nx = g.n[0];
    nxy = g.n[0] * g.n[1];
    N = nxy * (DIM == 3 ? g.n[2] : 1);
#pragma unroll
    for (int m = 0; m < kVCM; ++m) {
      c[m] = threadIdx.x + m * kVCWG;
      own[m] = c[m] < N;
      i[m] = c[m] % nx;
      const int r = c[m] / nx;
      j[m] = r % g.n[1];
      k[m] = DIM == 3 ? r / g.n[1] : 0;
      q[m] = own[m] ? g.ci(i[m], j[m], k[m]) : 0;
    }
  }
};

// stencil of cell (i,j,k) of a whole level held in LDS (su), with periodic
// wrap or the homogeneous wall rule
template <typename T, int DIM>
__device__ inline void vcwg_stencil(const Geom& g, const BCDesc& bc, const T* su, int c, int i, int j, int k,
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
        val = bc_ghost<T>(bc.type[d][s][0], T(0), su[c], s ? 1 : -1, T(g.dx));
      p[1 + 2 * d + (s ? 0 : 1)] = val;
    }
  }
}

template <typename T, int DIM>
__device__ void vcwg_relax(const VCWG<T, DIM>& W, const VCDesc<T>& D, const VCLevel<T>& lv, const T* sr, T* su,
                           int npass) {
  const Geom& g = W.g;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  VCArgs<T> A;
  A.g = g;
  A.bc = D.bc;
  A.alpha = D.alpha;
  A.beta = D.beta;
  A.a = lv.a;
  A.c = lv.c;
  for (int d = 0; d < 3; ++d) A.b[d] = lv.b[d];
  Faces1<T, DIM> F[kVCM];
  T aq[kVCM], cv[kVCM], lam[kVCM];
#pragma unroll
  for (int m = 0; m < kVCM; ++m) {
    if (!W.own[m]) continue;
    load_faces1<T, DIM>(A, W.i[m], W.j[m], W.k[m], F[m]);
    aq[m] = A.a ? A.a[W.q[m]] : T(1);
    cv[m] = A.c ? A.c[W.q[m]] : T(0);
    lam[m] = vc_lambda<T, DIM>(F[m], aq[m], A.c != nullptr, cv[m], A.alpha, A.beta, dxinv);
  }
  for (int p = 0; p < npass; ++p) {
#pragma unroll
    for (int m = 0; m < kVCM; ++m) {
      if (!W.own[m] || ((W.i[m] + W.j[m] + W.k[m] + p) & 1)) continue;
      T st[7];
      vcwg_stencil<T, DIM>(g, D.bc, su, W.c[m], W.i[m], W.j[m], W.k[m], W.nx, W.nxy, st);
      const T lof = vc_lof<T, DIM>(F[m], st, aq[m], A.c != nullptr, cv[m], A.alpha, A.beta, dxinv);
      su[W.c[m]] = st[0] - lam[m] * (lof - sr[W.c[m]]);
    }
    __syncthreads();
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(kVCWG) void vc_coarse_vcycle_kernel(const VCDesc<T>* __restrict__ Dp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const VCDesc<T>& D = *Dp;
  if (D.skip && *D.skip) return;
  T* su = reinterpret_cast<T*>(smem_raw);   // e of the level (kVCWGCells)
  T* sr = su + kVCWGCells;                  // r of the level
  const int l0 = D.l0, L = D.nlev;
  auto stage = [&](const VCLevel<T>& lv, const VCWG<T, DIM>& W, bool zero_e) {
#pragma unroll
    for (int m = 0; m < kVCM; ++m) {
      if (!W.own[m]) continue;
      su[W.c[m]] = zero_e ? T(0) : lv.e[W.q[m]];
      sr[W.c[m]] = lv.r[W.q[m]];
    }
    __syncthreads();
  };
  auto store = [&](const VCLevel<T>& lv, const VCWG<T, DIM>& W) {
#pragma unroll
    for (int m = 0; m < kVCM; ++m)
      if (W.own[m]) lv.e[W.q[m]] = su[W.c[m]];
    __syncthreads();
  };
  for (int l = l0; l < L - 1; ++l) {
    const VCLevel<T>& lv = D.lev[l];
    const VCLevel<T>& cv = D.lev[l + 1];
    VCWG<T, DIM> W;
    W.init(lv.g);
    stage(lv, W, true);
    vcwg_relax<T, DIM>(W, D, lv, sr, su, 2 * D.nu1);
    store(lv, W);
    // restricted residual (homogeneous), children in Fortran order
    const Geom& g = lv.g;
    const Geom& gcc = cv.g;
    const T dxinv = T(1) / (T(g.dx) * T(g.dx));
    const T denom = DIM == 3 ? T(8) : T(4);
    VCArgs<T> A;
    A.g = g;
    A.bc = D.bc;
    A.alpha = D.alpha;
    A.beta = D.beta;
    A.a = lv.a;
    A.c = lv.c;
    for (int d = 0; d < 3; ++d) A.b[d] = lv.b[d];
    const int NC = gcc.n[0] * gcc.n[1] * (DIM == 3 ? gcc.n[2] : 1);
    for (int cc = threadIdx.x; cc < NC; cc += kVCWG) {
      const int I = cc % gcc.n[0], rq = cc / gcc.n[0], J = rq % gcc.n[1], K = DIM == 3 ? rq / gcc.n[1] : 0;
      T s = T(0);
      for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
        for (int jj = 0; jj < 2; ++jj)
          for (int ii = 0; ii < 2; ++ii) {
            const int fi = 2 * I + ii, fj = 2 * J + jj, fk = DIM == 3 ? 2 * K + kk : 0;
            const int fc = fi + W.nx * fj + W.nxy * fk;
            Faces1<T, DIM> F;
            load_faces1<T, DIM>(A, fi, fj, fk, F);
            const long long q = g.ci(fi, fj, fk);
            const T aq = A.a ? A.a[q] : T(1);
            const T cvv = A.c ? A.c[q] : T(0);
            T st[7];
            vcwg_stencil<T, DIM>(g, D.bc, su, fc, fi, fj, fk, W.nx, W.nxy, st);
            s = s + (sr[fc] - vc_lof<T, DIM>(F, st, aq, A.c != nullptr, cvv, A.alpha, A.beta, dxinv)) / denom;
          }
      cv.r[gcc.ci(I, J, K)] = s;
    }
    __syncthreads();
  }
  {
    const VCLevel<T>& lv = D.lev[L - 1];
    VCWG<T, DIM> W;
    W.init(lv.g);
    stage(lv, W, true);
    vcwg_relax<T, DIM>(W, D, lv, sr, su, 2 * D.bottom);
    store(lv, W);
  }
  for (int l = L - 2; l >= l0; --l) {
    const VCLevel<T>& lv = D.lev[l];
    const VCLevel<T>& cv = D.lev[l + 1];
    VCWG<T, DIM> W;
    W.init(lv.g);
    stage(lv, W, false);
#pragma unroll
    for (int m = 0; m < kVCM; ++m) {
      if (!W.own[m]) continue;
      su[W.c[m]] = su[W.c[m]] + cv.e[cv.g.ci(W.i[m] / 2, W.j[m] / 2, DIM == 3 ? W.k[m] / 2 : 0)];
    }
    __syncthreads();
    vcwg_relax<T, DIM>(W, D, lv, sr, su, 2 * D.nu2);
    store(lv, W);
  }
}

template <int DIM>
struct VCRelaxTile;
template <>
struct VCRelaxTile<2> {
  static constexpr int TX = 32, TY = 16, TZ = 1, HMAX = 8, NT = 512;
};
template <>
struct VCRelaxTile<3> {
  static constexpr int TX = 16, TY = 8, TZ = 4, HMAX = 2, NT = 512;
};
template <int DIM>
struct VCTile;
template <>
struct VCTile<2> {
  static constexpr int TX = 32, TY = 8, TZ = 1;
};
template <>
struct VCTile<3> {
  static constexpr int TX = 16, TY = 8, TZ = 4;
};

template <int TX, int TY, int TZ>
int vc_tiles(const Geom& g, int& ntx, int& nty) {
  ntx = (g.n[0] + TX - 1) / TX;
  nty = (g.n[1] + TY - 1) / TY;
  const int ntz = g.dim == 3 ? (g.n[2] + TZ - 1) / TZ : 1;
  return ntx * nty * ntz;
}

template <typename T, int DIM>
hipError_t vc_relax_dim(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                        const Geom* gc, const T* pc, hipStream_t s, const Geom* gr = nullptr, T* resc = nullptr) {
  using D = VCRelaxTile<DIM>;
  int ntx, nty;
  const int nt = vc_tiles<D::TX, D::TY, D::TZ>(A.g, ntx, nty);
  const Geom gcc = gc ? *gc : A.g;
  const Geom grr = gr ? *gr : A.g;
  if (resc && npass + 1 > D::HMAX) return hipErrorInvalidValue;
  int done = 0;
  bool first_launch = true;
  while (done < npass || (first_launch && npass == 0 && (zero_init || pc || resc))) {
    first_launch = false;
    const int p = (npass - done) < D::HMAX ? (npass - done) : D::HMAX;
    const int h = p + (resc ? 1 : 0);
    size_t lds = size_t(D::TX + 2 * h) * (D::TY + 2 * h) * (DIM == 3 ? D::TZ + 2 * h : 1);
    if (resc) lds += size_t(D::TX) * D::TY * D::TZ;
    hipLaunchKernelGGL((vc_relax_tile_kernel<T, DIM, D::TX, D::TY, D::TZ, D::HMAX, D::NT>), dim3(nt), dim3(D::NT),
                       lds * sizeof(T), s, A, buf[*cur], buf[1 - *cur], rhs, p, done & 1,
                       (done == 0 && zero_init) ? 1 : 0, gcc, done == 0 ? pc : nullptr, ntx, nty, grr, resc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    *cur = 1 - *cur;
    done += p;
  }
  return hipSuccess;
}

template <typename T, int DIM, int MODE>
hipError_t vc_tile_dim(const VCArgs<T>& A, VCTileIO<T> io, hipStream_t s) {
  using D = VCTile<DIM>;
  const int nt = vc_tiles<D::TX, D::TY, D::TZ>(A.g, io.ntx, io.nty);
  size_t lds = size_t(D::TX + 2) * (D::TY + 2) * (DIM == 3 ? D::TZ + 2 : 1);
  if (MODE == kVCRestrict) lds += size_t(D::TX) * D::TY * D::TZ;
  hipLaunchKernelGGL((vc_tile_kernel<T, DIM, D::TX, D::TY, D::TZ, MODE>), dim3(nt), dim3(kVCThreads),
                     lds * sizeof(T), s, A, io);
  return hipGetLastError();
}

}  // namespace

template <typename T>
hipError_t vc_relax(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                    const Geom* gc, const T* pc, hipStream_t s) {
  return A.g.dim == 2 ? vc_relax_dim<T, 2>(A, buf, cur, rhs, npass, zero_init, gc, pc, s)
                      : vc_relax_dim<T, 3>(A, buf, cur, rhs, npass, zero_init, gc, pc, s);
}

template <typename T>
hipError_t vc_relax_restrict(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                             const Geom& gr, T* resc, hipStream_t s) {
  const int hmax = A.g.dim == 2 ? VCRelaxTile<2>::HMAX : VCRelaxTile<3>::HMAX;
  if (npass + 1 <= hmax)
    return A.g.dim == 2 ? vc_relax_dim<T, 2>(A, buf, cur, rhs, npass, zero_init, nullptr, nullptr, s, &gr, resc)
                        : vc_relax_dim<T, 3>(A, buf, cur, rhs, npass, zero_init, nullptr, nullptr, s, &gr, resc);
  hipError_t e = vc_relax<T>(A, buf, cur, rhs, npass, zero_init, nullptr, nullptr, s);
  if (e != hipSuccess) return e;
  return vc_restrict_residual<T>(A, gr, buf[*cur], rhs, resc, s);
}

template <typename T>
hipError_t vc_residual(const VCArgs<T>& A, const T* phi, const T* e, T* phi_out, const T* rhs, T* res,
                       unsigned long long* norm, int homog, hipStream_t s) {
  if (e && (!phi_out || phi_out == phi)) return hipErrorInvalidValue;
  VCTileIO<T> io{};
  io.phi = phi;
  io.e = e;
  io.phi_out = phi_out;
  io.rhs = rhs;
  io.out = res;
  io.norm = norm;
  io.homog = homog;
  return A.g.dim == 2 ? vc_tile_dim<T, 2, kVCResid>(A, io, s) : vc_tile_dim<T, 3, kVCResid>(A, io, s);
}

template <typename T>
hipError_t vc_apply_op(const VCArgs<T>& A, const T* phi, T* lhs, int homog, hipStream_t s) {
  VCTileIO<T> io{};
  io.phi = phi;
  io.out = lhs;
  io.homog = homog;
  return A.g.dim == 2 ? vc_tile_dim<T, 2, kVCApply>(A, io, s) : vc_tile_dim<T, 3, kVCApply>(A, io, s);
}

template <typename T>
hipError_t vc_restrict_residual(const VCArgs<T>& A, const Geom& gc, const T* phi, const T* rhs, T* resc,
                                hipStream_t s) {
  VCTileIO<T> io{};
  io.phi = phi;
  io.rhs = rhs;
  io.out = resc;
  io.gc = gc;
  io.homog = 1;
  return A.g.dim == 2 ? vc_tile_dim<T, 2, kVCRestrict>(A, io, s) : vc_tile_dim<T, 3, kVCRestrict>(A, io, s);
}

template <typename T>
hipError_t vc_coarse_vcycle(const VCDesc<T>* dev_desc, int dim, hipStream_t s) {
  const size_t lds = size_t(2) * kVCWGCells * sizeof(T);
  if (dim == 2)
    hipLaunchKernelGGL((vc_coarse_vcycle_kernel<T, 2>), dim3(1), dim3(kVCWG), lds, s, dev_desc);
  else
    hipLaunchKernelGGL((vc_coarse_vcycle_kernel<T, 3>), dim3(1), dim3(kVCWG), lds, s, dev_desc);
  return hipGetLastError();
}

bool vc_small_level(const Geom& g) { return g.nvalid() <= kVCWGCells; }

template hipError_t vc_relax<double>(const VCArgs<double>&, double* const[2], int*, const double*, int, bool,
                                     const Geom*, const double*, hipStream_t);
template hipError_t vc_relax_restrict<double>(const VCArgs<double>&, double* const[2], int*, const double*, int, bool,
                                              const Geom&, double*, hipStream_t);
template hipError_t vc_residual<double>(const VCArgs<double>&, const double*, const double*, double*,
                                        const double*, double*, unsigned long long*, int, hipStream_t);
template hipError_t vc_apply_op<double>(const VCArgs<double>&, const double*, double*, int, hipStream_t);
template hipError_t vc_restrict_residual<double>(const VCArgs<double>&, const Geom&, const double*, const double*,
                                                 double*, hipStream_t);
template hipError_t vc_coarse_vcycle<double>(const VCDesc<double>*, int, hipStream_t);

}  // namespace ml
