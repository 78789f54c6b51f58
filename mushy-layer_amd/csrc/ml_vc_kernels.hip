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
#include <hip/hip_ext.h>

#include "ml_common.h"
#include <algorithm>

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
// The body works on tile `tile` of the level (a batched launch maps its
// workgroups to (instance, tile) pairs).
template <typename T, int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__device__ __forceinline__ void vc_relax_tile_body(const VCArgs<T>& A, const T* __restrict__ in, T* __restrict__ out,
                                                   const T* __restrict__ rhs, int npass, int first, int zero_init,
                                                   const Geom& gc, const T* __restrict__ pc, int ntx, int nty,
                                                   const Geom& gr, T* __restrict__ resc, const VCFinal<T>& fin,
                                                   int tile) {
  if (A.skip && *A.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2 * HMAX) * (TY + 2 * HMAX) * (DIM == 3 ? TZ + 2 * HMAX : 1);
  constexpr int KMAX = (EMAX + NT - 1) / NT;
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  // with a restriction or a residual to produce, one more halo cell keeps
  // the core's neighbours exact after the last pass
  const bool final = fin.phi != nullptr;
  t.init_at(npass + ((resc || final) ? 1 : 0), ntx, nty, tile);
  T* su = reinterpret_cast<T*>(smem_raw);
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * NT, R[kk]);
  Faces1<T, DIM> F[KMAX];
  T aq[KMAX], cv[KMAX], lam[KMAX], rr[KMAX], v[KMAX], pv[KMAX], fv[KMAX], frhs[KMAX];
  // every global load of the tile is issued before the first use
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    v[kk] = T(0);
    pv[kk] = T(0);
    // phi near the core (for a wall ghost: of its interior neighbour)
    if (final && r.kind != 0 && r.dist <= 1) fv[kk] = fin.phi[r.q];
    if (r.kind != 1) continue;
    if (!zero_init) v[kk] = in[r.q];
    if (pc) pv[kk] = pc[gc.ci(r.i / 2, r.j / 2, DIM == 3 ? r.k / 2 : 0)];
    if (final && r.owner) frhs[kk] = fin.rhs[r.q];
  }
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1 || r.dist > t.H - 1) continue;
    load_faces1<T, DIM>(A, r.i, r.j, r.k, F[kk]);
    aq[kk] = A.a ? A.a[r.q] : T(1);
    cv[kk] = A.c ? A.c[r.q] : T(0);
    rr[kk] = rhs[r.q];
  }
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    T val = v[kk];
    if (pc) val = val + pv[kk];
    su[r.e] = val;
    if (r.dist > t.H - 1) continue;
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
  if (final) {
    // u = phi + e near the core, its inhomogeneous wall ghosts, then the
    // residual of the full equation at the owned cells (DBCOMPUTERES order,
    // as vc_tile_kernel<kVCResid> with e)
    T* sv = su + t.E;
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      const PosRec& r = R[kk];
      if (r.dist > 1) continue;
      if (r.kind == 1) {
        sv[r.e] = fv[kk] + su[r.e];
      } else if (r.kind == 2) {
        // the interior value u = phi + e of the ghost's neighbour
        const T ui = fv[kk] + su[r.eint];
        sv[r.e] = bc_ghost<T>(A.bc.type[r.gd][r.gs][0], T(A.bc.val[r.gd][r.gs][0]), ui, r.gs == 0 ? -1 : 1,
                              T(g.dx));
      }
    }
    __syncthreads();
    T mx = T(0);
    int nan = 0;
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      const PosRec& r = R[kk];
      if (!r.owner) continue;
      fin.phi_out[r.q] = sv[r.e];
      T st[7];
      gather_lds<T, DIM>(sv, r.e, t.EX, t.EXY, st);
      const T lof = vc_lof<T, DIM>(F[kk], st, aq[kk], A.c != nullptr, cv[kk], A.alpha, A.beta, dxinv);
      const T res = frhs[kk] - lof;
      fin.res[r.q] = res;
      const T ar = fabs(res);
      if (ar != ar) nan = 1;
      else if (ar > mx) mx = ar;
    }
    block_max_atomic<T>(mx, nan, ctl_slot(nullptr, fin.fuse), fin.fuse);
    return;
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

template <typename T, int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__global__ __launch_bounds__(NT) void vc_relax_tile_kernel(VCArgs<T> A, const T* __restrict__ in,
                                                           T* __restrict__ out, const T* __restrict__ rhs, int npass,
                                                           int first, int zero_init, Geom gc,
                                                           const T* __restrict__ pc, int ntx, int nty, Geom gr,
                                                           T* __restrict__ resc, VCFinal<T> fin) {
  vc_relax_tile_body<T, DIM, TX, TY, TZ, HMAX, NT>(A, in, out, rhs, npass, first, zero_init, gc, pc, ntx, nty, gr,
                                                   resc, fin, int(blockIdx.x));
}

// batched form: workgroup w relaxes tile map[w] & mask of job map[w] >> bits
template <typename T, int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__global__ __launch_bounds__(NT) void vc_relax_batch_kernel(const VCRelaxJob<T>* __restrict__ jobs,
                                                            const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const VCRelaxJob<T>& J = jobs[w >> kBatchTileBits];
  vc_relax_tile_body<T, DIM, TX, TY, TZ, HMAX, NT>(J.A, J.in, J.out, J.rhs, J.npass, J.first, J.zero_init, J.gc, J.pc,
                                                   J.ntx, J.nty, J.gr, J.resc, J.fin, int(w & kBatchTileMask));
}

// Residual / operator / restricted residual over a tile with a one-cell halo.
// kVCResid: u = phi (+ e when e != nullptr, written to phi_out), res = rhs - L u
//           (res optional), max-norm into *norm (optional);
// kVCApply: out = L u;  kVCRestrict: out (coarse) = child sums of (rhs - L u)/2^d.
template <typename T, int DIM, int TX, int TY, int TZ, int MODE>
__device__ __forceinline__ void vc_tile_body(const VCArgs<T>& A, const VCTileIO<T>& io, int tile) {
  if (A.skip && *A.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2) * (TY + 2) * (DIM == 3 ? TZ + 2 : 1);
  constexpr int KMAX = (EMAX + kVCThreads - 1) / kVCThreads;
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init_at(1, io.ntx, io.nty, tile);
  T* su = reinterpret_cast<T*>(smem_raw);
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * kVCThreads, R[kk]);
  // owner data of the residual / operator modes, loaded with the state
  Faces1<T, DIM> OF[KMAX];
  T oaq[KMAX], ocv[KMAX], orhs[KMAX], v[KMAX], ev[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    v[kk] = T(0);
    ev[kk] = T(0);
    if (r.kind != 1) continue;
    v[kk] = io.phi[r.q];
    if (io.e) ev[kk] = io.e[r.q];
  }
  if (MODE == kVCResid || MODE == kVCApply) {
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      const PosRec& r = R[kk];
      if (!r.owner) continue;
      load_faces1<T, DIM>(A, r.i, r.j, r.k, OF[kk]);
      oaq[kk] = A.a ? A.a[r.q] : T(1);
      ocv[kk] = A.c ? A.c[r.q] : T(0);
      if (MODE == kVCResid) orhs[kk] = io.rhs[r.q];
    }
  }
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    T val = v[kk];
    if (io.e) val = val + ev[kk];
    su[r.e] = val;
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
      T st[7];
      gather_lds<T, DIM>(su, r.e, t.EX, t.EXY, st);
      const T lof = vc_lof<T, DIM>(OF[kk], st, oaq[kk], A.c != nullptr, ocv[kk], A.alpha, A.beta, dxinv);
      if (MODE == kVCResid) {
        const T res = orhs[kk] - lof;
        if (io.out) io.out[r.q] = res;
        const T ar = fabs(res);
        if (ar != ar) nan = 1;
        else if (ar > mx) mx = ar;
      } else {
        io.out[r.q] = lof;
      }
    }
    if (MODE == kVCResid && (io.norm || io.fuse.ctl))
      block_max_atomic<T>(mx, nan, ctl_slot(io.norm, io.fuse), io.fuse);
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

template <typename T, int DIM, int TX, int TY, int TZ, int MODE>
__global__ __launch_bounds__(kVCThreads) void vc_tile_kernel(VCArgs<T> A, VCTileIO<T> io) {
  vc_tile_body<T, DIM, TX, TY, TZ, MODE>(A, io, int(blockIdx.x));
}

// batched residuals: workgroup w computes tile map[w] & mask of job map[w] >> bits
template <typename T, int DIM, int TX, int TY, int TZ>
__global__ __launch_bounds__(kVCThreads) void vc_resid_batch_kernel(const VCResidJob<T>* __restrict__ jobs,
                                                                    const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const VCResidJob<T>& J = jobs[w >> kBatchTileBits];
  vc_tile_body<T, DIM, TX, TY, TZ, kVCResid>(J.A, J.io, int(w & kBatchTileMask));
}

// batched 8-byte copies and zero fills: workgroup w handles chunk map[w] &
// mask of job map[w] >> bits
__global__ __launch_bounds__(256) void words_batch_kernel(const WordsJob* __restrict__ jobs,
                                                          const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const WordsJob& J = jobs[w >> kBatchTileBits];
  const long long c0 = (long long)(w & kBatchTileMask) * kWordsChunk;
  const long long c1 = c0 + kWordsChunk < J.n ? c0 + kWordsChunk : J.n;
  for (long long i = c0 + threadIdx.x; i < c1; i += 256) J.dst[i] = J.src ? J.src[i] : 0ull;
}

// ---------------------------------------------------------------------------
// Single-workgroup V-cycle over levels l0.. of a device descriptor.  The
// correction, right-hand side and operator data (faces, a, c, lambda) of
// every level live in LDS for the whole cycle (compact, x fastest, levels
// one after another); the operator data are loaded once at the start.  Each
// thread owns one cell of each colour of a level (even row length) or the
// cells t and t + 512, with their coordinates held in registers; levels
// whose owners fit one wavefront are relaxed by wave 0 alone.

constexpr int kVCL = 6;   // levels one launch handles: 1024 cells coarsen to one cell in 6 (2D), 4 (3D)
template <int DIM>
struct VCLevels {
  static constexpr int n = DIM == 3 ? 4 : kVCL;
};

// operator data of all levels in LDS; q = level offset + compact cell index
template <typename T, int DIM>
struct VCLds {
  T *se, *sr, *sres;   // corrections, right-hand sides, restriction scratch
  T *sF, *sA, *sC, *sL;   // faces [2*DIM][tot], a, c, lambda
  int tot;
  __device__ inline void faces(int q, Faces1<T, DIM>& F) const {
#pragma unroll
    for (int f = 0; f < 2 * DIM; ++f) F.v[f] = sF[f * tot + q];
  }
};

template <typename T>
size_t vc_coarse_lds_elems(int dim, long long tot) {
  return size_t(tot) * (2 + 2 * dim + 3) + kVCWGCells;
}

// homogeneous wall ghost of component 0 (bit 2d+s of bcm: type != 0)
template <typename T>
__device__ __forceinline__ T vc_wall(int bcm, int d, int s, T interior, T dx) {
  return bc_ghost<T>((bcm >> (2 * d + s)) & 1, T(0), interior, s ? 1 : -1, dx);
}

// A thread's cells of one level for the duration of a level visit:
// coordinates, colour, neighbour indices and operator data in registers.
template <typename T, int DIM>
struct VCCtx {
  int pk[2], c[2], col[2], wall[2];
  int nb[2][2 * DIM];
  Faces1<T, DIM> F[2];
  T aq[2], cv[2], lam[2], rr[2];
};

template <typename T, int DIM>
__device__ __forceinline__ void vcx_build(VCCtx<T, DIM>& X, const CLev& L, const int* per, const VCLds<T, DIM>& S) {
  COwn o;
  cown_set<kVCWG>(L, o);
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
    S.faces(q, X.F[m]);
    X.aq[m] = S.sA[q];
    X.cv[m] = S.sC[q];
    X.lam[m] = S.sL[q];
    X.rr[m] = S.sr[q];
  }
}

// stencil of cell m of the context from the level's corrections su
template <typename T, int DIM>
__device__ __forceinline__ void vcx_stencil(const VCCtx<T, DIM>& X, int m, int bcm, T dx, const T* su, T* p) {
  p[0] = su[X.c[m]];
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) {
    const int d = f >> 1, s = f & 1;
    T v = su[X.nb[m][f]];
    if ((X.wall[m] >> f) & 1) v = vc_wall<T>(bcm, d, s, v, dx);
    p[1 + 2 * d + (1 - s)] = v;
  }
}

template <typename T, int DIM>
__device__ __forceinline__ void vcx_relax(const VCCtx<T, DIM>& X, const CLev& L, int bcm, const VCLds<T, DIM>& S,
                                          bool hasc, T alpha, T beta, int npass) {
  const T dx = T(L.dx);
  const T dxinv = T(1) / (dx * dx);
  T* se = S.se + L.off;
  auto pass = [&](int p) {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      if (X.pk[m] < 0 || X.col[m] != (p & 1)) continue;
      T st[7];
      vcx_stencil<T, DIM>(X, m, bcm, dx, se, st);
      const T lof = vc_lof<T, DIM>(X.F[m], st, X.aq[m], hasc, X.cv[m], alpha, beta, dxinv);
      se[X.c[m]] = st[0] - X.lam[m] * (lof - X.rr[m]);
    }
  };
  // levels whose owners fit one wavefront: wave 0 alone with wavefront-scope
  // syncs (small is uniform, so the workgroup barrier is reached by all
  // waves or by none)
  const bool small = cown_active<kVCWG>(L) <= 64;
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

// restricted residual (homogeneous BCs, RESTRICTRESDB order) of level Lf into
// the right-hand side of level Lc, whose correction starts at zero
template <typename T, int DIM>
__device__ __forceinline__ void vcx_restrict(const VCCtx<T, DIM>& X, const CLev& Lf, const CLev& Lc, int bcm,
                                             const VCLds<T, DIM>& S, bool hasc, T alpha, T beta) {
  const T dx = T(Lf.dx);
  const T dxinv = T(1) / (dx * dx);
  const T denom = DIM == 3 ? T(8) : T(4);
  const T* se = S.se + Lf.off;
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    if (X.pk[m] < 0) continue;
    T st[7];
    vcx_stencil<T, DIM>(X, m, bcm, dx, se, st);
    const T lof = vc_lof<T, DIM>(X.F[m], st, X.aq[m], hasc, X.cv[m], alpha, beta, dxinv);
    S.sres[X.c[m]] = (X.rr[m] - lof) / denom;
  }
  __syncthreads();
  for (int cc = threadIdx.x; cc < Lc.N; cc += kVCWG) {
    int I, J, K;
    clev_coords(Lc, cc, I, J, K);
    T s = T(0);
    for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
      for (int jj = 0; jj < 2; ++jj)
        for (int ii = 0; ii < 2; ++ii)
          s = s + S.sres[(2 * I + ii) + Lf.nx * (2 * J + jj) + Lf.nxy * (DIM == 3 ? 2 * K + kk : 0)];
    S.sr[Lc.off + cc] = s;
    S.se[Lc.off + cc] = T(0);
  }
  __syncthreads();
}

// e_f += e_c(parent) (piecewise-constant prolongation)
template <typename T, int DIM>
__device__ __forceinline__ void vcx_prolong(const VCCtx<T, DIM>& X, const CLev& Lf, const CLev& Lc, T* se) {
#pragma unroll
  for (int m = 0; m < 2; ++m) {
    if (X.pk[m] < 0) continue;
    int i, j, k;
    cunpack(X.pk[m], i, j, k);
    const int c = Lf.off + X.c[m];
    se[c] = se[c] + se[Lc.off + (i / 2) + Lc.nx * (j / 2) + Lc.nxy * (DIM == 3 ? k / 2 : 0)];
  }
  __syncthreads();
}

// BiCGStab on the coarsest level L (held in LDS) from its relaxed
// correction: L e = r with homogeneous BCs, preconditioner preCond (e =
// lambda*r, then two red/black iterations).  Preconditioned BiCGStab with
// restarts; eps 1e-6, reps 1e-12, hang 1e-8, small 1e-30, 5 restarts, L2
// norms.  Dot products run redundantly in every participating wavefront in a
// fixed order: 64 interleaved lane sums over the compact cell index, then the
// pairwise tree s_t + s_{t+h}, h = P/2 .. 1 with P the smallest power of two
// >= min(N, 64).  Levels whose owners fit one
// wavefront run on wave 0 alone with wavefront-scope syncs.  wk holds 8
// vectors of L.N values.
template <typename T, int DIM>
__device__ __forceinline__ void vcx_bicgstab(const VCCtx<T, DIM>& X, const CLev& L, int bcm, const VCLds<T, DIM>& S, bool hasc,
                             T alpha, T beta, int imax, T* wk) {
  const int N = L.N;
  const bool small = cown_active<kVCWG>(L) <= 64 && N <= 64 * 2;
  if (small && threadIdx.x >= 64) {
    __syncthreads();
    return;
  }
  const int tid = threadIdx.x, stride = small ? 64 : kVCWG, lane = threadIdx.x & 63;
  auto sync = [&]() __attribute__((always_inline)) {
    if (small)
      wave_sync();
    else
      __syncthreads();
  };
  const T dx = T(L.dx);
  const T dxinv = T(1) / (dx * dx);
  T* phi = S.se + L.off;
  const T* rhs = S.sr + L.off;
  T *r = wk, *rt = wk + N, *e = wk + 2 * N, *p = wk + 3 * N, *pt = wk + 4 * N, *st = wk + 5 * N, *t = wk + 6 * N,
    *v = wk + 7 * N;
  auto lof_of = [&](int m, const T* u) __attribute__((always_inline)) {
    T stc[7];
    vcx_stencil<T, DIM>(X, m, bcm, dx, u, stc);
    return vc_lof<T, DIM>(X.F[m], stc, X.aq[m], hasc, X.cv[m], alpha, beta, dxinv);
  };
  auto apply = [&](T* out, const T* u) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < 2; ++m)
      if (X.pk[m] >= 0) out[X.c[m]] = lof_of(m, u);
    sync();
  };
  auto residual = [&](T* out, const T* u, const T* f) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < 2; ++m)
      if (X.pk[m] >= 0) out[X.c[m]] = f[X.c[m]] - lof_of(m, u);
    sync();
  };
  auto precond = [&](T* z, const T* f) __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < 2; ++m)
      if (X.pk[m] >= 0) z[X.c[m]] = f[X.c[m]] * X.lam[m];
    sync();
#pragma unroll 1
    for (int pass = 0; pass < 4; ++pass) {
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        if (X.pk[m] < 0 || X.col[m] != (pass & 1)) continue;
        T stc[7];
        vcx_stencil<T, DIM>(X, m, bcm, dx, z, stc);
        const T lof = vc_lof<T, DIM>(X.F[m], stc, X.aq[m], hasc, X.cv[m], alpha, beta, dxinv);
        z[X.c[m]] = stc[0] - X.lam[m] * (lof - f[X.c[m]]);
      }
      sync();
    }
  };
  // tree over the smallest power of two of lanes >= min(N, 64)
  int P2 = 1;
  while (P2 < 64 && P2 < N) P2 <<= 1;
  auto dot = [&](const T* x, const T* y) __attribute__((always_inline)) {
    T s = T(0);
    for (int c = lane; c < N; c += 64) s = s + x[c] * y[c];
    for (int h = P2 >> 1; h >= 1; h >>= 1) s = s + __shfl_down(s, h, 64);
    return __shfl(s, 0, 64);
  };
  const T eps = T(1e-6), reps = T(1e-12), hang = T(1e-8), tiny = T(1e-30);
  residual(r, phi, rhs);
  for (int c = tid; c < N; c += stride) {
    rt[c] = r[c];
    e[c] = T(0);
  }
  T norm0 = sqrt(dot(r, r)), norm1 = norm0;
  const T r0 = norm0;
  T rho0 = 0, rho1 = 0, alpha0 = 0, alpha1 = 0, omega0 = 0, omega1 = 0;
  bool init = true;
  int restarts = 0, recount = 0, i = 0;
  sync();
  while (i < imax && norm0 > eps * norm1 && norm1 > T(0)) {
    ++i;
    norm1 = norm0;
    alpha1 = alpha0;
    omega1 = omega0;
    rho1 = rho0;
    rho0 = dot(rt, r);
    if (rho0 == T(0)) break;
    if (init) {
      for (int c = tid; c < N; c += stride) p[c] = r[c];
      init = false;
    } else {
      const T bt = (rho0 / rho1) * (alpha1 / omega1);
      const T sv = -bt * omega1;
      for (int c = tid; c < N; c += stride) {
        T q = bt * p[c];
        q = q + sv * v[c];
        p[c] = q + T(1) * r[c];
      }
    }
    sync();
    precond(pt, p);
    apply(v, pt);
    const T m = dot(rt, v);
    alpha0 = rho0 / m;
    if (fabs(m) > tiny * fabs(rho0)) {
      const T na = -alpha0;
      for (int c = tid; c < N; c += stride) {
        r[c] = r[c] + na * v[c];
        e[c] = e[c] + alpha0 * pt[c];
      }
      sync();
      norm0 = sqrt(dot(r, r));
    } else {
      for (int c = tid; c < N; c += stride) r[c] = T(0);
      sync();
      norm0 = T(0);
    }
    if (norm0 > eps * r0 && norm0 > reps * r0) {
      precond(st, r);
      apply(t, st);
      omega0 = dot(t, r) / dot(t, t);
      sync();   // every wavefront has read r before it changes
      const T nw = -omega0;
      for (int c = tid; c < N; c += stride) {
        e[c] = e[c] + omega0 * st[c];
        r[c] = r[c] + nw * t[c];
      }
      sync();
      norm0 = sqrt(dot(r, r));
    }
    if (norm0 <= eps * r0 || norm0 <= reps * r0) break;
    if (omega0 == T(0) || norm0 > (T(1) - hang) * norm1) {
      if (recount == 0) {
        recount = 1;
      } else {
        recount = 0;
        for (int c = tid; c < N; c += stride) phi[c] = phi[c] + T(1) * e[c];
        if (restarts == 5) {
          for (int c = tid; c < N; c += stride) e[c] = T(0);
          break;
        }
        sync();
        residual(r, phi, rhs);
        norm0 = sqrt(dot(r, r));
        rho0 = alpha0 = omega0 = T(0);
        for (int c = tid; c < N; c += stride) {
          rt[c] = r[c];
          e[c] = T(0);
        }
        sync();
        ++restarts;
        init = true;
      }
    }
  }
  sync();
  for (int c = tid; c < N; c += stride) phi[c] = phi[c] + T(1) * e[c];
  __syncthreads();
}

// one cell's operator data (and right-hand side) read from global memory
template <typename T, int DIM>
struct VCPre {
  Faces1<T, DIM> F;
  T aq, cv, r;
};

template <typename T, int DIM>
__device__ __forceinline__ void vcpre_load(VCPre<T, DIM>& P, const VCLevel<T>& lv, const CLev& L, int c, bool rhs) {
  if (c >= L.N) return;
  int i, j, k;
  clev_coords(L, c, i, j, k);
  VCArgs<T> A;
  A.g = lv.g;
  for (int d = 0; d < 3; ++d) A.b[d] = lv.b[d];
  load_faces1<T, DIM>(A, i, j, k, P.F);
  const long long g = lv.g.ci(i, j, k);
  P.aq = lv.a ? lv.a[g] : T(1);
  P.cv = lv.c ? lv.c[g] : T(0);
  if (rhs) P.r = lv.r[g];
}

template <typename T, int DIM>
__device__ __forceinline__ void vcpre_store(const VCPre<T, DIM>& P, const VCLds<T, DIM>& S, const CLev& L, int c,
                                            bool hasc, T alpha, T beta, bool rhs) {
  if (c >= L.N) return;
  const int q = L.off + c;
#pragma unroll
  for (int f = 0; f < 2 * DIM; ++f) S.sF[f * S.tot + q] = P.F.v[f];
  S.sA[q] = P.aq;
  S.sC[q] = P.cv;
  const T dxinv = T(1) / (T(L.dx) * T(L.dx));
  S.sL[q] = vc_lambda<T, DIM>(P.F, P.aq, hasc, P.cv, alpha, beta, dxinv);
  if (rhs) {
    S.se[q] = T(0);
    S.sr[q] = P.r;
  }
}

template <typename T, int DIM>
__device__ __forceinline__ void vc_coarse_vcycle_body(const VCDesc<T>* __restrict__ Dp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int LL = VCLevels<DIM>::n;
  const VCDesc<T>& D = *Dp;
  if (D.skip && *D.skip) return;
  const int l0 = D.l0, nl = D.nlev - D.l0;
  const bool hasc = D.lev[l0].c != nullptr;
  const T alpha = D.alpha, beta = D.beta;
  int per[3];
  for (int d = 0; d < 3; ++d) per[d] = D.lev[l0].g.periodic[d];
  int bcm = 0;
  for (int d = 0; d < DIM; ++d)
    for (int sd = 0; sd < 2; ++sd) bcm |= (D.bc.type[d][sd][0] != 0 ? 1 : 0) << (2 * d + sd);
  CLev G[LL];
  int tot = 0;
#pragma unroll
  for (int l = 0; l < LL; ++l) {
    if (l >= nl) continue;
    clev_set(G[l], D.lev[l0 + l].g, tot);
    tot += G[l].N;
  }
  VCLds<T, DIM> S;
  S.tot = tot;
  S.se = reinterpret_cast<T*>(smem_raw);
  S.sr = S.se + tot;
  S.sF = S.sr + tot;
  S.sA = S.sF + 2 * DIM * tot;
  S.sC = S.sA + tot;
  S.sL = S.sC + tot;
  S.sres = S.sL + tot;
  // operator data of every level, and level l0's right-hand side: all
  // global loads first (level l0: cells t and t + 512, coarser levels: cell
  // t), then the LDS stores
  VCPre<T, DIM> P0[2], P1[LL - 1];
#pragma unroll
  for (int m = 0; m < 2; ++m) vcpre_load<T, DIM>(P0[m], D.lev[l0], G[0], threadIdx.x + m * kVCWG, true);
#pragma unroll
  for (int l = 1; l < LL; ++l)
    if (l < nl) vcpre_load<T, DIM>(P1[l - 1], D.lev[l0 + l], G[l], threadIdx.x, false);
#pragma unroll
  for (int m = 0; m < 2; ++m) vcpre_store<T, DIM>(P0[m], S, G[0], threadIdx.x + m * kVCWG, hasc, alpha, beta, true);
#pragma unroll
  for (int l = 1; l < LL; ++l)
    if (l < nl) vcpre_store<T, DIM>(P1[l - 1], S, G[l], threadIdx.x, hasc, alpha, beta, false);
  __syncthreads();
  // The level visits below are runtime loops over one copy of their code
  // (levels rebuilt from the descriptor, offsets by prefix sum): unrolled
  // per level, the kernel's code is several times the instruction cache and
  // a launch spends most of its time fetching it.
  auto level = [&](int l) __attribute__((always_inline)) {
    CLev L;
    int off = 0;
    for (int q = 0; q < l; ++q) off += int(D.lev[l0 + q].g.nvalid());
    clev_set(L, D.lev[l0 + l].g, off);
    return L;
  };
  // down: pre-smoothing and restriction; the coarsest level gets the bottom
  // smoothing
#pragma unroll 1
  for (int l = 0; l < nl; ++l) {
    const CLev Ll = level(l);
    VCCtx<T, DIM> X;
    vcx_build<T, DIM>(X, Ll, per, S);
    vcx_relax<T, DIM>(X, Ll, bcm, S, hasc, alpha, beta, 2 * (l == nl - 1 ? D.bottom : D.nu1));
    if (l + 1 < nl) vcx_restrict<T, DIM>(X, Ll, level(l + 1), bcm, S, hasc, alpha, beta);
  }
  // the bottom solver on the coarsest level (after its relaxations)
  if (D.bottom_solver == 1) {
    const Geom& gb = D.lev[D.nlev - 1].g;
    CLev Lb;
    clev_set(Lb, gb, tot - int(gb.nvalid()));
    VCCtx<T, DIM> X;
    vcx_build<T, DIM>(X, Lb, per, S);
    vcx_bicgstab<T, DIM>(X, Lb, bcm, S, hasc, alpha, beta, D.bottom_imax, S.sres + kVCWGCells);
  }
  // up: prolongation and post-smoothing
#pragma unroll 1
  for (int l = nl - 2; l >= 0; --l) {
    const CLev Ll = level(l);
    VCCtx<T, DIM> X;
    vcx_build<T, DIM>(X, Ll, per, S);
    vcx_prolong<T, DIM>(X, Ll, level(l + 1), S.se);
    vcx_relax<T, DIM>(X, Ll, bcm, S, hasc, alpha, beta, 2 * D.nu2);
  }
  const VCLevel<T>& lv = D.lev[l0];
  for (int c = threadIdx.x; c < G[0].N; c += kVCWG) {
    int i, j, k;
    clev_coords(G[0], c, i, j, k);
    lv.e[lv.g.ci(i, j, k)] = S.se[c];
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(kVCWG) void vc_coarse_vcycle_kernel(const VCDesc<T>* __restrict__ Dp) {
  vc_coarse_vcycle_body<T, DIM>(Dp);
}

// batched form: workgroup w runs the coarse V-cycle of descriptor descs[w]
template <typename T, int DIM>
__global__ __launch_bounds__(kVCWG) void vc_coarse_batch_kernel(const VCDesc<T>* const* __restrict__ descs) {
  // the descriptor address is uniform over the workgroup: keep it scalar
  const unsigned long long a = reinterpret_cast<unsigned long long>(descs[blockIdx.x]);
  const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(a)), hi = __builtin_amdgcn_readfirstlane(unsigned(a >> 32));
  vc_coarse_vcycle_body<T, DIM>(reinterpret_cast<const VCDesc<T>*>((static_cast<unsigned long long>(hi) << 32) | lo));
}

template <int DIM>
struct VCRelaxTile;
template <>
struct VCRelaxTile<2> {
  static constexpr int TX = 16, TY = 16, TZ = 1, HMAX = 9, NT = 1024;   // measured best of 11 shapes at 256^2
};
template <>
struct VCRelaxTile<3> {
  // 1024 threads: two staged positions per thread, no register spill (19 %
  // faster at 128^3 than 512 threads)
  static constexpr int TX = 16, TY = 8, TZ = 4, HMAX = 2, NT = 1024;
};
// batches with many cells are bandwidth-bound: a 32x32 core restages fewer
// halo cells per owned cell (2.4x at the 9-cell halo, 4.5x for 16x16)
struct VCRelaxBatchWide2 {
  static constexpr int TX = 32, TY = 32, TZ = 1, HMAX = 9, NT = 1024;
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

template <typename T, int DIM, typename D>
hipError_t vc_relax_cfg(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                        const Geom* gc, const T* pc, hipStream_t s, const Geom* gr, T* resc, const LaunchTimer* tm,
                        const VCFinal<T>* fin) {
  int ntx, nty;
  const int nt = vc_tiles<D::TX, D::TY, D::TZ>(A.g, ntx, nty);
  const Geom gcc = gc ? *gc : A.g;
  const Geom grr = gr ? *gr : A.g;
  if ((resc || fin) && npass + 1 > D::HMAX) return hipErrorInvalidValue;
  const VCFinal<T> none{};
  int done = 0;
  bool first_launch = true;
  while (done < npass || (first_launch && npass == 0 && (zero_init || pc || resc))) {
    first_launch = false;
    const int p = (npass - done) < D::HMAX ? (npass - done) : D::HMAX;
    const int h = p + ((resc || fin) ? 1 : 0);
    size_t lds = size_t(D::TX + 2 * h) * (D::TY + 2 * h) * (DIM == 3 ? D::TZ + 2 * h : 1);
    if (resc) lds += size_t(D::TX) * D::TY * D::TZ;
    if (fin) lds *= 2;   // e and u = phi + e
    const bool timed = tm && *tm->n < tm->cap;
    if (tm && !timed) ++*tm->missed;
    const bool zi = done == 0 && zero_init;
    const T* pcl = done == 0 ? pc : nullptr;
    // a timed launch carries its start and stop events in the dispatch
    hipExtLaunchKernelGGL((vc_relax_tile_kernel<T, DIM, D::TX, D::TY, D::TZ, D::HMAX, D::NT>), dim3(nt),
                          dim3(D::NT), lds * sizeof(T), s, timed ? tm->start[*tm->n] : nullptr,
                          timed ? tm->stop[*tm->n] : nullptr, 0, A, buf[*cur], buf[1 - *cur], rhs, p, done & 1,
                          zi ? 1 : 0, gcc, pcl, ntx, nty, grr, resc, fin ? *fin : none);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (timed) {
      // algorithmic bytes: state in (unless zero start), coarse correction,
      // rhs, a, c, faces once each; state out and restricted residual once
      const double cells = double(A.g.nvalid());
      const double coarse = cells / double(DIM == 3 ? 8 : 4);
      double faces = 0;
      for (int d = 0; d < DIM; ++d) faces += double(A.g.nface[d]);
      double b = cells * (zi ? 0 : 1) + (pcl ? coarse : 0) + cells + (A.a ? cells : 0) + (A.c ? cells : 0) + faces;
      b += cells + (resc ? coarse : 0);
      if (fin) b += 2 * cells + 2 * cells;   // phi and the full rhs in; phi + e and the residual out
      *tm->bytes += b * sizeof(T);
      ++*tm->n;
    }
    *cur = 1 - *cur;
    done += p;
  }
  return hipSuccess;
}

// ---- streaming relaxation for large levels -------------------------------
// A workgroup owns a block of columns (2D: TX cells in x; 3D: TX x TY cells
// in (x, y)) and a segment of planes along the last direction, and marches
// through them with a window of P+2 planes in LDS.  With plane R just
// loaded, colour pass p runs on plane R-1-p (p = 0..P-1), so plane R-P is
// final and is written: the passes form a pipeline along the march instead
// of a halo in that direction.  Pass p covers the lateral positions within
// P-1-p of the block and the planes within P-1-p of the segment, so the
// core ends exact after P passes, as in the tiled kernel.  Every input value
// of the extended block is read once per launch; a plane's coefficients
// stay in registers for the P steps that use them and are loaded one plane
// ahead of their first use, so their loads overlap the previous plane's
// passes.  Stencil, lambda and update are the tiled kernel's, so results
// are identical bit for bit.
template <typename T, int DIM>
struct SCoef {
  Faces1<T, DIM> F;
  T aq, cq, rr, lam;
};

template <typename T, int DIM, int TX, int TY, int P, int NT>
__global__ __launch_bounds__(NT) void vc_relax_stream_kernel(VCArgs<T> A, const T* __restrict__ in,
                                                             T* __restrict__ out, const T* __restrict__ rhs,
                                                             int first, int zero_init, Geom gc,
                                                             const T* __restrict__ pc, int nbx, int nby, int seg) {
  if (A.skip && *A.skip) return;
  constexpr int EX = TX + 2 * P;
  constexpr int EL = EX * (DIM == 3 ? TY + 2 * P : 1);
  constexpr int K = (EL + NT - 1) / NT;
  static_assert(P >= 1 && P <= 2, "one or two colour passes per launch");
  constexpr int W = 4;   // planes in the window (P+2 <= 4; a power of two)
  constexpr int MD = DIM - 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* su = reinterpret_cast<T*>(smem_raw);
  const Geom& g = A.g;
  const int nz = g.n[MD];
  const bool zper = g.periodic[MD] != 0;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  const bool hasc = A.c != nullptr;
  int b = blockIdx.x;
  const int bx = b % nbx;
  b /= nbx;
  const int by = DIM == 3 ? b % nby : 0;
  const int bz = DIM == 3 ? b / nby : b;
  const int x0 = bx * TX, y0 = by * TY;
  const int z0 = bz * seg;
  const int z1 = z0 + seg < nz ? z0 + seg : nz;
  SPos S[K];
#pragma unroll
  for (int kk = 0; kk < K; ++kk) {
    const int e = threadIdx.x + kk * NT;
    if (e < EL) {
      spos_init<DIM, TX, TY, P>(g, x0, y0, e, S[kk]);
      S[kk].gtype = 0;
      for (int d = 0; d < DIM - 1; ++d)
        for (int sd = 0; sd < 2; ++sd)
          if (S[kk].kind == 2 && S[kk].gd == d && S[kk].gs == sd) S[kk].gtype = A.bc.type[d][sd][0];
    } else {
      S[kk].kind = 0;
      S[kk].dist = P;
      S[kk].owner = 0;
    }
  }
  auto zwrap = [&](int z) { return z < 0 ? z + nz : (z >= nz ? z - nz : z); };
  auto real_plane = [&](int z) { return zper || (z >= 0 && z < nz); };
  // 32-bit element offsets: a position's lateral part, computed once, plus a
  // plane part common to the workgroup (levels below 2^29 cells per array)
  const unsigned csz = unsigned(DIM == 3 ? g.sz : g.sy);   // cell stride between planes
  unsigned fpl[DIM], fup[DIM];                              // face plane stride, offset to the upper face
#pragma unroll
  for (int d = 0; d < DIM; ++d) fpl[d] = unsigned(g.fn[d][0] * (DIM == 3 ? g.fn[d][1] : 1));
  fup[0] = 1;
  if (DIM == 3) fup[1] = unsigned(g.fn[1][0]);
  fup[MD] = fpl[MD];
  const unsigned cgsz = unsigned(DIM == 3 ? gc.sz : gc.sy);
  unsigned cl[K], fl[K][DIM], ccl[K];
#pragma unroll
  for (int kk = 0; kk < K; ++kk) {
    const SPos& s = S[kk];
    cl[kk] = unsigned(s.gx + 1) + (DIM == 3 ? unsigned(g.sy) * unsigned(s.gy + 1) : 0u);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      fl[kk][d] = unsigned(s.gx) + (DIM == 3 ? unsigned(g.fn[d][0]) * unsigned(s.gy) : 0u);
    ccl[kk] = unsigned(s.gx / 2 + 1) + (DIM == 3 ? unsigned(gc.sy) * unsigned(s.gy / 2 + 1) : 0u);
  }
  auto ld = [](const T* p, unsigned i) -> T {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (i * unsigned(sizeof(T))));
  };
  // register sets of P+3 planes (R-P .. R+2 at step R), addressed by
  // compile-time indices: plane zl+m lives in set m % NS.  Loads go straight
  // into their set and are used only at that plane's commit two steps later,
  // so no instruction touches a load in flight.
  constexpr int NS = P + 3;
  SCoef<T, DIM> cs[NS][K];
  T sv[NS][K], scv[NS][K];   // state and coarse correction as loaded
  // the same loads for every position and plane (in-domain addresses; an
  // absent array is replaced by a duplicate read of rhs), so that the waits
  // for a set's loads can count the loads issued after it exactly
  const T* vin = zero_init ? rhs : in;
  const T* ain = A.a ? A.a : rhs;
  const T* cin = hasc ? A.c : rhs;
  const T* cvp = pc ? pc : rhs;
  auto prefetch = [&](int R, SCoef<T, DIM>(&pf)[K], T(&v)[K], T(&cv)[K]) {
    const int zw = zper ? zwrap(R) : (R < 0 ? 0 : (R >= nz ? nz - 1 : R));
    const unsigned qz = unsigned(zw + 1) * csz, cz = unsigned(zw / 2 + 1) * cgsz;
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      const unsigned q = cl[kk] + qz;   // a ghost reads its interior neighbour
      v[kk] = ld(vin, q);
      cv[kk] = ld(cvp, pc ? ccl[kk] + cz : q);
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const unsigned f = fl[kk][d] + unsigned(zw) * fpl[d];
        pf[kk].F.v[2 * d] = ld(A.b[d], f);
        pf[kk].F.v[2 * d + 1] = ld(A.b[d], f + fup[d]);
      }
      pf[kk].aq = ld(ain, q);
      pf[kk].cq = ld(cin, q);
      pf[kk].rr = ld(rhs, q);
    }
  };
  const int zl = (!zper && z0 - P < 0) ? 0 : z0 - P;
  const int zend = z1 - 1 + P;   // last step: plane z1 - 1 is final
  // window slot of plane R (the bottom wall's ghost plane -1 is slot 3)
  auto slot = [&](int R) { return su + ((R - zl) & (W - 1)) * EL; };
  auto step = [&](auto J, int R) {
    constexpr int j = decltype(J)::value;
    // A. commit plane R (or the top wall's ghost plane)
    T* pl = slot(R);
    if (real_plane(R)) {
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const SPos& s = S[kk];
        const int e = threadIdx.x + kk * NT;
        if (s.kind == 0) continue;
        T val = zero_init ? T(0) : sv[j][kk];
        if (pc) val = val + scv[j][kk];
        if (s.kind == 1) {
          pl[e] = val;
          if (!zper && R == 0) slot(-1)[e] = bc_ghost<T>(A.bc.type[MD][0][0], T(0), val, -1, T(g.dx));
          if (!A.a) cs[j][kk].aq = T(1);
          if (!hasc) cs[j][kk].cq = T(0);
          if (s.dist <= P - 1)
            cs[j][kk].lam = vc_lambda<T, DIM>(cs[j][kk].F, cs[j][kk].aq, hasc, cs[j][kk].cq, A.alpha, A.beta, dxinv);
        } else {
          pl[e] = bc_ghost<T>(s.gtype, T(0), val, s.gs == 0 ? -1 : 1, T(g.dx));
        }
      }
    } else if (!zper && R == nz) {
      const T* below = slot(nz - 1);
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const int e = threadIdx.x + kk * NT;
        if (S[kk].kind == 1) pl[e] = bc_ghost<T>(A.bc.type[MD][1][0], T(0), below[e], 1, T(g.dx));
      }
    }
    // B. loads of plane R+2 into the set of plane R+2-NS, no longer in use
    constexpr int jn = (j + 2) % NS;
    prefetch(R + 2, cs[jn], sv[jn], scv[jn]);   // beyond the segment: a clamped plane, unused
    __syncthreads();
    // C. the colour passes; pass p on plane R-1-p uses set (j-1-p) mod NS
    auto pass = [&](auto PP) {
      constexpr int p = decltype(PP)::value;
      constexpr int jp = (j - 1 - p + 2 * NS) % NS;
      const int z = R - 1 - p;
      const int reach = P - 1 - p;
      const bool run = z >= z0 - reach && z < z1 + reach && real_plane(z);
      if (run) {
        const int zw = zwrap(z);
        const int colour = (first + p) & 1;
        T* pz = slot(z);
        const T* pm = slot(z - 1);
        const T* pp = slot(z + 1);
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
          const SPos& s = S[kk];
          if (s.kind != 1 || s.dist > reach || ((s.gx + s.gy + zw + colour) & 1)) continue;
          const int e = threadIdx.x + kk * NT;
          T st[7];
          st[0] = pz[e];
          st[1] = pz[e + 1];
          st[2] = pz[e - 1];
          if (DIM == 3) {
            st[3] = pz[e + EX];
            st[4] = pz[e - EX];
            st[5] = pp[e];
            st[6] = pm[e];
          } else {
            st[3] = pp[e];
            st[4] = pm[e];
          }
          const SCoef<T, DIM>& C = cs[jp][kk];
          const T lof = vc_lof<T, DIM>(C.F, st, C.aq, hasc, C.cq, A.alpha, A.beta, dxinv);
          const T nv = st[0] - C.lam * (lof - C.rr);
          pz[e] = nv;
          // wall ghosts of this cell: lateral, then along the march
          const int gl[2] = {s.gx, s.gy};
#pragma unroll
          for (int d = 0; d < (DIM == 3 ? 2 : 1); ++d) {
            if (g.periodic[d]) continue;
            const int st1 = d == 0 ? 1 : EX;
            if (gl[d] == 0) pz[e - st1] = bc_ghost<T>(A.bc.type[d][0][0], T(0), nv, -1, T(g.dx));
            if (gl[d] == g.n[d] - 1) pz[e + st1] = bc_ghost<T>(A.bc.type[d][1][0], T(0), nv, 1, T(g.dx));
          }
          if (!zper) {
            if (zw == 0) slot(-1)[e] = bc_ghost<T>(A.bc.type[MD][0][0], T(0), nv, -1, T(g.dx));
            if (zw == nz - 1) slot(nz)[e] = bc_ghost<T>(A.bc.type[MD][1][0], T(0), nv, 1, T(g.dx));
          }
        }
      }
      __syncthreads();
    };
    pass(IC<0>{});
    if constexpr (P > 1) pass(IC<1>{});
    // D. plane R-P is final
    const int zs = R - P;
    if (zs >= z0 && zs < z1) {
      const T* ps = slot(zs);
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const SPos& s = S[kk];
        if (s.owner)
          *reinterpret_cast<T*>(reinterpret_cast<char*>(out) + ((cl[kk] + unsigned(zs + 1) * csz) * unsigned(sizeof(T)))) =
              ps[threadIdx.x + kk * NT];
      }
    }
  };
  prefetch(zl, cs[0], sv[0], scv[0]);
  prefetch(zl + 1, cs[1], sv[1], scv[1]);
  for (int R = zl; R <= zend; R += NS) {
    step(IC<0>{}, R);
    if (R + 1 <= zend) step(IC<1>{}, R + 1);
    if (R + 2 <= zend) step(IC<2>{}, R + 2);
    if constexpr (NS > 3) {
      if (R + 3 <= zend) step(IC<3 % NS>{}, R + 3);
    }
    if constexpr (NS > 4) {
      if (R + 4 <= zend) step(IC<4 % NS>{}, R + 4);
    }
  }
}

template <int DIM>
struct VCStream;
template <>
struct VCStream<2> {
  static constexpr int NT = 256, TY = 1, SEGMIN = 8;
};
template <>
struct VCStream<3> {
  static constexpr int NT = 256, TY = 12, SEGMIN = 8;
};

// levels on which plain relaxation launches stream (2D above 512^2, 3D
// above 96^3); the tiled kernel keeps the fused restriction / residual
constexpr long long kStreamCells2 = 512 * 512;
constexpr long long kStreamCells3 = 96LL * 96 * 96;
inline bool vc_streams(const Geom& g) {
  long long nf = 0;
  for (int d = 0; d < g.dim; ++d) nf = nf > g.nface[d] ? nf : g.nface[d];
  const long long most = g.ncell > nf ? g.ncell : nf;
  return g.nvalid() > (g.dim == 2 ? kStreamCells2 : kStreamCells3) && g.n[g.dim - 1] >= 8 &&
         most < (1LL << 29);   // byte offsets of every array within 32 bits
}

template <typename T, int DIM, int P>
hipError_t vc_stream_launch(const VCArgs<T>& A, const T* in, T* out, const T* rhs, int first, bool zero_init,
                            const Geom& gc, const T* pc, hipStream_t s, const LaunchTimer* tm) {
  using D = VCStream<DIM>;
  constexpr int TX = DIM == 2 ? D::NT - 2 * P : D::TY;
  constexpr int TY = DIM == 2 ? 1 : D::TY;
  constexpr int EL = (TX + 2 * P) * (DIM == 3 ? TY + 2 * P : 1);
  const Geom& g = A.g;
  const int nbx = (g.n[0] + TX - 1) / TX;
  const int nby = DIM == 3 ? (g.n[1] + TY - 1) / TY : 1;
  const size_t lds = size_t(4) * EL * sizeof(T);
  auto kern = vc_relax_stream_kernel<T, DIM, TX, TY, P, D::NT>;
  // segments sized so that the whole grid is resident at once (one round of
  // workgroups; measured best at 2048^2: a grid that fills every CU's
  // workgroup slots beats both fewer, longer segments and a second round)
  static int resident = 0;
  if (!resident) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ==
                                                 hipSuccess &&
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void*>(kern), D::NT, lds) ==
            hipSuccess &&
        cus > 0 && per > 0)
      resident = cus * per;
    else
      resident = 768;
  }
  const int nz = g.n[DIM - 1];
  const int columns = nbx * nby;
  const int per_column = resident / columns > 1 ? resident / columns : 1;
  int seg = (nz + per_column - 1) / per_column;
  if (seg < D::SEGMIN) seg = D::SEGMIN;
  const int nbz = (nz + seg - 1) / seg;
  const bool timed = tm && *tm->n < tm->cap;
  if (tm && !timed) ++*tm->missed;
  hipExtLaunchKernelGGL(kern, dim3(nbx * nby * nbz), dim3(D::NT), lds, s, timed ? tm->start[*tm->n] : nullptr,
                        timed ? tm->stop[*tm->n] : nullptr, 0, A, in, out, rhs, first, zero_init ? 1 : 0, gc, pc, nbx,
                        nby, seg);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (timed) {
    const double cells = double(g.nvalid());
    double faces = 0;
    for (int d = 0; d < DIM; ++d) faces += double(g.nface[d]);
    double by = cells * (zero_init ? 0 : 1) + (pc ? cells / double(DIM == 3 ? 8 : 4) : 0) + cells +
                (A.a ? cells : 0) + (A.c ? cells : 0) + faces + cells;
    *tm->bytes += by * sizeof(T);
    ++*tm->n;
  }
  return hipSuccess;
}

// npass colour passes by streaming launches of up to 2 passes each
template <typename T, int DIM>
hipError_t vc_relax_stream(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                           const Geom* gc, const T* pc, hipStream_t s, const LaunchTimer* tm) {
  const Geom gcc = gc ? *gc : A.g;
  int done = 0;
  while (done < npass) {
    const int p = (npass - done) >= 2 ? 2 : (npass - done);
    const bool zi = done == 0 && zero_init;
    const T* pcl = done == 0 ? pc : nullptr;
    hipError_t e = p == 2 ? vc_stream_launch<T, DIM, 2>(A, buf[*cur], buf[1 - *cur], rhs, done & 1, zi, gcc, pcl, s, tm)
                          : vc_stream_launch<T, DIM, 1>(A, buf[*cur], buf[1 - *cur], rhs, done & 1, zi, gcc, pcl, s,
                                                        tm);
    if (e != hipSuccess) return e;
    *cur = 1 - *cur;
    done += p;
  }
  return hipSuccess;
}

// tiles of the 2D relaxation on large levels (short halo, wide rows;
// measured best of 6 shapes at 2048^2)
struct VCRelaxWide2 {
  static constexpr int TX = 64, TY = 8, TZ = 1, HMAX = 2, NT = 256;
};
constexpr long long kWideCells = 512 * 512;   // levels above this use the wide tiles

inline bool vc_wide(const Geom& g) { return g.dim == 2 && g.nvalid() > kWideCells; }

// passes one relaxation launch can do on level g
inline int vc_hmax(const Geom& g) {
  return vc_wide(g) ? VCRelaxWide2::HMAX : (g.dim == 2 ? VCRelaxTile<2>::HMAX : VCRelaxTile<3>::HMAX);
}

template <typename T, int DIM>
hipError_t vc_relax_dim(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                        const Geom* gc, const T* pc, hipStream_t s, const Geom* gr = nullptr, T* resc = nullptr,
                        const LaunchTimer* tm = nullptr, const VCFinal<T>* fin = nullptr) {
  if (!resc && !fin && npass > 0 && vc_streams(A.g))
    return vc_relax_stream<T, DIM>(A, buf, cur, rhs, npass, zero_init, gc, pc, s, tm);
  if constexpr (DIM == 2) {
    if (vc_wide(A.g))
      return vc_relax_cfg<T, 2, VCRelaxWide2>(A, buf, cur, rhs, npass, zero_init, gc, pc, s, gr, resc, tm, fin);
  }
  return vc_relax_cfg<T, DIM, VCRelaxTile<DIM>>(A, buf, cur, rhs, npass, zero_init, gc, pc, s, gr, resc, tm, fin);
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
                    const Geom* gc, const T* pc, hipStream_t s, const LaunchTimer* tm) {
  return A.g.dim == 2 ? vc_relax_dim<T, 2>(A, buf, cur, rhs, npass, zero_init, gc, pc, s, nullptr, nullptr, tm)
                      : vc_relax_dim<T, 3>(A, buf, cur, rhs, npass, zero_init, gc, pc, s, nullptr, nullptr, tm);
}

template <typename T>
hipError_t vc_relax_restrict(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                             const Geom& gr, T* resc, hipStream_t s, const LaunchTimer* tm) {
  const int hmax = vc_hmax(A.g);
  if (npass + 1 <= hmax)
    return A.g.dim == 2 ? vc_relax_dim<T, 2>(A, buf, cur, rhs, npass, zero_init, nullptr, nullptr, s, &gr, resc, tm)
                        : vc_relax_dim<T, 3>(A, buf, cur, rhs, npass, zero_init, nullptr, nullptr, s, &gr, resc, tm);
  hipError_t e = vc_relax<T>(A, buf, cur, rhs, npass, zero_init, nullptr, nullptr, s, tm);
  if (e != hipSuccess) return e;
  return vc_restrict_residual<T>(A, gr, buf[*cur], rhs, resc, s);
}

template <typename T>
hipError_t vc_relax_final(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, const Geom* gc,
                          const T* pc, const VCFinal<T>& fin, bool* fused, hipStream_t s, const LaunchTimer* tm) {
  const int hmax = vc_hmax(A.g);
  *fused = npass >= 1 && npass + 1 <= hmax && fin.phi && fin.phi_out != fin.phi;
  if (!*fused) return hipSuccess;
  return A.g.dim == 2 ? vc_relax_dim<T, 2>(A, buf, cur, rhs, npass, false, gc, pc, s, nullptr, nullptr, tm, &fin)
                      : vc_relax_dim<T, 3>(A, buf, cur, rhs, npass, false, gc, pc, s, nullptr, nullptr, tm, &fin);
}

template <typename T>
hipError_t vc_residual(const VCArgs<T>& A, const T* phi, const T* e, T* phi_out, const T* rhs, T* res,
                       unsigned long long* norm, int homog, hipStream_t s, const CtlFuse& fuse) {
  if (e && (!phi_out || phi_out == phi)) return hipErrorInvalidValue;
  VCTileIO<T> io{};
  io.phi = phi;
  io.e = e;
  io.phi_out = phi_out;
  io.rhs = rhs;
  io.out = res;
  io.norm = norm;
  io.fuse = fuse;
  io.homog = homog;
  return A.g.dim == 2 ? vc_tile_dim<T, 2, kVCResid>(A, io, s) : vc_tile_dim<T, 3, kVCResid>(A, io, s);
}

int vc_resid_tiles(const Geom& g, int* ntx, int* nty) {
  return g.dim == 2 ? vc_tiles<VCTile<2>::TX, VCTile<2>::TY, VCTile<2>::TZ>(g, *ntx, *nty)
                    : vc_tiles<VCTile<3>::TX, VCTile<3>::TY, VCTile<3>::TZ>(g, *ntx, *nty);
}

template <typename T>
hipError_t vc_resid_batch(int dim, const VCResidJob<T>* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  if (dim == 2) {
    using D = VCTile<2>;
    const size_t lds = size_t(D::TX + 2) * (D::TY + 2) * sizeof(T);
    hipLaunchKernelGGL((vc_resid_batch_kernel<T, 2, D::TX, D::TY, D::TZ>), dim3(nblocks), dim3(kVCThreads), lds, s,
                       jobs, map);
  } else {
    using D = VCTile<3>;
    const size_t lds = size_t(D::TX + 2) * (D::TY + 2) * (D::TZ + 2) * sizeof(T);
    hipLaunchKernelGGL((vc_resid_batch_kernel<T, 3, D::TX, D::TY, D::TZ>), dim3(nblocks), dim3(kVCThreads), lds, s,
                       jobs, map);
  }
  return hipGetLastError();
}

hipError_t words_batch(const WordsJob* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL(words_batch_kernel, dim3(nblocks), dim3(256), 0, s, jobs, map);
  return hipGetLastError();
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

// phi = rhs * lambda on the valid cells (the first step of preCond)
template <typename T, int DIM>
__global__ __launch_bounds__(256) void vc_lambda_mult_kernel(VCArgs<T> A, const T* __restrict__ rhs,
                                                             T* __restrict__ phi) {
  const Geom& g = A.g;
  const long long n = g.nvalid();
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const int i = int(c % g.n[0]);
    const long long r = c / g.n[0];
    const int j = int(r % g.n[1]);
    const int k = int(r / g.n[1]);
    Faces1<T, DIM> F;
    load_faces1<T, DIM>(A, i, j, k, F);
    const long long q = g.ci(i, j, k);
    const T aq = A.a ? A.a[q] : T(1);
    const T cv = A.c ? A.c[q] : T(0);
    phi[q] = rhs[q] * vc_lambda<T, DIM>(F, aq, A.c != nullptr, cv, A.alpha, A.beta, dxinv);
  }
}

// (x, y) over the valid cells in the order of the coarse V-cycle's BiCGStab:
// one wavefront, 64 interleaved lane sums over the compact cell index, then
// the pairwise tree s_t + s_{t+h}, h = P/2 .. 1 (P the smallest power of two
// >= min(cells, 64))
template <typename T>
__global__ __launch_bounds__(64) void vc_dot_kernel(Geom g, const T* __restrict__ x, const T* __restrict__ y,
                                                    T* __restrict__ out) {
  const long long n = g.nvalid();
  T s = T(0);
  for (long long c = threadIdx.x; c < n; c += 64) {
    const int i = int(c % g.n[0]);
    const long long r = c / g.n[0];
    const long long q = g.ci(i, int(r % g.n[1]), int(r / g.n[1]));
    s = s + x[q] * y[q];
  }
  int P2 = 1;
  while (P2 < 64 && P2 < n) P2 <<= 1;
  for (int h = P2 >> 1; h >= 1; h >>= 1) s = s + __shfl_down(s, h, 64);
  if (threadIdx.x == 0) *out = s;
}

template <typename T>
hipError_t vc_lambda_mult(const VCArgs<T>& A, const T* rhs, T* phi, hipStream_t s) {
  const long long n = A.g.nvalid();
  const int blocks = int(std::min<long long>((n + 255) / 256, 4096));
  if (A.g.dim == 2)
    hipLaunchKernelGGL((vc_lambda_mult_kernel<T, 2>), dim3(blocks), dim3(256), 0, s, A, rhs, phi);
  else
    hipLaunchKernelGGL((vc_lambda_mult_kernel<T, 3>), dim3(blocks), dim3(256), 0, s, A, rhs, phi);
  return hipGetLastError();
}

template <typename T>
hipError_t vc_dot(const Geom& g, const T* x, const T* y, T* out, hipStream_t s) {
  hipLaunchKernelGGL((vc_dot_kernel<T>), dim3(1), dim3(64), 0, s, g, x, y, out);
  return hipGetLastError();
}

// ---- batched form --------------------------------------------------------
bool vc_batch_level_ok(const Geom& g) { return !vc_wide(g); }

int vc_batch_hmax(int dim) { return dim == 2 ? VCRelaxTile<2>::HMAX : VCRelaxTile<3>::HMAX; }

int vc_batch_tiles(const Geom& g, int shape, int* ntx, int* nty) {
  if (g.dim == 3) return vc_tiles<VCRelaxTile<3>::TX, VCRelaxTile<3>::TY, VCRelaxTile<3>::TZ>(g, *ntx, *nty);
  return shape == 1 ? vc_tiles<VCRelaxBatchWide2::TX, VCRelaxBatchWide2::TY, 1>(g, *ntx, *nty)
                    : vc_tiles<VCRelaxTile<2>::TX, VCRelaxTile<2>::TY, VCRelaxTile<2>::TZ>(g, *ntx, *nty);
}

template <class D, int DIM>
static size_t vc_batch_lds_dim(int npass, bool resc, bool fin) {
  const int h = npass + ((resc || fin) ? 1 : 0);
  size_t lds = size_t(D::TX + 2 * h) * (D::TY + 2 * h) * (DIM == 3 ? D::TZ + 2 * h : 1);
  if (resc) lds += size_t(D::TX) * D::TY * D::TZ;
  if (fin) lds *= 2;
  return lds;
}

size_t vc_batch_lds(int dim, int shape, int npass, bool resc, bool fin) {
  const size_t e = dim == 3 ? vc_batch_lds_dim<VCRelaxTile<3>, 3>(npass, resc, fin)
                            : (shape == 1 ? vc_batch_lds_dim<VCRelaxBatchWide2, 2>(npass, resc, fin)
                                          : vc_batch_lds_dim<VCRelaxTile<2>, 2>(npass, resc, fin));
  return e * sizeof(double);
}

template <typename T>
hipError_t vc_relax_batch(int dim, int shape, const VCRelaxJob<T>* jobs, const unsigned* map, int nblocks, size_t lds,
                          hipStream_t s, const LaunchTimer* tm) {
  if (nblocks <= 0) return hipSuccess;
  const bool timed = tm && *tm->n < tm->cap;
  if (tm && !timed) ++*tm->missed;
  hipEvent_t e0 = timed ? tm->start[*tm->n] : nullptr, e1 = timed ? tm->stop[*tm->n] : nullptr;
  if (dim == 2 && shape == 1) {
    using D = VCRelaxBatchWide2;
    hipExtLaunchKernelGGL((vc_relax_batch_kernel<T, 2, D::TX, D::TY, D::TZ, D::HMAX, D::NT>), dim3(nblocks),
                          dim3(D::NT), lds, s, e0, e1, 0, jobs, map);
  } else if (dim == 2) {
    using D = VCRelaxTile<2>;
    hipExtLaunchKernelGGL((vc_relax_batch_kernel<T, 2, D::TX, D::TY, D::TZ, D::HMAX, D::NT>), dim3(nblocks),
                          dim3(D::NT), lds, s, e0, e1, 0, jobs, map);
  } else {
    using D = VCRelaxTile<3>;
    hipExtLaunchKernelGGL((vc_relax_batch_kernel<T, 3, D::TX, D::TY, D::TZ, D::HMAX, D::NT>), dim3(nblocks),
                          dim3(D::NT), lds, s, e0, e1, 0, jobs, map);
  }
  if (timed) ++*tm->n;
  return hipGetLastError();
}

__global__ void ctl_gather_kernel(SolveCtl* const* __restrict__ ctls, SolveCtl* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = *ctls[i];
}

hipError_t ctl_gather(SolveCtl* const* ctls, SolveCtl* out, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(ctl_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, s, ctls, out, n);
  return hipGetLastError();
}

template <typename T>
hipError_t vc_coarse_batch(int dim, const VCDesc<T>* const* descs, int n, size_t lds, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (dim == 2)
    hipLaunchKernelGGL((vc_coarse_batch_kernel<T, 2>), dim3(n), dim3(kVCWG), lds, s, descs);
  else
    hipLaunchKernelGGL((vc_coarse_batch_kernel<T, 3>), dim3(n), dim3(kVCWG), lds, s, descs);
  return hipGetLastError();
}

template <typename T>
size_t vc_coarse_lds_bytes(const Geom* g, int l0, int nlev, bool bicgstab) {
  long long tot = 0;
  for (int l = l0; l < nlev; ++l) tot += g[l].nvalid();
  // BiCGStab: 8 vectors of the coarsest level after the restriction scratch
  const size_t extra = bicgstab ? size_t(8 * g[nlev - 1].nvalid()) : 0;
  return (vc_coarse_lds_elems<T>(g[l0].dim, tot) + extra) * sizeof(T);
}

template <typename T>
hipError_t vc_coarse_vcycle(const VCDesc<T>* dev_desc, int dim, size_t lds, hipStream_t s) {
  if (dim == 2)
    hipLaunchKernelGGL((vc_coarse_vcycle_kernel<T, 2>), dim3(1), dim3(kVCWG), lds, s, dev_desc);
  else
    hipLaunchKernelGGL((vc_coarse_vcycle_kernel<T, 3>), dim3(1), dim3(kVCWG), lds, s, dev_desc);
  return hipGetLastError();
}

bool vc_small_level(const Geom& g) { return g.nvalid() <= kVCWGCells; }
int vc_coarse_levels_max(int dim) { return dim == 3 ? VCLevels<3>::n : VCLevels<2>::n; }

template hipError_t vc_relax<double>(const VCArgs<double>&, double* const[2], int*, const double*, int, bool,
                                     const Geom*, const double*, hipStream_t, const LaunchTimer*);
template hipError_t vc_relax_restrict<double>(const VCArgs<double>&, double* const[2], int*, const double*, int, bool,
                                              const Geom&, double*, hipStream_t, const LaunchTimer*);
template hipError_t vc_relax_final<double>(const VCArgs<double>&, double* const[2], int*, const double*, int,
                                           const Geom*, const double*, const VCFinal<double>&, bool*, hipStream_t,
                                           const LaunchTimer*);
template hipError_t vc_residual<double>(const VCArgs<double>&, const double*, const double*, double*,
                                        const double*, double*, unsigned long long*, int, hipStream_t,
                                        const CtlFuse&);
template hipError_t vc_apply_op<double>(const VCArgs<double>&, const double*, double*, int, hipStream_t);
template hipError_t vc_lambda_mult<double>(const VCArgs<double>&, const double*, double*, hipStream_t);
template hipError_t vc_dot<double>(const Geom&, const double*, const double*, double*, hipStream_t);
template hipError_t vc_restrict_residual<double>(const VCArgs<double>&, const Geom&, const double*, const double*,
                                                 double*, hipStream_t);
template hipError_t vc_coarse_vcycle<double>(const VCDesc<double>*, int, size_t, hipStream_t);
template size_t vc_coarse_lds_bytes<double>(const Geom*, int, int, bool);
template hipError_t vc_relax_batch<double>(int, int, const VCRelaxJob<double>*, const unsigned*, int, size_t,
                                           hipStream_t, const LaunchTimer*);
template hipError_t vc_coarse_batch<double>(int, const VCDesc<double>* const*, int, size_t, hipStream_t);
template hipError_t vc_resid_batch<double>(int, const VCResidJob<double>*, const unsigned*, int, hipStream_t);

}  // namespace ml
