// ml_hc_kernels.hip — HIP kernels for gfx950 of the nonlinear (H, Theta)
// operator of AMRNonLinearMultiCompOp: fused closure + red/black smoothing,
// residual with max-norm, FAS restriction, operator application, and a
// single-workgroup kernel that runs the coarse part of a FAS V-cycle.
// Every kernel recomputes the phase-diagram closure from (H, Theta) staged
// in LDS, so theta/Theta_l are never written to HBM.
//
// Arithmetic follows the reference's evaluation order with contraction off
// (-ffp-contract=off), so the FP64 instantiation reproduces the CPU oracle
// bit for bit.
#include <hip/hip_ext.h>

#include "ml_common.h"
#include "ml_kernels.h"
#include "ml_tile.h"

namespace ml {

namespace {

constexpr int kThreads = 256;
constexpr int kWGThreads = 1024;   // single-workgroup kernel: one thread per cell ...
constexpr int kWGCells = 1024;     // ... of the largest level it holds

enum : int { kResid = 1, kApply = 2, kRestrict = 3 };

// Closure of every in-domain position within ext, then the derived-variable
// BC at wall ghosts within ext (computeDiffusedVar,
// src/AMRNonLinearMultiCompOp.cpp:57-98).
template <typename T, int DIM, int TX, int TY, int TZ>
__device__ inline void tile_diffused_var(const HCArgs<T>& A, const Tile<DIM, TX, TY, TZ>& t,
                                         const T* sphi, T* sphl, int ext, int homog) {
  const Geom& g = A.g;
  const int E = t.E;
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int ex, ey, ez;
    t.coords(e, ex, ey, ez);
    if (!t.within(ex, ey, ez, ext)) continue;
    const Pos P = t.pos(g, ex, ey, ez);
    if (P.kind != 1) continue;
    T th, cl;
    closure<T>(sphi[e], sphi[E + e], A.cp, th, cl);
    sphl[e] = th;
    sphl[E + e] = cl;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int ex, ey, ez;
    t.coords(e, ex, ey, ez);
    if (!t.within(ex, ey, ez, ext)) continue;
    const Pos P = t.pos(g, ex, ey, ez);
    if (P.kind != 2) continue;
    const int stride = P.gd == 0 ? 1 : (P.gd == 1 ? t.EX : t.EXY);
    const int eint = P.gs == 0 ? e + stride : e - stride;
    const int isign = P.gs == 0 ? -1 : 1;
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const T x1 = sphl[n * E + eint];
      T gv = x1;
      if (A.bc.apply) {
        const T v = homog ? T(0) : T(A.bc.val[P.gd][P.gs][n]);
        gv = bc_ghost<T>(A.bc.type[P.gd][P.gs][n], v, x1, isign, T(g.dx));
      }
      sphl[n * E + e] = gv;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Red/black relaxation of one tile: `npass` colour passes with halo npass
// staged in LDS (levelGSRB, src/AMRNonLinearMultiCompOp.cpp:815-949, and
// NONLINEARSMOOTHING, src/AMRNonLinearMultiCompOpF.ChF:125-195: phi_l of
// every cell as of the start of the pass, frozen within it).  Each thread
// owns fixed LDS positions whose classification is computed once; a wall
// ghost's derived value comes from its interior neighbour's state.  When
// io.pc is set, the staged state first receives the linear prolongation of
// the coarse correction pc - pc0 (prolongIncrement fused with the
// post-smoother).  Input and output arrays are distinct.
template <typename T, int DIM, int TX, int TY, int TZ, int HMAX, int NT>
__global__ __launch_bounds__(NT) void hc_relax_tile_kernel(HCArgs<T> A, TileIO<T> io) {
  if (A.skip && *A.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2 * HMAX) * (TY + 2 * HMAX) * (DIM == 3 ? TZ + 2 * HMAX : 1);
  constexpr int KMAX = (EMAX + NT - 1) / NT;
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init(io.npass, io.ntx, io.nty);
  const int E = t.E;
  T* sphi = reinterpret_cast<T*>(smem_raw);
  T* sphl = sphi + 2 * E;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  const T scale = dxinv;
  const Geom gc = io.gc;

  PosRec R[KMAX];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    pos_record<DIM, TX, TY, TZ>(g, t, threadIdx.x + kk * NT, R[kk]);
  }

  // stage (H, Theta), with the prolonged coarse correction when requested
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1) continue;
    T h = io.phi[r.q], c = io.phi[r.q + g.ncell];
    if (io.pc) {
      h = prolong_value<T, DIM>(gc, r.i, r.j, r.k, io.pc, io.pc0, h);
      c = prolong_value<T, DIM>(gc, r.i, r.j, r.k, io.pc + gc.ncell, io.pc0 ? io.pc0 + gc.ncell : nullptr, c);
    }
    sphi[r.e] = h;
    sphi[E + r.e] = c;
  }
  __syncthreads();

  // closure of every staged cell and wall ghost (computeDiffusedVar,
  // src/AMRNonLinearMultiCompOp.cpp:57-98); ghosts from their interior
  // neighbour
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind == 0) continue;
    const int es = r.kind == 1 ? r.e : r.eint;
    T th, cl;
    closure<T>(sphi[es], sphi[E + es], A.cp, th, cl);
    if (r.kind == 2 && A.bc.apply) {
      const int isign = r.gs == 0 ? -1 : 1;
      th = bc_ghost<T>(A.bc.type[r.gd][r.gs][0], T(A.bc.val[r.gd][r.gs][0]), th, isign, T(g.dx));
      cl = bc_ghost<T>(A.bc.type[r.gd][r.gs][1], T(A.bc.val[r.gd][r.gs][1]), cl, isign, T(g.dx));
    }
    sphl[r.e] = th;
    sphl[E + r.e] = cl;
  }
  __syncthreads();

  // Each pass updates one colour and recomputes the closure of exactly the
  // cells it updated (and of their wall ghosts): the other colour's phi_l is
  // unchanged, and in this pass a cell reads only its own and the other
  // colour's phi_l, so one barrier per pass suffices.
  // operator data of every position this thread may update, in registers
  CellFaces<T, DIM> F[KMAX];
  T rr[KMAX][2], aa[KMAX][2], lam[KMAX][2];
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (r.kind != 1 || r.dist > t.H - 1) continue;
    load_faces<T, DIM>(A, r.i, r.j, r.k, F[kk]);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      rr[kk][n] = io.rhs[r.q + n * g.ncell];
      aa[kk][n] = A.a ? A.a[r.q + n * g.ncell] : T(1);
      lam[kk][n] = lambda_of<T, DIM>(F[kk], n, aa[kk][n], A.alpha, A.beta, scale);
    }
  }
  for (int p = 0; p < io.npass; ++p) {
    const int gext = t.H - p;
    const int colour = (io.first + p) & 1;
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      const PosRec& r = R[kk];
      if (!(r.kind == 1 && r.dist <= gext - 1 && !((r.i + r.j + r.k + colour) & 1))) continue;
      const int e = r.e;
      T nv[2];
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        T st[7];
        gather_lds<T, DIM>(sphl + n * E, e, t.EX, t.EXY, st);
        const T ph = sphi[n * E + e];
        const T lof = lof_of<T, DIM>(F[kk], n, st, aa[kk][n], ph, A.alpha, A.beta, dxinv);
        nv[n] = ph - lam[kk][n] * (lof - rr[kk][n]);
      }
      sphi[e] = nv[0];
      sphi[E + e] = nv[1];
      // after the last pass only the state is written out
      if (p == io.npass - 1) continue;
      T th, cl;
      closure<T>(nv[0], nv[1], A.cp, th, cl);
      sphl[e] = th;
      sphl[E + e] = cl;
      // wall ghosts of this cell (it is their interior neighbour)
      const int iv[3] = {r.i, r.j, r.k};
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        if (g.periodic[d]) continue;
        const int stride = d == 0 ? 1 : (d == 1 ? t.EX : t.EXY);
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
          if (iv[d] != (sd ? g.n[d] - 1 : 0)) continue;
          const int ge = sd ? e + stride : e - stride;
          T gth = th, gcl = cl;
          if (A.bc.apply) {
            const int isign = sd ? 1 : -1;
            gth = bc_ghost<T>(A.bc.type[d][sd][0], T(A.bc.val[d][sd][0]), th, isign, T(g.dx));
            gcl = bc_ghost<T>(A.bc.type[d][sd][1], T(A.bc.val[d][sd][1]), cl, isign, T(g.dx));
          }
          sphl[ge] = gth;
          sphl[E + ge] = gcl;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int kk = 0; kk < KMAX; ++kk) {
    const PosRec& r = R[kk];
    if (!r.owner) continue;
    io.phi_out[r.q] = sphi[r.e];
    io.phi_out[r.q + g.ncell] = sphi[E + r.e];
  }
}

// Residual (with max-norm), operator application, and FAS restriction of
// one tile with a one-cell halo.
template <typename T, int DIM, int TX, int TY, int TZ, int MODE>
__device__ __forceinline__ void hc_tile_body(const HCArgs<T>& A, const TileIO<T>& io, int tile) {
  if (A.skip && *A.skip) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init_at(1, io.ntx, io.nty, tile);
  const int E = t.E;
  T* sphi = reinterpret_cast<T*>(smem_raw);
  T* sphl = sphi + 2 * E;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));

  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int ex, ey, ez;
    t.coords(e, ex, ey, ez);
    const Pos P = t.pos(g, ex, ey, ez);
    if (P.kind == 1) {
      const long long q = g.ci(P.i, P.j, P.k);
      sphi[e] = io.phi[q];
      sphi[E + e] = io.phi[q + g.ncell];
    }
  }
  __syncthreads();
  tile_diffused_var<T, DIM, TX, TY, TZ>(A, t, sphi, sphl, 1, io.homog);

  if (MODE == kResid || MODE == kApply) {
    T mx = T(0);
    int nan = 0;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      int ex, ey, ez;
      t.coords(e, ex, ey, ez);
      if (!t.within(ex, ey, ez, 0)) continue;
      const int gi = t.x0 - t.H + ex, gj = t.y0 - t.H + ey, gk = DIM == 3 ? t.z0 - t.H + ez : 0;
      if (gi >= g.n[0] || gj >= g.n[1] || (DIM == 3 && gk >= g.n[2])) continue;
      const long long q0 = g.ci(gi, gj, gk);
      CellFaces<T, DIM> F;
      load_faces<T, DIM>(A, gi, gj, gk, F);
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const long long q = q0 + n * g.ncell;
        T st[7];
        gather_lds<T, DIM>(sphl + n * E, e, t.EX, t.EXY, st);
        const T aq = A.a ? A.a[q] : T(1);
        const T lof = lof_of<T, DIM>(F, n, st, aq, sphi[n * E + e], A.alpha, A.beta, dxinv);
        if (MODE == kResid) {
          const T r = io.rhs[q] - lof;
          if (io.out) io.out[q] = r;
          const T ar = fabs(r);
          if (ar != ar) nan = 1;
          else if (ar > mx) mx = ar;
        } else {
          io.out[q] = io.add ? io.add[q] + lof : lof;
        }
      }
    }
    if (MODE == kResid && io.norm) block_max_atomic<T>(mx, nan, io.norm, io.fuse);
    return;
  }

  if (MODE == kRestrict) {
    // RESTRICTRESNLVC (src/AMRNonLinearMultiCompOpF.ChF:452-542):
    // res_c(i/2) = 0 + sum over children in Fortran order of (rhs - L phi)/2^d;
    // with io.phic also RESTRICT (:555-585) of the state into phic/phic0.
    T* sres = sphl + 2 * E;
    constexpr int NC = TX * TY * (DIM == 3 ? TZ : 1);
    const T denom = DIM == 3 ? T(8) : T(4);
    const Geom gc = io.gc;
    for (int c = threadIdx.x; c < NC; c += blockDim.x) {
      const int cx = c % TX, cr = c / TX, cy = cr % TY, cz = cr / TY;
      const int gi = t.x0 + cx, gj = t.y0 + cy, gk = DIM == 3 ? t.z0 + cz : 0;
      if (gi >= g.n[0] || gj >= g.n[1] || (DIM == 3 && gk >= g.n[2])) continue;
      const int e = (cx + t.H) + t.EX * (cy + t.H) + t.EXY * (DIM == 3 ? cz + t.H : 0);
      const long long q0 = g.ci(gi, gj, gk);
      CellFaces<T, DIM> F;
      load_faces<T, DIM>(A, gi, gj, gk, F);
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const long long q = q0 + n * g.ncell;
        T st[7];
        gather_lds<T, DIM>(sphl + n * E, e, t.EX, t.EXY, st);
        const T aq = A.a ? A.a[q] : T(1);
        const T lof = lof_of<T, DIM>(F, n, st, aq, sphi[n * E + e], A.alpha, A.beta, dxinv);
        sres[n * NC + c] = (io.rhs[q] - lof) / denom;
      }
    }
    __syncthreads();
    constexpr int CX = TX / 2, CY = TY / 2, CZ = DIM == 3 ? TZ / 2 : 1;
    constexpr int NCC = CX * CY * CZ;
    for (int c = threadIdx.x; c < NCC; c += blockDim.x) {
      const int cx = c % CX, cr = c / CX, cy = cr % CY, cz = cr / CY;
      const int I = t.x0 / 2 + cx, J = t.y0 / 2 + cy, K = DIM == 3 ? t.z0 / 2 + cz : 0;
      if (I >= gc.n[0] || J >= gc.n[1] || (DIM == 3 && K >= gc.n[2])) continue;
      const long long qc = gc.ci(I, J, K);
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        T s = T(0), sp = T(0);
        for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
          for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) {
              const int f = (2 * cx + ii) + TX * ((2 * cy + jj) + TY * (DIM == 3 ? 2 * cz + kk : 0));
              s = s + sres[n * NC + f];
              const int e = (2 * cx + ii + t.H) + t.EX * (2 * cy + jj + t.H) +
                            t.EXY * (DIM == 3 ? 2 * cz + kk + t.H : 0);
              sp = sp + sphi[n * E + e] / denom;
            }
        io.out[qc + n * gc.ncell] = s;
        if (io.phic) {
          io.phic[qc + n * gc.ncell] = sp;
          io.phic0[qc + n * gc.ncell] = sp;
        }
      }
    }
  }
}

template <typename T, int DIM, int TX, int TY, int TZ, int MODE>
__global__ __launch_bounds__(kThreads) void hc_tile_kernel(HCArgs<T> A, TileIO<T> io) {
  hc_tile_body<T, DIM, TX, TY, TZ, MODE>(A, io, int(blockIdx.x));
}

// batched operator applications: workgroup w applies tile map[w] & mask of job map[w] >> bits
template <typename T, int DIM, int TX, int TY, int TZ>
__global__ __launch_bounds__(kThreads) void hc_apply_batch_kernel(const HCApplyJob<T>* __restrict__ jobs,
                                                                  const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const HCApplyJob<T>& J = jobs[w >> kBatchTileBits];
  hc_tile_body<T, DIM, TX, TY, TZ, kApply>(J.A, J.io, int(w & kBatchTileMask));
}

// ---------------------------------------------------------------------------
// Single-workgroup V-cycle over the small levels (at most kWGCells cells):
// thread t owns cell t of the current level; phi and phi_l of that level live
// in LDS; neighbours are found by periodic wrap or the derived-variable wall
// rule.  A relaxation pass updates one colour and recomputes the closure of
// exactly the cells it updated (the other colour's phi_l is unchanged since
// its own update), so each pass needs a single barrier and half the closures
// of a full recomputation, with the same values.
// The level's geometry is copied into registers: reads through the
// descriptor pointer would be repeated after every barrier or wavefront fence
// of the relaxation passes.
template <typename T, int DIM>
struct WGLevel {
  const LevelDesc<T>* L;
  int N, nx, nxy;
  int n[3], per[3];
  T dx;
  long long ncell;
  int c, i, j, k;
  long long q;
  bool own;
  __device__ inline void init(const LevelDesc<T>* lv) {
    L = lv;
    const Geom& g = L->g;
    for (int d = 0; d < 3; ++d) {
      n[d] = g.n[d];
      per[d] = g.periodic[d];
    }
    dx = T(g.dx);
    ncell = g.ncell;
    nx = n[0];
    nxy = n[0] * n[1];
    N = nxy * (DIM == 3 ? n[2] : 1);
    c = threadIdx.x;
    own = c < N;
    int r, kk;
    divmod_pos(c, nx, r, i);
    divmod_pos(r, n[1], kk, j);
    k = DIM == 3 ? kk : 0;
    q = own ? g.ci(i, j, k) : 0;
  }
};

// The descriptor's scalars that every pass uses, in registers for the same
// reason.
template <typename T>
struct WGConst {
  const BCDesc* bc;   // read at wall cells only
  ClosureP<T> cp;
  T alpha, beta;
  __device__ inline void init(const CoarseDesc<T>& D) {
    bc = &D.bc;
    cp = D.cp;
    alpha = D.alpha;
    beta = D.beta;
  }
};

template <typename T, int DIM>
__device__ inline void wg_stencil(const WGLevel<T, DIM>& W, const BCDesc& bc, const T* pl, int n, T* p) {
  // without branches on the cell's position: the cells of one colour lie on
  // different sides of the walls, so branches would diverge in a wavefront.
  // Across a wall the load reads the cell itself and the ghost formula's
  // result is selected; the values are those of the branching form.
  const int c = W.c;
  const int iv[3] = {W.i, W.j, W.k};
  const T x0 = pl[c];
  p[0] = x0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int nd = W.n[d];
    const int st = d == 0 ? 1 : (d == 1 ? W.nx : W.nxy);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int v = iv[d] + (s ? 1 : -1);
      const bool inside = v >= 0 && v < nd;
      const bool wall = !inside && !W.per[d];
      const int off = inside ? (s ? st : -st) : (wall ? 0 : (s ? -(nd - 1) * st : (nd - 1) * st));
      const T nb = pl[c + off];
      const T gh = bc_ghost<T>(bc.type[d][s][n], T(bc.val[d][s][n]), x0, s ? 1 : -1, W.dx);
      p[1 + 2 * d + (s ? 0 : 1)] = (wall && bc.apply) ? gh : nb;
    }
  }
}

// closure of the thread's own cell from its LDS state
template <typename T, int DIM>
__device__ inline void wg_closure_own(const ClosureP<T>& cp, const WGLevel<T, DIM>& W, const T* sphi, T* sphl) {
  if (!W.own) return;
  T th, cl;
  closure<T>(sphi[W.c], sphi[W.N + W.c], cp, th, cl);
  sphl[W.c] = th;
  sphl[W.N + W.c] = cl;
}

template <typename T, int DIM>
__device__ inline void wg_faces(const LevelDesc<T>* L, int i, int j, int k, CellFaces<T, DIM>& F, T* aq) {
  load_faces<T, DIM>(L->g, L->b, i, j, k, F);
  const long long q = L->g.ci(i, j, k);
  aq[0] = L->a ? L->a[q] : T(1);
  aq[1] = L->a ? L->a[q + L->g.ncell] : T(1);
}

// L(phi) of the thread's cell (both components) from phi/phi_l in LDS
template <typename T, int DIM>
__device__ inline void wg_lof(const WGConst<T>& K, const WGLevel<T, DIM>& W, const CellFaces<T, DIM>& F,
                              const T* aq, const T* sphi, const T* sphl, T* lof) {
  const T dxinv = T(1) / (W.dx * W.dx);
#pragma unroll
  for (int n = 0; n < 2; ++n) {
    T st[7];
    wg_stencil<T, DIM>(W, *K.bc, sphl + n * W.N, n, st);
    lof[n] = lof_of<T, DIM>(F, n, st, aq[n], sphi[n * W.N + W.c], K.alpha, K.beta, dxinv);
  }
}

// Per-level operator data of the thread's cell, held in registers.
template <typename T, int DIM>
struct WGCellOp {
  CellFaces<T, DIM> F;
  T aq[2], lam[2];
  __device__ inline void load(const WGConst<T>& K, const WGLevel<T, DIM>& W) {
    if (!W.own) return;
    wg_faces<T, DIM>(W.L, W.i, W.j, W.k, F, aq);
    const T scale = T(1) / (W.dx * W.dx);
    lam[0] = lambda_of<T, DIM>(F, 0, aq[0], K.alpha, K.beta, scale);
    lam[1] = lambda_of<T, DIM>(F, 1, aq[1], K.alpha, K.beta, scale);
  }
};

// npass colour passes (first colour red) on the level whose phi is staged in
// sphi and whose phi_l (closure of sphi) is staged in sphl; rr = rhs of the
// thread's cell.  Leaves sphi/sphl consistent and writes phi to global memory.
// With fixed_stop, the passes end after the first red/black iteration that
// leaves every cell's state unchanged to the bit: each later iteration would
// start from the same state and repeat it exactly, so the result is that of
// all npass passes.
template <typename T, int DIM>
__device__ void wg_relax(const WGConst<T>& K, const WGLevel<T, DIM>& W, const WGCellOp<T, DIM>& O,
                         const T* rr, T* sphi, T* sphl, int npass, bool fixed_stop = false) {
  bool changed = false;
  auto pass = [&](int p) {
    if (W.own && !((W.i + W.j + W.k + p) & 1)) {
      T lof[2];
      wg_lof<T, DIM>(K, W, O.F, O.aq, sphi, sphl, lof);
      const int c = W.c;
      const T h0 = sphi[c], t0 = sphi[W.N + c];
      const T h = h0 - O.lam[0] * (lof[0] - rr[0]);
      const T t = t0 - O.lam[1] * (lof[1] - rr[1]);
      changed = changed || bits_of(h) != bits_of(h0) || bits_of(t) != bits_of(t0);
      sphi[c] = h;
      sphi[W.N + c] = t;
      T th, cl;
      closure<T>(h, t, K.cp, th, cl);
      sphl[c] = th;
      sphl[W.N + c] = cl;
    }
  };
  if (W.N <= 64) {
    // every cell is owned by wave 0: it runs the passes alone, synchronising
    // at wavefront scope; the other waves wait once
    if (threadIdx.x < 64)
      for (int p = 0; p < npass; ++p) {
        pass(p);
        wave_sync();
        if (fixed_stop && (p & 1)) {
          if (!__any(changed)) break;
          changed = false;
        }
      }
    __syncthreads();
  } else {
    for (int p = 0; p < npass; ++p) {
      pass(p);
      if (fixed_stop && (p & 1)) {
        if (!__syncthreads_or(changed)) break;
        changed = false;
      } else {
        __syncthreads();
      }
    }
  }
  if (W.own) {
    W.L->phi[W.q] = sphi[W.c];
    W.L->phi[W.q + W.ncell] = sphi[W.N + W.c];
  }
}

// stage the thread's (phi) into LDS and compute every closure
template <typename T, int DIM>
__device__ inline void wg_stage(const WGConst<T>& K, const WGLevel<T, DIM>& W, const T* ph, T* sphi, T* sphl) {
  if (W.own) {
    sphi[W.c] = ph[0];
    sphi[W.N + W.c] = ph[1];
  }
  __syncthreads();
  wg_closure_own<T, DIM>(K.cp, W, sphi, sphl);
  __syncthreads();
}

// Coarse part of the FAS V-cycle (levels l0 .. nlev-1) in one workgroup:
//   rhs_l0 += L(phi_l0) when apply_first (the restriction into l0 was done by
//   the finer level's kernel); then for l = l0 .. nlev-2: relax nu1; restrict
//   phi and rhs - L phi; rhs_{l+1} += L(phi_{l+1}); bottom relax; then back
//   up: phi_l += P(phi_{l+1} - phi0_{l+1}); relax nu2.
// Levels are visited with their state in registers/LDS; global memory keeps
// phi, rhs and phi0 of every level for the way back up and for the finer
// tiled levels.
template <typename T, int DIM>
__global__ __launch_bounds__(kWGThreads) void hc_coarse_vcycle_kernel(const CoarseDesc<T>* __restrict__ Dp,
                                                                      int apply_first) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const CoarseDesc<T>& D = *Dp;
  if (D.skip && *D.skip) return;
  WGConst<T> K;
  K.init(D);
  const int nu1 = D.nu1, nu2 = D.nu2, bottom = D.bottom;
  T* sphi = reinterpret_cast<T*>(smem_raw);   // 2*kWGCells
  T* sphl = sphi + 2 * kWGCells;              // 2*kWGCells
  T* se = sphl + 2 * kWGCells;                // coarse correction, 2*kWGCells/4
  const int l0 = D.l0, L = D.nlev;
  const T denom = DIM == 3 ? T(8) : T(4);

  // state of the current level's cell in registers
  T ph[2] = {T(0), T(0)}, rr[2] = {T(0), T(0)};
  // `staged`: sphi/sphl already hold the current level's state and closures
  // and `Oc` its operator data (loaded with its FAS right-hand side), so the
  // level visit that follows neither reloads nor restages them
  WGCellOp<T, DIM> Oc;
  bool staged = false;
  {
    WGLevel<T, DIM> W;
    W.init(&D.lev[l0]);
    if (W.own) {
      ph[0] = W.L->phi[W.q];
      ph[1] = W.L->phi[W.q + W.L->g.ncell];
      rr[0] = W.L->rhs[W.q];
      rr[1] = W.L->rhs[W.q + W.L->g.ncell];
    }
    if (apply_first) {
      Oc.load(K, W);
      wg_stage<T, DIM>(K, W, ph, sphi, sphl);
      if (W.own) {
        T lof[2];
        wg_lof<T, DIM>(K, W, Oc.F, Oc.aq, sphi, sphl, lof);
        rr[0] = rr[0] + lof[0];
        rr[1] = rr[1] + lof[1];
        W.L->rhs[W.q] = rr[0];
        W.L->rhs[W.q + W.L->g.ncell] = rr[1];
      }
      __syncthreads();
      staged = true;
    }
  }

  for (int l = l0; l < L - 1; ++l) {
    WGLevel<T, DIM> W;
    W.init(&D.lev[l]);
    if (!staged) {
      Oc.load(K, W);
      wg_stage<T, DIM>(K, W, ph, sphi, sphl);
    }
    const WGCellOp<T, DIM>& O = Oc;   // reloaded below for the coarser level after its last use
    wg_relax<T, DIM>(K, W, O, rr, sphi, sphl, 2 * nu1);
    // restriction (RESTRICTRESNLVC / RESTRICT): each fine cell's
    // (rhs - L phi)/2^d and phi/2^d, then child sums in Fortran order
    T tr[2] = {T(0), T(0)}, tp[2] = {T(0), T(0)};
    if (W.own) {
      T lof[2];
      wg_lof<T, DIM>(K, W, O.F, O.aq, sphi, sphl, lof);
      tr[0] = (rr[0] - lof[0]) / denom;
      tr[1] = (rr[1] - lof[1]) / denom;
      tp[0] = sphi[W.c] / denom;
      tp[1] = sphi[W.N + W.c] / denom;
    }
    __syncthreads();
    if (W.own) {
      sphl[W.c] = tr[0];
      sphl[W.N + W.c] = tr[1];
      sphi[W.c] = tp[0];
      sphi[W.N + W.c] = tp[1];
    }
    __syncthreads();
    WGLevel<T, DIM> Wc;
    Wc.init(&D.lev[l + 1]);
    if (Wc.own) {
      T s[2] = {T(0), T(0)}, sp[2] = {T(0), T(0)};
      for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
        for (int jj = 0; jj < 2; ++jj)
          for (int ii = 0; ii < 2; ++ii) {
            const int fc = (2 * Wc.i + ii) + W.nx * (2 * Wc.j + jj) + W.nxy * (DIM == 3 ? 2 * Wc.k + kk : 0);
#pragma unroll
            for (int n = 0; n < 2; ++n) {
              s[n] = s[n] + sphl[n * W.N + fc];
              sp[n] = sp[n] + sphi[n * W.N + fc];
            }
          }
      const long long qc = Wc.q;
      const long long ncc = Wc.L->g.ncell;
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        Wc.L->phi[qc + n * ncc] = sp[n];
        Wc.L->phi0[qc + n * ncc] = sp[n];
      }
      ph[0] = sp[0];
      ph[1] = sp[1];
      rr[0] = s[0];
      rr[1] = s[1];
    }
    __syncthreads();
    // FAS coarse right-hand side rhs_{l+1} += L(phi_{l+1})
    {
      Oc.load(K, Wc);
      wg_stage<T, DIM>(K, Wc, ph, sphi, sphl);
      if (Wc.own) {
        T lof[2];
        wg_lof<T, DIM>(K, Wc, Oc.F, Oc.aq, sphi, sphl, lof);
        rr[0] = rr[0] + lof[0];
        rr[1] = rr[1] + lof[1];
        Wc.L->rhs[Wc.q] = rr[0];
        Wc.L->rhs[Wc.q + Wc.L->g.ncell] = rr[1];
      }
      __syncthreads();
      staged = true;
    }
  }
  {
    WGLevel<T, DIM> W;
    W.init(&D.lev[L - 1]);
    if (!staged) {
      Oc.load(K, W);
      wg_stage<T, DIM>(K, W, ph, sphi, sphl);
    }
    wg_relax<T, DIM>(K, W, Oc, rr, sphi, sphl, 2 * bottom, true);
  }
  for (int l = L - 2; l >= l0; --l) {
    // coarse correction phi_{l+1} - phi0_{l+1} into LDS (compact layout)
    WGLevel<T, DIM> Wc;
    Wc.init(&D.lev[l + 1]);
    if (Wc.own) {
      const long long ncc = Wc.L->g.ncell;
      se[Wc.c] = sphi[Wc.c] - Wc.L->phi0[Wc.q];
      se[Wc.N + Wc.c] = sphi[Wc.N + Wc.c] - Wc.L->phi0[Wc.q + ncc];
    }
    __syncthreads();
    WGLevel<T, DIM> W;
    W.init(&D.lev[l]);
    WGCellOp<T, DIM> O;
    O.load(K, W);
    if (W.own) {
      const long long nc = W.L->g.ncell;
      rr[0] = W.L->rhs[W.q];
      rr[1] = W.L->rhs[W.q + nc];
      ph[0] = prolong_value_lds<T, DIM>(Wc.L->g, W.i, W.j, W.k, se, W.L->phi[W.q]);
      ph[1] = prolong_value_lds<T, DIM>(Wc.L->g, W.i, W.j, W.k, se + Wc.N, W.L->phi[W.q + nc]);
    }
    __syncthreads();
    wg_stage<T, DIM>(K, W, ph, sphi, sphl);
    wg_relax<T, DIM>(K, W, O, rr, sphi, sphl, 2 * nu2);
  }
}

// closure of valid cells and wall ghosts (computeDiffusedVar, LevelData form)
template <typename T, int DIM>
__global__ void hc_diffused_var_kernel(HCArgs<T> A, const T* __restrict__ phi, T* __restrict__ out,
                                       int homog) {
  const Geom& g = A.g;
  const int gx = g.n[0] + 2, gy = g.n[1] + 2, gz = DIM == 3 ? g.n[2] + 2 : 1;
  const long long total = (long long)gx * gy * gz;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < total;
       c += (long long)gridDim.x * blockDim.x) {
    const int i = int(c % gx) - 1;
    const long long r = c / gx;
    const int j = int(r % gy) - 1;
    const int k = DIM == 3 ? int(r / gy) - 1 : 0;
    const Pos P = classify<DIM>(g, i, j, k);
    const long long q = g.ci(i, j, k);
    if (P.kind == 0) continue;
    const long long qs = g.ci(P.i, P.j, P.k);
    T th, cl;
    closure<T>(phi[qs], phi[qs + g.ncell], A.cp, th, cl);
    if (P.kind == 2 && A.bc.apply) {
      const int isign = P.gs ? 1 : -1;
      const T v0 = homog ? T(0) : T(A.bc.val[P.gd][P.gs][0]);
      const T v1 = homog ? T(0) : T(A.bc.val[P.gd][P.gs][1]);
      th = bc_ghost<T>(A.bc.type[P.gd][P.gs][0], v0, th, isign, T(g.dx));
      cl = bc_ghost<T>(A.bc.type[P.gd][P.gs][1], v1, cl, isign, T(g.dx));
    }
    out[q] = th;
    out[q + g.ncell] = cl;
  }
}

// ---- streaming relaxation for large levels -------------------------------
// The (H, Theta) counterpart of the linear operator's streaming kernel: a
// workgroup owns a block of columns and a segment of planes along the last
// direction and marches through them with a window of 4 planes in LDS
// holding (H, Theta) and their closure (theta, Theta_l).  With plane R just
// loaded (and its closure taken), colour pass p runs on plane R-1-p, so
// plane R-P is final and is written.  A cell's pass reads its own state and
// the closure of its other-colour neighbours, all as of the start of that
// pass, exactly as in the tiled kernel; an updated cell's closure (and its
// wall ghosts') is refreshed at once, which its same-colour cells never read
// (after the launch's last pass nothing reads it, so it is skipped there).
// Per plane and position the face coefficients, rhs (and aCoef) stay in
// registers in compile-time-indexed sets, loaded two planes ahead; lambda is
// formed where it is used.  Every input is read once per launch.
template <typename T, int DIM>
struct HCoef {
  CellFaces<T, DIM> F;
  T rr[2], aq[2], v[2];   // rhs, aCoef, state as loaded
};

template <typename T, int DIM, int TX, int TY, int P, int NT, bool HASA>
__global__ __launch_bounds__(NT) void hc_relax_stream_kernel(HCArgs<T> A, const T* __restrict__ in,
                                                             T* __restrict__ out, const T* __restrict__ rhs,
                                                             int first, int nbx, int nby, int seg) {
  if (A.skip && *A.skip) return;
  constexpr int EX = TX + 2 * P;
  constexpr int EL = EX * (DIM == 3 ? TY + 2 * P : 1);
  constexpr int K = (EL + NT - 1) / NT;
  static_assert(P >= 1 && P <= 2, "one or two colour passes per launch");
  constexpr int W = 4;   // planes in the window
  constexpr int MD = DIM - 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* su = reinterpret_cast<T*>(smem_raw);
  const Geom& g = A.g;
  const int nz = g.n[MD];
  const bool zper = g.periodic[MD] != 0;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  const T dxg = T(g.dx);
  int b = blockIdx.x;
  const int bx = b % nbx;
  b /= nbx;
  const int by = DIM == 3 ? b % nby : 0;
  const int bz = DIM == 3 ? b / nby : b;
  const int x0 = bx * TX, y0 = by * TY;
  const int z0 = bz * seg;
  const int z1 = z0 + seg < nz ? z0 + seg : nz;
  SPos S[K];
  int gt[K][2];   // a lateral ghost's BC type and value per component
  T gval[K][2];
#pragma unroll
  for (int kk = 0; kk < K; ++kk) {
    const int e = threadIdx.x + kk * NT;
    gt[kk][0] = gt[kk][1] = 0;
    gval[kk][0] = gval[kk][1] = T(0);
    if (e < EL) {
      spos_init<DIM, TX, TY, P>(g, x0, y0, e, S[kk]);
      for (int d = 0; d < DIM - 1; ++d)
        for (int sd = 0; sd < 2; ++sd)
          if (S[kk].kind == 2 && S[kk].gd == d && S[kk].gs == sd)
            for (int n = 0; n < 2; ++n) {
              gt[kk][n] = A.bc.type[d][sd][n];
              gval[kk][n] = T(A.bc.val[d][sd][n]);
            }
    } else {
      S[kk].kind = 0;
      S[kk].dist = P;
      S[kk].owner = 0;
    }
  }
  // derived-variable ghost beyond wall (d, sd) of component n from the
  // interior closure value x (computeDiffusedVar's BC)
  auto ghost = [&](int d, int sd, int n, T x) {
    return A.bc.apply ? bc_ghost<T>(A.bc.type[d][sd][n], T(A.bc.val[d][sd][n]), x, sd ? 1 : -1, dxg) : x;
  };
  auto zwrap = [&](int z) { return z < 0 ? z + nz : (z >= nz ? z - nz : z); };
  auto real_plane = [&](int z) { return zper || (z >= 0 && z < nz); };
  // 32-bit element offsets (levels below 2^28 cells per component)
  const unsigned csz = unsigned(DIM == 3 ? g.sz : g.sy);
  const unsigned cn = unsigned(g.ncell);
  unsigned fpl[DIM], fup[DIM], fn[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    fpl[d] = unsigned(g.fn[d][0] * (DIM == 3 ? g.fn[d][1] : 1));
    fn[d] = unsigned(g.nface[d]);
  }
  fup[0] = 1;
  if (DIM == 3) fup[1] = unsigned(g.fn[1][0]);
  fup[MD] = fpl[MD];
  unsigned cl[K], fl[K][DIM];
#pragma unroll
  for (int kk = 0; kk < K; ++kk) {
    const SPos& s = S[kk];
    cl[kk] = unsigned(s.gx + 1) + (DIM == 3 ? unsigned(g.sy) * unsigned(s.gy + 1) : 0u);
#pragma unroll
    for (int d = 0; d < DIM; ++d)
      fl[kk][d] = unsigned(s.gx) + (DIM == 3 ? unsigned(g.fn[d][0]) * unsigned(s.gy) : 0u);
  }
  auto ld = [](const T* p, unsigned i) -> T {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + (i * unsigned(sizeof(T))));
  };
  constexpr int NS = P + 3;
  HCoef<T, DIM> cs[NS][K];
  auto prefetch = [&](int R, HCoef<T, DIM>(&pf)[K]) {
    const int zw = zper ? zwrap(R) : (R < 0 ? 0 : (R >= nz ? nz - 1 : R));
    const unsigned qz = unsigned(zw + 1) * csz;
#pragma unroll
    for (int kk = 0; kk < K; ++kk) {
      const unsigned q = cl[kk] + qz;   // a ghost reads its interior neighbour
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        pf[kk].v[n] = ld(in, q + unsigned(n) * cn);
        pf[kk].rr[n] = ld(rhs, q + unsigned(n) * cn);
        if (HASA) pf[kk].aq[n] = ld(A.a, q + unsigned(n) * cn);
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const unsigned f = fl[kk][d] + unsigned(zw) * fpl[d] + unsigned(n) * fn[d];
          pf[kk].F.v[n][2 * d] = ld(A.b[d], f);
          pf[kk].F.v[n][2 * d + 1] = ld(A.b[d], f + fup[d]);
        }
      }
    }
  };
  const int zl = (!zper && z0 - P < 0) ? 0 : z0 - P;
  const int zend = z1 - 1 + P;
  // window slot of plane R: (H, Theta, theta, Theta_l) of EL positions each
  // (the bottom wall's ghost plane -1 is slot 3)
  auto slot = [&](int R) { return su + ((R - zl) & (W - 1)) * (4 * EL); };
  auto step = [&](auto J, int R) {
    constexpr int j = decltype(J)::value;
    // A. commit plane R: state, closure, lateral ghosts (or the top wall's
    // ghost plane)
    T* pl = slot(R);
    if (real_plane(R)) {
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const SPos& s = S[kk];
        const int e = threadIdx.x + kk * NT;
        if (s.kind == 0) continue;
        const T h = cs[j][kk].v[0], c = cs[j][kk].v[1];
        T th, clv;
        closure<T>(h, c, A.cp, th, clv);
        if (s.kind == 1) {
          pl[e] = h;
          pl[EL + e] = c;
          pl[2 * EL + e] = th;
          pl[3 * EL + e] = clv;
          if (!zper && R == 0) {
            T* gp = slot(-1);
            gp[2 * EL + e] = ghost(MD, 0, 0, th);
            gp[3 * EL + e] = ghost(MD, 0, 1, clv);
          }
        } else {
          const int isign = s.gs == 0 ? -1 : 1;
          pl[2 * EL + e] = A.bc.apply ? bc_ghost<T>(gt[kk][0], gval[kk][0], th, isign, dxg) : th;
          pl[3 * EL + e] = A.bc.apply ? bc_ghost<T>(gt[kk][1], gval[kk][1], clv, isign, dxg) : clv;
        }
      }
    } else if (!zper && R == nz) {
      const T* below = slot(nz - 1);
#pragma unroll
      for (int kk = 0; kk < K; ++kk) {
        const int e = threadIdx.x + kk * NT;
        if (S[kk].kind != 1) continue;
        pl[2 * EL + e] = ghost(MD, 1, 0, below[2 * EL + e]);
        pl[3 * EL + e] = ghost(MD, 1, 1, below[3 * EL + e]);
      }
    }
    // B. loads of plane R+2 into the set of plane R+2-NS, no longer in use
    constexpr int jn = (j + 2) % NS;
    prefetch(R + 2, cs[jn]);
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
          const HCoef<T, DIM>& C = cs[jp][kk];
          T nv[2];
#pragma unroll
          for (int n = 0; n < 2; ++n) {
            const T* lz = pz + (2 + n) * EL;
            T st[7];
            st[0] = lz[e];
            st[1] = lz[e + 1];
            st[2] = lz[e - 1];
            if (DIM == 3) {
              st[3] = lz[e + EX];
              st[4] = lz[e - EX];
              st[5] = pp[(2 + n) * EL + e];
              st[6] = pm[(2 + n) * EL + e];
            } else {
              st[3] = pp[(2 + n) * EL + e];
              st[4] = pm[(2 + n) * EL + e];
            }
            const T ph = pz[n * EL + e];
            const T aq = HASA ? C.aq[n] : T(1);
            const T lam = lambda_of<T, DIM>(C.F, n, aq, A.alpha, A.beta, dxinv);
            const T lof = lof_of<T, DIM>(C.F, n, st, aq, ph, A.alpha, A.beta, dxinv);
            nv[n] = ph - lam * (lof - C.rr[n]);
          }
          pz[e] = nv[0];
          pz[EL + e] = nv[1];
          // the last pass's closure is never read: its cells' neighbours
          // (the other colour) are not updated again in this launch
          if constexpr (p == P - 1) continue;
          T th, clv;
          closure<T>(nv[0], nv[1], A.cp, th, clv);
          pz[2 * EL + e] = th;
          pz[3 * EL + e] = clv;
          // wall ghosts of this cell: lateral, then along the march
          const int glc[2] = {s.gx, s.gy};
#pragma unroll
          for (int d = 0; d < (DIM == 3 ? 2 : 1); ++d) {
            if (g.periodic[d]) continue;
            const int st1 = d == 0 ? 1 : EX;
            if (glc[d] == 0) {
              pz[2 * EL + e - st1] = ghost(d, 0, 0, th);
              pz[3 * EL + e - st1] = ghost(d, 0, 1, clv);
            }
            if (glc[d] == g.n[d] - 1) {
              pz[2 * EL + e + st1] = ghost(d, 1, 0, th);
              pz[3 * EL + e + st1] = ghost(d, 1, 1, clv);
            }
          }
          if (!zper) {
            if (zw == 0) {
              T* gp = slot(-1);
              gp[2 * EL + e] = ghost(MD, 0, 0, th);
              gp[3 * EL + e] = ghost(MD, 0, 1, clv);
            }
            if (zw == nz - 1) {
              T* gp = slot(nz);
              gp[2 * EL + e] = ghost(MD, 1, 0, th);
              gp[3 * EL + e] = ghost(MD, 1, 1, clv);
            }
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
        if (!S[kk].owner) continue;
        const unsigned q = cl[kk] + unsigned(zs + 1) * csz;
        const int e = threadIdx.x + kk * NT;
        *reinterpret_cast<T*>(reinterpret_cast<char*>(out) + q * unsigned(sizeof(T))) = ps[e];
        *reinterpret_cast<T*>(reinterpret_cast<char*>(out) + (q + cn) * unsigned(sizeof(T))) = ps[EL + e];
      }
    }
  };
  prefetch(zl, cs[0]);
  prefetch(zl + 1, cs[1]);
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

}  // namespace

// ---------------------------------------------------------------- launchers

template <typename T, int DIM>
struct TileDims;
template <typename T>
struct TileDims<T, 2> {
  static constexpr int TX = 32, TY = 8, TZ = 1;
};
template <typename T>
struct TileDims<T, 3> {
  static constexpr int TX = 16, TY = 8, TZ = 4;
};

template <typename T, int DIM>
static int tiles_of(const Geom& g, int& ntx, int& nty) {
  using D = TileDims<T, DIM>;
  ntx = (g.n[0] + D::TX - 1) / D::TX;
  nty = (g.n[1] + D::TY - 1) / D::TY;
  const int ntz = DIM == 3 ? (g.n[2] + D::TZ - 1) / D::TZ : 1;
  return ntx * nty * ntz;
}

template <typename T, int DIM>
static size_t tile_lds(int halo, bool restrict_mode) {
  using D = TileDims<T, DIM>;
  const size_t E = size_t(D::TX + 2 * halo) * (D::TY + 2 * halo) * (DIM == 3 ? D::TZ + 2 * halo : 1);
  size_t n = 4 * E;
  if (restrict_mode) n += 2 * size_t(D::TX) * D::TY * (DIM == 3 ? D::TZ : 1);
  return n * sizeof(T);
}

template <typename T, int DIM, int MODE>
static hipError_t launch_tile(const HCArgs<T>& A, TileIO<T> io, hipStream_t s) {
  using D = TileDims<T, DIM>;
  const int nt = tiles_of<T, DIM>(A.g, io.ntx, io.nty);
  const size_t lds = tile_lds<T, DIM>(1, MODE == kRestrict);
  if (lds > kMaxLds) return hipErrorInvalidValue;
  hipLaunchKernelGGL((hc_tile_kernel<T, DIM, D::TX, D::TY, D::TZ, MODE>), dim3(nt), dim3(kThreads), lds, s, A, io);
  return hipGetLastError();
}

// Passes per relax launch: small grids are latency bound, so up to four
// red/black sweeps share one launch (halo 8); large grids use one sweep per
// launch (halo 2) to keep the redundant halo work small.
constexpr int kTilesForDeepHalo = 1024;

template <typename T, int DIM>
static int relax_passes_per_launch(const Geom& g) {
  int ntx, nty;
  const int nt = tiles_of<T, DIM>(g, ntx, nty);
  return (DIM == 2 && nt <= kTilesForDeepHalo) ? 8 : 2;
}

// tiles of the one-sweep 2D relaxation launch on large levels (measured
// best of 5 shapes at 2048^2)
struct HCRelaxWide2 {
  static constexpr int TX = 32, TY = 8, TZ = 1, NT = 256;
};

// tiles of the 3D relaxation launch (measured best of 5 shapes at 128^3;
// the others spill registers)
struct HCRelaxTile3 {
  static constexpr int TX = 16, TY = 8, TZ = 4, NT = 512;
};

template <typename T, int DIM>
static hipError_t launch_relax_tile(const HCArgs<T>& A, TileIO<T> io, hipStream_t s, const LaunchTimer* tm) {
  using D = TileDims<T, DIM>;
  using DW = HCRelaxWide2;
  using D3 = HCRelaxTile3;
  int nt = tiles_of<T, DIM>(A.g, io.ntx, io.nty);
  const int pmax = relax_passes_per_launch<T, DIM>(A.g);
  if (io.npass < 1 || io.npass > pmax) return hipErrorInvalidValue;
  size_t lds = tile_lds<T, DIM>(io.npass, false);
  if (DIM == 2 && pmax == 2) {
    io.ntx = (A.g.n[0] + DW::TX - 1) / DW::TX;
    io.nty = (A.g.n[1] + DW::TY - 1) / DW::TY;
    nt = io.ntx * io.nty;
    lds = 4 * size_t(DW::TX + 2 * io.npass) * (DW::TY + 2 * io.npass) * sizeof(T);
  }
  if (DIM == 3) {
    io.ntx = (A.g.n[0] + D3::TX - 1) / D3::TX;
    io.nty = (A.g.n[1] + D3::TY - 1) / D3::TY;
    nt = io.ntx * io.nty * ((A.g.n[2] + D3::TZ - 1) / D3::TZ);
    lds = 4 * size_t(D3::TX + 2 * io.npass) * (D3::TY + 2 * io.npass) * (D3::TZ + 2 * io.npass) * sizeof(T);
  }
  if (lds > kMaxLds) return hipErrorInvalidValue;
  const bool timed = tm && *tm->n < tm->cap;
  if (tm && !timed) ++*tm->missed;
  hipEvent_t e0 = timed ? tm->start[*tm->n] : nullptr, e1 = timed ? tm->stop[*tm->n] : nullptr;
  if constexpr (DIM == 2) {
    if (pmax == 8)
      hipExtLaunchKernelGGL((hc_relax_tile_kernel<T, DIM, D::TX, D::TY, D::TZ, 8, 512>), dim3(nt), dim3(512), lds,
                            s, e0, e1, 0, A, io);
    else
      hipExtLaunchKernelGGL((hc_relax_tile_kernel<T, DIM, DW::TX, DW::TY, DW::TZ, 2, DW::NT>), dim3(nt), dim3(DW::NT),
                            lds, s, e0, e1, 0, A, io);
  } else {
    hipExtLaunchKernelGGL((hc_relax_tile_kernel<T, DIM, D3::TX, D3::TY, D3::TZ, 2, D3::NT>), dim3(nt), dim3(D3::NT),
                          lds, s, e0, e1, 0, A, io);
  }
  const hipError_t err = hipGetLastError();
  if (err == hipSuccess && timed) {
    // algorithmic bytes: state (2 comps) in and out, rhs, a, faces (2 comps)
    // once each; the coarse state (and its initial copy) when prolonged on
    // load
    const double cells = double(A.g.nvalid());
    double faces = 0;
    for (int d = 0; d < DIM; ++d) faces += double(A.g.nface[d]);
    double b = 2 * cells + 2 * cells + (A.a ? 2 * cells : 0) + 2 * faces + 2 * cells;
    if (io.pc) b += 2 * io.gc.nvalid() * (io.pc0 ? 2 : 1);
    *tm->bytes += b * sizeof(T);
    ++*tm->n;
  }
  return err;
}

// streaming launch shapes: 2D columns of 252 cells (256 threads), 3D blocks
// of 12 x 12 columns
template <int DIM>
struct HCStream;
template <>
struct HCStream<2> {
  static constexpr int NT = 256, TY = 1, SEGMIN = 8;
};
template <>
struct HCStream<3> {
  static constexpr int NT = 256, TY = 12, SEGMIN = 8;
};

// levels whose plain relaxations stream (2D above 512^2 cells, 3D above
// 96^3), with 32-bit offsets for both components of every array
constexpr long long kHCStreamCells2 = 512 * 512;
constexpr long long kHCStreamCells3 = 96LL * 96 * 96;
inline bool hc_streams(const Geom& g) {
  long long nf = 0;
  for (int d = 0; d < g.dim; ++d) nf = nf > g.nface[d] ? nf : g.nface[d];
  const long long most = 2 * (g.ncell > nf ? g.ncell : nf);
  return g.nvalid() > (g.dim == 2 ? kHCStreamCells2 : kHCStreamCells3) && g.n[g.dim - 1] >= 8 && most < (1LL << 29);
}

template <typename T, int DIM, int P, bool HASA>
hipError_t hc_stream_launch(const HCArgs<T>& A, const T* in, T* out, const T* rhs, int first, hipStream_t s,
                            const LaunchTimer* tm) {
  using D = HCStream<DIM>;
  constexpr int TX = DIM == 2 ? D::NT - 2 * P : D::TY;
  constexpr int TY = DIM == 2 ? 1 : D::TY;
  constexpr int EL = (TX + 2 * P) * (DIM == 3 ? TY + 2 * P : 1);
  const Geom& g = A.g;
  const int nbx = (g.n[0] + TX - 1) / TX;
  const int nby = DIM == 3 ? (g.n[1] + TY - 1) / TY : 1;
  const size_t lds = size_t(4) * 4 * EL * sizeof(T);
  auto kern = hc_relax_stream_kernel<T, DIM, TX, TY, P, D::NT, HASA>;
  // segments sized so that the whole grid is resident at once
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
                        timed ? tm->stop[*tm->n] : nullptr, 0, A, in, out, rhs, first, nbx, nby, seg);
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess && timed) {
    const double cells = double(g.nvalid());
    double faces = 0;
    for (int d = 0; d < DIM; ++d) faces += double(g.nface[d]);
    const double by = 2 * cells + 2 * cells + (HASA ? 2 * cells : 0) + 2 * faces + 2 * cells;
    *tm->bytes += by * sizeof(T);
    ++*tm->n;
  }
  return e;
}

template <typename T, int DIM>
hipError_t hc_relax_stream(const HCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, int first,
                           hipStream_t s, const LaunchTimer* tm) {
  int done = 0;
  while (done < npass) {
    const int p = (npass - done) >= 2 ? 2 : 1;
    const int f = (first + done) & 1;
    hipError_t e;
    if (A.a)
      e = p == 2 ? hc_stream_launch<T, DIM, 2, true>(A, buf[*cur], buf[1 - *cur], rhs, f, s, tm)
                 : hc_stream_launch<T, DIM, 1, true>(A, buf[*cur], buf[1 - *cur], rhs, f, s, tm);
    else
      e = p == 2 ? hc_stream_launch<T, DIM, 2, false>(A, buf[*cur], buf[1 - *cur], rhs, f, s, tm)
                 : hc_stream_launch<T, DIM, 1, false>(A, buf[*cur], buf[1 - *cur], rhs, f, s, tm);
    if (e != hipSuccess) return e;
    *cur = 1 - *cur;
    done += p;
  }
  return hipSuccess;
}

template <typename T>
hipError_t hc_relax_tiles(const HCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, int first,
                          const Geom* gc, const T* pc, const T* pc0, hipStream_t s, const LaunchTimer* tm) {
  // plain relaxations of large levels stream along the last direction
  if (!pc && npass > 0 && hc_streams(A.g))
    return A.g.dim == 2 ? hc_relax_stream<T, 2>(A, buf, cur, rhs, npass, first, s, tm)
                        : hc_relax_stream<T, 3>(A, buf, cur, rhs, npass, first, s, tm);
  const int pmax = A.g.dim == 2 ? relax_passes_per_launch<T, 2>(A.g) : relax_passes_per_launch<T, 3>(A.g);
  int done = 0;
  while (done < npass) {
    TileIO<T> io{};
    io.phi = buf[*cur];
    io.phi_out = buf[1 - *cur];
    io.rhs = rhs;
    io.npass = (npass - done) < pmax ? (npass - done) : pmax;
    io.first = (first + done) & 1;
    if (done == 0 && pc) {
      io.gc = *gc;
      io.pc = pc;
      io.pc0 = pc0;
    }
    hipError_t e = A.g.dim == 2 ? launch_relax_tile<T, 2>(A, io, s, tm) : launch_relax_tile<T, 3>(A, io, s, tm);
    if (e != hipSuccess) return e;
    *cur = 1 - *cur;
    done += io.npass;
  }
  if (npass == 0 && pc) {
    // a prolongation with no smoothing after it (per cell, in place)
    return prolong_linear<T>(A.g, 2, buf[*cur], pc, pc0, s);
  }
  return hipSuccess;
}

template <typename T>
hipError_t hc_relax(const HCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, int first,
                    hipStream_t s, const LaunchTimer* tm) {
  if (npass <= 0) return hipSuccess;
  return hc_relax_tiles<T>(A, buf, cur, rhs, npass, first, nullptr, nullptr, nullptr, s, tm);
}

template <typename T>
hipError_t hc_coarse_vcycle(const CoarseDesc<T>* dev_desc, int dim, int apply_first, hipStream_t s) {
  const size_t lds = (size_t(4) * kWGCells + size_t(2) * (kWGCells / 4)) * sizeof(T);
  if (dim == 2)
    hipLaunchKernelGGL((hc_coarse_vcycle_kernel<T, 2>), dim3(1), dim3(kWGThreads), lds, s, dev_desc, apply_first);
  else
    hipLaunchKernelGGL((hc_coarse_vcycle_kernel<T, 3>), dim3(1), dim3(kWGThreads), lds, s, dev_desc, apply_first);
  return hipGetLastError();
}

template <typename T>
hipError_t hc_residual(const HCArgs<T>& A, T* phi, const T* rhs, T* res, unsigned long long* norm,
                       int homog, hipStream_t s, const CtlFuse& fuse) {
  if (fuse.ctl && !norm) return hipErrorInvalidValue;
  TileIO<T> io{};
  io.fuse = fuse;
  io.phi = phi;
  io.rhs = rhs;
  io.out = res;
  io.norm = norm;
  io.homog = homog;
  return A.g.dim == 2 ? launch_tile<T, 2, kResid>(A, io, s) : launch_tile<T, 3, kResid>(A, io, s);
}

template <typename T>
hipError_t hc_apply_op(const HCArgs<T>& A, T* phi, T* lhs, const T* add, int homog, hipStream_t s) {
  TileIO<T> io{};
  io.phi = phi;
  io.out = lhs;
  io.add = add;
  io.homog = homog;
  return A.g.dim == 2 ? launch_tile<T, 2, kApply>(A, io, s) : launch_tile<T, 3, kApply>(A, io, s);
}

int hc_apply_tiles(const Geom& g, int* ntx, int* nty) {
  return g.dim == 2 ? tiles_of<double, 2>(g, *ntx, *nty) : tiles_of<double, 3>(g, *ntx, *nty);
}

template <typename T>
hipError_t hc_apply_batch(int dim, const HCApplyJob<T>* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  if (dim == 2) {
    using D = TileDims<T, 2>;
    const size_t lds = tile_lds<T, 2>(1, false);
    hipLaunchKernelGGL((hc_apply_batch_kernel<T, 2, D::TX, D::TY, D::TZ>), dim3(nblocks), dim3(kThreads), lds, s,
                       jobs, map);
  } else {
    using D = TileDims<T, 3>;
    const size_t lds = tile_lds<T, 3>(1, false);
    hipLaunchKernelGGL((hc_apply_batch_kernel<T, 3, D::TX, D::TY, D::TZ>), dim3(nblocks), dim3(kThreads), lds, s,
                       jobs, map);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t hc_restrict_residual(const HCArgs<T>& A, const Geom& gc, T* phi, const T* rhs, T* resc,
                                int homog, hipStream_t s) {
  return hc_restrict_all<T>(A, gc, phi, rhs, resc, nullptr, nullptr, homog, s);
}

template <typename T>
hipError_t hc_restrict_all(const HCArgs<T>& A, const Geom& gc, T* phi, const T* rhs, T* resc, T* phic, T* phic0,
                           int homog, hipStream_t s) {
  TileIO<T> io{};
  io.phi = phi;
  io.rhs = rhs;
  io.out = resc;
  io.gc = gc;
  io.phic = phic;
  io.phic0 = phic0;
  io.homog = homog;
  return A.g.dim == 2 ? launch_tile<T, 2, kRestrict>(A, io, s) : launch_tile<T, 3, kRestrict>(A, io, s);
}

template <typename T>
hipError_t hc_diffused_var(const HCArgs<T>& A, const T* phi, T* out, int homog, hipStream_t s) {
  const long long total = A.g.ncell;
  const int blocks = int((total + 255) / 256);
  if (A.g.dim == 2)
    hipLaunchKernelGGL((hc_diffused_var_kernel<T, 2>), dim3(blocks), dim3(256), 0, s, A, phi, out, homog);
  else
    hipLaunchKernelGGL((hc_diffused_var_kernel<T, 3>), dim3(blocks), dim3(256), 0, s, A, phi, out, homog);
  return hipGetLastError();
}

template hipError_t hc_relax<double>(const HCArgs<double>&, double* const[2], int*, const double*, int, int,
                                     hipStream_t, const LaunchTimer*);
template hipError_t hc_relax_tiles<double>(const HCArgs<double>&, double* const[2], int*, const double*, int, int,
                                           const Geom*, const double*, const double*, hipStream_t,
                                           const LaunchTimer*);
template hipError_t hc_coarse_vcycle<double>(const CoarseDesc<double>*, int, int, hipStream_t);
template hipError_t hc_residual<double>(const HCArgs<double>&, double*, const double*, double*,
                                        unsigned long long*, int, hipStream_t, const CtlFuse&);
template hipError_t hc_apply_op<double>(const HCArgs<double>&, double*, double*, const double*, int, hipStream_t);
template hipError_t hc_apply_batch<double>(int, const HCApplyJob<double>*, const unsigned*, int, hipStream_t);
template hipError_t hc_restrict_residual<double>(const HCArgs<double>&, const Geom&, double*, const double*,
                                                 double*, int, hipStream_t);
template hipError_t hc_restrict_all<double>(const HCArgs<double>&, const Geom&, double*, const double*, double*,
                                            double*, double*, int, hipStream_t);
template hipError_t hc_diffused_var<double>(const HCArgs<double>&, const double*, double*, int, hipStream_t);

}  // namespace ml
