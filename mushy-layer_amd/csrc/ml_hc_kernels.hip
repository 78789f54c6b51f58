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
#include "ml_common.h"
#include "ml_kernels.h"

namespace ml {

namespace {

constexpr int kThreads = 256;
constexpr int kWGThreads = 1024;   // single-workgroup kernels

enum : int { kResid = 1, kApply = 2, kRestrict = 3 };

// Position classification of an (unwrapped) index vector of a tile's
// extended region: in-domain (after periodic wrap), one-cell ghost of a
// non-periodic wall (exactly one direction outside), or invalid.
struct Pos {
  int i, j, k;   // wrapped coordinates; for a ghost: the interior neighbour
  int kind;      // 0 invalid, 1 in-domain, 2 wall ghost
  int gd, gs;    // ghost direction and side (0 lo, 1 hi)
};

template <int DIM>
__device__ inline Pos classify(const Geom& g, int ci, int cj, int ck) {
  int c[3] = {ci, cj, ck};
  Pos P;
  P.gd = 0;
  P.gs = 0;
  int nout = 0, far = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int n = g.n[d];
    if (c[d] < 0 || c[d] >= n) {
      if (g.periodic[d]) {
        c[d] = ((c[d] % n) + n) % n;
      } else if (c[d] == -1) {
        P.gd = d; P.gs = 0; c[d] = 0; ++nout;
      } else if (c[d] == n) {
        P.gd = d; P.gs = 1; c[d] = n - 1; ++nout;
      } else {
        far = 1;
      }
    }
  }
  if (DIM == 2) c[2] = 0;
  P.i = c[0];
  P.j = c[1];
  P.k = c[2];
  P.kind = (far || nout > 1) ? 0 : (nout == 1 ? 2 : 1);
  return P;
}

// Face coefficients around a cell, both components: bf[n][2d+s] is the
// lo (s = 0) / hi (s = 1) face of direction d.
template <typename T, int DIM>
struct CellFaces {
  T v[2][2 * DIM];
};

template <typename T, int DIM>
__device__ inline void load_faces(const HCArgs<T>& A, int i, int j, int k, CellFaces<T, DIM>& F) {
  const Geom& g = A.g;
#pragma unroll
  for (int n = 0; n < 2; ++n)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const T* bd = A.b[d] + (long long)n * g.nface[d];
      F.v[n][2 * d] = bd[g.fi(d, i, j, k)];
      F.v[n][2 * d + 1] = bd[g.fi(d, i + (d == 0), j + (d == 1), k + (d == 2))];
    }
}

// lambda = 1/(alpha*a + sum_d scale*beta*(b_hi + b_lo))
// (src/AMRNonLinearMultiCompOp.cpp:540-573, SUMFACESNLVC).
template <typename T, int DIM>
__device__ inline T lambda_of(const CellFaces<T, DIM>& F, int n, T aq, T alpha, T beta, T scale) {
  T lam = aq * alpha;
#pragma unroll
  for (int d = 0; d < DIM; ++d) lam = lam + scale * beta * (F.v[n][2 * d + 1] + F.v[n][2 * d]);
  return T(1) / lam;
}

// The L(phi) bracket (src/AMRNonLinearMultiCompOpF.ChF:125-160) from the
// seven stencil values p[0] = centre, p[2d+1] = +e_d, p[2d+2] = -e_d.
template <typename T, int DIM>
__device__ inline T lof_of(const CellFaces<T, DIM>& F, int n, const T* p, T aq, T phic, T alpha, T beta,
                           T dxinv) {
  const T p0 = p[0];
  T s = F.v[n][1] * (p[1] - p0) - F.v[n][0] * (p0 - p[2]) + F.v[n][3] * (p[3] - p0) - F.v[n][2] * (p0 - p[4]);
  if (DIM == 3) s = s + F.v[n][5] * (p[5] - p0) - F.v[n][4] * (p0 - p[6]);
  return alpha * aq * phic - beta * s * dxinv;
}

// Extended-region bookkeeping of one tile.
template <int DIM, int TX, int TY, int TZ>
struct Tile {
  int H, EX, EY, EZ, EXY, E;
  int x0, y0, z0;
  __device__ inline void init(int halo, int ntx, int nty) {
    H = halo;
    EX = TX + 2 * H;
    EY = TY + 2 * H;
    EZ = DIM == 3 ? TZ + 2 * H : 1;
    EXY = EX * EY;
    E = EXY * EZ;
    const int b = blockIdx.x;
    const int tx = b % ntx, r = b / ntx, ty = r % nty, tz = r / nty;
    x0 = tx * TX;
    y0 = ty * TY;
    z0 = DIM == 3 ? tz * TZ : 0;
  }
  __device__ inline void coords(int e, int& ex, int& ey, int& ez) const {
    ex = e % EX;
    const int r = e / EX;
    ey = r % EY;
    ez = r / EY;
  }
  // within the core grown by `ext` cells
  __device__ inline bool within(int ex, int ey, int ez, int ext) const {
    const int ox = ex - H, oy = ey - H, oz = ez - (DIM == 3 ? H : 0);
    bool ok = ox >= -ext && ox < TX + ext && oy >= -ext && oy < TY + ext;
    if (DIM == 3) ok = ok && oz >= -ext && oz < TZ + ext;
    return ok;
  }
  __device__ inline Pos pos(const Geom& g, int ex, int ey, int ez) const {
    return classify<DIM>(g, x0 - H + ex, y0 - H + ey, DIM == 3 ? z0 - H + ez : 0);
  }
};

// Gather the 2*DIM+1 phi_l stencil values of LDS position e.
template <typename T, int DIM>
__device__ inline void gather_lds(const T* pl, int e, int EX, int EXY, T* p) {
  p[0] = pl[e];
  p[1] = pl[e + 1];
  p[2] = pl[e - 1];
  p[3] = pl[e + EX];
  p[4] = pl[e - EX];
  if (DIM == 3) {
    p[5] = pl[e + EXY];
    p[6] = pl[e - EXY];
  }
}

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
// NONLINEARSMOOTHING, src/AMRNonLinearMultiCompOpF.ChF:125-195: phi_l
// recomputed before each pass, frozen within it).  The right-hand side and
// face coefficients of the cells a thread updates are loaded into registers
// before the closure of the pass, so their latency hides behind it.  When
// io.pc is set, the staged state first receives the linear prolongation of
// the coarse correction pc - pc0 (prolongIncrement fused with the
// post-smoother).
template <typename T, int DIM, int TX, int TY, int TZ, int HMAX>
__global__ __launch_bounds__(kThreads) void hc_relax_tile_kernel(HCArgs<T> A, TileIO<T> io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int EMAX = (TX + 2 * HMAX) * (TY + 2 * HMAX) * (DIM == 3 ? TZ + 2 * HMAX : 1);
  constexpr int KMAX = (EMAX + kThreads - 1) / kThreads;
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init(io.npass, io.ntx, io.nty);
  const int E = t.E;
  T* sphi = reinterpret_cast<T*>(smem_raw);
  T* sphl = sphi + 2 * E;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  const T scale = dxinv;
  const Geom gc = io.gc;

  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int ex, ey, ez;
    t.coords(e, ex, ey, ez);
    const Pos P = t.pos(g, ex, ey, ez);
    if (P.kind == 1) {
      const long long q = g.ci(P.i, P.j, P.k);
      T h = io.phi[q], c = io.phi[q + g.ncell];
      if (io.pc) {
        h = prolong_value<T, DIM>(gc, P.i, P.j, P.k, io.pc, io.pc0, h);
        c = prolong_value<T, DIM>(gc, P.i, P.j, P.k, io.pc + gc.ncell, io.pc0 + gc.ncell, c);
      }
      sphi[e] = h;
      sphi[E + e] = c;
    }
  }
  __syncthreads();

  for (int p = 0; p < io.npass; ++p) {
    const int gext = t.H - p;
    const int colour = (io.first + p) & 1;
    // prefetch the data of the cells this thread updates in this pass
    CellFaces<T, DIM> F[KMAX];
    T rr[KMAX][2], aa[KMAX][2];
    bool upd[KMAX];
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      const int e = threadIdx.x + kk * kThreads;
      upd[kk] = false;
      if (e >= E) continue;
      int ex, ey, ez;
      t.coords(e, ex, ey, ez);
      if (!t.within(ex, ey, ez, gext - 1)) continue;
      const Pos P = t.pos(g, ex, ey, ez);
      if (P.kind != 1 || ((P.i + P.j + P.k + colour) & 1)) continue;
      upd[kk] = true;
      const long long q = g.ci(P.i, P.j, P.k);
      load_faces<T, DIM>(A, P.i, P.j, P.k, F[kk]);
      rr[kk][0] = io.rhs[q];
      rr[kk][1] = io.rhs[q + g.ncell];
      aa[kk][0] = A.a ? A.a[q] : T(1);
      aa[kk][1] = A.a ? A.a[q + g.ncell] : T(1);
    }
    tile_diffused_var<T, DIM, TX, TY, TZ>(A, t, sphi, sphl, gext, 0);
#pragma unroll
    for (int kk = 0; kk < KMAX; ++kk) {
      if (!upd[kk]) continue;
      const int e = threadIdx.x + kk * kThreads;
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        T st[7];
        gather_lds<T, DIM>(sphl + n * E, e, t.EX, t.EXY, st);
        const T ph = sphi[n * E + e];
        const T lof = lof_of<T, DIM>(F[kk], n, st, aa[kk][n], ph, A.alpha, A.beta, dxinv);
        const T lam = lambda_of<T, DIM>(F[kk], n, aa[kk][n], A.alpha, A.beta, scale);
        sphi[n * E + e] = ph - lam * (lof - rr[kk][n]);
      }
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int ex, ey, ez;
    t.coords(e, ex, ey, ez);
    if (!t.within(ex, ey, ez, 0)) continue;
    const int gi = t.x0 - t.H + ex, gj = t.y0 - t.H + ey, gk = DIM == 3 ? t.z0 - t.H + ez : 0;
    if (gi >= g.n[0] || gj >= g.n[1] || (DIM == 3 && gk >= g.n[2])) continue;
    const long long q = g.ci(gi, gj, gk);
    io.phi[q] = sphi[e];
    io.phi[q + g.ncell] = sphi[E + e];
  }
}

// Residual (with max-norm), operator application, and FAS restriction of
// one tile with a one-cell halo.
template <typename T, int DIM, int TX, int TY, int TZ, int MODE>
__global__ __launch_bounds__(kThreads) void hc_tile_kernel(HCArgs<T> A, TileIO<T> io) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init(1, io.ntx, io.nty);
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
    if (MODE == kResid && io.norm) block_max_atomic<T>(mx, nan, io.norm);
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

// ---------------------------------------------------------------------------
// Single-workgroup building blocks over a whole (small) level: cell c of the
// level is owned by thread c (N <= kWGThreads), phi and phi_l live in LDS,
// neighbours are found by periodic wrap or the derived-variable wall rule.
template <typename T, int DIM>
struct WGLevel {
  const LevelDesc<T>* L;
  int N, i, j, k;
  bool own;
  __device__ inline void init(const LevelDesc<T>* lv) {
    L = lv;
    const Geom& g = L->g;
    N = g.n[0] * g.n[1] * (DIM == 3 ? g.n[2] : 1);
    const int c = threadIdx.x;
    own = c < N;
    i = c % g.n[0];
    const int r = c / g.n[0];
    j = r % g.n[1];
    k = DIM == 3 ? r / g.n[1] : 0;
  }
};

template <typename T, int DIM>
__device__ inline void wg_stencil(const Geom& g, const BCDesc& bc, const T* pl, int c, int i, int j, int k, int n,
                                  int nx, int nxy, T* p) {
  const int iv[3] = {i, j, k};
  p[0] = pl[c];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int nd = g.n[d];
    const int st = d == 0 ? 1 : (d == 1 ? nx : nxy);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int v = iv[d] + (s ? 1 : -1);
      T val;
      if (v >= 0 && v < nd) {
        val = pl[c + (s ? st : -st)];
      } else if (g.periodic[d]) {
        val = pl[c + (s ? -(nd - 1) * st : (nd - 1) * st)];
      } else {
        const T x1 = pl[c];
        val = x1;
        if (bc.apply) val = bc_ghost<T>(bc.type[d][s][n], T(bc.val[d][s][n]), x1, s ? 1 : -1, T(g.dx));
      }
      p[1 + 2 * d + (s ? 0 : 1)] = val;
    }
  }
}

template <typename T, int DIM>
__device__ inline void wg_closure(const ClosureP<T>& cp, const T* sphi, T* sphl, int N) {
  const int c = threadIdx.x;
  if (c < N) {
    T th, cl;
    closure<T>(sphi[c], sphi[N + c], cp, th, cl);
    sphl[c] = th;
    sphl[N + c] = cl;
  }
}

// L(phi) of the owned cell (both components) from phi/phi_l in LDS.
template <typename T, int DIM>
__device__ inline void wg_lof(const CoarseDesc<T>& D, const WGLevel<T, DIM>& W, const CellFaces<T, DIM>& F,
                              const T* aq, const T* sphi, const T* sphl, T* lof) {
  const Geom& g = W.L->g;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  const int c = threadIdx.x;
#pragma unroll
  for (int n = 0; n < 2; ++n) {
    T st[7];
    wg_stencil<T, DIM>(g, D.bc, sphl + n * W.N, c, W.i, W.j, W.k, n, g.n[0], g.n[0] * g.n[1], st);
    lof[n] = lof_of<T, DIM>(F, n, st, aq[n], sphi[n * W.N + c], D.alpha, D.beta, dxinv);
  }
}

template <typename T, int DIM>
__device__ inline void wg_load_cell(const CoarseDesc<T>& D, const WGLevel<T, DIM>& W, CellFaces<T, DIM>& F,
                                    T* aq) {
  HCArgs<T> A;
  A.g = W.L->g;
  for (int d = 0; d < 3; ++d) A.b[d] = W.L->b[d];
  load_faces<T, DIM>(A, W.i, W.j, W.k, F);
  const long long q = W.L->g.ci(W.i, W.j, W.k);
  aq[0] = W.L->a ? W.L->a[q] : T(1);
  aq[1] = W.L->a ? W.L->a[q + W.L->g.ncell] : T(1);
}

// relax: npass colour passes of the whole level, phi kept in LDS on exit.
template <typename T, int DIM>
__device__ void wg_relax(const CoarseDesc<T>& D, const WGLevel<T, DIM>& W, T* sphi, T* sphl, int npass,
                         bool load) {
  const Geom& g = W.L->g;
  const int c = threadIdx.x;
  const long long q = W.own ? g.ci(W.i, W.j, W.k) : 0;
  CellFaces<T, DIM> F;
  T aq[2] = {T(1), T(1)}, rr[2] = {T(0), T(0)}, lam[2] = {T(0), T(0)};
  if (W.own) {
    wg_load_cell<T, DIM>(D, W, F, aq);
    rr[0] = W.L->rhs[q];
    rr[1] = W.L->rhs[q + g.ncell];
    const T scale = T(1) / (T(g.dx) * T(g.dx));
    lam[0] = lambda_of<T, DIM>(F, 0, aq[0], D.alpha, D.beta, scale);
    lam[1] = lambda_of<T, DIM>(F, 1, aq[1], D.alpha, D.beta, scale);
    if (load) {
      sphi[c] = W.L->phi[q];
      sphi[W.N + c] = W.L->phi[q + g.ncell];
    }
  }
  __syncthreads();
  for (int p = 0; p < npass; ++p) {
    wg_closure<T, DIM>(D.cp, sphi, sphl, W.N);
    __syncthreads();
    if (W.own && !((W.i + W.j + W.k + p) & 1)) {
      T lof[2];
      wg_lof<T, DIM>(D, W, F, aq, sphi, sphl, lof);
      sphi[c] = sphi[c] - lam[0] * (lof[0] - rr[0]);
      sphi[W.N + c] = sphi[W.N + c] - lam[1] * (lof[1] - rr[1]);
    }
    __syncthreads();
  }
  if (W.own) {
    W.L->phi[q] = sphi[c];
    W.L->phi[q + g.ncell] = sphi[W.N + c];
  }
  __syncthreads();
}

// Coarse part of the FAS V-cycle (levels l0 .. nlev-1) in one workgroup:
//   rhs_l0 += L(phi_l0)  (FAS coarse rhs; the restriction was done by the
//   finer level's kernel);  then for l = l0 .. nlev-2: relax nu1; restrict
//   phi and rhs - L phi; rhs_{l+1} += L(phi_{l+1});  bottom relax;  then back
//   up: phi_l += P(phi_{l+1} - phi0_{l+1}); relax nu2.
template <typename T, int DIM>
__global__ __launch_bounds__(kWGThreads) void hc_coarse_vcycle_kernel(const CoarseDesc<T>* __restrict__ Dp,
                                                                      int apply_first) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const CoarseDesc<T>& D = *Dp;
  T* sphi = reinterpret_cast<T*>(smem_raw);
  T* sphl = sphi + 2 * kWGThreads;
  const int l0 = D.l0, L = D.nlev;
  const int c = threadIdx.x;
  const T denom = DIM == 3 ? T(8) : T(4);

  // rhs_l += L_l(phi_l): the FAS coarse right-hand side of level l
  auto apply_add = [&](int l) {
    WGLevel<T, DIM> W;
    W.init(&D.lev[l]);
    const Geom& g = W.L->g;
    const long long q = W.own ? g.ci(W.i, W.j, W.k) : 0;
    if (W.own) {
      sphi[c] = W.L->phi[q];
      sphi[W.N + c] = W.L->phi[q + g.ncell];
    }
    __syncthreads();
    wg_closure<T, DIM>(D.cp, sphi, sphl, W.N);
    __syncthreads();
    if (W.own) {
      CellFaces<T, DIM> F;
      T aq[2], lof[2];
      wg_load_cell<T, DIM>(D, W, F, aq);
      wg_lof<T, DIM>(D, W, F, aq, sphi, sphl, lof);
      W.L->rhs[q] = W.L->rhs[q] + lof[0];
      W.L->rhs[q + g.ncell] = W.L->rhs[q + g.ncell] + lof[1];
    }
    __syncthreads();
  };

  if (apply_first) apply_add(l0);

  for (int l = l0; l < L - 1; ++l) {
    WGLevel<T, DIM> W;
    W.init(&D.lev[l]);
    wg_relax<T, DIM>(D, W, sphi, sphl, 2 * D.nu1, true);
    // restrict phi and rhs - L(phi) (phi is still in LDS)
    wg_closure<T, DIM>(D.cp, sphi, sphl, W.N);
    __syncthreads();
    {
      const Geom& g = W.L->g;
      const LevelDesc<T>& C = D.lev[l + 1];
      const Geom& gcc = C.g;
      const int NC = gcc.n[0] * gcc.n[1] * (DIM == 3 ? gcc.n[2] : 1);
      if (c < NC) {
        const int I = c % gcc.n[0], r = c / gcc.n[0], J = r % gcc.n[1], K = DIM == 3 ? r / gcc.n[1] : 0;
        T s[2] = {T(0), T(0)}, sp[2] = {T(0), T(0)};
        const T dxinv = T(1) / (T(g.dx) * T(g.dx));
        for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
          for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) {
              const int fi = 2 * I + ii, fj = 2 * J + jj, fk = DIM == 3 ? 2 * K + kk : 0;
              const int fc = fi + g.n[0] * (fj + g.n[1] * fk);
              HCArgs<T> A;
              A.g = g;
              for (int d = 0; d < 3; ++d) A.b[d] = W.L->b[d];
              CellFaces<T, DIM> F;
              load_faces<T, DIM>(A, fi, fj, fk, F);
              const long long q = g.ci(fi, fj, fk);
#pragma unroll
              for (int n = 0; n < 2; ++n) {
                T st[7];
                wg_stencil<T, DIM>(g, D.bc, sphl + n * W.N, fc, fi, fj, fk, n, g.n[0], g.n[0] * g.n[1], st);
                const T aq = W.L->a ? W.L->a[q + n * g.ncell] : T(1);
                const T lof = lof_of<T, DIM>(F, n, st, aq, sphi[n * W.N + fc], D.alpha, D.beta, dxinv);
                s[n] = s[n] + (W.L->rhs[q + n * g.ncell] - lof) / denom;
                sp[n] = sp[n] + sphi[n * W.N + fc] / denom;
              }
            }
        const long long qc = gcc.ci(I, J, K);
#pragma unroll
        for (int n = 0; n < 2; ++n) {
          C.rhs[qc + n * gcc.ncell] = s[n];
          C.phi[qc + n * gcc.ncell] = sp[n];
          C.phi0[qc + n * gcc.ncell] = sp[n];
        }
      }
    }
    __syncthreads();
    apply_add(l + 1);
  }
  {
    WGLevel<T, DIM> W;
    W.init(&D.lev[L - 1]);
    wg_relax<T, DIM>(D, W, sphi, sphl, 2 * D.bottom, true);
  }
  for (int l = L - 2; l >= l0; --l) {
    WGLevel<T, DIM> W;
    W.init(&D.lev[l]);
    const LevelDesc<T>& C = D.lev[l + 1];
    if (W.own) {
      const long long q = W.L->g.ci(W.i, W.j, W.k);
      const T h = prolong_value<T, DIM>(C.g, W.i, W.j, W.k, C.phi, C.phi0, W.L->phi[q]);
      const T cc = prolong_value<T, DIM>(C.g, W.i, W.j, W.k, C.phi + C.g.ncell, C.phi0 + C.g.ncell,
                                         W.L->phi[q + W.L->g.ncell]);
      sphi[c] = h;
      sphi[W.N + c] = cc;
    }
    __syncthreads();
    wg_relax<T, DIM>(D, W, sphi, sphl, 2 * D.nu2, false);
  }
}

// Whole level in one workgroup for the per-operator relax entry point.
template <typename T, int DIM>
__global__ __launch_bounds__(kWGThreads) void hc_relax_small_kernel(const CoarseDesc<T>* __restrict__ Dp, int l,
                                                                    int npass) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const CoarseDesc<T>& D = *Dp;
  T* sphi = reinterpret_cast<T*>(smem_raw);
  T* sphl = sphi + 2 * kWGThreads;
  WGLevel<T, DIM> W;
  W.init(&D.lev[l]);
  wg_relax<T, DIM>(D, W, sphi, sphl, npass, true);
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

// passes per relax launch: halo 2 (one red/black sweep) in 2D and 3D
template <typename T>
int max_tile_passes(int dim) {
  (void)dim;
  return 2;
}

template <typename T>
bool small_level_fits(const Geom& g) {
  return g.nvalid() <= kWGThreads;
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

template <typename T, int DIM>
static hipError_t launch_relax_tile(const HCArgs<T>& A, TileIO<T> io, hipStream_t s) {
  using D = TileDims<T, DIM>;
  const int nt = tiles_of<T, DIM>(A.g, io.ntx, io.nty);
  if (io.npass < 1 || io.npass > 2) return hipErrorInvalidValue;
  const size_t lds = tile_lds<T, DIM>(io.npass, false);
  if (lds > kMaxLds) return hipErrorInvalidValue;
  hipLaunchKernelGGL((hc_relax_tile_kernel<T, DIM, D::TX, D::TY, D::TZ, 2>), dim3(nt), dim3(kThreads), lds, s, A,
                     io);
  return hipGetLastError();
}

template <typename T>
hipError_t hc_relax_tiles(const HCArgs<T>& A, T* phi, const T* rhs, int npass, int first, const Geom* gc,
                          const T* pc, const T* pc0, hipStream_t s) {
  int done = 0;
  while (done < npass) {
    TileIO<T> io{};
    io.phi = phi;
    io.rhs = rhs;
    io.npass = (npass - done) < 2 ? (npass - done) : 2;
    io.first = (first + done) & 1;
    if (done == 0 && pc) {
      io.gc = *gc;
      io.pc = pc;
      io.pc0 = pc0;
    }
    hipError_t e = A.g.dim == 2 ? launch_relax_tile<T, 2>(A, io, s) : launch_relax_tile<T, 3>(A, io, s);
    if (e != hipSuccess) return e;
    done += io.npass;
  }
  return hipSuccess;
}

template <typename T>
hipError_t hc_relax(const HCArgs<T>& A, T* phi, const T* rhs, int npass, int first, hipStream_t s) {
  if (npass <= 0) return hipSuccess;
  return hc_relax_tiles<T>(A, phi, rhs, npass, first, nullptr, nullptr, nullptr, s);
}

template <typename T>
hipError_t hc_relax_small(const CoarseDesc<T>* dev_desc, int dim, int level, int npass, hipStream_t s) {
  const size_t lds = size_t(4) * kWGThreads * sizeof(T);
  if (dim == 2)
    hipLaunchKernelGGL((hc_relax_small_kernel<T, 2>), dim3(1), dim3(kWGThreads), lds, s, dev_desc, level, npass);
  else
    hipLaunchKernelGGL((hc_relax_small_kernel<T, 3>), dim3(1), dim3(kWGThreads), lds, s, dev_desc, level, npass);
  return hipGetLastError();
}

template <typename T>
hipError_t hc_coarse_vcycle(const CoarseDesc<T>* dev_desc, int dim, int apply_first, hipStream_t s) {
  const size_t lds = size_t(4) * kWGThreads * sizeof(T);
  if (dim == 2)
    hipLaunchKernelGGL((hc_coarse_vcycle_kernel<T, 2>), dim3(1), dim3(kWGThreads), lds, s, dev_desc, apply_first);
  else
    hipLaunchKernelGGL((hc_coarse_vcycle_kernel<T, 3>), dim3(1), dim3(kWGThreads), lds, s, dev_desc, apply_first);
  return hipGetLastError();
}

template <typename T>
hipError_t hc_residual(const HCArgs<T>& A, T* phi, const T* rhs, T* res, unsigned long long* norm,
                       int homog, hipStream_t s) {
  TileIO<T> io{};
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

template hipError_t hc_relax<double>(const HCArgs<double>&, double*, const double*, int, int, hipStream_t);
template hipError_t hc_relax_tiles<double>(const HCArgs<double>&, double*, const double*, int, int, const Geom*,
                                           const double*, const double*, hipStream_t);
template hipError_t hc_relax_small<double>(const CoarseDesc<double>*, int, int, int, hipStream_t);
template hipError_t hc_coarse_vcycle<double>(const CoarseDesc<double>*, int, int, hipStream_t);
template hipError_t hc_residual<double>(const HCArgs<double>&, double*, const double*, double*,
                                        unsigned long long*, int, hipStream_t);
template hipError_t hc_apply_op<double>(const HCArgs<double>&, double*, double*, const double*, int, hipStream_t);
template hipError_t hc_restrict_residual<double>(const HCArgs<double>&, const Geom&, double*, const double*,
                                                 double*, int, hipStream_t);
template hipError_t hc_restrict_all<double>(const HCArgs<double>&, const Geom&, double*, const double*, double*,
                                            double*, double*, int, hipStream_t);
template hipError_t hc_diffused_var<double>(const HCArgs<double>&, const double*, double*, int, hipStream_t);
template int max_tile_passes<double>(int);
template bool small_level_fits<double>(const Geom&);

}  // namespace ml
