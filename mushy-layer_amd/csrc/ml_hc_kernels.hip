// ml_hc_kernels.hip — HIP kernels for gfx950 of the nonlinear (H, Theta)
// operator of AMRNonLinearMultiCompOp: fused closure + red/black smoothing,
// residual with max-norm, FAS restricted residual, operator application, and
// the level transfers.  Every kernel recomputes the phase-diagram closure from
// (H, Theta) tiles staged in LDS, so theta/Theta_l are never written to HBM.
//
// Arithmetic follows the reference's evaluation order with contraction off
// (-ffp-contract=off), so the FP64 instantiation reproduces the CPU oracle
// bit for bit.
#include "ml_common.h"
#include "ml_kernels.h"

namespace ml {

namespace {

constexpr int kThreads = 256;

enum : int { kRelax = 0, kResid = 1, kApply = 2, kRestrict = 3 };

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

// lambda = 1/(alpha*a + sum_d scale*beta*(b_hi + b_lo))
// (src/AMRNonLinearMultiCompOp.cpp:540-573, SUMFACESNLVC).
template <typename T, int DIM>
__device__ inline T hc_lambda(const HCArgs<T>& A, int n, int i, int j, int k, long long q) {
  const Geom& g = A.g;
  const T scale = T(1) / (T(g.dx) * T(g.dx));
  T lam = (A.a ? A.a[q] : T(1)) * A.alpha;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const T* bd = A.b[d] + (long long)n * g.nface[d];
    T s = bd[g.fi(d, i + (d == 0), j + (d == 1), k + (d == 2))] + bd[g.fi(d, i, j, k)];
    lam = lam + scale * A.beta * s;
  }
  return T(1) / lam;
}

// The L(phi) bracket (src/AMRNonLinearMultiCompOpF.ChF:125-160): stencil of
// phi_l staged in LDS at position e with strides (1, EX, EXY).
template <typename T, int DIM>
__device__ inline T hc_lof(const HCArgs<T>& A, int n, int i, int j, int k, long long q, T phic,
                           const T* pl, int e, int EX, int EXY, T dxinv) {
  const Geom& g = A.g;
  const T p0 = pl[e];
  const T p1 = pl[e + 1], p2 = pl[e - 1];
  const T p3 = pl[e + EX], p4 = pl[e - EX];
  const T* b0 = A.b[0] + (long long)n * g.nface[0];
  const T* b1 = A.b[1] + (long long)n * g.nface[1];
  T s = b0[g.fi(0, i + 1, j, k)] * (p1 - p0) - b0[g.fi(0, i, j, k)] * (p0 - p2) +
        b1[g.fi(1, i, j + 1, k)] * (p3 - p0) - b1[g.fi(1, i, j, k)] * (p0 - p4);
  if (DIM == 3) {
    const T p5 = pl[e + EXY], p6 = pl[e - EXY];
    const T* b2 = A.b[2] + (long long)n * g.nface[2];
    s = s + b2[g.fi(2, i, j, k + 1)] * (p5 - p0) - b2[g.fi(2, i, j, k)] * (p0 - p6);
  }
  const T aq = A.a ? A.a[q] : T(1);
  return A.alpha * aq * phic - A.beta * s * dxinv;
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

template <typename T, int DIM, int TX, int TY, int TZ, int MODE>
__global__ __launch_bounds__(kThreads) void hc_tile_kernel(HCArgs<T> A, T* __restrict__ phi,
                                                           const T* __restrict__ rhs, T* out,
                                                           const T* add, unsigned long long* norm,
                                                           Geom gc, int npass, int first,
                                                           int homog, int ntx, int nty) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const Geom& g = A.g;
  Tile<DIM, TX, TY, TZ> t;
  t.init(MODE == kRelax ? npass : 1, ntx, nty);
  const int E = t.E;
  T* sphi = reinterpret_cast<T*>(smem_raw);
  T* sphl = sphi + 2 * E;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));

  // stage (H, Theta) over the extended tile
  for (int e = threadIdx.x; e < E; e += blockDim.x) {
    int ex, ey, ez;
    t.coords(e, ex, ey, ez);
    const Pos P = t.pos(g, ex, ey, ez);
    if (P.kind == 1) {
      const long long q = g.ci(P.i, P.j, P.k);
      sphi[e] = phi[q];
      sphi[E + e] = phi[q + g.ncell];
    }
  }
  __syncthreads();

  if (MODE == kRelax) {
    // NONLINEARSMOOTHING passes (src/AMRNonLinearMultiCompOpF.ChF:125-195):
    // phi_l recomputed before each pass and frozen within it.
    for (int p = 0; p < npass; ++p) {
      const int gext = t.H - p;
      const int colour = (first + p) & 1;
      tile_diffused_var<T, DIM, TX, TY, TZ>(A, t, sphi, sphl, gext, 0);
      for (int e = threadIdx.x; e < E; e += blockDim.x) {
        int ex, ey, ez;
        t.coords(e, ex, ey, ez);
        if (!t.within(ex, ey, ez, gext - 1)) continue;
        const Pos P = t.pos(g, ex, ey, ez);
        if (P.kind != 1 || ((P.i + P.j + P.k + colour) & 1)) continue;
        const long long q0 = g.ci(P.i, P.j, P.k);
#pragma unroll
        for (int n = 0; n < 2; ++n) {
          const long long q = q0 + n * g.ncell;
          const T ph = sphi[n * E + e];
          const T lof = hc_lof<T, DIM>(A, n, P.i, P.j, P.k, q, ph, sphl + n * E, e, t.EX, t.EXY, dxinv);
          const T lam = hc_lambda<T, DIM>(A, n, P.i, P.j, P.k, q);
          sphi[n * E + e] = ph - lam * (lof - rhs[q]);
        }
      }
      __syncthreads();
    }
    // write back the core
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
      int ex, ey, ez;
      t.coords(e, ex, ey, ez);
      if (!t.within(ex, ey, ez, 0)) continue;
      const int gi = t.x0 - t.H + ex, gj = t.y0 - t.H + ey, gk = DIM == 3 ? t.z0 - t.H + ez : 0;
      if (gi >= g.n[0] || gj >= g.n[1] || (DIM == 3 && gk >= g.n[2])) continue;
      const long long q = g.ci(gi, gj, gk);
      phi[q] = sphi[e];
      phi[q + g.ncell] = sphi[E + e];
    }
    return;
  }

  tile_diffused_var<T, DIM, TX, TY, TZ>(A, t, sphi, sphl, 1, homog);

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
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const long long q = q0 + n * g.ncell;
        const T lof = hc_lof<T, DIM>(A, n, gi, gj, gk, q, sphi[n * E + e], sphl + n * E, e, t.EX, t.EXY, dxinv);
        if (MODE == kResid) {
          const T r = rhs[q] - lof;
          if (out) out[q] = r;
          const T ar = fabs(r);
          if (ar != ar) nan = 1;
          else if (ar > mx) mx = ar;
        } else {
          out[q] = add ? add[q] + lof : lof;
        }
      }
    }
    if (MODE == kResid && norm) block_max_atomic<T>(mx, nan, norm);
    return;
  }

  if (MODE == kRestrict) {
    // RESTRICTRESNLVC (src/AMRNonLinearMultiCompOpF.ChF:452-542):
    // res_c(i/2) = 0 + sum over children in Fortran order of (rhs - L phi)/2^d
    T* sres = sphl + 2 * E;
    constexpr int NC = TX * TY * (DIM == 3 ? TZ : 1);
    const T denom = DIM == 3 ? T(8) : T(4);
    for (int c = threadIdx.x; c < NC; c += blockDim.x) {
      const int cx = c % TX, cr = c / TX, cy = cr % TY, cz = cr / TY;
      const int gi = t.x0 + cx, gj = t.y0 + cy, gk = DIM == 3 ? t.z0 + cz : 0;
      if (gi >= g.n[0] || gj >= g.n[1] || (DIM == 3 && gk >= g.n[2])) continue;
      const int e = (cx + t.H) + t.EX * (cy + t.H) + t.EXY * (DIM == 3 ? cz + t.H : 0);
      const long long q0 = g.ci(gi, gj, gk);
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const long long q = q0 + n * g.ncell;
        const T lof = hc_lof<T, DIM>(A, n, gi, gj, gk, q, sphi[n * E + e], sphl + n * E, e, t.EX, t.EXY, dxinv);
        sres[n * NC + c] = (rhs[q] - lof) / denom;
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
        T s = T(0);
        for (int kk = 0; kk < (DIM == 3 ? 2 : 1); ++kk)
          for (int jj = 0; jj < 2; ++jj)
            for (int ii = 0; ii < 2; ++ii) {
              const int f = (2 * cx + ii) + TX * ((2 * cy + jj) + TY * (DIM == 3 ? 2 * cz + kk : 0));
              s = s + sres[n * NC + f];
            }
        out[qc + n * gc.ncell] = s;
      }
    }
  }
}

// Whole level in one workgroup (coarse MG levels and the relax bottom solve):
// valid cells staged in LDS, neighbours by periodic wrap or the derived-var
// wall rule, `npass` colour passes separated by workgroup barriers.
template <typename T, int DIM>
__global__ __launch_bounds__(1024) void hc_relax_small_kernel(HCArgs<T> A, T* __restrict__ phi,
                                                              const T* __restrict__ rhs, int npass,
                                                              int first) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const Geom& g = A.g;
  const int nx = g.n[0], ny = g.n[1], nz = DIM == 3 ? g.n[2] : 1;
  const int N = nx * ny * nz;
  T* sphi = reinterpret_cast<T*>(smem_raw);
  T* sphl = sphi + 2 * N;
  const T dxinv = T(1) / (T(g.dx) * T(g.dx));
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    const int i = c % nx, r = c / nx, j = r % ny, k = r / ny;
    const long long q = g.ci(i, j, k);
    sphi[c] = phi[q];
    sphi[N + c] = phi[q + g.ncell];
  }
  __syncthreads();
  for (int p = 0; p < npass; ++p) {
    const int colour = (first + p) & 1;
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
      T th, cl;
      closure<T>(sphi[c], sphi[N + c], A.cp, th, cl);
      sphl[c] = th;
      sphl[N + c] = cl;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
      const int i = c % nx, r = c / nx, j = r % ny, k = r / ny;
      if ((i + j + k + colour) & 1) continue;
      const int iv[3] = {i, j, k};
      const long long q0 = g.ci(i, j, k);
#pragma unroll
      for (int n = 0; n < 2; ++n) {
        const T* pl = sphl + n * N;
        T nb[3][2];
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const int nd = g.n[d];
          const int st = d == 0 ? 1 : (d == 1 ? nx : nx * ny);
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            const int v = iv[d] + (s ? 1 : -1);
            if (v >= 0 && v < nd) {
              nb[d][s] = pl[c + (s ? st : -st)];
            } else if (g.periodic[d]) {
              nb[d][s] = pl[c + (s ? -(nd - 1) * st : (nd - 1) * st)];
            } else {
              const T x1 = pl[c];
              T gv = x1;
              if (A.bc.apply)
                gv = bc_ghost<T>(A.bc.type[d][s][n], T(A.bc.val[d][s][n]), x1, s ? 1 : -1, T(g.dx));
              nb[d][s] = gv;
            }
          }
        }
        const long long q = q0 + n * g.ncell;
        const T p0 = pl[c];
        const T* b0 = A.b[0] + (long long)n * g.nface[0];
        const T* b1 = A.b[1] + (long long)n * g.nface[1];
        T sum = b0[g.fi(0, i + 1, j, k)] * (nb[0][1] - p0) - b0[g.fi(0, i, j, k)] * (p0 - nb[0][0]) +
                b1[g.fi(1, i, j + 1, k)] * (nb[1][1] - p0) - b1[g.fi(1, i, j, k)] * (p0 - nb[1][0]);
        if (DIM == 3) {
          const T* b2 = A.b[2] + (long long)n * g.nface[2];
          sum = sum + b2[g.fi(2, i, j, k + 1)] * (nb[2][1] - p0) - b2[g.fi(2, i, j, k)] * (p0 - nb[2][0]);
        }
        const T aq = A.a ? A.a[q] : T(1);
        const T ph = sphi[n * N + c];
        const T lof = A.alpha * aq * ph - A.beta * sum * dxinv;
        const T lam = hc_lambda<T, DIM>(A, n, i, j, k, q);
        sphi[n * N + c] = ph - lam * (lof - rhs[q]);
      }
    }
    __syncthreads();
  }
  for (int c = threadIdx.x; c < N; c += blockDim.x) {
    const int i = c % nx, r = c / nx, j = r % ny, k = r / ny;
    const long long q = g.ci(i, j, k);
    phi[q] = sphi[c];
    phi[q + g.ncell] = sphi[N + c];
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

template <typename T>
int max_tile_passes(int dim) {
  int p = 1;
  while (p < 64) {
    size_t lds = dim == 2 ? tile_lds<T, 2>(p + 1, false) : tile_lds<T, 3>(p + 1, false);
    if (lds > kMaxLds) break;
    ++p;
  }
  return p;
}

template <typename T>
bool small_level_fits(const Geom& g) {
  return size_t(4) * g.nvalid() * sizeof(T) <= kMaxLds;
}

template <typename T, int DIM, int MODE>
static hipError_t launch_tile(const HCArgs<T>& A, T* phi, const T* rhs, T* out, const T* add,
                              unsigned long long* norm, const Geom& gc, int npass, int first,
                              int homog, hipStream_t s) {
  using D = TileDims<T, DIM>;
  int ntx, nty;
  const int nt = tiles_of<T, DIM>(A.g, ntx, nty);
  const size_t lds = tile_lds<T, DIM>(MODE == kRelax ? npass : 1, MODE == kRestrict);
  if (lds > kMaxLds) return hipErrorInvalidValue;
  hipLaunchKernelGGL((hc_tile_kernel<T, DIM, D::TX, D::TY, D::TZ, MODE>), dim3(nt), dim3(kThreads), lds, s,
                     A, phi, rhs, out, add, norm, gc, npass, first, homog, ntx, nty);
  return hipGetLastError();
}

template <typename T>
hipError_t hc_relax(const HCArgs<T>& A, T* phi, const T* rhs, int npass, int first, hipStream_t s) {
  if (npass <= 0) return hipSuccess;
  if (small_level_fits<T>(A.g)) {
    const size_t lds = size_t(4) * A.g.nvalid() * sizeof(T);
    int threads = 64;
    while (threads < 1024 && threads < A.g.nvalid()) threads *= 2;
    if (A.g.dim == 2)
      hipLaunchKernelGGL((hc_relax_small_kernel<T, 2>), dim3(1), dim3(threads), lds, s, A, phi, rhs, npass, first);
    else
      hipLaunchKernelGGL((hc_relax_small_kernel<T, 3>), dim3(1), dim3(threads), lds, s, A, phi, rhs, npass, first);
    return hipGetLastError();
  }
  const int pmax = max_tile_passes<T>(A.g.dim);
  Geom dummy = A.g;
  int done = 0;
  while (done < npass) {
    const int p = (npass - done) < pmax ? (npass - done) : pmax;
    const int f = (first + done) & 1;
    hipError_t e = A.g.dim == 2
                       ? launch_tile<T, 2, kRelax>(A, phi, rhs, nullptr, nullptr, nullptr, dummy, p, f, 0, s)
                       : launch_tile<T, 3, kRelax>(A, phi, rhs, nullptr, nullptr, nullptr, dummy, p, f, 0, s);
    if (e != hipSuccess) return e;
    done += p;
  }
  return hipSuccess;
}

template <typename T>
hipError_t hc_residual(const HCArgs<T>& A, T* phi, const T* rhs, T* res, unsigned long long* norm,
                       int homog, hipStream_t s) {
  Geom dummy = A.g;
  return A.g.dim == 2 ? launch_tile<T, 2, kResid>(A, phi, rhs, res, nullptr, norm, dummy, 0, 0, homog, s)
                      : launch_tile<T, 3, kResid>(A, phi, rhs, res, nullptr, norm, dummy, 0, 0, homog, s);
}

template <typename T>
hipError_t hc_apply_op(const HCArgs<T>& A, T* phi, T* lhs, const T* add, int homog, hipStream_t s) {
  Geom dummy = A.g;
  return A.g.dim == 2 ? launch_tile<T, 2, kApply>(A, phi, nullptr, lhs, add, nullptr, dummy, 0, 0, homog, s)
                      : launch_tile<T, 3, kApply>(A, phi, nullptr, lhs, add, nullptr, dummy, 0, 0, homog, s);
}

template <typename T>
hipError_t hc_restrict_residual(const HCArgs<T>& A, const Geom& gc, T* phi, const T* rhs, T* resc,
                                int homog, hipStream_t s) {
  return A.g.dim == 2 ? launch_tile<T, 2, kRestrict>(A, phi, rhs, resc, nullptr, nullptr, gc, 0, 0, homog, s)
                      : launch_tile<T, 3, kRestrict>(A, phi, rhs, resc, nullptr, nullptr, gc, 0, 0, homog, s);
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
template hipError_t hc_residual<double>(const HCArgs<double>&, double*, const double*, double*,
                                        unsigned long long*, int, hipStream_t);
template hipError_t hc_apply_op<double>(const HCArgs<double>&, double*, double*, const double*, int, hipStream_t);
template hipError_t hc_restrict_residual<double>(const HCArgs<double>&, const Geom&, double*, const double*,
                                                 double*, int, hipStream_t);
template hipError_t hc_diffused_var<double>(const HCArgs<double>&, const double*, double*, int, hipStream_t);
template int max_tile_passes<double>(int);
template bool small_level_fits<double>(const Geom&);

}  // namespace ml
