// ml_tile.h — device helpers shared by the tiled stencil kernels: position
// classification against periodic / wall boundaries, per-thread position
// records of a tile's extended region, face-coefficient gathers and the
// L(phi) bracket of the reference's Fortran stencils.
#pragma once

#include "ml_common.h"
#include "ml_kernels.h"

namespace ml {

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
__device__ inline void load_faces(const Geom& g, const T* const* b, int i, int j, int k, CellFaces<T, DIM>& F) {
#pragma unroll
  for (int n = 0; n < 2; ++n)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const T* bd = b[d] + (long long)n * g.nface[d];
      F.v[n][2 * d] = bd[g.fi(d, i, j, k)];
      F.v[n][2 * d + 1] = bd[g.fi(d, i + (d == 0), j + (d == 1), k + (d == 2))];
    }
}

template <typename T, int DIM>
__device__ inline void load_faces(const HCArgs<T>& A, int i, int j, int k, CellFaces<T, DIM>& F) {
  load_faces<T, DIM>(A.g, A.b, i, j, k, F);
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
  __device__ inline void init(int halo, int ntx, int nty) { init_at(halo, ntx, nty, blockIdx.x); }
  // tile b of the level (b != blockIdx.x when several grids share a launch)
  __device__ inline void init_at(int halo, int ntx, int nty, int b) {
    H = halo;
    EX = TX + 2 * H;
    EY = TY + 2 * H;
    EZ = DIM == 3 ? TZ + 2 * H : 1;
    EXY = EX * EY;
    E = EXY * EZ;
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

// Per-thread record of one LDS position of a tile's extended region.
struct PosRec {
  long long q;   // global cell index (in-domain) or of the interior neighbour (ghost)
  int e;         // LDS position
  int kind;      // 0 unused/invalid, 1 in-domain, 2 wall ghost
  int dist;      // distance outside the core (0 inside)
  int eint;      // ghost: LDS position of the interior neighbour
  int gd, gs;    // ghost direction / side
  int i, j, k;   // wrapped coordinates
  bool owner;    // core cell inside the domain: written back by this tile
};


// Fill the record of LDS position e of tile t.
template <int DIM, int TX, int TY, int TZ>
__device__ inline void pos_record(const Geom& g, const Tile<DIM, TX, TY, TZ>& t, int e, PosRec& r) {
  r.e = e;
  r.kind = 0;
  r.owner = false;
  r.dist = 1 << 20;
  if (e >= t.E) return;
  int ex, ey, ez;
  t.coords(e, ex, ey, ez);
  const int ox = ex - t.H, oy = ey - t.H, oz = DIM == 3 ? ez - t.H : 0;
  const int dxo = ox < 0 ? -ox : (ox >= TX ? ox - TX + 1 : 0);
  const int dyo = oy < 0 ? -oy : (oy >= TY ? oy - TY + 1 : 0);
  const int dzo = DIM == 3 ? (oz < 0 ? -oz : (oz >= TZ ? oz - TZ + 1 : 0)) : 0;
  r.dist = dxo > dyo ? (dxo > dzo ? dxo : dzo) : (dyo > dzo ? dyo : dzo);
  const int gi = t.x0 + ox, gj = t.y0 + oy, gk = DIM == 3 ? t.z0 + oz : 0;
  const Pos P = classify<DIM>(g, gi, gj, gk);
  r.kind = P.kind;
  r.i = P.i;
  r.j = P.j;
  r.k = P.k;
  r.gd = P.gd;
  r.gs = P.gs;
  r.q = g.ci(P.i, P.j, P.k);
  const int stride = P.gd == 0 ? 1 : (P.gd == 1 ? t.EX : t.EXY);
  r.eint = P.gs == 0 ? e + stride : e - stride;
  r.owner = r.dist == 0 && P.kind == 1 && gi < g.n[0] && gj < g.n[1] && (DIM == 2 || gk < g.n[2]);
}

// ---- single-workgroup kernels: one level held compactly in LDS (x fastest)
struct CLev {
  int n[3];
  int nx, nxy, N, off;   // off: the level's offset in the workgroup's LDS arrays
  double dx;
};

// q = x / d, r = x % d for x >= 0, d > 0: a shift and a mask when d is a
// power of two (every level of a power-of-two grid), integer division
// otherwise.  d is uniform across the workgroup, so the branch is too.
__device__ __forceinline__ void divmod_pos(int x, int d, int& q, int& r) {
  if ((d & (d - 1)) == 0) {
    const int s = __builtin_ctz(d);
    q = x >> s;
    r = x & (d - 1);
  } else {
    q = x / d;
    r = x - q * d;
  }
}

__device__ inline void clev_coords(const CLev& L, int c, int& i, int& j, int& k) {
  int r;
  divmod_pos(c, L.nx, r, i);
  divmod_pos(r, L.n[1], k, j);
}

__device__ inline void clev_set(CLev& L, const Geom& g, int off) {
  for (int d = 0; d < 3; ++d) L.n[d] = g.n[d];
  L.nx = g.n[0];
  L.nxy = g.n[0] * g.n[1];
  L.N = L.nxy * g.n[2];
  L.off = off;
  L.dx = g.dx;
}

// wavefront-scope barrier: orders the LDS accesses of the lanes of one wave
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// cells of a level owned by this thread: packed (i | j << 10 | k << 20), -1
// when none.  Even row length: the pair (2s + par, j, k), (2s + 1 - par, j, k)
// of opposite colours; otherwise cells t and t + NT.
struct COwn {
  int p[2];
};

__device__ inline int cpack(int i, int j, int k) { return i | (j << 10) | (k << 20); }
__device__ inline void cunpack(int v, int& i, int& j, int& k) {
  i = v & 1023;
  j = (v >> 10) & 1023;
  k = v >> 20;
}

// threads owning cells of level L
template <int NT>
__device__ inline int cown_active(const CLev& L) {
  return (L.nx & 1) == 0 ? (L.N >> 1) : (L.N < NT ? L.N : NT);
}

template <int NT>
__device__ inline void cown_set(const CLev& L, COwn& o) {
  const int t = threadIdx.x;
  o.p[0] = o.p[1] = -1;
  if ((L.nx & 1) == 0) {
    const int h = L.nx >> 1;
    if (t < (L.N >> 1)) {
      int s, r, j, k;
      divmod_pos(t, h, r, s);
      divmod_pos(r, L.n[1], k, j);
      const int par = (j + k) & 1;
      o.p[0] = cpack(2 * s + par, j, k);
      o.p[1] = cpack(2 * s + 1 - par, j, k);
    }
  } else {
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      const int c = t + m * NT;
      if (c >= L.N) continue;
      int i, j, k;
      clev_coords(L, c, i, j, k);
      o.p[m] = cpack(i, j, k);
    }
  }
}


// neighbour compact indices of cell (i,j,k) of level L, face f = 2d + s
// (s = 1: +e_d): periodic images wrap; across a wall the cell itself, with
// bit f of `wall` set
template <int DIM>
__device__ __forceinline__ void clev_nbrs(const CLev& L, const int* per, int i, int j, int k, int* nb, int& wall) {
  const int iv[3] = {i, j, k};
  const int c = i + L.nx * j + L.nxy * k;
  wall = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int nd = L.n[d];
    const int st = d == 0 ? 1 : (d == 1 ? L.nx : L.nxy);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const int v = iv[d] + (s ? 1 : -1);
      if (v >= 0 && v < nd) {
        nb[2 * d + s] = c + (s ? st : -st);
      } else if (per[d]) {
        nb[2 * d + s] = c + (s ? -(nd - 1) * st : (nd - 1) * st);
      } else {
        nb[2 * d + s] = c;
        wall |= 1 << (2 * d + s);
      }
    }
  }
}

// ---- streaming (plane-marching) kernels ----------------------------------
// A workgroup owns a block of columns (2D: TX cells in x; 3D: TX x TY cells
// in (x, y)) extended by a lateral halo of P cells, and marches along the
// last direction.  IC<N> passes a compile-time index to a generic lambda.
template <int N>
struct IC {
  static constexpr int value = N;
};

// lateral position of the extended block: global coordinates (wrapped in
// periodic directions), kind 1 cell / 2 wall ghost / 0 unused, Chebyshev
// distance to the core, owner flag, and for a ghost its wall (direction,
// side) and the LDS offset of its interior neighbour
struct SPos {
  int gx, gy;
  int kind, dist, owner;
  int gd, gs, eint;
  int gtype;   // a ghost's boundary type (resolved once: a per-thread index
               // into the kernel arguments would be a memory load)
};

template <int DIM, int TX, int TY, int P>
__device__ inline void spos_init(const Geom& g, int x0, int y0, int e, SPos& s) {
  constexpr int EX = TX + 2 * P;
  const int ex = e % EX, ey = e / EX;
  const int lc[2] = {ex - P, DIM == 3 ? ey - P : 0};
  const int org[2] = {x0, y0};
  const int nl = DIM == 3 ? 2 : 1;
  s.kind = 1;
  s.dist = 0;
  s.owner = 1;
  s.gd = -1;
  s.gs = 0;
  s.eint = e;
  int gv[2] = {0, 0};
  int outside = 0;
  for (int d = 0; d < nl; ++d) {
    const int ext = d == 0 ? TX : TY;
    const int l = lc[d];
    const int dd = l < 0 ? -l : (l >= ext ? l - ext + 1 : 0);
    if (dd > s.dist) s.dist = dd;
    int v = org[d] + l;
    const int n = g.n[d];
    if (v >= n || v < 0) s.owner = 0;
    if (v >= 0 && v < n) {
      gv[d] = v;
    } else if (g.periodic[d]) {
      gv[d] = v < 0 ? v + n : v - n;
    } else if (v == -1 || v == n) {
      ++outside;
      s.gd = d;
      s.gs = v == n ? 1 : 0;
      gv[d] = v == n ? n - 1 : 0;   // the interior neighbour
      s.eint = e + (v == n ? -1 : 1) * (d == 0 ? 1 : EX);
    } else {
      outside += 2;
    }
  }
  if (s.dist > 0) s.owner = 0;
  if (outside == 1) s.kind = 2;
  else if (outside > 1) s.kind = 0;
  // every position gets in-domain coordinates (unused ones clamped), so the
  // prefetch can load unconditionally
  for (int d = 0; d < nl; ++d) gv[d] = gv[d] < 0 ? 0 : (gv[d] >= g.n[d] ? g.n[d] - 1 : gv[d]);
  s.gx = gv[0];
  s.gy = gv[1];
}

}  // namespace ml
