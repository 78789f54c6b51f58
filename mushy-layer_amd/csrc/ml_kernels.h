// ml_kernels.h — launch interfaces of the HIP kernels (internal to
// libmlhip.so; the public boundary is include/mlhip.h).
#pragma once

#include <hip/hip_runtime.h>
#include <vector>

#include "ml_common.h"

namespace ml {

// LDS budget per workgroup used by the tiled kernels.
constexpr size_t kMaxLds = 64 * 1024;

// Arguments of the nonlinear (H, Theta) operator on one MG level.
template <typename T>
struct HCArgs {
  Geom g;
  ClosureP<T> cp;
  BCDesc bc;
  T alpha, beta;
  const T* a;      // aCoef (2 comps) or nullptr for aCoef == 1
  const T* b[3];   // face coefficients (2 comps each)
  const int* skip = nullptr;   // device flag: the kernel returns at once when set
};

// Device-side iteration control of a multigrid solve.  The host launches
// several iterations ahead; every kernel of an iteration returns at once when
// `done` is set, so the iterations that run, and their results, are exactly
// those of a host loop that checks the norm after each iteration.
struct SolveCtl {
  int done;        // read by every kernel of the loop (skip flag)
  int iter;        // iterations accepted so far
  int status;      // exit-status bits 1/2/4/8 (0 while running)
  int need_host;   // refined form: a trial step was rejected (host backtracks)
  int max_iter;
  double eps, hang, thresh;
  double r0, prev, norm;
};

// Stopping rule after an accepted iteration with residual max-norm `norm`.
__device__ inline void ctl_finish_iteration(SolveCtl* c, double norm) {
  int status = 0;
  if (norm <= c->eps * c->r0) status |= 1;
  if (norm <= c->thresh) status |= 8;
  if (status == 0 && norm > (1.0 - c->hang) * c->prev) status |= 4;
  c->prev = norm;
  c->norm = norm;
  if (status != 0 || c->iter >= c->max_iter) {
    c->status = status != 0 ? status : 2;
    c->done = 1;
  }
}

// Iteration iter+1 wrote its norm into slots[iter+1]: apply the stopping
// rule; with refined != 0 a norm that does not decrease is not accepted but
// handed to the host (need_host, done).  One thread.
__device__ inline void ctl_check_step(SolveCtl* c, const unsigned long long* slots, int refined) {
  if (c->done) return;
  const int it = c->iter + 1;
  const unsigned long long bits =
      __hip_atomic_load(const_cast<unsigned long long*>(slots + it), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const double norm = __longlong_as_double((long long)bits);
  if (refined && !(norm < c->prev)) {
    c->need_host = 1;
    c->norm = norm;
    c->done = 1;
    return;
  }
  c->iter = it;
  ctl_finish_iteration(c, norm);
}

// The stopping rule applied by the last workgroup of the kernel that
// computes the residual norm (ctl == nullptr: none).  `arrive` counts the
// workgroups that have added their norm; the last one resets it.  A kernel
// given no norm pointer writes the norm into slots[ctl->iter + 1] (the
// iteration it completes), so the same launch serves every iteration.
struct CtlFuse {
  SolveCtl* ctl;
  const unsigned long long* slots;
  unsigned int* arrive;
  int refined;
  unsigned int total;   // workgroups that report (0: the whole grid)
};

// Max-reduction of a non-negative value over the workgroup, then one
// atomicMax on its IEEE bit pattern (monotone for non-negative doubles; a
// NaN anywhere wins); with f.ctl set, the grid's last workgroup then applies
// the stopping rule.

// the norm slot of a residual kernel: `norm`, or the iteration's slot
__device__ inline unsigned long long* ctl_slot(unsigned long long* norm, const CtlFuse& f) {
  return norm ? norm : const_cast<unsigned long long*>(f.slots) + f.ctl->iter + 1;
}
template <typename T>
__device__ inline void block_max_atomic(T v, int nan, unsigned long long* out, CtlFuse f = CtlFuse{}) {
  double d = nan ? __longlong_as_double(0x7ff8000000000000LL) : double(v);
  for (int off = 32; off > 0; off >>= 1) {
    double o = __shfl_xor(d, off, 64);
    d = (o != o || d != d) ? __longlong_as_double(0x7ff8000000000000LL) : (o > d ? o : d);
  }
  __shared__ double wmax[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wmax[w] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 1; i < nw; ++i) {
      const double o = wmax[i];
      m = (o != o || m != m) ? __longlong_as_double(0x7ff8000000000000LL) : (o > m ? o : m);
    }
    const unsigned long long bits = (unsigned long long)__double_as_longlong(m);
    if (!f.ctl) {
      atomicMax(out, bits);
    } else {
      // Device-scope atomics are performed at the device's coherence point.
      // The arrival is counted only once this workgroup's maximum has been
      // performed there: the increment depends on the value the max returns
      // (the sign bit of a stored norm, always 0), so the wave waits for it.
      // The last arrival then reads every maximum with device-scope atomic
      // loads; no fence (and no L2 write-back) is needed in any workgroup.
      const unsigned long long before =
          __hip_atomic_fetch_max(out, bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned int inc = 1u + unsigned(before >> 63);
      const unsigned int total = f.total ? f.total : gridDim.x * gridDim.y * gridDim.z;
      if (__hip_atomic_fetch_add(f.arrive, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == total - 1) {
        __hip_atomic_store(f.arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ctl_check_step(f.ctl, f.slots, f.refined);
      }
    }
  }
}

// HIP events around the launches of one kernel family (measurement):
// launch n is bracketed by start[n], stop[n] while n < cap, and its
// algorithmic bytes (every input read once, every output written once) are
// added to *bytes.
struct LaunchTimer {
  hipEvent_t* start;
  hipEvent_t* stop;
  int cap;
  int* n;
  double* bytes;
  long long* missed;   // launches beyond cap (not timed)
};

// Inputs/outputs of the tiled kernels.
template <typename T>
struct TileIO {
  T* phi;
  T* phi_out;     // relax: the relaxed state is written here (never == phi)
  const T* rhs;
  T* out;
  const T* add;
  unsigned long long* norm;
  CtlFuse fuse;   // residual: stopping rule applied by the last workgroup
  Geom gc;        // coarse level (restriction outputs / prolongation inputs)
  T* phic;        // restricted state (optional)
  T* phic0;       // copy of the restricted state (FAS initial coarse state)
  const T* pc;    // prolongation-on-load: coarse state ...
  const T* pc0;   // ... minus its initial value (nullable)
  int npass, first, homog, ntx, nty;
};

// grid-stride indices: of a plain launch, or of workgroup b of the nb that
// serve one job of a batched launch
struct GridIx {
  __device__ long long start() const { return blockIdx.x * (long long)blockDim.x + threadIdx.x; }
  __device__ long long stride() const { return (long long)gridDim.x * blockDim.x; }
};
struct JobIx {
  long long b, nb;
  __device__ long long start() const { return b * blockDim.x + threadIdx.x; }
  __device__ long long stride() const { return nb * blockDim.x; }
};

// Per-operator arguments of the batched Darcy-Brinkman step kernels
// (flow_batch, hc_bcoef_batch): one job per operator, nbc workgroups over
// its cells and nbf[d] over its faces of direction d.
struct FlowJob {
  Geom g;
  ClosureP<double> cp;
  const double* phi;
  double* vars;
  const double* vold;
  const double* vnew;
  const double* pi;
  const double* chi;
  double* c;
  double coeff, hele_shaw, buoyT, buoyS, limit, dt, kcond, DH, DC;
  int law, scaleP;
  double* src[3];
  const double* ustar[3];
  double* u[3];
  double* rhs;
  double* b[3];
  long long nbc, nbf[3];
};
// stages of flow_batch (ml_flow_kernels.hip), each over every job:
enum {
  kFlowEnthalpy = 0,   // vars = enthalpy variables of phi
  kFlowDarcy = 1,      // c = Darcy coefficient of chi
  kFlowUstarSrc = 2,   // src = U* source from vold, vnew, pi
  kFlowVelocity = 3,   // u = ustar + grad pi (zero where solid)
  kFlowDivergence = 4, // rhs = div(u)/dt
  kFlowPressFaces = 5, // b[d] = pressure faces, stage 5 + d
  kFlowCorrect = 8     // u -= grad pi (zero where solid)
};
hipError_t flow_batch(int dim, int stage, const FlowJob* jobs, const unsigned* map, int nblocks, hipStream_t s);
// b[d] = HC face coefficients of chi for every job (ml_transfer_kernels.hip)
hipError_t hc_bcoef_batch(int dim, int d, const FlowJob* jobs, const unsigned* map, int nblocks, hipStream_t s);

// one batched operator application (hc_apply_op) of one operator
template <typename T>
struct HCApplyJob {
  HCArgs<T> A;
  TileIO<T> io;
};

// One MG level as seen by the single-workgroup kernels (device memory).
template <typename T>
struct LevelDesc {
  Geom g;
  T* phi;
  T* rhs;
  T* phi0;
  T* tmp;
  const T* a;
  const T* b[3];
};

template <typename T>
struct CoarseDesc {
  int l0, nlev;
  int nu1, nu2, bottom;
  ClosureP<T> cp;
  BCDesc bc;
  T alpha, beta;
  const int* skip;   // device flag: the kernel returns at once when set
  LevelDesc<T> lev[kMaxLevels];
};

__host__ __device__ inline int wrap_clamp(int v, int n, int periodic) {
  if (v >= 0 && v < n) return v;
  if (periodic) return ((v % n) + n) % n;
  return v < 0 ? 0 : n - 1;
}

// Linear prolongation (ratio 2) of the coarse correction pc - pc0 (pc alone
// when pc0 is null) to fine cell (i,j,k), added to p: PROLONG_JP weights
// (src/AMRProjectionOpF.ChF:15-77), coarse ghosts by periodic image or the
// no-flux rule (nearest interior index).
template <typename T, int DIM>
__device__ inline T prolong_value(const Geom& gc, int i, int j, int k, const T* pc, const T* pc0, T p) {
  const T den = T(1) / (DIM == 3 ? T(64) : T(16));
  const T fx1 = T(3) * den, fx2 = T(9) * den, fx3 = T(27) * den, f0 = T(1) * den;
  const int ic = i / 2, jc = j / 2, kc = DIM == 3 ? k / 2 : 0;
  const int ox = 2 * (i % 2) - 1, oy = 2 * (j % 2) - 1, oz = DIM == 3 ? 2 * (k % 2) - 1 : 0;
  const int X0 = ic, X1 = wrap_clamp(ic + ox, gc.n[0], gc.periodic[0]);
  const int Y0 = jc, Y1 = wrap_clamp(jc + oy, gc.n[1], gc.periodic[1]);
  const int Z0 = kc, Z1 = DIM == 3 ? wrap_clamp(kc + oz, gc.n[2], gc.periodic[2]) : 0;
  auto e = [&](int x, int y, int z) -> T {
    const long long q = gc.ci(x, y, z);
    return pc0 ? pc[q] - pc0[q] : pc[q];
  };
  if (DIM == 2) {
    p = p + fx2 * e(X0, Y0, 0) + f0 * e(X1, Y1, 0);
    p = p + fx1 * (e(X1, Y0, 0) + e(X0, Y1, 0));
  } else {
    p = p + fx3 * e(X0, Y0, Z0) + f0 * e(X1, Y1, Z1);
    p = p + fx2 * (e(X1, Y0, Z0) + e(X0, Y1, Z0) + e(X0, Y0, Z1));
    p = p + fx1 * (e(X1, Y1, Z0) + e(X0, Y1, Z1) + e(X1, Y0, Z1));
  }
  return p;
}

// The same prolongation in two steps, so that the loads of several cells can
// be in flight together: the coarse values (pc alone) in the order the
// combination uses them, then the weighted sum added to p.
template <typename T, int DIM>
struct ProlongCorners {
  T v[DIM == 3 ? 8 : 4];
};

template <typename T, int DIM>
__device__ inline void prolong_fetch(const Geom& gc, int i, int j, int k, const T* pc, ProlongCorners<T, DIM>& P) {
  const int ic = i / 2, jc = j / 2, kc = DIM == 3 ? k / 2 : 0;
  const int ox = 2 * (i % 2) - 1, oy = 2 * (j % 2) - 1, oz = DIM == 3 ? 2 * (k % 2) - 1 : 0;
  const int X0 = ic, X1 = wrap_clamp(ic + ox, gc.n[0], gc.periodic[0]);
  const int Y0 = jc, Y1 = wrap_clamp(jc + oy, gc.n[1], gc.periodic[1]);
  const int Z0 = kc, Z1 = DIM == 3 ? wrap_clamp(kc + oz, gc.n[2], gc.periodic[2]) : 0;
  if (DIM == 2) {
    P.v[0] = pc[gc.ci(X0, Y0, 0)];
    P.v[1] = pc[gc.ci(X1, Y1, 0)];
    P.v[2] = pc[gc.ci(X1, Y0, 0)];
    P.v[3] = pc[gc.ci(X0, Y1, 0)];
  } else {
    P.v[0] = pc[gc.ci(X0, Y0, Z0)];
    P.v[1] = pc[gc.ci(X1, Y1, Z1)];
    P.v[2] = pc[gc.ci(X1, Y0, Z0)];
    P.v[3] = pc[gc.ci(X0, Y1, Z0)];
    P.v[4] = pc[gc.ci(X0, Y0, Z1)];
    P.v[5] = pc[gc.ci(X1, Y1, Z0)];
    P.v[6] = pc[gc.ci(X0, Y1, Z1)];
    P.v[7] = pc[gc.ci(X1, Y0, Z1)];
  }
}

template <typename T, int DIM>
__device__ inline T prolong_combine(const ProlongCorners<T, DIM>& P, T p) {
  const T den = T(1) / (DIM == 3 ? T(64) : T(16));
  const T fx1 = T(3) * den, fx2 = T(9) * den, fx3 = T(27) * den, f0 = T(1) * den;
  if (DIM == 2) {
    p = p + fx2 * P.v[0] + f0 * P.v[1];
    p = p + fx1 * (P.v[2] + P.v[3]);
  } else {
    p = p + fx3 * P.v[0] + f0 * P.v[1];
    p = p + fx2 * (P.v[2] + P.v[3] + P.v[4]);
    p = p + fx1 * (P.v[5] + P.v[6] + P.v[7]);
  }
  return p;
}

// The same prolongation from a coarse correction held in a compact
// (ghost-free, x fastest) array e of level gc.
template <typename T, int DIM>
__device__ inline T prolong_value_lds(const int* n, const int* per, int i, int j, int k, const T* e, T p) {
  const T den = T(1) / (DIM == 3 ? T(64) : T(16));
  const T fx1 = T(3) * den, fx2 = T(9) * den, fx3 = T(27) * den, f0 = T(1) * den;
  const int ic = i / 2, jc = j / 2, kc = DIM == 3 ? k / 2 : 0;
  const int ox = 2 * (i % 2) - 1, oy = 2 * (j % 2) - 1, oz = DIM == 3 ? 2 * (k % 2) - 1 : 0;
  const int X0 = ic, X1 = wrap_clamp(ic + ox, n[0], per[0]);
  const int Y0 = jc, Y1 = wrap_clamp(jc + oy, n[1], per[1]);
  const int Z0 = kc, Z1 = DIM == 3 ? wrap_clamp(kc + oz, n[2], per[2]) : 0;
  const int nx = n[0], nxy = n[0] * n[1];
  auto E = [&](int x, int y, int z) -> T { return e[x + nx * y + nxy * z]; };
  if (DIM == 2) {
    p = p + fx2 * E(X0, Y0, 0) + f0 * E(X1, Y1, 0);
    p = p + fx1 * (E(X1, Y0, 0) + E(X0, Y1, 0));
  } else {
    p = p + fx3 * E(X0, Y0, Z0) + f0 * E(X1, Y1, Z1);
    p = p + fx2 * (E(X1, Y0, Z0) + E(X0, Y1, Z0) + E(X0, Y0, Z1));
    p = p + fx1 * (E(X1, Y1, Z0) + E(X0, Y1, Z1) + E(X1, Y0, Z1));
  }
  return p;
}

template <typename T, int DIM>
__device__ inline T prolong_value_lds(const Geom& gc, int i, int j, int k, const T* e, T p) {
  return prolong_value_lds<T, DIM>(gc.n, gc.periodic, i, j, k, e, p);
}

// ---- FP32-refined form (ml_fp32r.hip)
// Linearised (H, Theta) operator J e = alpha*a*e - beta*div(b grad(M e)) on
// one MG level: M is the per-cell 2x2 Jacobian of the closure
// (theta, Theta_l)(H, Theta), components (M00, M01, M10, M11).
template <typename T>
struct LinLevel {
  Geom g;
  const T* b[3];   // face coefficients (2 comps)
  const T* a;      // aCoef (2 comps) or nullptr
  T* M;            // Jacobian (4 comps)
  T* e;            // correction (2 comps)
  T* r;            // right-hand side (2 comps)
  const int* skip;   // device flag: the kernel returns at once when set
};

struct LinDesc {
  int l0, nlev, nu1, nu2, bottom;
  BCDesc bc;
  double alpha, beta;
  const int* skip;
  LinLevel<float> lev[kMaxLevels];
  const double* b64[3];   // coarsest level, FP64 faces for the bottom solve
  const double* a64;
};

// fine level: phi_out = phi + omega*e (when e != nullptr; phi_out != phi),
// then r32 = rhs - L(phi_out), M32 = closure Jacobian, max-norm of r into
// *norm
hipError_t hc_residual_jacobian(const HCArgs<double>& A, const double* phi, double* phi_out, const double* rhs,
                                const float* e, double omega, float* r32, float* M32, unsigned long long* norm,
                                hipStream_t s, const CtlFuse& fuse = CtlFuse{});
// red/black block relaxation of J e = r on a tiled level, ping-ponging
// between ebuf[0] and ebuf[1] (*cur indexes the current correction; L.e is
// ignored); zero_init starts from e = 0; pc (coarse correction, level gc) is
// prolonged into e first
hipError_t lin_relax(const LinLevel<float>& L, float* const ebuf[2], int* cur, const BCDesc& bc, float alpha,
                     float beta, int npass, bool zero_init, const Geom* gc, const float* pc, hipStream_t s);
// lin_relax from a zero start followed by r_c = R(r - J e), M_c = R(M) into C
// (one launch when the passes fit the tile halo)
hipError_t lin_relax_restrict(const LinLevel<float>& L, float* const ebuf[2], int* cur, const LinLevel<float>& C,
                              const BCDesc& bc, float alpha, float beta, int npass, bool zero_init, hipStream_t s);
// r_c = R(r - J e), M_c = R(M) on the next coarser level
hipError_t lin_restrict(const LinLevel<float>& L, const LinLevel<float>& C, const BCDesc& bc, float alpha,
                        float beta, hipStream_t s);
// single-workgroup linear V-cycle over levels l0.. of the device descriptor
hipError_t lin_coarse_vcycle(const LinDesc* dev_desc, int dim, size_t lds, hipStream_t s);
// dynamic LDS of lin_coarse_vcycle for levels l0..nlev-1 of g, and the
// largest number of levels it handles
size_t lin_coarse_lds_bytes(const Geom* g, int l0, int nlev);
size_t lin_coarse_static_lds();
int lin_coarse_levels_max(int dim);
// cast of FP64 face coefficients to FP32
hipError_t cast_f64_f32(long long n, const double* x, float* y, hipStream_t s);
// several casts in one launch
struct CastJobs {
  int n;
  long long begin[3 * kMaxLevels + 1];
  const double* src[3 * kMaxLevels];
  float* dst[3 * kMaxLevels];
};
hipError_t cast_many(const CastJobs& J, hipStream_t s);

// ---- batched form: independent instances share every launch ------------
// A batched tiled launch covers one level of each of several instances;
// workgroup w works on tile (map[w] & kBatchTileMask) of job
// (map[w] >> kBatchTileBits).
constexpr int kBatchTileBits = 20;
constexpr unsigned kBatchTileMask = (1u << kBatchTileBits) - 1;

// one instance's level in a batched relaxation launch of J e = r
struct BatchLinJob {
  LinLevel<float> L;   // L.skip: the instance's done flag
  const float* ein;
  float* eout;
  BCDesc bc;
  float alpha, beta;
  Geom gc;             // prolongation source level (pc != nullptr)
  const float* pc;
  Geom gr;             // restriction target level (rc != nullptr)
  float* rc;
  float* Mc;
  int ntx, nty;
};

// one instance in a batched FP64 residual + Jacobian launch
struct BatchResJob {
  HCArgs<double> A;    // level 0; A.skip == &ctl->done
  double* pb[2];       // state pair: iteration k reads pb[(k - 1) & 1] and writes the other
  const double* rhs;
  const float* e;      // level-0 correction
  float* r32;
  float* M32;
  unsigned long long* slots;   // residual norm of iteration k in slots[k]
  SolveCtl* ctl;
  unsigned int* arrive;
  unsigned int ntiles;
  int ntx, nty;
};

// one instance in the step set-up / wrap-up launches (ml_batch.hip)
struct BatchSetupJob {
  Geom g;                 // level 0
  ClosureP<double> cp;
  double kcond, DH, DC;   // bCoef parameters
  double* phi;            // state (in; the wrap-up copies the final state back)
  double* alt;            // its partner, a full copy
  double* rhs;            // backward-Euler rhs: phi*aCoef (aCoef == 1), + dt*src with a source
  const double* src;      // source (nullable)
  double dt;
  double* b64[3];         // level-0 faces
  float* b32[3];
  unsigned long long* slots;   // norm slots, zeroed
  int nslots;
  SolveCtl* ctl;
};

// one instance in the coefficient-hierarchy launch: job jb covers the faces
// of direction dir[jb] on level level[jb] >= 1
struct BatchCoarsenJob {
  Geom g0;
  int njobs;
  long long begin[3 * kMaxLevels + 1];
  short level[3 * kMaxLevels];
  short dir[3 * kMaxLevels];
  const double* b0[3];
  double* b64[3 * kMaxLevels];
  float* b32[3 * kMaxLevels];
};

long long batch_chunk_items();   // items of one workgroup of the set-up launches
long long batch_setup_items(const BatchSetupJob& J);
hipError_t batch_setup(const BatchSetupJob* jobs, const unsigned* map, int nblocks, hipStream_t s);
hipError_t batch_coarsen(const BatchCoarsenJob* jobs, const unsigned* map, int nblocks, hipStream_t s);
hipError_t batch_ctl_init(SolveCtl* ctl, const unsigned long long* slots, int nslots, int n, int max_iter, double eps,
                          double hang, double thresh, hipStream_t s);
hipError_t batch_finish(const BatchSetupJob* jobs, const unsigned* map, int nblocks, hipStream_t s);

// tiles of the relaxation / residual launches of a 2D level (batched form)
int lin_relax_tiles2(const Geom& g, int& ntx, int& nty);
int lin_residual_tiles2(const Geom& g, int& ntx, int& nty);
// level-0 tiles above which a single solve changes its relaxation tiles
long long lin_deep_tiles_max2();
// tile shapes of the batched relaxation: count, shape of variant v, and
// the tiles of level g under it
int lin_relax_batch_tiles();
int lin_relax_batch_tile_shape(int v, int* tx, int* ty, int* nt);
int lin_relax_batch_tiles_of(int v, const Geom& g, int& ntx, int& nty);
// tm (may be null): launch timing, `bytes` the launch's algorithmic bytes
hipError_t lin_relax_batch(const BatchLinJob* jobs, const unsigned* map, int nblocks, int npass, bool zero_init,
                           bool restrict_, int variant, hipStream_t s, const LaunchTimer* tm = nullptr,
                           double bytes = 0);
// initial: residual of the state pb[0] into slots[0], no correction, no
// stopping rule; else iteration ctl->iter + 1 with the stopping rule
hipError_t hc_residual_jacobian_batch(const BatchResJob* jobs, const unsigned* map, int nblocks, bool initial,
                                      hipStream_t s);
// backtracking of the listed instances (their steps were rejected): the
// residual at step length omega into slots[ctl->iter + 1]; zero / gather
// those slots (out[j] for list[j])
hipError_t hc_residual_jacobian_batch_trial(const BatchResJob* jobs, const unsigned* map, int nblocks, double omega,
                                            hipStream_t s);
hipError_t batch_trial_slots(const BatchResJob* jobs, const int* list, int n, bool gather, unsigned long long* out,
                             hipStream_t s);
// the coarse V-cycles of n instances, one workgroup each
hipError_t lin_coarse_vcycle_batch(const LinDesc* descs, int n, size_t lds, hipStream_t s);

// ---- nonlinear (H, Theta) operator (ml_hc_kernels.hip)
// Tiled red/black relaxation.  Every launch reads buf[*cur] and writes
// buf[1 - *cur] (tiles read their halo from the input while other tiles
// store), then flips *cur.
template <typename T>
hipError_t hc_relax(const HCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, int first,
                    hipStream_t s, const LaunchTimer* tm = nullptr);
// relax tiles with the linear prolongation of pc - pc0 (level gc) fused on
// load; npass == 0 prolongs in place into buf[*cur]
template <typename T>
hipError_t hc_relax_tiles(const HCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, int first,
                          const Geom* gc, const T* pc, const T* pc0, hipStream_t s, const LaunchTimer* tm = nullptr);
template <typename T>
hipError_t hc_coarse_vcycle(const CoarseDesc<T>* dev_desc, int dim, int apply_first, hipStream_t s);
template <typename T>
hipError_t hc_restrict_all(const HCArgs<T>& A, const Geom& gc, T* phi, const T* rhs, T* resc, T* phic, T* phic0,
                           int homog, hipStream_t s);
template <typename T>
hipError_t hc_residual(const HCArgs<T>& A, T* phi, const T* rhs, T* res, unsigned long long* norm,
                       int homog, hipStream_t s, const CtlFuse& fuse = CtlFuse{});
template <typename T>
hipError_t hc_apply_op(const HCArgs<T>& A, T* phi, T* lhs, const T* add, int homog, hipStream_t s);
template <typename T>
hipError_t hc_restrict_residual(const HCArgs<T>& A, const Geom& gc, T* phi, const T* rhs, T* resc,
                                int homog, hipStream_t s);
template <typename T>
hipError_t hc_diffused_var(const HCArgs<T>& A, const T* phi, T* out, int homog, hipStream_t s);

// ---- transfers, norms, coefficients (ml_transfer_kernels.hip)
template <typename T>
hipError_t restrict_r(const Geom& gf, int ncomp, const T* phif, T* phic, T* phic0, hipStream_t s);
template <typename T>
hipError_t prolong_linear(const Geom& gf, int ncomp, T* phif, const T* phic, const T* phic0, hipStream_t s);
template <typename T>
hipError_t prolong_constant(const Geom& gf, int ncomp, T* phif, const T* corr, hipStream_t s);
template <typename T>
hipError_t norm_max(const Geom& g, int ncomp, const T* x, unsigned long long* out, hipStream_t s);
hipError_t zero_u64(unsigned long long* p, int n, hipStream_t s);

// ctl from the initial residual norm in slots[0]
hipError_t ctl_init(SolveCtl* ctl, const unsigned long long* slots, int max_iter, double eps, double hang,
                    double thresh, hipStream_t s);
// after iteration iter+1 wrote its norm into slots[iter+1]: apply the
// stopping rule; with refined != 0 a norm that does not decrease is not
// accepted but handed to the host (need_host, done)
hipError_t ctl_check(SolveCtl* ctl, const unsigned long long* slots, int refined, hipStream_t s);

// The MG coefficient hierarchy: for every level l = 1..nlev-1, the
// arithmetic average by ratio 2^l of the finest faces bf[d] into bc[l][d]
// and (when af != nullptr) of the finest cell field af into ac[l], in the
// summation order of the CPU restatement (mlr_average_faces/cells).
constexpr int kMaxCoarsenJobs = 96;
template <typename T>
struct CoarsenJobs {
  int njobs;
  long long begin[kMaxCoarsenJobs + 1];
  short level[kMaxCoarsenJobs];
  short kind[kMaxCoarsenJobs];     // 0..2 face direction, 3 cells
  T* dst[kMaxCoarsenJobs];
};
// batched hc_apply_op over (job, tile) maps of the residual tile (hc_apply_tiles)
int hc_apply_tiles(const Geom& g, int* ntx, int* nty);
template <typename T>
hipError_t hc_apply_batch(int dim, const HCApplyJob<T>* jobs, const unsigned* map, int nblocks, hipStream_t s);
template <typename T>
hipError_t coarsen_hierarchy(const Geom& gf, int ncomp, int nlev, const T* const* bf, T* const* const* bc, const T* af,
                             T* const* ac, hipStream_t s, bool faces = true);
// One average of a hierarchy (level, kind as in CoarsenJobs; count = coarse
// items times components) for the batched form, which averages the
// hierarchies of many operators in two launches (light items one thread
// each, heavy items one workgroup each; the same arithmetic as
// coarsen_hierarchy)
template <typename T>
struct CoarsenRec {
  Geom gf;
  int level, kind;
  const T* src;
  T* dst;
  long long count;
};
template <typename T>
void coarsen_records(const Geom& gf, int ncomp, int nlev, const T* const* bf, T* const* const* bc, const T* af,
                     T* const* ac, bool faces, std::vector<CoarsenRec<T>>& light, std::vector<CoarsenRec<T>>& heavy);
// lmap[w] = light record << 32 | chunk of 256 items, hmap[w] = heavy record << 32 | item
template <typename T>
hipError_t coarsen_recs(int dim, const CoarsenRec<T>* light, const unsigned long long* lmap, int lblocks,
                        const CoarsenRec<T>* heavy, const unsigned long long* hmap, int hblocks, hipStream_t s);
hipError_t porosity_field(const Geom& g, const ClosureP<double>& cp, const double* phi, double* chi,
                          hipStream_t s);
hipError_t hc_bcoef(const Geom& g, double k, double DH, double DC, const double* chi, double* const* b,
                    hipStream_t s);
template <typename T>
hipError_t be_rhs(const Geom& g, int ncomp, T* y, const T* x, const T* acoef, T dt, const T* src, hipStream_t s);
hipError_t darcy_ccoef(const Geom& g, double coeff, const double* chi, double* c, hipStream_t s);
hipError_t porosity_faces(const Geom& g, double scale, const double* chi, double* const* b, hipStream_t s);

// ---- state update and velocity closure (ml_flow_kernels.hip) ----
hipError_t enthalpy_vars(const Geom& g, const ClosureP<double>& cp, const double* phi, double* vars, hipStream_t s);
hipError_t ustar_src(const Geom& g, double buoyT, double buoyS, int scaleP, const double* vold, const double* vnew,
                     const double* pi, double* const* src, hipStream_t s);
hipError_t cc_project_rhs(const Geom& g, double dt, double limit, int scaleP, const double* vold, const double* vnew,
                          const double* pi, const double* const* ustar, double* const* u, double* rhs, hipStream_t s);
hipError_t pressure_faces(const Geom& g, const double* vold, const double* vnew, double* const* b, hipStream_t s);
hipError_t cc_correct(const Geom& g, double dt, double limit, int scaleP, const double* vold, const double* vnew,
                      const double* pi, double* const* u, hipStream_t s);
hipError_t darcy_ccoef_law(const Geom& g, double coeff, int law, double hele_shaw, const double* chi, double* c,
                           hipStream_t s);

}  // namespace ml

namespace ml {

// ---- linear variable-coefficient operator (ml_vc_kernels.hip)
// L u = alpha*a*u - beta*c*u - beta*div(b grad u), one component; a, c
// nullable (a == 1, no c term); BC component 0 of bc, always applied.
template <typename T>
struct VCArgs {
  Geom g;
  BCDesc bc;
  T alpha, beta;
  const T* a;
  const T* c;
  const T* b[3];
  const int* skip = nullptr;   // device flag: the kernel returns at once when set
};

template <typename T>
struct VCTileIO {
  const T* phi;
  const T* e;          // residual: u = phi + e (nullable)
  T* phi_out;          // ... written here when e != nullptr
  const T* rhs;
  T* out;
  unsigned long long* norm;
  CtlFuse fuse;        // residual: stopping rule applied by the last workgroup
  Geom gc;
  int homog, ntx, nty;
};

// The residual of the full equation after the last relaxation of an
// iteration, fused into that launch: phi_out = phi + e, res = rhs - L(phi_out)
// (inhomogeneous BCs), its max-norm into slots[ctl->iter + 1] with the
// stopping rule (phi == nullptr: none).
template <typename T>
struct VCFinal {
  const T* phi;
  T* phi_out;
  const T* rhs;
  T* res;
  CtlFuse fuse;
};

template <typename T>
struct VCLevel {
  Geom g;
  const T* a;
  const T* c;
  const T* b[3];
  T* e;   // correction
  T* r;   // residual right-hand side
};

// one (instance, level) relaxation of a batched launch (the arguments of
// vc_relax_tile_kernel)
template <typename T>
struct VCRelaxJob {
  VCArgs<T> A;
  const T* in;
  T* out;
  const T* rhs;
  int npass, first, zero_init, ntx, nty;
  Geom gc;
  const T* pc;
  Geom gr;
  T* resc;
  VCFinal<T> fin;
};

// one batched residual (vc_residual's kVCResid tile pass) of one operator level
template <typename T>
struct VCResidJob {
  VCArgs<T> A;
  VCTileIO<T> io;
};

// dst[0..n) = src[0..n) as 8-byte words, or zeros when src is null
struct WordsJob {
  const unsigned long long* src;
  unsigned long long* dst;
  long long n;
};
constexpr int kWordsChunk = 4096;   // words per workgroup of words_batch

template <typename T>
struct VCDesc {
  int l0, nlev, nu1, nu2, bottom;
  int bottom_solver, bottom_imax;   // BiCGStab on the coarsest level after `bottom` relaxations (solver 1)
  BCDesc bc;
  T alpha, beta;
  const int* skip;
  VCLevel<T> lev[kMaxLevels];
};

// red/black relaxation of L e = r with homogeneous BCs, ping-ponging between
// buf[0] and buf[1] (*cur indexes the current state); zero_init starts from
// e = 0; pc (level gc) is added by piecewise-constant prolongation first
template <typename T>
hipError_t vc_relax(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                    const Geom* gc, const T* pc, hipStream_t s, const LaunchTimer* tm = nullptr);
// vc_relax followed in the same launch by the residual of the full equation
// (VCFinal); false in *fused when the passes do not fit one launch (then
// nothing was launched)
template <typename T>
hipError_t vc_relax_final(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, const Geom* gc,
                          const T* pc, const VCFinal<T>& fin, bool* fused, hipStream_t s,
                          const LaunchTimer* tm = nullptr);
// relaxation followed by the restricted residual R(rhs - L e) (homogeneous
// BCs) of the relaxed state into resc on level gr, fused in one launch when
// the passes fit the tile halo
template <typename T>
hipError_t vc_relax_restrict(const VCArgs<T>& A, T* const buf[2], int* cur, const T* rhs, int npass, bool zero_init,
                             const Geom& gr, T* resc, hipStream_t s, const LaunchTimer* tm = nullptr);
// res = rhs - L(phi [+ e]) (res nullable), max-norm into *norm (nullable);
// with e the updated state phi + e is written to phi_out (!= phi)
template <typename T>
hipError_t vc_residual(const VCArgs<T>& A, const T* phi, const T* e, T* phi_out, const T* rhs, T* res,
                       unsigned long long* norm, int homog, hipStream_t s, const CtlFuse& fuse = CtlFuse{});
template <typename T>
hipError_t vc_apply_op(const VCArgs<T>& A, const T* phi, T* lhs, int homog, hipStream_t s);
template <typename T>
hipError_t vc_restrict_residual(const VCArgs<T>& A, const Geom& gc, const T* phi, const T* rhs, T* resc,
                                hipStream_t s);
template <typename T>
hipError_t vc_coarse_vcycle(const VCDesc<T>* dev_desc, int dim, size_t lds, hipStream_t s);
// phi = rhs * lambda on the valid cells (preCond's first step)
template <typename T>
hipError_t vc_lambda_mult(const VCArgs<T>& A, const T* rhs, T* phi, hipStream_t s);
// *out = (x, y) over the valid cells, in the BiCGStab bottom's fixed order
template <typename T>
hipError_t vc_dot(const Geom& g, const T* x, const T* y, T* out, hipStream_t s);
// batched form: relaxation launches whose workgroups are (job, tile) pairs
// (map[w] = job << kBatchTileBits | tile), every job on the <= 512^2-cell
// tile configuration (2D shape 0: 16x16 cores, shape 1: 32x32 cores for
// large batches); one coarse V-cycle workgroup per descriptor
bool vc_batch_level_ok(const Geom& g);
int vc_batch_hmax(int dim);
int vc_batch_tiles(const Geom& g, int shape, int* ntx, int* nty);
size_t vc_batch_lds(int dim, int shape, int npass, bool resc, bool fin);
template <typename T>
hipError_t vc_relax_batch(int dim, int shape, const VCRelaxJob<T>* jobs, const unsigned* map, int nblocks, size_t lds,
                          hipStream_t s, const LaunchTimer* tm = nullptr);
template <typename T>
hipError_t vc_coarse_batch(int dim, const VCDesc<T>* const* descs, int n, size_t lds, hipStream_t s);
// out[i] = *ctls[i] (the iteration control blocks of many solves, for one copy to the host)
hipError_t ctl_gather(SolveCtl* const* ctls, SolveCtl* out, int n, hipStream_t s);
// batched residuals over (job, tile) maps of the residual tile (vc_resid_tiles)
int vc_resid_tiles(const Geom& g, int* ntx, int* nty);
template <typename T>
hipError_t vc_resid_batch(int dim, const VCResidJob<T>* jobs, const unsigned* map, int nblocks, hipStream_t s);
// batched copies / zero fills over (job, chunk of kWordsChunk words) maps
hipError_t words_batch(const WordsJob* jobs, const unsigned* map, int nblocks, hipStream_t s);
// ctl_init of n control blocks (ctls[i] from slots[i][0]), one thread each
hipError_t ctl_init_many(SolveCtl* const* ctls, const unsigned long long* const* slots, int n, int max_iter, double eps,
                         double hang, double thresh, hipStream_t s);
// dynamic LDS of vc_coarse_vcycle for levels l0..nlev-1 of g
template <typename T>
size_t vc_coarse_lds_bytes(const Geom* g, int l0, int nlev, bool bicgstab = false);
bool vc_small_level(const Geom& g);
// levels (the small level and all coarser) one vc_coarse_vcycle launch handles
int vc_coarse_levels_max(int dim);

hipError_t vec_axby(long long n, double* lhs, const double* x, const double* y, double a, double b, hipStream_t s);
hipError_t vec_incr(long long n, double* lhs, const double* x, double sc, hipStream_t s);
hipError_t vec_mult(long long n, double* lhs, const double* x, hipStream_t s);
hipError_t vec_set(long long n, double* lhs, double v, hipStream_t s);
}  // namespace ml
