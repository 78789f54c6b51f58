// ml_flow_kernels.hip — the level-wide state update around the implicit
// solves and the cell-centred velocity closure of a Darcy-Brinkman step:
//
//   enthalpy_vars_kernel   updateEnthalpyVariables (src/phaseDiagram.cpp:92-177)
//   ustar_src_kernel       computeUstarSrc (srcSubcycle/AMRLevelMushyLayerVel.cpp
//                          :1724-1853): buoyancy and porosity * grad P
//   cc_velocity_kernel     removeGradPressure + setVelZero before the
//                          projection (computeCCvelocity, :673-917)
//   cc_divergence_kernel   the projection's right-hand side div(u)/dt
//                          (Projector::levelProject, srcSubcycle/Projector.cpp
//                          :1379-1397; util/Divergence.cpp:349-390)
//   pressure_faces_kernel  the projection's face coefficient (arithmetic face
//                          porosity at the half time)
//   cc_correct_kernel      applyCCcorrection (srcSubcycle/Projector.cpp
//                          :1632-1683) + setVelZero
//   darcy_ccoef_law_kernel the U* Darcy coefficient for every permeability
//                          law (src/MushyLayerParams.cpp:1031-1090)
//
// Every kernel is one pass over the valid cells (one thread per cell,
// grid-stride), reading neighbours straight from HBM: these are single
// streaming passes with no reuse beyond a 3-point stencil per direction,
// so they run at the bandwidth of their compulsory bytes.  Ghost values are
// formed in registers, never stored: the pressure beyond a wall is the
// second-order extrapolation 3 p1 - 3 p2 + p3 (EXTRAPBC order 2,
// BCutil/BCFuncF.ChF:97-156), the velocity beyond a wall 2*0 - u1 (no-slip,
// Dirichlet order 1), periodic sides wrap.
//
// Fields: `vars` = (theta, Theta_l, Theta_s, porosity), 4 components in
// cell layout; velocity components, sources and the pressure are separate
// 1-component cell arrays.  Time-centred values at t^n + dt/2 are
// 0.5*x^n + 0.5*x^{n+1}.  Operation order follows the oracle
// (oracle/mlref_flow.c) term by term, so FP64 results are bit-identical.
#include <hip/hip_runtime.h>

#include "ml_kernels.h"
#include "mlhip.h"

namespace ml {
namespace {

// flat index (i fastest) -> (i, j, k)
__device__ inline void unflatten(long long c, int nx, int ny, int& i, int& j, int& k) {
  i = int(c % nx);
  const long long r = c / nx;
  j = int(r % ny);
  k = int(r / ny);
}

__device__ inline void cell_of(const Geom& g, long long t, int& i, int& j, int& k) { unflatten(t, g.n[0], g.n[1], i, j, k); }

// workgroups of 256 for a grid-stride pass over n items
inline int grid_for(long long n) {
  long long b = (n + 255) / 256;
  if (b > 65535 * 4) b = 65535 * 4;
  if (b < 1) b = 1;
  return int(b);
}

// pressure at (i,j,k) + s e_d, s = +-1: interior, periodic image, or the
// order-2 extrapolation from the three cells inside the wall
__device__ inline double p_at(const Geom& g, const double* __restrict__ p, int d, int i, int j, int k, int s) {
  int a[3] = {i, j, k};
  const int v = a[d] + s;
  const int n = g.n[d];
  if (v >= 0 && v < n) {
    a[d] = v;
    return p[g.ci(a[0], a[1], a[2])];
  }
  if (g.periodic[d]) {
    a[d] = v < 0 ? n - 1 : 0;
    return p[g.ci(a[0], a[1], a[2])];
  }
  const int isign = v < 0 ? -1 : 1;
  int b[3] = {a[0], a[1], a[2]};
  b[d] = v - isign;
  const double p1 = p[g.ci(b[0], b[1], b[2])];
  b[d] = v - 2 * isign;
  const double p2 = p[g.ci(b[0], b[1], b[2])];
  b[d] = v - 3 * isign;
  const double p3 = p[g.ci(b[0], b[1], b[2])];
  return 3 * p1 - 3 * p2 + p3;
}

// velocity component at (i,j,k) + s e_d: interior, periodic image, or the
// no-slip ghost 2*0 - u1 (u1 the cell itself, next to the wall)
__device__ inline double u_at(const Geom& g, const double* __restrict__ u, double u1, int d, int i, int j, int k,
                              int s) {
  int a[3] = {i, j, k};
  const int v = a[d] + s;
  const int n = g.n[d];
  if (v >= 0 && v < n) {
    a[d] = v;
    return u[g.ci(a[0], a[1], a[2])];
  }
  if (g.periodic[d]) {
    a[d] = v < 0 ? n - 1 : 0;
    return u[g.ci(a[0], a[1], a[2])];
  }
  const double zero = 0.0;
  return 2.0 * zero - u1;
}

__device__ inline double centred(double xo, double xn) { return 0.5 * xo + 0.5 * xn; }

// GRADCC (util/GradientF.ChF:194-219)
__device__ inline double grad_cc(const Geom& g, const double* __restrict__ p, double factor, int d, int i, int j,
                                 int k) {
  return factor * (p_at(g, p, d, i, j, k, 1) - p_at(g, p, d, i, j, k, -1));
}

// ---------------------------------------------------------------------------
// CALCULATEPOROSITY / CALCULATECL / CALCULATECS / CALCULATET
// (src/EnthalpyVariablesF.ChF:13-287) after CALCULATE_BOUNDING_ENERGY
// (:657-698), one cell
__device__ inline void enthalpy_vars_point(double H, double C, const ClosureP<double>& p, double* out) {
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
  double T = NAN, Cl = NAN, Cs = NAN, chi = NAN;
  if (H <= hs) {
    chi = 0;
    Cl = p.Te;
    Cs = C;
    T = H / p.c;
  } else if (H > hs && H <= he) {
    chi = (H - p.te * p.c) / (p.St + p.te * (1 - p.c));
    Cl = p.Te;
    const double pm = porosity_mush<double>(C, H, p);
    Cs = C / (1 - pm);
    T = p.te;
  } else if (H > he && H < hl) {
    const double pm = porosity_mush<double>(C, H, p);
    chi = pm;
    Cl = (C + p.CR * (1 - pm)) / (pm + p.pc * (1 - pm));
    Cs = (C * p.pc - p.CR * pm) / (pm + p.pc * (1 - pm));
    T = -(C + p.CR * (1 - pm)) / (pm + p.pc * (1 - pm));
  } else if (H >= hl) {
    chi = 1;
    Cl = C;
    Cs = 0;
    T = H - p.St;
  }
  out[0] = T;
  out[1] = Cl;
  out[2] = Cs;
  out[3] = chi;
}

template <class IX>
__device__ __forceinline__ void enthalpy_vars_body(Geom g, ClosureP<double> cp, const double* __restrict__ phi,
                                     double* __restrict__ vars, IX ix) {
  const long long N = g.nvalid();
  for (long long t = ix.start(); t < N; t += ix.stride()) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const long long q = g.ci(i, j, k);
    double o[4];
    enthalpy_vars_point(phi[q], phi[q + g.ncell], cp, o);
    for (int c = 0; c < 4; ++c) vars[q + c * g.ncell] = o[c];
  }
}

__global__ void enthalpy_vars_kernel(Geom g, ClosureP<double> cp, const double* __restrict__ phi,
                                     double* __restrict__ vars) {
  enthalpy_vars_body(g, cp, phi, vars, GridIx{});
}

struct Vec3 {
  double* p[3];
};
struct CVec3 {
  const double* p[3];
};

// src_d = (0 + buoyancy_d) - P_src_d (fillBuoyancy
// srcSubcycle/AMRLevelMushyLayerFill.cpp:870-885; fillPressureSrcTerm :787-807)
template <class IX>
__device__ __forceinline__ void ustar_src_body(Geom g, double buoyT, double buoyS, int scaleP, const double* __restrict__ vold,
                                 const double* __restrict__ vnew, const double* __restrict__ pi, Vec3 src, IX ix) {
  const long long N = g.nvalid();
  const double factor = 0.5 / g.dx;
  const int gd = g.dim - 1;
  for (long long t = ix.start(); t < N; t += ix.stride()) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const long long q = g.ci(i, j, k);
    const double th = centred(vold[q], vnew[q]);
    const double thl = centred(vold[q + g.ncell], vnew[q + g.ncell]);
    const double chi = centred(vold[q + 3 * g.ncell], vnew[q + 3 * g.ncell]);
    for (int d = 0; d < g.dim; ++d) {
      double buoy = 0.0;
      if (d == gd) {
        buoy = buoy + buoyT * th;
        buoy = buoy + (-buoyS) * thl;
        buoy = buoy + 1.0 * 0.0;
        buoy = buoy * chi;
      }
      double gp = grad_cc(g, pi, factor, d, i, j, k);
      if (scaleP) gp = gp * chi;
      double s = 0.0;
      s = s + buoy;
      s = s - gp;
      src.p[d][q] = s;
    }
  }
}

__global__ void ustar_src_kernel(Geom g, double buoyT, double buoyS, int scaleP, const double* __restrict__ vold,
                                 const double* __restrict__ vnew, const double* __restrict__ pi, Vec3 src) {
  ustar_src_body(g, buoyT, buoyS, scaleP, vold, vnew, pi, src, GridIx{});
}

// u_d = ustar_d + (P_src_d * dt) (removeGradPressure,
// srcSubcycle/AMRLevelMushyLayerVel.cpp:654-671); u = 0 where the new
// porosity <= limit (SETZEROVELOCITY, src/utils.ChF:73-96)
template <class IX>
__device__ __forceinline__ void cc_velocity_body(Geom g, double dt, double limit, int scaleP, const double* __restrict__ vold,
                                   const double* __restrict__ vnew, const double* __restrict__ pi, CVec3 ustar,
                                   Vec3 u, IX ix) {
  const long long N = g.nvalid();
  const double factor = 0.5 / g.dx;
  for (long long t = ix.start(); t < N; t += ix.stride()) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const long long q = g.ci(i, j, k);
    const double chi = centred(vold[q + 3 * g.ncell], vnew[q + 3 * g.ncell]);
    const bool solid = vnew[q + 3 * g.ncell] <= limit;
    for (int d = 0; d < g.dim; ++d) {
      double gp = grad_cc(g, pi, factor, d, i, j, k);
      if (scaleP) gp = gp * chi;
      gp = gp * dt;
      u.p[d][q] = solid ? 0.0 : ustar.p[d][q] + gp;
    }
  }
}

__global__ void cc_velocity_kernel(Geom g, double dt, double limit, int scaleP, const double* __restrict__ vold,
                                   const double* __restrict__ vnew, const double* __restrict__ pi, CVec3 ustar,
                                   Vec3 u) {
  cc_velocity_body(g, dt, limit, scaleP, vold, vnew, pi, ustar, u, GridIx{});
}

// rhs = (sum_d (one/dx)*(u_face(i+1/2) - u_face(i-1/2))) * (1/dt), face
// values the arithmetic mean of the two cells (DIVERGENCE,
// util/DivergenceF.ChF:22-49)
template <class IX>
__device__ __forceinline__ void cc_divergence_body(Geom g, double dt, CVec3 u, double* __restrict__ rhs, IX ix) {
  const long long N = g.nvalid();
  const double one_on_dx = 1.0 / g.dx;
  const double dtScale = 1.0 / dt;
  for (long long t = ix.start(); t < N; t += ix.stride()) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const long long q = g.ci(i, j, k);
    double div = 0.0;
    for (int d = 0; d < g.dim; ++d) {
      const double* ud = u.p[d];
      const double uc = ud[q];
      const double up = u_at(g, ud, uc, d, i, j, k, 1);
      const double um = u_at(g, ud, uc, d, i, j, k, -1);
      const double ehi = 0.5 * (uc + up);
      const double elo = 0.5 * (um + uc);
      div = div + one_on_dx * (ehi - elo);
    }
    rhs[q] = div * dtScale;
  }
}

__global__ void cc_divergence_kernel(Geom g, double dt, CVec3 u, double* __restrict__ rhs) {
  cc_divergence_body(g, dt, u, rhs, GridIx{});
}

// b = -(face porosity at the half time): arithmetic mean inside and across
// periodic sides, the adjacent cell on walls
template <int DIM, class IX>
__device__ __forceinline__ void pressure_faces_body(Geom g, int d, const double* __restrict__ vold, const double* __restrict__ vnew,
                                      double* __restrict__ b, IX ix) {
  const long long N = g.nface[d];
  const long long c3 = 3 * g.ncell;
  for (long long t = ix.start(); t < N; t += ix.stride()) {
    int i, j, k;
    unflatten(t, int(g.fn[d][0]), int(g.fn[d][1]), i, j, k);
    int lo[3] = {i, j, DIM == 3 ? k : 0}, hi[3] = {i, j, DIM == 3 ? k : 0};
    lo[d] -= 1;
    const int nd = g.n[d];
    double v;
    if (lo[d] < 0 && !g.periodic[d]) {
      const long long q = g.ci(hi[0], hi[1], hi[2]);
      v = centred(vold[q + c3], vnew[q + c3]);
    } else if (hi[d] >= nd && !g.periodic[d]) {
      const long long q = g.ci(lo[0], lo[1], lo[2]);
      v = centred(vold[q + c3], vnew[q + c3]);
    } else {
      if (lo[d] < 0) lo[d] += nd;
      if (hi[d] >= nd) hi[d] -= nd;
      const long long ql = g.ci(lo[0], lo[1], lo[2]), qh = g.ci(hi[0], hi[1], hi[2]);
      const double cl = centred(vold[ql + c3], vnew[ql + c3]);
      const double ch = centred(vold[qh + c3], vnew[qh + c3]);
      v = 0.5 * (cl + ch);
    }
    b[t] = -v;
  }
}

template <int DIM>
__global__ void pressure_faces_kernel(Geom g, int d, const double* __restrict__ vold, const double* __restrict__ vnew,
                                      double* __restrict__ b) {
  pressure_faces_body<DIM>(g, d, vold, vnew, b, GridIx{});
}

// u_d = u_d - (GRADCC(pi)_d * chi) * (1/(1/dt)); u = 0 where the new
// porosity <= limit
template <class IX>
__device__ __forceinline__ void cc_correct_body(Geom g, double dt, double limit, int scaleP, const double* __restrict__ vold,
                                  const double* __restrict__ vnew, const double* __restrict__ pi, Vec3 u, IX ix) {
  const long long N = g.nvalid();
  const double factor = 0.5 / g.dx;
  const double dtScale = 1.0 / dt;
  const double scale = 1 / dtScale;
  for (long long t = ix.start(); t < N; t += ix.stride()) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const long long q = g.ci(i, j, k);
    const double chi = centred(vold[q + 3 * g.ncell], vnew[q + 3 * g.ncell]);
    const bool solid = vnew[q + 3 * g.ncell] <= limit;
    for (int d = 0; d < g.dim; ++d) {
      double gp = grad_cc(g, pi, factor, d, i, j, k);
      if (scaleP) gp = gp * chi;
      gp = gp * scale;
      const double un = u.p[d][q] - gp;
      u.p[d][q] = solid ? 0.0 : un;
    }
  }
}

__global__ void cc_correct_kernel(Geom g, double dt, double limit, int scaleP, const double* __restrict__ vold,
                                  const double* __restrict__ vnew, const double* __restrict__ pi, Vec3 u) {
  cc_correct_body(g, dt, limit, scaleP, vold, vnew, pi, u, GridIx{});
}

// c = (chi / Pi(chi)) * coeff (srcSubcycle/AMRLevelMushyLayerInit.cpp:665-689)
// with MushyLayerParams::calculatePermeability (src/MushyLayerParams.cpp
// :1031-1090; pow(x, 3.0) and pow(x, 2.0) as the once-rounded cube and
// square); hele_shaw > 0: harmonic combination with the Hele-Shaw cell
// permeability 1/(1/hele_shaw + 1/Pi)
template <class IX>
__device__ __forceinline__ void darcy_ccoef_law_body(Geom g, double coeff, int law, double hele_shaw,
                                       const double* __restrict__ chi, double* __restrict__ c, IX ix) {
  const long long N = g.nvalid();
  for (long long t = ix.start(); t < N; t += ix.stride()) {
    int i, j, k;
    cell_of(g, t, i, j, k);
    const long long q = g.ci(i, j, k);
    const double liquid = chi[q];
    const double solid = 1 - liquid;
    double perm;
    switch (law) {
      case ML_PERM_PURE_FLUID: perm = 1; break;
      case ML_PERM_CUBIC: perm = cube_rn(liquid); break;
      case ML_PERM_LOG: perm = -pow(liquid, 2.0) * log(solid); break;
      case ML_PERM_POROSITY: perm = liquid; break;
      default: perm = cube_rn(liquid) / (solid * solid); break;
    }
    if (hele_shaw > 0) perm = 1 / (1 / hele_shaw + 1 / perm);
    double v = liquid;
    v = v / perm;
    c[q] = v * coeff;
  }
}

__global__ void darcy_ccoef_law_kernel(Geom g, double coeff, int law, double hele_shaw,
                                       const double* __restrict__ chi, double* __restrict__ c) {
  darcy_ccoef_law_body(g, coeff, law, hele_shaw, chi, c, GridIx{});
}

template <int DIM, int STAGE>
__global__ __launch_bounds__(256) void flow_batch_kernel(const FlowJob* __restrict__ jobs,
                                                         const unsigned* __restrict__ map) {
  const unsigned w = map[blockIdx.x];
  const FlowJob& J = jobs[w >> kBatchTileBits];
  const long long b = w & kBatchTileMask;
  const JobIx cells{b, J.nbc};
  if (STAGE == kFlowEnthalpy) enthalpy_vars_body(J.g, J.cp, J.phi, J.vars, cells);
  if (STAGE == kFlowDarcy) darcy_ccoef_law_body(J.g, J.coeff, J.law, J.hele_shaw, J.chi, J.c, cells);
  if (STAGE == kFlowUstarSrc)
    ustar_src_body(J.g, J.buoyT, J.buoyS, J.scaleP, J.vold, J.vnew, J.pi, Vec3{{J.src[0], J.src[1], J.src[2]}}, cells);
  if (STAGE == kFlowVelocity)
    cc_velocity_body(J.g, J.dt, J.limit, J.scaleP, J.vold, J.vnew, J.pi, CVec3{{J.ustar[0], J.ustar[1], J.ustar[2]}},
                     Vec3{{J.u[0], J.u[1], J.u[2]}}, cells);
  if (STAGE == kFlowDivergence) cc_divergence_body(J.g, J.dt, CVec3{{J.u[0], J.u[1], J.u[2]}}, J.rhs, cells);
  if (STAGE >= kFlowPressFaces && STAGE < kFlowPressFaces + 3)
    pressure_faces_body<DIM>(J.g, STAGE - kFlowPressFaces, J.vold, J.vnew, J.b[STAGE - kFlowPressFaces],
                             JobIx{b, J.nbf[STAGE - kFlowPressFaces]});
  if (STAGE == kFlowCorrect)
    cc_correct_body(J.g, J.dt, J.limit, J.scaleP, J.vold, J.vnew, J.pi, Vec3{{J.u[0], J.u[1], J.u[2]}}, cells);
}

template <int DIM>
hipError_t flow_batch_dim(int stage, const FlowJob* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  const dim3 gr(nblocks), bl(256);
  switch (stage) {
    case kFlowEnthalpy: hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowEnthalpy>), gr, bl, 0, s, jobs, map); break;
    case kFlowDarcy: hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowDarcy>), gr, bl, 0, s, jobs, map); break;
    case kFlowUstarSrc: hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowUstarSrc>), gr, bl, 0, s, jobs, map); break;
    case kFlowVelocity: hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowVelocity>), gr, bl, 0, s, jobs, map); break;
    case kFlowDivergence: hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowDivergence>), gr, bl, 0, s, jobs, map); break;
    case kFlowPressFaces: hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowPressFaces>), gr, bl, 0, s, jobs, map); break;
    case kFlowPressFaces + 1:
      hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowPressFaces + 1>), gr, bl, 0, s, jobs, map);
      break;
    case kFlowPressFaces + 2:
      hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowPressFaces + 2>), gr, bl, 0, s, jobs, map);
      break;
    case kFlowCorrect: hipLaunchKernelGGL((flow_batch_kernel<DIM, kFlowCorrect>), gr, bl, 0, s, jobs, map); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t flow_batch(int dim, int stage, const FlowJob* jobs, const unsigned* map, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  return dim == 2 ? flow_batch_dim<2>(stage, jobs, map, nblocks, s) : flow_batch_dim<3>(stage, jobs, map, nblocks, s);
}

hipError_t enthalpy_vars(const Geom& g, const ClosureP<double>& cp, const double* phi, double* vars, hipStream_t s) {
  hipLaunchKernelGGL(enthalpy_vars_kernel, dim3(grid_for(g.nvalid())), dim3(256), 0, s, g, cp, phi, vars);
  return hipGetLastError();
}

hipError_t ustar_src(const Geom& g, double buoyT, double buoyS, int scaleP, const double* vold, const double* vnew,
                     const double* pi, double* const* src, hipStream_t s) {
  Vec3 o{{src[0], src[1], g.dim == 3 ? src[2] : nullptr}};
  hipLaunchKernelGGL(ustar_src_kernel, dim3(grid_for(g.nvalid())), dim3(256), 0, s, g, buoyT, buoyS, scaleP, vold,
                     vnew, pi, o);
  return hipGetLastError();
}

hipError_t cc_project_rhs(const Geom& g, double dt, double limit, int scaleP, const double* vold, const double* vnew,
                          const double* pi, const double* const* ustar, double* const* u, double* rhs,
                          hipStream_t s) {
  CVec3 us{{ustar[0], ustar[1], g.dim == 3 ? ustar[2] : nullptr}};
  Vec3 uo{{u[0], u[1], g.dim == 3 ? u[2] : nullptr}};
  CVec3 ui{{u[0], u[1], g.dim == 3 ? u[2] : nullptr}};
  hipLaunchKernelGGL(cc_velocity_kernel, dim3(grid_for(g.nvalid())), dim3(256), 0, s, g, dt, limit, scaleP, vold,
                     vnew, pi, us, uo);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(cc_divergence_kernel, dim3(grid_for(g.nvalid())), dim3(256), 0, s, g, dt, ui, rhs);
  return hipGetLastError();
}

hipError_t pressure_faces(const Geom& g, const double* vold, const double* vnew, double* const* b, hipStream_t s) {
  for (int d = 0; d < g.dim; ++d) {
    const int blocks = grid_for(g.nface[d]);
    if (g.dim == 2)
      hipLaunchKernelGGL((pressure_faces_kernel<2>), dim3(blocks), dim3(256), 0, s, g, d, vold, vnew, b[d]);
    else
      hipLaunchKernelGGL((pressure_faces_kernel<3>), dim3(blocks), dim3(256), 0, s, g, d, vold, vnew, b[d]);
  }
  return hipGetLastError();
}

hipError_t cc_correct(const Geom& g, double dt, double limit, int scaleP, const double* vold, const double* vnew,
                      const double* pi, double* const* u, hipStream_t s) {
  Vec3 uo{{u[0], u[1], g.dim == 3 ? u[2] : nullptr}};
  hipLaunchKernelGGL(cc_correct_kernel, dim3(grid_for(g.nvalid())), dim3(256), 0, s, g, dt, limit, scaleP, vold,
                     vnew, pi, uo);
  return hipGetLastError();
}

hipError_t darcy_ccoef_law(const Geom& g, double coeff, int law, double hele_shaw, const double* chi, double* c,
                           hipStream_t s) {
  hipLaunchKernelGGL(darcy_ccoef_law_kernel, dim3(grid_for(g.nvalid())), dim3(256), 0, s, g, coeff, law, hele_shaw,
                     chi, c);
  return hipGetLastError();
}

}  // namespace ml
