// ml_adjoint.hip — the adjoint of one backward-Euler step of the (H, Theta)
// operator (a build extension: the reference has no adjoint).
//
// Forward step n (the single-instance FP64 solve):
//   R(phi^n) = A(phi^n; b(chi(phi^{n-1})), St) - rhs = 0,
//   A(phi) = alpha phi - beta/dx^2 S_b(phi_l(phi; St)),
//   rhs = phi^{n-1} + dt src(F, b),   alpha = 1, beta = -dt,
// with S_b the face-coefficient stencil of the reference's operator
// (NLVCCOMPUTEOP, src/AMRNonLinearMultiCompOpF.ChF:221-313), phi_l the
// closure (src/EnthalpyVariablesF.ChF:419-467, :657-698), b the HC face
// coefficients of the porosity at t^{n-1}
// (srcSubcycle/AMRLevelMushyLayerInit.cpp:914-953, geometric faces
// src/CellToEdge2F.ChF:20-46) and src the bottom heat flux F as the source
// that is exactly the Neumann ghost g = phi_1 - dx F (BCutil/BCFuncF.ChF:80-88)
// on the theta component: src_H = b_H,face F / dx in the bottom cells.
//
// Given gbar = dL/dphi^n, the step's adjoint solves J^T lambda = gbar with
// J = dA/dphi^n = alpha I - beta/dx^2 S M (M the closure Jacobian, S
// symmetric with the homogeneous wall rule), then
//   dL/dphi^{n-1} = lambda + (dL/dchi^{n-1}) dchi/dphi^{n-1},
//   dL/dchi_c     = -sum_faces mu_f db_f/dchi_c,
//   mu_f          = lambda^T dA/db_f,
//   dL/dF_x      += dt lambda_H(bottom cell x) b_H,face / dx,
//   dL/dSt       += beta/dx^2 (S lambda)^T dphi_l/dSt + (dL/dchi)^T dchi/dSt.
// The solve is BiCGStab preconditioned by the transposed smoother (block
// red/black Gauss-Seidel on J^T in the reverse colour order of the forward
// smoother) or by the 2x2 block-diagonal of J^T, all vectors on the device,
// dot products reduced per workgroup in a fixed order and summed on the host
// in a fixed order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "ml_internal.h"

namespace ml {

struct AdjWork {
  Geom g{};
  size_t n2 = 0;   // doubles of a 2-component cell array
  double *M = nullptr, *pl = nullptr, *pst = nullptr, *X = nullptr;
  double *x = nullptr, *r = nullptr, *rh = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr,
         *y = nullptr, *z = nullptr, *rhs = nullptr;
  double* part = nullptr;    // per-workgroup partial sums (device)
  double* hpart = nullptr;   // pinned mirror
  int nblk = 0;
  std::vector<void*> owned;
};

namespace {

constexpr int kT = 256;
constexpr int kMaxBlk = 512;

struct AdjGeo {
  Geom g;
  BCDesc bc;
  ClosureP<double> cp;
  double alpha, bd;   // bd = beta / dx^2
  double DC;          // salt diffusion coefficient: b_Theta = -chi_face DC
  const double* b[3];
};

__device__ inline void cell_of(const Geom& g, long long c, int& i, int& j) {
  i = int(c % g.n[0]);
  j = int(c / g.n[0]);
}

// d(theta, Theta_l)/dSt at fixed (H, Theta), regime as the closure decides
__device__ void closure_dst(double H, double C, const ClosureP<double>& p, double* d) {
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
  d[0] = d[1] = 0.0;
  if (H > he && H < hl) {
    if (p.St == 0.0) return;
    const double chi = porosity_mush<double>(C, H, p);
    const double Cv = (C >= -p.CR) ? C : -p.CR;
    const double a = p.CR * (p.c - 1.0) + p.St * (p.pc - 1.0);
    const double b = p.CR * (1.0 - 2.0 * p.c) + H * (1.0 - p.pc) - Cv * (p.c - 1.0) - p.pc * p.St;
    const double Fchi = 2.0 * a * chi + b;
    const double chiS = -((p.pc - 1.0) * chi * chi - p.pc * chi) / Fchi;
    const double D = chi + p.pc * (1.0 - chi);
    const double Tl = (C + p.CR * (1.0 - chi)) / D;
    const double dTl = -(p.CR + Tl * (1.0 - p.pc)) / D;
    d[1] = dTl * chiS;
    d[0] = -d[1];
  } else if (H >= hl) {
    d[0] = -1.0;
  }
}

// porosity and its derivatives in H, Theta and St
__device__ void porosity_jac(double H, double C, const ClosureP<double>& p, double* X) {
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
  X[0] = porosity<double>(H, C, p);
  X[1] = X[2] = X[3] = 0.0;
  if (H > hs && H <= he) {
    const double den = p.St + p.te * (1.0 - p.c);
    X[1] = 1.0 / den;
    X[3] = -X[0] / den;
  } else if (H > he && H < hl && p.St != 0.0) {
    const double chi = X[0];
    const double Cv = (C >= -p.CR) ? C : -p.CR;
    const double a = p.CR * (p.c - 1.0) + p.St * (p.pc - 1.0);
    const double b = p.CR * (1.0 - 2.0 * p.c) + H * (1.0 - p.pc) - Cv * (p.c - 1.0) - p.pc * p.St;
    const double Fchi = 2.0 * a * chi + b;
    X[1] = -(chi * (1.0 - p.pc) + p.pc) / Fchi;
    X[2] = (C >= -p.CR) ? -(p.c - chi * (p.c - 1.0)) / Fchi : 0.0;
    X[3] = -((p.pc - 1.0) * chi * chi - p.pc * chi) / Fchi;
  }
}

// the closure Jacobian M (4 comps), phi_l (2) and dphi_l/dSt (2) of the new state
__global__ void adj_state_kernel(Geom g, ClosureP<double> cp, const double* __restrict__ phi, double* M, double* pl,
                                 double* pst) {
  const long long N = g.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, c, i, j);
    const long long q = g.ci(i, j, 0);
    const double H = phi[q], C = phi[q + g.ncell];
    double m[4], th, cl, d[2];
    closure_jacobian(H, C, cp, m);
    closure<double>(H, C, cp, th, cl);
    closure_dst(H, C, cp, d);
#pragma unroll
    for (int k = 0; k < 4; ++k) M[q + k * g.ncell] = m[k];
    pl[q] = th;
    pl[q + g.ncell] = cl;
    pst[q] = d[0];
    pst[q + g.ncell] = d[1];
  }
}

// porosity of the previous state and its derivatives (4 comps)
__global__ void adj_prev_kernel(Geom g, ClosureP<double> cp, const double* __restrict__ phi, double* X) {
  const long long N = g.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, c, i, j);
    const long long q = g.ci(i, j, 0);
    double x[4];
    porosity_jac(phi[q], phi[q + g.ncell], cp, x);
#pragma unroll
    for (int k = 0; k < 4; ++k) X[q + k * g.ncell] = x[k];
  }
}

// neighbour of cell (i, j) across face (d, side): index of the neighbour
// (periodic wrap) or -1 at a wall
__device__ inline long long nbr(const Geom& g, int i, int j, int d, int side) {
  int ii = i + (d == 0 ? (side ? 1 : -1) : 0);
  int jj = j + (d == 1 ? (side ? 1 : -1) : 0);
  const int v = d == 0 ? ii : jj;
  if (v < 0 || v >= g.n[d]) {
    if (!g.periodic[d]) return -1;
    if (d == 0) ii = (ii + g.n[0]) % g.n[0];
    else jj = (jj + g.n[1]) % g.n[1];
  }
  return g.ci(ii, jj, 0);
}

// face coefficient of component n on face (d, side) of cell (i, j)
__device__ inline double face_b(const AdjGeo& G, int i, int j, int d, int side, int n) {
  const Geom& g = G.g;
  const int fi = i + (d == 0 && side ? 1 : 0), fj = j + (d == 1 && side ? 1 : 0);
  return G.b[d][g.fi(d, fi, fj, 0) + n * g.nface[d]];
}

// homogeneous wall rule of component n on wall (d, side): ghost = s x_1
__device__ inline double wall_s(const BCDesc& bc, int d, int side, int n) {
  if (!bc.apply) return 1.0;
  return bc.type[d][side][n] == 0 ? -1.0 : 1.0;
}

// (S x) of component n at cell (i, j), q its index
__device__ inline double stencil(const AdjGeo& G, const double* x, int i, int j, long long q, int n) {
  const Geom& g = G.g;
  const double xc = x[q + n * g.ncell];
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < 2; ++d)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const long long nb = nbr(g, i, j, d, side);
      const double xn = nb >= 0 ? x[nb + n * g.ncell] : wall_s(G.bc, d, side, n) * xc;
      s += face_b(G, i, j, d, side, n) * (xn - xc);
    }
  return s;
}

// y = J^T x on valid cells
__global__ void adj_apply_kernel(AdjGeo G, const double* __restrict__ M, const double* __restrict__ x, double* y) {
  const Geom& g = G.g;
  const long long N = g.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, c, i, j);
    const long long q = g.ci(i, j, 0);
    const double w0 = stencil(G, x, i, j, q, 0), w1 = stencil(G, x, i, j, q, 1);
    const double m0 = M[q], m1 = M[q + g.ncell], m2 = M[q + 2 * g.ncell], m3 = M[q + 3 * g.ncell];
    y[q] = G.alpha * x[q] - G.bd * (m0 * w0 + m2 * w1);
    y[q + g.ncell] = G.alpha * x[q + g.ncell] - G.bd * (m1 * w0 + m3 * w1);
  }
}

// z = P^{-1} r with P the 2x2 diagonal block of J^T at each cell
__global__ void adj_jacobi_kernel(AdjGeo G, const double* __restrict__ M, const double* __restrict__ r, double* z) {
  const Geom& g = G.g;
  const long long N = g.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, c, i, j);
    const long long q = g.ci(i, j, 0);
    double sd[2] = {0.0, 0.0};
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
      for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const double bf = face_b(G, i, j, d, side, n);
          sd[n] += nbr(g, i, j, d, side) >= 0 ? -bf : bf * (wall_s(G.bc, d, side, n) - 1.0);
        }
    const double m0 = M[q], m1 = M[q + g.ncell], m2 = M[q + 2 * g.ncell], m3 = M[q + 3 * g.ncell];
    const double p00 = G.alpha - G.bd * m0 * sd[0], p01 = -G.bd * m2 * sd[1];
    const double p10 = -G.bd * m1 * sd[0], p11 = G.alpha - G.bd * m3 * sd[1];
    const double det = p00 * p11 - p01 * p10;
    const double r0 = r[q], r1 = r[q + g.ncell];
    z[q] = (p11 * r0 - p01 * r1) / det;
    z[q + g.ncell] = (p00 * r1 - p10 * r0) / det;
  }
}

// One colour of the transposed smoother: block Gauss-Seidel on J^T
// z = P^{-1} r, updating the cells with (i + j) % 2 == colour from the
// other colour's current z:  z_q = D_q^{-1} (r_q + bd M_q^T w_q), w_q,n the
// face-weighted sum of the interior neighbours' z_n (the wall terms sit in
// D_q, the 2x2 diagonal block of J^T).  The forward smoother relaxes red
// then black; its transpose relaxes black then red.
__global__ void adj_gs_kernel(AdjGeo G, const double* __restrict__ M, const double* __restrict__ r, double* z,
                              int colour) {
  const Geom& g = G.g;
  const long long N = g.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, c, i, j);
    if (((i + j) & 1) != colour) continue;
    const long long q = g.ci(i, j, 0);
    double sd[2] = {0.0, 0.0}, wn[2] = {0.0, 0.0};
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
      for (int d = 0; d < 2; ++d)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          const double bf = face_b(G, i, j, d, side, n);
          const long long nb = nbr(g, i, j, d, side);
          if (nb >= 0) {
            sd[n] += -bf;
            wn[n] += bf * z[nb + n * g.ncell];
          } else {
            sd[n] += bf * (wall_s(G.bc, d, side, n) - 1.0);
          }
        }
    const double m0 = M[q], m1 = M[q + g.ncell], m2 = M[q + 2 * g.ncell], m3 = M[q + 3 * g.ncell];
    const double p00 = G.alpha - G.bd * m0 * sd[0], p01 = -G.bd * m2 * sd[1];
    const double p10 = -G.bd * m1 * sd[0], p11 = G.alpha - G.bd * m3 * sd[1];
    const double det = p00 * p11 - p01 * p10;
    const double r0 = r[q] + G.bd * (m0 * wn[0] + m2 * wn[1]);
    const double r1 = r[q + g.ncell] + G.bd * (m1 * wn[0] + m3 * wn[1]);
    z[q] = (p11 * r0 - p01 * r1) / det;
    z[q + g.ncell] = (p00 * r1 - p10 * r0) / det;
  }
}

// out = a x + b y on valid cells (y may be null)
__global__ void adj_axpby_kernel(Geom g, double a, const double* __restrict__ x, double b, const double* y, double* out) {
  const long long N = g.nvalid();
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, c, i, j);
    const long long q = g.ci(i, j, 0);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const long long k = q + n * g.ncell;
      out[k] = y ? a * x[k] + b * y[k] : a * x[k];
    }
  }
}

// per-workgroup partial sums of a.b and c.d over valid cells (c may be
// null), in a fixed order
__global__ __launch_bounds__(kT) void adj_dot_kernel(Geom g, const double* __restrict__ a, const double* __restrict__ b,
                                                     const double* c, const double* d, double* part) {
  __shared__ double s0[kT], s1[kT];
  const long long N = g.nvalid();
  double u = 0.0, w = 0.0;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < N; t += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, t, i, j);
    const long long q = g.ci(i, j, 0);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
      const long long k = q + n * g.ncell;
      u += a[k] * b[k];
      if (c) w += c[k] * d[k];
    }
  }
  s0[threadIdx.x] = u;
  s1[threadIdx.x] = w;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      s0[threadIdx.x] += s0[threadIdx.x + h];
      s1[threadIdx.x] += s1[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

// gradients after the solve: dL/dphi^{n-1} (valid cells), the heat-flux
// gradient (bottom faces, accumulated) and per-workgroup partial sums of
// the St gradient
__global__ __launch_bounds__(kT) void adj_grad_kernel(AdjGeo G, const double* __restrict__ lam,
                                                      const double* __restrict__ pl, const double* __restrict__ pst,
                                                      const double* __restrict__ X, double dt, double* gold,
                                                      double* gflux, double* part) {
  __shared__ double s0[kT];
  const Geom& g = G.g;
  const long long N = g.nvalid();
  double st = 0.0;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, c, i, j);
    const long long q = g.ci(i, j, 0);
    const long long C1 = g.ncell;
    // dL/dchi_c = -sum_f mu_f db_f/dchi_c over the Theta faces (b_H does
    // not depend on chi when the conductivity ratio is 1)
    const double chic = X[q];
    double dLdchi = 0.0;
#pragma unroll
    for (int d = 0; d < 2; ++d)
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        const long long nb = nbr(g, i, j, d, side);
        const double pc = pl[q + C1], lc = lam[q + C1];
        double mu, dcf;
        if (nb >= 0) {
          mu = -G.bd * (pl[nb + C1] - pc) * (lc - lam[nb + C1]);
          const double chin = X[nb];
          const double cf = sqrt(fabs(chic * chin));
          dcf = cf > 0.0 ? (chic * chin >= 0.0 ? chin : -chin) / (2.0 * cf) : 0.0;
        } else {
          // wall: the face porosity is the cell's, the ghost follows the BC
          double ghost = pc;
          if (G.bc.apply) {
            const int isign = side == 0 ? -1 : 1;
            ghost = bc_ghost<double>(G.bc.type[d][side][1], G.bc.val[d][side][1], pc, isign, g.dx);
          }
          mu = -G.bd * (ghost - pc) * lc;
          dcf = chic > 0.0 ? 1.0 : 0.0;
        }
        dLdchi -= mu * (-G.DC * dcf);
      }
    gold[q] = lam[q] + dLdchi * X[q + C1];
    gold[q + C1] = lam[q + C1] + dLdchi * X[q + 2 * C1];
    // St: beta/dx^2 (S lambda)^T dphi_l/dSt + dL/dchi dchi/dSt
    st += G.bd * (stencil(G, lam, i, j, q, 0) * pst[q] + stencil(G, lam, i, j, q, 1) * pst[q + C1]) +
          dLdchi * X[q + 3 * C1];
    // heat flux through the bottom faces: src_H = b_H F / dx
    if (j == 0 && gflux) gflux[i] += dt * lam[q] * face_b(G, i, 0, 1, 0, 0) / g.dx;
  }
  s0[threadIdx.x] = st;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) s0[threadIdx.x] += s0[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s0[0];
}

// src = 0 except the theta component of the bottom cells: b_H,face F / dx
__global__ void flux_source_kernel(Geom g, const double* __restrict__ bbot, const double* __restrict__ F, double* src) {
  const long long N = 2 * g.ncell;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < N; t += (long long)gridDim.x * blockDim.x)
    src[t] = 0.0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n[0]; i += gridDim.x * blockDim.x)
    src[g.ci(i, 0, 0)] = bbot[g.fi(1, i, 0, 0)] * F[i] / g.dx;
}

// g_phi += g_chi dchi/dphi on valid cells; per-workgroup partial sums of
// g_chi dchi/dSt (the porosity's own dependence on St)
__global__ __launch_bounds__(kT) void porosity_vjp_kernel(Geom g, ClosureP<double> cp, const double* __restrict__ phi,
                                                          const double* __restrict__ gchi, double* gphi, double* part) {
  __shared__ double s0[kT];
  const long long N = g.nvalid();
  double st = 0.0;
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
    int i, j;
    cell_of(g, c, i, j);
    const long long q = g.ci(i, j, 0);
    double x[4];
    porosity_jac(phi[q], phi[q + g.ncell], cp, x);
    gphi[q] += gchi[q] * x[1];
    gphi[q + g.ncell] += gchi[q] * x[2];
    st += gchi[q] * x[3];
  }
  s0[threadIdx.x] = st;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) s0[threadIdx.x] += s0[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s0[0];
}

int blocks_for(long long n) {
  long long b = (n + kT - 1) / kT;
  if (b > kMaxBlk) b = kMaxBlk;
  return int(b < 1 ? 1 : b);
}

void adj_free(AdjWork* w) {
  if (!w) return;
  for (void* p : w->owned) (void)hipFree(p);
  if (w->hpart) (void)hipHostFree(w->hpart);
  delete w;
}

int adj_init(ml_hc_op op) {
  if (op->adj) return ML_OK;
  ml_ctx ctx = op->ctx;
  AdjWork* w = new AdjWork();
  op->adj = std::unique_ptr<AdjWork, void (*)(AdjWork*)>(w, adj_free);
  w->g = op->lev[0].g;
  w->n2 = 2 * size_t(w->g.ncell);
  double** arrs[] = {&w->pl, &w->pst, &w->x, &w->r, &w->rh, &w->p, &w->v, &w->s, &w->t, &w->y, &w->z, &w->rhs};
  for (double** a : arrs) {
    void* v = nullptr;
    ML_HIP(ctx, hipMalloc(&v, sizeof(double) * w->n2));
    w->owned.push_back(v);
    ML_HIP(ctx, hipMemsetAsync(v, 0, sizeof(double) * w->n2, ctx->stream));
    *a = static_cast<double*>(v);
  }
  for (double** a : {&w->M, &w->X}) {
    void* v = nullptr;
    ML_HIP(ctx, hipMalloc(&v, sizeof(double) * 2 * w->n2));
    w->owned.push_back(v);
    ML_HIP(ctx, hipMemsetAsync(v, 0, sizeof(double) * 2 * w->n2, ctx->stream));
    *a = static_cast<double*>(v);
  }
  void* v = nullptr;
  ML_HIP(ctx, hipMalloc(&v, sizeof(double) * 2 * kMaxBlk));
  w->owned.push_back(v);
  w->part = static_cast<double*>(v);
  ML_HIP(ctx, hipHostMalloc(&v, sizeof(double) * 2 * kMaxBlk, hipHostMallocDefault));
  w->hpart = static_cast<double*>(v);
  w->nblk = blocks_for(w->g.nvalid());
  return ML_OK;
}

// a.b (and c.d): partial sums per workgroup, added on the host in order
int dots(ml_ctx ctx, AdjWork* w, const double* a, const double* b, const double* c, const double* d, double* ab,
         double* cd) {
  hipLaunchKernelGGL(adj_dot_kernel, dim3(w->nblk), dim3(kT), 0, ctx->stream, w->g, a, b, c, d, w->part);
  ML_HIP(ctx, hipGetLastError());
  ML_HIP(ctx, hipMemcpyAsync(w->hpart, w->part, sizeof(double) * 2 * w->nblk, hipMemcpyDeviceToHost, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double s0 = 0.0, s1 = 0.0;
  for (int k = 0; k < w->nblk; ++k) {
    s0 += w->hpart[2 * k];
    s1 += w->hpart[2 * k + 1];
  }
  *ab = s0;
  if (cd) *cd = s1;
  return ML_OK;
}

}  // namespace

void adj_work_free(AdjWork* w) { adj_free(w); }

}  // namespace ml

using namespace ml;

extern "C" {

int ml_hc_flux_source(ml_hc_op op, const double* flux, double* src) {
  if (!op || !flux || !src) return ML_EINVAL;
  if (op->dim != 2) return fail(op->ctx, ML_ENOTSUP, "ml_hc_flux_source: 2D only");
  const Geom& g = op->lev[0].g;
  hipLaunchKernelGGL(flux_source_kernel, dim3(1), dim3(kT), 0, op->ctx->stream, g, op->lev[0].b[1], flux, src);
  ML_HIP(op->ctx, hipGetLastError());
  return ML_OK;
}

int ml_hc_porosity_vjp(ml_hc_op op, const double* phi, const double* g_chi, double* g_phi, double* g_stefan) {
  if (!op || !phi || !g_chi || !g_phi) return ML_EINVAL;
  ml_ctx ctx = op->ctx;
  if (op->dim != 2) return fail(ctx, ML_ENOTSUP, "ml_hc_porosity_vjp: 2D only");
  int rc = adj_init(op);
  if (rc) return rc;
  AdjWork* w = op->adj.get();
  hipLaunchKernelGGL(porosity_vjp_kernel, dim3(w->nblk), dim3(kT), 0, ctx->stream, w->g, op->cp, phi, g_chi, g_phi,
                     w->part);
  ML_HIP(ctx, hipGetLastError());
  ML_HIP(ctx, hipMemcpyAsync(w->hpart, w->part, sizeof(double) * w->nblk, hipMemcpyDeviceToHost, ctx->stream));
  ML_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double st = 0.0;
  for (int k = 0; k < w->nblk; ++k) st += w->hpart[k];
  if (g_stefan) *g_stefan += st;
  return ML_OK;
}

int ml_hc_adjoint_step(ml_hc_op op, const double* phi_new, const double* phi_old, double dt, const double* g_new,
                       double* g_old, double* g_flux, double* g_stefan, const ml_solve_opts* opts,
                       ml_solve_stats* stats) {
  if (!op || !phi_new || !phi_old || !g_new || !g_old || !(dt > 0)) return ML_EINVAL;
  ml_ctx ctx = op->ctx;
  if (op->dim != 2) return fail(ctx, ML_ENOTSUP, "ml_hc_adjoint_step: 2D only");
  if (op->params.heatConductivityRatio != 1.0)
    return fail(ctx, ML_ENOTSUP, "ml_hc_adjoint_step: conductivity ratio 1 only (b_H independent of porosity)");
  if (op->lev[0].a) return fail(ctx, ML_ENOTSUP, "ml_hc_adjoint_step: aCoef == 1 only");
  ml_solve_opts o;
  if (opts)
    o = *opts;
  else
    ml_default_opts(&o);
  int rc = adj_init(op);
  if (rc) return rc;
  AdjWork* w = op->adj.get();
  hipStream_t s = ctx->stream;
  const Geom& g = w->g;
  const int nb = w->nblk;
  // face coefficients of the previous state's porosity (the step's
  // operator), the closure data of the new state, porosity data of the old
  ML_HIP(ctx, porosity_field(g, op->cp, phi_old, w->y, s));
  ML_HIP(ctx, hc_bcoef(g, op->params.heatConductivityRatio, op->params.heatDiffusionCoeff, op->params.saltDiffusionCoeff,
                       w->y, op->lev[0].b, s));
  hipLaunchKernelGGL(adj_state_kernel, dim3(nb), dim3(kT), 0, s, g, op->cp, phi_new, w->M, w->pl, w->pst);
  ML_HIP(ctx, hipGetLastError());
  hipLaunchKernelGGL(adj_prev_kernel, dim3(nb), dim3(kT), 0, s, g, op->cp, phi_old, w->X);
  ML_HIP(ctx, hipGetLastError());
  AdjGeo G{};
  G.g = g;
  G.bc = op->bc;
  G.cp = op->cp;
  G.alpha = 1.0;
  G.bd = -dt / (g.dx * g.dx);
  G.DC = op->params.saltDiffusionCoeff;
  for (int d = 0; d < 3; ++d) G.b[d] = op->lev[0].b[d];
  auto apply = [&](const double* x, double* y) -> int {
    hipLaunchKernelGGL(adj_apply_kernel, dim3(nb), dim3(kT), 0, s, G, w->M, x, y);
    ML_HIP(ctx, hipGetLastError());
    return ML_OK;
  };
  // preconditioner: o.nu1 sweeps of the transposed smoother (block
  // Gauss-Seidel on J^T, black then red) from z = 0; nu1 == 0: the 2x2
  // block-diagonal (Jacobi)
  // (an odd periodic length would put two cells of one colour side by side
  // across the seam: the block diagonal is used there)
  bool odd_seam = false;
  for (int d = 0; d < 2; ++d) odd_seam = odd_seam || (g.periodic[d] && (g.n[d] & 1));
  const int sweeps = (o.nu1 > 0 && !odd_seam) ? o.nu1 : 0;
  auto prec = [&](const double* r, double* z) -> int {
    if (sweeps == 0) {
      hipLaunchKernelGGL(adj_jacobi_kernel, dim3(nb), dim3(kT), 0, s, G, w->M, r, z);
      ML_HIP(ctx, hipGetLastError());
      return ML_OK;
    }
    hipLaunchKernelGGL(adj_axpby_kernel, dim3(nb), dim3(kT), 0, s, g, 0.0, r, 0.0, nullptr, z);
    ML_HIP(ctx, hipGetLastError());
    for (int k = 0; k < sweeps; ++k)
      for (int colour = 1; colour >= 0; --colour) {
        hipLaunchKernelGGL(adj_gs_kernel, dim3(nb), dim3(kT), 0, s, G, w->M, r, z, colour);
        ML_HIP(ctx, hipGetLastError());
      }
    return ML_OK;
  };
  auto axpby = [&](double a, const double* x, double b, const double* y, double* out) -> int {
    hipLaunchKernelGGL(adj_axpby_kernel, dim3(nb), dim3(kT), 0, s, g, a, x, b, y, out);
    ML_HIP(ctx, hipGetLastError());
    return ML_OK;
  };
  // BiCGStab on J^T lambda = g_new (right block-Jacobi preconditioning),
  // from lambda = 0
  double bb;
  if ((rc = axpby(1.0, g_new, 0.0, nullptr, w->rhs))) return rc;
  if ((rc = dots(ctx, w, w->rhs, w->rhs, nullptr, nullptr, &bb, nullptr))) return rc;
  const double bnorm = std::sqrt(bb);
  if ((rc = axpby(0.0, w->rhs, 0.0, nullptr, w->x))) return rc;
  if ((rc = axpby(1.0, w->rhs, 0.0, nullptr, w->r))) return rc;
  if ((rc = axpby(1.0, w->rhs, 0.0, nullptr, w->rh))) return rc;
  if ((rc = axpby(0.0, w->rhs, 0.0, nullptr, w->p))) return rc;
  if ((rc = axpby(0.0, w->rhs, 0.0, nullptr, w->v))) return rc;
  double rho = 1.0, alpha = 1.0, omega = 1.0, rnorm = bnorm;
  int it = 0, status = bnorm == 0.0 ? 8 : 0;
  const double tol = o.eps * bnorm;
  while (status == 0 && it < o.max_iter) {
    ++it;
    double rho1;
    if ((rc = dots(ctx, w, w->rh, w->r, nullptr, nullptr, &rho1, nullptr))) return rc;
    if (rho1 == 0.0) {
      status = 4;
      break;
    }
    const double beta = (rho1 / rho) * (alpha / omega);
    rho = rho1;
    // p = r + beta (p - omega v)
    if ((rc = axpby(1.0, w->p, -omega, w->v, w->p))) return rc;
    if ((rc = axpby(1.0, w->r, beta, w->p, w->p))) return rc;
    if ((rc = prec(w->p, w->y))) return rc;
    if ((rc = apply(w->y, w->v))) return rc;
    double rv;
    if ((rc = dots(ctx, w, w->rh, w->v, nullptr, nullptr, &rv, nullptr))) return rc;
    alpha = rho / rv;
    // s = r - alpha v
    if ((rc = axpby(1.0, w->r, -alpha, w->v, w->s))) return rc;
    if ((rc = axpby(1.0, w->x, alpha, w->y, w->x))) return rc;
    double ss;
    if ((rc = dots(ctx, w, w->s, w->s, nullptr, nullptr, &ss, nullptr))) return rc;
    if (std::sqrt(ss) <= tol) {
      rnorm = std::sqrt(ss);
      status = 1;
      break;
    }
    if ((rc = prec(w->s, w->z))) return rc;
    if ((rc = apply(w->z, w->t))) return rc;
    double ts, tt;
    if ((rc = dots(ctx, w, w->t, w->s, w->t, w->t, &ts, &tt))) return rc;
    omega = tt > 0.0 ? ts / tt : 0.0;
    if ((rc = axpby(1.0, w->x, omega, w->z, w->x))) return rc;
    if ((rc = axpby(1.0, w->s, -omega, w->t, w->r))) return rc;
    double rr;
    if ((rc = dots(ctx, w, w->r, w->r, nullptr, nullptr, &rr, nullptr))) return rc;
    rnorm = std::sqrt(rr);
    if (rnorm <= tol) status = 1;
    if (omega == 0.0 && status == 0) status = 4;
  }
  if (status == 0) status = 2;
  // gradients
  hipLaunchKernelGGL(adj_grad_kernel, dim3(nb), dim3(kT), 0, s, G, w->x, w->pl, w->pst, w->X, dt, g_old, g_flux,
                     w->part);
  ML_HIP(ctx, hipGetLastError());
  ML_HIP(ctx, hipMemcpyAsync(w->hpart, w->part, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  ML_HIP(ctx, hipStreamSynchronize(s));
  double gst = 0.0;
  for (int k = 0; k < nb; ++k) gst += w->hpart[k];
  if (g_stefan) *g_stefan += gst;
  if (stats) {
    stats->exit_status = status;
    stats->iterations = it;
    stats->initial_residual = bnorm;
    stats->final_residual = rnorm;
  }
  return ML_OK;
}

}  // extern "C"
