// ml_common.h — device/host shared definitions of the MI355X mushy-layer
// solve: level geometry in Chombo FAB layout, closure parameters, boundary
// descriptors and the (H, Theta) -> (theta, Theta_l) phase-diagram closure.
//
// Layout of a cell array (one ghost layer, component outermost):
//   idx(i,j,k,c) = (i+1) + sy*(j+1) + sz*(k+gz) + c*ncell
// gz = 1 in 3D, 0 in 2D.  Face arrays of direction d have dims n + e_d, no
// ghosts, face i between cells i-1 and i, component outermost.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ml {

constexpr int kMaxLevels = 24;

struct Geom {
  int dim;
  int n[3];         // valid cells (n[2] == 1 in 2D)
  int periodic[3];
  double dx;
  long long sy, sz, ncell;   // cell strides, doubles per component
  long long fn[3][3];        // face dims [dir][axis]
  long long nface[3];

  __host__ __device__ inline long long ci(int i, int j, int k) const {
    return (long long)(i + 1) + sy * (long long)(j + 1) + sz * (long long)(k + (dim == 3 ? 1 : 0));
  }
  __host__ __device__ inline long long fi(int d, int i, int j, int k) const {
    return (long long)i + fn[d][0] * ((long long)j + fn[d][1] * (long long)k);
  }
  __host__ __device__ inline long long nvalid() const {
    return (long long)n[0] * n[1] * n[2];
  }
};

// the bit pattern of a value (exact comparison, NaN payloads and signed
// zeros included)
__device__ __forceinline__ unsigned long long bits_of(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ unsigned bits_of(float x) { return __float_as_uint(x); }

__host__ __device__ inline Geom make_geom(int dim, const int n[3], const int periodic[3], double dx) {
  Geom g{};
  g.dim = dim;
  for (int d = 0; d < 3; ++d) {
    g.n[d] = (d < dim) ? n[d] : 1;
    g.periodic[d] = (d < dim) ? periodic[d] : 0;
  }
  g.dx = dx;
  g.sy = g.n[0] + 2;
  g.sz = (long long)(g.n[0] + 2) * (g.n[1] + 2);
  g.ncell = g.sz * (dim == 3 ? g.n[2] + 2 : 1);
  for (int d = 0; d < 3; ++d) {
    g.fn[d][0] = g.n[0] + (d == 0);
    g.fn[d][1] = g.n[1] + (d == 1);
    g.fn[d][2] = dim == 3 ? g.n[2] + (d == 2) : 1;
    g.nface[d] = g.fn[d][0] * g.fn[d][1] * g.fn[d][2];
  }
  return g;
}

__host__ __device__ inline Geom coarsen_geom(const Geom& f, int ratio) {
  int n[3] = {f.n[0] / ratio, f.n[1] / ratio, f.dim == 3 ? f.n[2] / ratio : 1};
  return make_geom(f.dim, n, f.periodic, f.dx * ratio);
}

// Closure constants (src/MushyLayerParams.h).
template <typename T>
struct ClosureP {
  T St, CR, pc, c, te, Te;
};

// Derived-variable (or scalar) boundary conditions per [dir][side][comp]:
// type 0 Dirichlet order 1, 1 Neumann.  `apply` = m_apply_bcs_to_diagnostic_var.
struct BCDesc {
  int type[3][2][2];
  double val[3][2][2];
  int apply;
};

// ---------------------------------------------------------------- closure
// CALCULATEPOROSITYMUSH (src/EnthalpyVariablesF.ChF:103-150).
template <typename T>
__host__ __device__ inline T porosity_mush(T C, T H, const ClosureP<T>& p) {
  T validC = (C >= -p.CR) ? C : -p.CR;
  if (p.St == T(0)) return T(1);
  T a = p.CR * (p.c - T(1)) + p.St * (p.pc - T(1));
  T b = p.CR * (T(1) - T(2) * p.c) + H * (T(1) - p.pc) - validC * (p.c - T(1)) - p.pc * p.St;
  T q = (p.CR + validC) * p.c + p.pc * H;
  return (-b - sqrt(b * b - T(4) * a * q)) / (T(2) * a);
}

// CALCULATE_BOUNDING_ENERGY (src/EnthalpyVariablesF.ChF:657-698) followed by
// CALCULATE_T_CL (src/EnthalpyVariablesF.ChF:419-467), one cell.
template <typename T>
__host__ __device__ inline void closure(T H, T C, const ClosureP<T>& p, T& theta, T& Cl) {
  T hs;
  if (C < -p.CR)
    hs = T(0);
  else
    hs = p.c * ((-C - p.CR) / p.pc);
  if (hs < T(0)) hs = T(0);
  hs = hs + p.te * p.c;
  T chie = (p.CR + C) / (p.CR + p.Te);
  T he = chie * (p.St + p.te * (T(1) - p.c)) + p.c * p.te;
  T hl = p.St - C + p.te + p.Te;
  if (H <= hs) {
    theta = H / p.c;
    Cl = p.Te;
  } else if (H > hs && H <= he) {
    theta = p.te;
    Cl = p.Te;
  } else if (H > he && H < hl) {
    T chi = porosity_mush(C, H, p);
    // one division: (-x)/y and -(x/y) are the same IEEE value
    Cl = (C + p.CR * (T(1) - chi)) / (chi + p.pc * (T(1) - chi));
    theta = -Cl;
  } else if (H >= hl) {
    theta = H - p.St;
    Cl = C;
  } else {
    theta = T(NAN);
    Cl = T(NAN);
  }
}

// Jacobian of the closure at (H, C), regime decided exactly as the closure
// decides it.  M[0] = dtheta/dH, M[1] = dtheta/dTheta, M[2] = dTheta_l/dH,
// M[3] = dTheta_l/dTheta.
__host__ __device__ inline void closure_jacobian(double H, double C, const ClosureP<double>& p, double* M) {
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

// CALCULATEPOROSITY_MULTICOMP (src/EnthalpyVariablesF.ChF:56-94).
template <typename T>
__host__ __device__ inline T porosity(T H, T C, const ClosureP<T>& p) {
  T hs;
  if (C < -p.CR)
    hs = T(0);
  else
    hs = p.c * ((-C - p.CR) / p.pc);
  if (hs < T(0)) hs = T(0);
  hs = hs + p.te * p.c;
  T chie = (p.CR + C) / (p.CR + p.Te);
  T he = chie * (p.St + p.te * (T(1) - p.c)) + p.c * p.te;
  T hl = p.St - C + p.te + p.Te;
  if (H <= hs) return T(0);
  if (H > hs && H <= he) return (H - p.te * p.c) / (p.St + p.te * (T(1) - p.c));
  if (H > he && H < hl) return porosity_mush(C, H, p);
  if (H >= hl) return T(1);
  return T(NAN);
}

// x^3 rounded once: x*x = p + e and p*x = h + l exactly (FMA remainders), so
// x^3 = h + (l + e*x) to within a tiny fraction of an ulp before the final
// rounding.  Host and device (and the oracle) evaluate the same expression,
// so they agree bit for bit; it is the value of a correctly rounded
// pow(x, 3.0) (MushyLayerParams::calculatePermeability,
// src/MushyLayerParams.cpp:1031-1090).
__host__ __device__ inline double cube_rn(double x) {
  const double p = x * x;
  const double e = fma(x, x, -p);
  const double h = p * x;
  const double l = fma(p, x, -h);
  return h + (l + e * x);
}

// Boundary ghost value from the adjacent interior value x1:
// Dirichlet g = 2v - x1 (BCutil/BCFuncF.ChF:40-46),
// Neumann  g = x1 + isign*dx*v (BCutil/BCFuncF.ChF:80-88).
template <typename T>
__host__ __device__ inline T bc_ghost(int type, T v, T x1, int isign, T dx) {
  if (type == 0) return T(2) * v - x1;
  return x1 + (T(isign) * dx) * v;
}

}  // namespace ml
