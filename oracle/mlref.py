"""ctypes wrapper of the CPU oracle ``libmlref.so``.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``.  The product package never
imports this module.

Arrays follow the Chombo FAB layout used throughout (one ghost layer,
component outermost); see ``oracle/mlref.h``.  Cell arrays are numpy float64
arrays of shape ``(ncomp, nz+2|1, ny+2, nx+2)``; face arrays of direction d
have shape ``(ncomp, fz, fy, fx)`` with ``f = n + e_d``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Grid(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("periodic", C.c_int * 3), ("dx", C.c_double)]


class Params(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "stefan", "compositionRatio", "waterDistributionCoeff", "specificHeatRatio",
        "heatConductivityRatio", "thetaEutectic", "ThetaEutectic", "heatDiffusionCoeff",
        "saltDiffusionCoeff")]


class BC(C.Structure):
    _fields_ = [("type", C.c_int * 2 * 2 * 3), ("val", C.c_double * 2 * 2 * 3), ("apply", C.c_int)]


class AdvectBC(C.Structure):
    _fields_ = [("type", C.c_int * 2 * 2 * 3), ("val", C.c_double * 2 * 2 * 3)]


class Opts(C.Structure):
    _fields_ = [("nu1", C.c_int), ("nu2", C.c_int), ("bottom", C.c_int), ("max_iter", C.c_int),
                ("eps", C.c_double), ("hang", C.c_double), ("norm_thresh", C.c_double),
                ("bottom_solver", C.c_int), ("bottom_max_iter", C.c_int)]


BOTTOM_RELAX, BOTTOM_BICGSTAB = 0, 1


class Stats(C.Structure):
    _fields_ = [("exit_status", C.c_int), ("iterations", C.c_int),
                ("initial_residual", C.c_double), ("final_residual", C.c_double)]


def build():
    """Compile libmlref.so with the committed Makefile (gcc, -O2 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libmlref.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_dp = C.POINTER(C.c_double)


def _declare(L):
    G = C.POINTER(Grid)
    P = C.POINTER(Params)
    B = C.POINTER(BC)
    pp = C.POINTER(_dp)
    sig = {
        "mlr_ncell": (C.c_long, [G]),
        "mlr_nface": (C.c_long, [G, C.c_int]),
        "mlr_mg_depth": (C.c_int, [G]),
        "mlr_bounding_energy": (None, [C.c_double, C.c_double, P, _dp, _dp, _dp]),
        "mlr_porosity_mush": (C.c_double, [C.c_double, C.c_double, P]),
        "mlr_closure_point": (None, [C.c_double, C.c_double, P, _dp, _dp]),
        "mlr_porosity_point": (C.c_double, [C.c_double, C.c_double, P]),
        "mlr_exchange": (None, [G, C.c_int, _dp]),
        "mlr_noflux_bc": (None, [G, C.c_int, _dp]),
        "mlr_scalar_bc": (None, [G, C.c_int, B, C.c_int, _dp]),
        "mlr_hc_diffused_var": (None, [G, P, B, _dp, _dp, C.c_int]),
        "mlr_lambda": (None, [G, C.c_int, C.c_double, _dp, C.c_double, pp, _dp, _dp]),
        "mlr_hc_gsrb": (None, [G, P, B, C.c_double, _dp, C.c_double, pp, _dp, _dp, _dp, _dp]),
        "mlr_hc_residual": (None, [G, P, B, C.c_double, _dp, C.c_double, pp, _dp, _dp, _dp, _dp]),
        "mlr_hc_apply_op": (None, [G, P, B, C.c_double, _dp, C.c_double, pp, _dp, _dp, _dp]),
        "mlr_hc_restrict_residual": (None, [G, P, B, C.c_double, _dp, C.c_double, pp, _dp, _dp, _dp, _dp]),
        "mlr_restrict_r": (None, [G, C.c_int, _dp, _dp]),
        "mlr_prolong_linear": (None, [G, C.c_int, _dp, _dp]),
        "mlr_prolong_constant": (None, [G, C.c_int, _dp, _dp]),
        "mlr_average_cells": (None, [G, C.c_int, _dp, C.c_int, _dp]),
        "mlr_average_faces": (None, [G, C.c_int, pp, C.c_int, pp]),
        "mlr_norm_max": (C.c_double, [G, C.c_int, _dp]),
        "mlr_porosity_field": (None, [G, P, _dp, _dp]),
        "mlr_geometric_faces": (None, [G, _dp, pp]),
        "mlr_hc_bcoef": (None, [G, P, _dp, pp]),
        "mlr_default_opts": (None, [C.POINTER(Opts)]),
        "mlr_set_threads": (C.c_int, [C.c_int]),
        "mlr_hc_fas_solve": (C.c_int, [G, P, B, C.c_double, _dp, C.c_double, pp, _dp, _dp,
                                       C.POINTER(Opts), C.POINTER(Stats)]),
        "mlr_hc_be_step": (C.c_int, [G, P, B, _dp, pp, C.c_double, _dp, _dp, _dp,
                                     C.POINTER(Opts), C.POINTER(Stats)]),
        "mlr_lin_gsrb": (None, [G, C.c_int, B, C.c_double, _dp, C.c_double, _dp, pp, _dp, _dp, _dp]),
        "mlr_lin_residual": (None, [G, C.c_int, B, C.c_double, _dp, C.c_double, _dp, pp, _dp, _dp, _dp,
                                    C.c_int]),
        "mlr_lin_apply_op": (None, [G, C.c_int, B, C.c_double, _dp, C.c_double, _dp, pp, _dp, _dp, C.c_int]),
        "mlr_lin_restrict_residual": (None, [G, C.c_int, B, C.c_double, _dp, C.c_double, _dp, pp, _dp, _dp,
                                             _dp]),
        "mlr_enthalpy_vars_point": (None, [C.c_double, C.c_double, P, _dp]),
        "mlr_enthalpy_vars": (None, [G, P, _dp, _dp]),
        "mlr_ustar_src": (None, [G, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, pp]),
        "mlr_cc_project_rhs": (None, [G, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, pp, pp, _dp]),
        "mlr_pressure_faces": (None, [G, _dp, _dp, pp]),
        "mlr_cc_correct": (None, [G, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp, pp]),
        "mlr_hc_jacobi": (None, [G, P, B, C.c_double, _dp, C.c_double, pp, _dp, _dp, _dp, _dp, _dp]),
        "mlr_hc_precond": (None, [G, P, B, C.c_double, _dp, C.c_double, pp, _dp, _dp, _dp, _dp]),
        "mlr_hc_get_flux": (None, [G, P, B, C.c_double, pp, _dp, pp, _dp]),
        "mlr_darcy_ccoef": (None, [G, _dp, C.c_double, C.c_int, C.c_double, _dp]),
        "mlr_edge_to_cell": (None, [G, pp, pp]),
        "mlr_face_velocity": (None, [G, pp, pp]),
        "mlr_advect_src": (C.c_int, [G, C.POINTER(AdvectBC), C.c_double, C.c_double, _dp, pp, _dp, _dp, pp, pp]),
        "mlr_lin_mg_solve": (C.c_int, [G, C.c_int, B, C.c_double, _dp, C.c_double, _dp, pp, _dp, _dp, C.c_int,
                                       C.POINTER(Opts), C.POINTER(Stats)]),
        "mlr_lin_dot": (C.c_double, [G, C.c_int, _dp, _dp]),
        "mlr_lin_precond": (None, [G, C.c_int, B, C.c_double, _dp, C.c_double, _dp, pp, _dp, _dp, _dp]),
        "mlr_lin_bicgstab": (C.c_int, [G, C.c_int, B, C.c_double, _dp, C.c_double, _dp, pp, _dp, _dp, _dp,
                                       C.c_int, C.POINTER(C.c_int), _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


# ---------------------------------------------------------------- helpers

def make_grid(n, dx=None, periodic=(0, 0, 0)):
    n = list(n)
    dim = len(n)
    g = Grid()
    g.dim = dim
    for d in range(3):
        g.n[d] = n[d] if d < dim else 1
        g.periodic[d] = int(periodic[d]) if d < len(periodic) else 0
    g.dx = float(dx if dx is not None else 1.0 / n[0])
    return g


def cell_shape(g, ncomp):
    if g.dim == 3:
        return (ncomp, g.n[2] + 2, g.n[1] + 2, g.n[0] + 2)
    return (ncomp, 1, g.n[1] + 2, g.n[0] + 2)


def face_shape(g, d, ncomp):
    n = [g.n[0], g.n[1], g.n[2] if g.dim == 3 else 1]
    n[d] += 1
    return (ncomp, n[2], n[1], n[0])


def valid(g, x):
    """View of the valid cells of a cell array: (ncomp, nz, ny, nx)."""
    if g.dim == 3:
        return x[:, 1:-1, 1:-1, 1:-1]
    return x[:, :, 1:-1, 1:-1]


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def ptrs(arrs):
    arr = (_dp * 3)()
    for d, a in enumerate(arrs):
        arr[d] = ptr(a)
    return arr


def default_params(**kw):
    """Reference defaults (src/MushyLayerParams.cpp): pc = 1e-5, c = k = 1,
    theta_e = Theta_e = 0, D_H = 1, D_Theta = 1/Le (Le given as lewis)."""
    p = Params()
    p.stefan = kw.get("stefan", 5.0)
    p.compositionRatio = kw.get("compositionRatio", 2.0)
    p.waterDistributionCoeff = kw.get("waterDistributionCoeff", 1e-5)
    p.specificHeatRatio = kw.get("specificHeatRatio", 1.0)
    p.heatConductivityRatio = kw.get("heatConductivityRatio", 1.0)
    p.thetaEutectic = kw.get("thetaEutectic", 0.0)
    p.ThetaEutectic = kw.get("ThetaEutectic", 0.0)
    p.heatDiffusionCoeff = kw.get("heatDiffusionCoeff", 1.0)
    p.saltDiffusionCoeff = kw.get("saltDiffusionCoeff", 1.0 / kw.get("lewis", 200.0))
    return p


def make_bc(spec=None, apply=1):
    """spec: dict {(dir, side, comp): (type, value)}; default Neumann 0."""
    b = BC()
    for d in range(3):
        for s in range(2):
            for c in range(2):
                b.type[d][s][c] = 1
                b.val[d][s][c] = 0.0
    for (d, s, c), (t, v) in (spec or {}).items():
        b.type[d][s][c] = t
        b.val[d][s][c] = v
    b.apply = apply
    return b


def default_opts(**kw):
    o = Opts()
    lib().mlr_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


# ------------------------------------------------------------ operations

def closure_point(H, C_, p):
    t = C.c_double()
    cl = C.c_double()
    lib().mlr_closure_point(H, C_, C.byref(p), C.byref(t), C.byref(cl))
    return t.value, cl.value


def porosity_point(H, C_, p):
    return lib().mlr_porosity_point(H, C_, C.byref(p))


def bounding_energy(H, C_, p):
    hs, he, hl = C.c_double(), C.c_double(), C.c_double()
    lib().mlr_bounding_energy(H, C_, C.byref(p), C.byref(hs), C.byref(he), C.byref(hl))
    return hs.value, he.value, hl.value


def hc_bcoef(g, p, chi):
    chi = np.ascontiguousarray(chi, dtype=np.float64).copy()
    b = [np.zeros(face_shape(g, d, 2)) for d in range(g.dim)]
    lib().mlr_hc_bcoef(C.byref(g), C.byref(p), ptr(chi), ptrs(b))
    return b


def porosity_field(g, p, phi):
    chi = np.zeros(cell_shape(g, 1))
    lib().mlr_porosity_field(C.byref(g), C.byref(p), ptr(phi), ptr(chi))
    return chi


def lam(g, ncomp, alpha, a, beta, b, c=None):
    out = np.zeros(cell_shape(g, ncomp))
    lib().mlr_lambda(C.byref(g), ncomp, alpha, ptr(a), beta, ptrs(b), ptr(c), ptr(out))
    return out


def hc_gsrb(g, p, bc, alpha, a, beta, b, phi, rhs, iters=1):
    lm = lam(g, 2, alpha, a, beta, b)
    scratch = np.zeros(cell_shape(g, 2))
    for _ in range(iters):
        lib().mlr_hc_gsrb(C.byref(g), C.byref(p), C.byref(bc), alpha, ptr(a), beta, ptrs(b),
                          ptr(lm), ptr(phi), ptr(rhs), ptr(scratch))
    return phi


def hc_residual(g, p, bc, alpha, a, beta, b, phi, rhs):
    res = np.zeros(cell_shape(g, 2))
    scratch = np.zeros(cell_shape(g, 2))
    lib().mlr_hc_residual(C.byref(g), C.byref(p), C.byref(bc), alpha, ptr(a), beta, ptrs(b),
                          ptr(phi), ptr(rhs), ptr(res), ptr(scratch))
    return res


def hc_apply_op(g, p, bc, alpha, a, beta, b, phi):
    lhs = np.zeros(cell_shape(g, 2))
    scratch = np.zeros(cell_shape(g, 2))
    lib().mlr_hc_apply_op(C.byref(g), C.byref(p), C.byref(bc), alpha, ptr(a), beta, ptrs(b),
                          ptr(phi), ptr(lhs), ptr(scratch))
    return lhs


def hc_jacobi(g, p, bc, alpha, a, beta, b, phi, rhs, iters=1):
    """levelJacobi x iters (phi in place)."""
    lm = lam(g, 2, alpha, a, beta, b)
    res = np.zeros(cell_shape(g, 2))
    scratch = np.zeros(cell_shape(g, 2))
    for _ in range(iters):
        lib().mlr_hc_jacobi(C.byref(g), C.byref(p), C.byref(bc), alpha, ptr(a), beta, ptrs(b), ptr(lm), ptr(phi),
                            ptr(rhs), ptr(res), ptr(scratch))
    return phi


def hc_precond(g, p, bc, alpha, a, beta, b, rhs):
    """preCond: phi = lambda*rhs then two GSRB iterations."""
    lm = lam(g, 2, alpha, a, beta, b)
    phi = np.zeros(cell_shape(g, 2))
    scratch = np.zeros(cell_shape(g, 2))
    lib().mlr_hc_precond(C.byref(g), C.byref(p), C.byref(bc), alpha, ptr(a), beta, ptrs(b), ptr(lm), ptr(phi),
                         ptr(rhs), ptr(scratch))
    return phi


def hc_get_flux(g, p, bc, beta, b, phi):
    """getFlux on every face: list of face arrays (2 components)."""
    flux = [np.zeros(face_shape(g, d, 2)) for d in range(g.dim)]
    scratch = np.zeros(cell_shape(g, 2))
    lib().mlr_hc_get_flux(C.byref(g), C.byref(p), C.byref(bc), beta, ptrs(b), ptr(phi), ptrs(flux), ptr(scratch))
    return flux


def hc_diffused_var(g, p, bc, phi, homogeneous=0):
    out = np.zeros(cell_shape(g, 2))
    lib().mlr_hc_diffused_var(C.byref(g), C.byref(p), C.byref(bc), ptr(phi), ptr(out), homogeneous)
    return out


def hc_restrict_residual(g, p, bc, alpha, a, beta, b, phi, rhs):
    gc = coarsen(g, 2)
    resc = np.zeros(cell_shape(gc, 2))
    scratch = np.zeros(cell_shape(g, 2))
    lib().mlr_hc_restrict_residual(C.byref(g), C.byref(p), C.byref(bc), alpha, ptr(a), beta,
                                   ptrs(b), ptr(phi), ptr(rhs), ptr(resc), ptr(scratch))
    return resc


def coarsen(g, r):
    n = [g.n[d] // r for d in range(g.dim)]
    gc = make_grid(n, g.dx * r, [g.periodic[d] for d in range(3)])
    return gc


def restrict_r(g, ncomp, phif):
    gc = coarsen(g, 2)
    out = np.zeros(cell_shape(gc, ncomp))
    lib().mlr_restrict_r(C.byref(g), ncomp, ptr(phif), ptr(out))
    return out


def prolong_linear(g, ncomp, phif, corrc):
    corrc = corrc.copy()
    lib().mlr_prolong_linear(C.byref(g), ncomp, ptr(phif), ptr(corrc))
    return phif


def prolong_constant(g, ncomp, phif, corrc):
    lib().mlr_prolong_constant(C.byref(g), ncomp, ptr(phif), ptr(corrc))
    return phif


def average_cells(g, ncomp, af, ratio):
    gc = coarsen(g, ratio)
    out = np.zeros(cell_shape(gc, ncomp))
    lib().mlr_average_cells(C.byref(g), ncomp, ptr(np.ascontiguousarray(af)), ratio, ptr(out))
    return out


def average_faces(g, ncomp, bf, ratio):
    gc = coarsen(g, ratio)
    out = [np.zeros(face_shape(gc, d, ncomp)) for d in range(g.dim)]
    lib().mlr_average_faces(C.byref(g), ncomp, ptrs(bf), ratio, ptrs(out))
    return out


def norm_max(g, ncomp, x):
    return lib().mlr_norm_max(C.byref(g), ncomp, ptr(x))


def hc_fas_solve(g, p, bc, alpha, a, beta, b, phi, rhs, opts=None):
    st = Stats()
    o = opts or default_opts()
    lib().mlr_hc_fas_solve(C.byref(g), C.byref(p), C.byref(bc), alpha, ptr(a), beta, ptrs(b),
                           ptr(phi), ptr(rhs), C.byref(o), C.byref(st))
    return st


def hc_be_step(g, p, bc, a, b, dt, phi_new, phi_old, src=None, opts=None):
    st = Stats()
    o = opts or default_opts()
    lib().mlr_hc_be_step(C.byref(g), C.byref(p), C.byref(bc), ptr(a), ptrs(b), dt, ptr(phi_new),
                         ptr(phi_old), ptr(src), C.byref(o), C.byref(st))
    return st


# ------------------------------------------------- linear VC Helmholtz / Poisson

def lin_gsrb(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs, iters=1):
    lm = lam(g, ncomp, alpha, a, beta, b, c)
    for _ in range(iters):
        lib().mlr_lin_gsrb(C.byref(g), ncomp, C.byref(bc), alpha, ptr(a), beta, ptr(c), ptrs(b), ptr(lm),
                           ptr(phi), ptr(rhs))
    return phi


def lin_residual(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs, homogeneous=0):
    res = np.zeros(cell_shape(g, ncomp))
    lib().mlr_lin_residual(C.byref(g), ncomp, C.byref(bc), alpha, ptr(a), beta, ptr(c), ptrs(b), ptr(phi),
                           ptr(rhs), ptr(res), int(homogeneous))
    return res


def lin_apply_op(g, ncomp, bc, alpha, a, beta, c, b, phi, homogeneous=0):
    lhs = np.zeros(cell_shape(g, ncomp))
    lib().mlr_lin_apply_op(C.byref(g), ncomp, C.byref(bc), alpha, ptr(a), beta, ptr(c), ptrs(b), ptr(phi),
                           ptr(lhs), int(homogeneous))
    return lhs


def lin_restrict_residual(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs):
    resc = np.zeros(cell_shape(coarsen(g, 2), ncomp))
    lib().mlr_lin_restrict_residual(C.byref(g), ncomp, C.byref(bc), alpha, ptr(a), beta, ptr(c), ptrs(b),
                                    ptr(phi), ptr(rhs), ptr(resc))
    return resc


def lin_dot(g, ncomp, x, y):
    return lib().mlr_lin_dot(C.byref(g), ncomp, ptr(x), ptr(y))


def lin_precond(g, ncomp, bc, alpha, a, beta, c, b, rhs):
    """preCond: lambda*rhs, then two GSRB iterations (homogeneous BCs)."""
    lm = lam(g, ncomp, alpha, a, beta, b, c)
    phi = np.zeros(cell_shape(g, ncomp))
    lib().mlr_lin_precond(C.byref(g), ncomp, C.byref(bc), alpha, ptr(a), beta, ptr(c), ptrs(b), ptr(lm), ptr(phi),
                          ptr(rhs))
    return phi


def lin_bicgstab(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs, imax=80):
    """BiCGStab on L phi = rhs (homogeneous BCs), phi updated in place:
    (exit status, iterations, final L2 residual norm)."""
    lm = lam(g, ncomp, alpha, a, beta, b, c)
    it = C.c_int(0)
    fn = C.c_double(0)
    status = lib().mlr_lin_bicgstab(C.byref(g), ncomp, C.byref(bc), alpha, ptr(a), beta, ptr(c), ptrs(b), ptr(lm),
                                    ptr(phi), ptr(rhs), int(imax), C.byref(it), C.byref(fn))
    return status, it.value, fn.value


def lin_mg_solve(g, ncomp, bc, alpha, a, beta, c, b, phi, rhs, prolong_linear=0, opts=None):
    st = Stats()
    o = opts or default_opts()
    lib().mlr_lin_mg_solve(C.byref(g), ncomp, C.byref(bc), alpha, ptr(a), beta, ptr(c), ptrs(b), ptr(phi),
                           ptr(rhs), int(prolong_linear), C.byref(o), C.byref(st))
    return st


def enthalpy_vars_point(H, C_, p):
    out = (C.c_double * 4)()
    lib().mlr_enthalpy_vars_point(H, C_, C.byref(p), out)
    return tuple(out)


def enthalpy_vars(g, p, phi):
    """(theta, Theta_l, Theta_s, porosity) of every valid cell (4 components)."""
    out = np.zeros(cell_shape(g, 4))
    lib().mlr_enthalpy_vars(C.byref(g), C.byref(p), ptr(phi), ptr(out))
    return out


def ustar_src(g, buoyT, buoyS, scaleP, vold, vnew, pi):
    """U* source per velocity component (list of 1-component arrays); fills
    the ghosts of pi."""
    src = [np.zeros(cell_shape(g, 1)) for _ in range(g.dim)]
    lib().mlr_ustar_src(C.byref(g), buoyT, buoyS, int(scaleP), ptr(vold), ptr(vnew), ptr(pi), ptrs(src))
    return src


def cc_project_rhs(g, dt, limit, scaleP, vold, vnew, pi, ustar):
    """(velocity before the projection, projection right-hand side)."""
    u = [np.zeros(cell_shape(g, 1)) for _ in range(g.dim)]
    rhs = np.zeros(cell_shape(g, 1))
    lib().mlr_cc_project_rhs(C.byref(g), dt, limit, int(scaleP), ptr(vold), ptr(vnew), ptr(pi), ptrs(ustar),
                             ptrs(u), ptr(rhs))
    return u, rhs


def pressure_faces(g, vold, vnew):
    b = [np.zeros(face_shape(g, d, 1)) for d in range(g.dim)]
    lib().mlr_pressure_faces(C.byref(g), ptr(vold), ptr(vnew), ptrs(b))
    return b


def cc_correct(g, dt, limit, scaleP, vold, vnew, pi, u):
    """Projection correction of the velocity components u (in place)."""
    lib().mlr_cc_correct(C.byref(g), dt, limit, int(scaleP), ptr(vold), ptr(vnew), ptr(pi), ptrs(u))
    return u


PERM_LAWS = {"pure_fluid": 0, "cubic": 1, "kozeny_carman": 2, "log": 3, "porosity": 4}


def darcy_ccoef(g, chi, coeff, law="kozeny_carman", hele_shaw=0.0):
    """U* Darcy coefficient (chi/Pi(chi))*coeff of a 1-component cell field."""
    c = np.zeros(cell_shape(g, 1))
    lib().mlr_darcy_ccoef(C.byref(g), ptr(np.ascontiguousarray(chi)), coeff, PERM_LAWS.get(law, law), hele_shaw,
                          ptr(c))
    return c


# advection BC types (AdvectIBC numbering, BCutil/AdvectIBC.H:137-144)
ADV_DIRICHLET, ADV_EXTRAP, ADV_INFLOW_OUTFLOW, ADV_NEUMANN = 0, 1, 2, 4


def make_advect_bc(spec=None, hc_bc=None):
    """Advection BC of (theta, Theta_l).  hc_bc: the HC operator's BC, converted
    as PhysBCUtil::convertBCType does (Dirichlet -> Dirichlet, Neumann ->
    Neumann, same values; BCutil/PhysBCUtil.cpp:3380-3434, 4059-4087);
    spec: {(dir, side, comp): (type, value)} overrides.  Default extrapolation."""
    b = AdvectBC()
    for d in range(3):
        for s in range(2):
            for c in range(2):
                t, v = ADV_EXTRAP, 0.0
                if hc_bc is not None:
                    t = ADV_DIRICHLET if hc_bc.type[d][s][c] == 0 else ADV_NEUMANN
                    v = hc_bc.val[d][s][c]
                b.type[d][s][c] = t
                b.val[d][s][c] = v
    for (d, s, c), (t, v) in (spec or {}).items():
        b.type[d][s][c] = t
        b.val[d][s][c] = v
    return b


def edge_to_cell(g, uf):
    vc = [np.zeros(cell_shape(g, 1)) for _ in range(g.dim)]
    lib().mlr_edge_to_cell(C.byref(g), ptrs([np.ascontiguousarray(u) for u in uf]), ptrs(vc))
    return vc


def face_velocity(g, u):
    """Face velocities (CellToEdge, zero on walls) of the velocity components u."""
    uf = [np.zeros(face_shape(g, d, 1)) for d in range(g.dim)]
    lib().mlr_face_velocity(C.byref(g), ptrs([np.ascontiguousarray(x) for x in u]), ptrs(uf))
    return uf


def advect_src(g, abc, dt, coeff, W, uf, S=None, src=None, whalf=False):
    """Advective source of (H, Theta): returns (src - div(flux), flux[d],
    whalf[d] or None).  W: 2-component (theta, Theta_l) cell array; uf: face
    velocities; S: 2-component diffusive source or None."""
    out = np.zeros(cell_shape(g, 2)) if src is None else np.array(src, dtype=np.float64, copy=True)
    flux = [np.zeros(face_shape(g, d, 2)) for d in range(g.dim)]
    wh = [np.zeros(face_shape(g, d, 2)) for d in range(g.dim)] if whalf else None
    W = np.ascontiguousarray(W[:2])
    rc = lib().mlr_advect_src(C.byref(g), C.byref(abc), dt, coeff, ptr(W),
                              ptrs([np.ascontiguousarray(u) for u in uf]),
                              ptr(np.ascontiguousarray(S)) if S is not None else None, ptr(out), ptrs(flux),
                              ptrs(wh) if wh is not None else None)
    if rc != 0:
        raise ValueError("unsupported advection boundary condition")
    return out, flux, wh


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops; returns the count in use."""
    return lib().mlr_set_threads(int(n))
