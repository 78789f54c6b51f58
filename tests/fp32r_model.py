"""Numpy model of the FP32-refined (H, Theta) solve, for tests: the same
outer defect correction (FP64 nonlinear residual from the oracle, closure
Jacobian M) and the same inner linear V-cycle on
    J e = alpha e - beta div(b grad(M e)) = r
(2x2 point-block red/black Gauss-Seidel, restriction of r - J e and of M,
PROLONG_JP linear prolongation, an exact solve on a coarsest level of at
most 16 unknowns and `bottom` block GS iterations on a larger one), evaluated here in FP64 (or FP32 with dtype=np.float32), with the
same step-length backtracking on the residual max-norm.

It documents the algorithm the GPU form implements and lets tests separate
"the algorithm does this" from "the kernels do this".  Valid-cell arrays
have shape (2, nz, ny, nx) (nz = 1 in 2D).
"""
import numpy as np

from oracle import mlref as O


def closure_jacobian(H, C, p):
    """Per-cell dtheta/dH, dtheta/dTheta, dTheta_l/dH, dTheta_l/dTheta of the
    closure (c = 1, theta_e = Theta_e = 0), regime decided as the closure
    decides it; returns M of shape (4,) + H.shape."""
    St, CR, pc = p.stefan, p.compositionRatio, p.waterDistributionCoeff
    assert p.specificHeatRatio == 1.0 and p.thetaEutectic == 0.0 and p.ThetaEutectic == 0.0
    hs = np.where(C < -CR, 0.0, np.maximum((-C - CR) / pc, 0.0))
    he = (CR + C) / CR * St
    hl = St - C
    M = np.zeros((4,) + H.shape)
    solid = H <= hs
    mush = (H > he) & (H < hl)
    liq = H >= hl
    M[0][solid] = 1.0
    M[0][liq] = 1.0
    M[3][liq] = 1.0
    Cv = np.maximum(C, -CR)
    a = St * (pc - 1.0)
    b = -CR + H * (1 - pc) - pc * St
    q = CR + Cv + pc * H
    with np.errstate(invalid="ignore", divide="ignore"):
        chi = (-b - np.sqrt(np.maximum(b * b - 4 * a * q, 0))) / (2 * a)
        D = chi + pc * (1 - chi)
        Tl = (C + CR * (1 - chi)) / D
        dTl = -(CR + Tl * (1 - pc)) / D
        F = 2 * a * chi + b
        chiH = -(chi * (1 - pc) + pc) / F
        chiC = np.where(C >= -CR, -1.0 / F, 0.0)
        M2 = dTl * chiH
        M3 = 1.0 / D + dTl * chiC
    M[2][mush] = M2[mush]
    M[3][mush] = M3[mush]
    M[0][mush] = -M2[mush]
    M[1][mush] = -M3[mush]
    return M


class Level:
    def __init__(self, g, b, bc, periodic, dtype):
        self.g = g
        self.dim = g.dim
        self.dx = g.dx
        # faces as (2, fz, fy, fx) valid arrays
        self.b = [np.asarray(x, dtype=dtype) for x in b]
        self.bc = bc
        self.periodic = periodic
        self.dtype = dtype


def _pad(u, lev, n):
    """One ghost layer around u (nz, ny, nx): periodic wrap or the homogeneous
    derived-variable wall rule (Dirichlet: -u, Neumann: +u)."""
    dim = lev.dim
    out = u
    for d in range(dim):
        ax = 2 - d
        if lev.periodic[d]:
            lo = np.take(out, [-1], axis=ax)
            hi = np.take(out, [0], axis=ax)
        else:
            slo = -1.0 if (lev.bc.apply and lev.bc.type[d][0][n] == 0) else 1.0
            shi = -1.0 if (lev.bc.apply and lev.bc.type[d][1][n] == 0) else 1.0
            lo = slo * np.take(out, [0], axis=ax)
            hi = shi * np.take(out, [-1], axis=ax)
        out = np.concatenate([lo, out, hi], axis=ax)
    if dim == 2:
        out = out[:, :, :]
    return out


def _div_b_grad(lev, u, n):
    """sum_d b_hi (u_hi - u) - b_lo (u - u_lo) (no 1/dx^2) of component n."""
    up = _pad(u, lev, n)
    dim = lev.dim
    core = (slice(1, -1) if dim == 3 else slice(None), slice(1, -1), slice(1, -1))
    s = np.zeros_like(u)
    for d in range(dim):
        ax = 2 - d
        bd = lev.b[d][n]
        hi_sl = [slice(None)] * 3
        lo_sl = [slice(None)] * 3
        hi_sl[ax] = slice(1, None)
        lo_sl[ax] = slice(0, -1)
        shift_hi = list(core)
        shift_lo = list(core)
        shift_hi[ax] = slice(2, None)
        shift_lo[ax] = slice(0, -2)
        s = s + bd[tuple(hi_sl)] * (up[tuple(shift_hi)] - u) - bd[tuple(lo_sl)] * (u - up[tuple(shift_lo)])
    return s


def apply_J(lev, M, e, alpha, beta):
    u = np.stack([M[0] * e[0] + M[1] * e[1], M[2] * e[0] + M[3] * e[1]])
    dxinv = 1.0 / (lev.dx * lev.dx)
    return np.stack([alpha * e[n] - beta * _div_b_grad(lev, u[n], n) * dxinv for n in range(2)])


def smooth(lev, M, e, r, alpha, beta, passes):
    nz, ny, nx = e.shape[1:]
    kk, jj, ii = np.mgrid[0:nz, 0:ny, 0:nx]
    dxinv = 1.0 / (lev.dx * lev.dx)
    sig = []
    for n in range(2):
        s = 0.0
        for d in range(lev.dim):
            ax = 2 - d
            bd = lev.b[d][n]
            hi_sl = [slice(None)] * 3
            lo_sl = [slice(None)] * 3
            hi_sl[ax] = slice(1, None)
            lo_sl[ax] = slice(0, -1)
            s = s + (bd[tuple(hi_sl)] + bd[tuple(lo_sl)])
        sig.append(beta * s * dxinv)
    D00 = alpha + sig[0] * M[0]
    D01 = sig[0] * M[1]
    D10 = sig[1] * M[2]
    D11 = alpha + sig[1] * M[3]
    det = D00 * D11 - D01 * D10
    for p in range(passes):
        mask = ((ii + jj + kk + p) % 2) == 0
        res = r - apply_J(lev, M, e, alpha, beta)
        d0 = (D11 * res[0] - D01 * res[1]) / det
        d1 = (D00 * res[1] - D10 * res[0]) / det
        e[0][mask] += d0[mask]
        e[1][mask] += d1[mask]
    return e


def _restrict(x, dim):
    if dim == 2:
        return 0.25 * (x[..., 0::2, 0::2] + x[..., 1::2, 0::2] + x[..., 0::2, 1::2] + x[..., 1::2, 1::2])
    s = 0.0
    for kk in (0, 1):
        for jj in (0, 1):
            for ii in (0, 1):
                s = s + x[..., kk::2, jj::2, ii::2]
    return 0.125 * s


def _prolong(lev_f, ec):
    gf = lev_f.g
    gc = O.coarsen(gf, 2)
    f = np.zeros(O.cell_shape(gf, 2))
    c = np.zeros(O.cell_shape(gc, 2))
    O.valid(gc, c)[...] = ec
    O.prolong_linear(gf, 2, f, c)
    return O.valid(gf, f).copy()


def direct(L, M, r, alpha, beta):
    """Exact solve of J e = r on a tiny level (dense assembly from unit
    vectors)."""
    shape = r.shape
    U = r.size
    A = np.zeros((U, U))
    for t in range(U):
        unit = np.zeros(U)
        unit[t] = 1.0
        A[:, t] = apply_J(L, M, unit.reshape(shape), alpha, beta).ravel()
    return np.linalg.solve(A, r.ravel()).reshape(shape)


def vcycle(levs, Ms, l, e, r, alpha, beta, nu1, nu2, bottom):
    L = levs[l]
    if l == len(levs) - 1:
        if r.size <= 16:
            return direct(L, Ms[l], r, alpha, beta)
        return smooth(L, Ms[l], e, r, alpha, beta, 2 * bottom)
    e = smooth(L, Ms[l], e, r, alpha, beta, 2 * nu1)
    rc = _restrict(r - apply_J(L, Ms[l], e, alpha, beta), L.dim)
    ec = vcycle(levs, Ms, l + 1, np.zeros_like(rc), rc, alpha, beta, nu1, nu2, bottom)
    e = e + _prolong(L, ec)
    return smooth(L, Ms[l], e, r, alpha, beta, 2 * nu2)


def solve(g, p, bc, alpha, beta, b, phi, rhs, nu1=1, nu2=4, bottom=40, max_iter=100, eps=1e-10,
          norm_thresh=1e-10, hang=1e-10, dtype=np.float64, max_halvings=6):
    """The refined solve on host arrays in FAB layout (phi updated in place).
    Returns (exit_status, iterations, residual history)."""
    periodic = [g.periodic[d] for d in range(g.dim)]
    nlev = O.lib().mlr_mg_depth(g)
    levs = []
    for l in range(nlev):
        gl = g if l == 0 else O.coarsen(g, 1 << l)
        bl = b if l == 0 else O.average_faces(g, 2, b, 1 << l)
        levs.append(Level(gl, [x.reshape(x.shape) for x in bl], bc, periodic, dtype))
    res = O.hc_residual(g, p, bc, alpha, None, beta, b, phi, rhs)
    r0 = O.norm_max(g, 2, res)
    norm, prev = r0, r0
    hist = [r0]
    it, status = 0, 0
    if r0 <= norm_thresh:
        status = 8
    while status == 0 and it < max_iter:
        v = O.valid(g, phi)
        Ms = [closure_jacobian(v[0], v[1], p)]
        for l in range(1, nlev):
            Ms.append(_restrict(Ms[-1], g.dim))
        r = O.valid(g, res).astype(dtype)
        e = vcycle(levs, Ms, 0, np.zeros_like(r), r, alpha, beta, nu1, nu2, bottom).astype(np.float64)
        it += 1
        # step-length backtracking: accept phi + omega*e once the residual
        # max-norm decreases, or after max_halvings halvings
        omega = 1.0
        for k in range(max_halvings + 1):
            trial = phi.copy()
            O.valid(g, trial)[...] += omega * e
            res = O.hc_residual(g, p, bc, alpha, None, beta, b, trial, rhs)
            norm = O.norm_max(g, 2, res)
            if norm < prev or k == max_halvings:
                break
            omega *= 0.5
        phi[...] = trial
        hist.append(norm)
        if norm <= eps * r0:
            status |= 1
        if norm <= norm_thresh:
            status |= 8
        if status == 0 and norm > (1.0 - hang) * prev:
            status |= 4
        prev = norm
    if status == 0:
        status = 2
    return status, it, hist
