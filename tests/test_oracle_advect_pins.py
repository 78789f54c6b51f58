"""Pins of the oracle's advective source (mlr_advect_src): the unsplit PLM
Godunov predictor (Chombo PatchGodunov, restated as build spec), the
boundary faces of AdvectIBC::primBC (BCutil/AdvectIBC.cpp:130-260,
BCutil/SolidF.ChF:14-109), the flux of computeScalarAdvectiveFlux
(srcSubcycle/AMRLevelMushyLayer.cpp:1165-1198) and the MAC divergence
(util/DivergenceF.ChF:22-47).

Each check holds by the mathematics, not by the oracle's own formula:
  * linear fields, constant velocity: every interior face value is the exact
    upwind value W(x_f - u dt/2) + dt/2 S (slopes and transverse corrections
    are exact for linear data);
  * bilinear fields: the closed forms of the corner-transport-upwind
    predictor -- in 2D the transverse term is taken at the upwind cell
    centre, in 3D the y*z cross term of an x face carries (1/3) nu_y nu_z dx^2
    (the 1/3 of the intermediate corner coupling: (1/2)*(1/3) from each of
    the two transverse directions);
  * 1D transport: the face values equal the flux-limiter (Sweby) form of
    MUSCL-Hancock with the monotonised-central limiter
    phi(r) = max(0, min(2r, (1+r)/2, 2)), and one step W - dt div(F) is TVD;
  * conservation: with periodic sides and zero normal velocity on the walls
    the divergence sums to zero;
  * the wall faces follow Dirichlet / Neumann / extrapolation /
    inflow-outflow as primBC states them, and the flux carries the
    advection coefficient once per component, as the reference's loop does.
"""
import numpy as np
import pytest

from oracle import mlref as O


def _coords(g):
    """Cell-centre coordinates on the full (ghosted) cell array."""
    dx = g.dx
    n = [g.n[d] for d in range(g.dim)]
    ax = [(np.arange(-1, n[d] + 1) + 0.5) * dx for d in range(g.dim)]
    if g.dim == 2:
        Y, X = np.meshgrid(ax[1], ax[0], indexing="ij")
        return X[None], Y[None], None
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X, Y, Z


def _faces_const(g, vel):
    return [np.full(O.face_shape(g, d, 1), float(vel[d])) for d in range(g.dim)]


def _face_coords(g, d):
    """Coordinates of the face centres of direction d, face-array shape."""
    dx = g.dx
    n = [g.n[e] for e in range(g.dim)]
    ax = []
    for e in range(g.dim):
        if e == d:
            ax.append(np.arange(n[e] + 1) * dx)
        else:
            ax.append((np.arange(n[e]) + 0.5) * dx)
    if g.dim == 2:
        Y, X = np.meshgrid(ax[1], ax[0], indexing="ij")
        return X[None], Y[None], None
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X, Y, Z


def _inner(g, d, margin=2):
    """Mask of faces of direction d whose stencil stays off every wall."""
    shp = O.face_shape(g, d, 1)[1:]
    m = np.ones(shp, dtype=bool)
    n = [g.n[e] for e in range(g.dim)]
    for e in range(g.dim):
        ax = (len(shp) - 1) - e
        idx = np.arange(shp[ax])
        if e == d:
            keep = (idx >= margin) & (idx <= n[e] - margin)
        else:
            keep = (idx >= margin) & (idx <= n[e] - 1 - margin)
        sh = [1] * len(shp)
        sh[ax] = -1
        m &= keep.reshape(sh)
    return m


def _W2(f0, f1):
    return np.ascontiguousarray(np.stack([f0, f1]))


@pytest.mark.parametrize("vel", [(0.7, -0.4), (-0.3, 0.9), (0.5, 0.0), (0.0, -0.6)])
def test_linear_exact_2d(vel):
    n = 24
    g = O.make_grid([n, n], periodic=(0, 0, 0))
    dt = 0.4 * g.dx
    X, Y, _ = _coords(g)
    coef = [(0.3, 1.7, -2.1), (-0.8, -0.6, 1.3)]
    W = _W2(coef[0][0] + coef[0][1] * X + coef[0][2] * Y, coef[1][0] + coef[1][1] * X + coef[1][2] * Y)
    S = np.zeros_like(W)
    S[0] = 0.37
    S[1] = -1.1
    uf = _faces_const(g, vel)
    _, _, wh = O.advect_src(g, O.make_advect_bc(), dt, 1.0, W, uf, S=S, whalf=True)
    for d in range(2):
        FX, FY, _ = _face_coords(g, d)
        m = _inner(g, d)
        for c in range(2):
            a0, ax_, ay = coef[c]
            exact = a0 + ax_ * (FX - vel[0] * dt / 2) + ay * (FY - vel[1] * dt / 2) + 0.5 * dt * S[c].flat[0]
            np.testing.assert_allclose(wh[d][c][m], exact[m], rtol=0, atol=1e-13)


@pytest.mark.parametrize("vel", [(0.3, -0.5, 0.2), (-0.6, 0.4, -0.7)])
def test_linear_exact_3d(vel):
    n = 12
    g = O.make_grid([n, n, n], periodic=(0, 0, 0))
    dt = 0.3 * g.dx
    X, Y, Z = _coords(g)
    coef = [(0.2, 1.1, -0.7, 0.9), (1.0, -0.4, 0.6, -1.3)]
    W = np.stack([cf[0] + cf[1] * X + cf[2] * Y + cf[3] * Z for cf in coef])
    uf = _faces_const(g, vel)
    _, _, wh = O.advect_src(g, O.make_advect_bc(), dt, 1.0, W, uf, whalf=True)
    for d in range(3):
        FX, FY, FZ = _face_coords(g, d)
        m = _inner(g, d)
        for c in range(2):
            cf = coef[c]
            exact = (cf[0] + cf[1] * (FX - vel[0] * dt / 2) + cf[2] * (FY - vel[1] * dt / 2) +
                     cf[3] * (FZ - vel[2] * dt / 2))
            np.testing.assert_allclose(wh[d][c][m], exact[m], rtol=0, atol=1e-13)


@pytest.mark.parametrize("vel", [(0.6, 0.5), (-0.45, 0.7), (0.0, -0.3)])
def test_bilinear_ctu_2d(vel):
    """W = x*y: the x-face value is x_f y - (nu_x/2) dx y - (nu_y/2) dx x_up,
    x_up the centre of the upwind cell (the face centre when u == 0)."""
    n = 20
    g = O.make_grid([n, n], periodic=(0, 0, 0))
    dt = 0.5 * g.dx
    dx = g.dx
    X, Y, _ = _coords(g)
    W = _W2(X * Y, 2.0 * X * Y)
    uf = _faces_const(g, vel)
    _, _, wh = O.advect_src(g, O.make_advect_bc(), dt, 1.0, W, uf, whalf=True)
    nu = [v * dt / dx for v in vel]
    FX, FY, _ = _face_coords(g, 0)
    xup = FX - (0.5 * dx if vel[0] > 0 else -0.5 * dx if vel[0] < 0 else 0.0)
    exact = FX * FY - nu[0] / 2 * dx * FY - nu[1] / 2 * dx * xup
    m = _inner(g, 0)
    np.testing.assert_allclose(wh[0][0][m], exact[m], rtol=0, atol=1e-13)
    np.testing.assert_allclose(wh[0][1][m], 2 * exact[m], rtol=0, atol=2e-13)


@pytest.mark.parametrize("vel", [(0.4, 0.5, 0.3), (-0.2, -0.6, 0.45), (0.7, 0.35, -0.5)])
def test_corner_coupling_3d(vel):
    """W = y*z: the x-face value is y z - (nu_y/2) dx z - (nu_z/2) dx y
    + (1/3) nu_y nu_z dx^2 for every sign of the velocity."""
    n = 12
    g = O.make_grid([n, n, n], periodic=(0, 0, 0))
    dt = 0.4 * g.dx
    dx = g.dx
    X, Y, Z = _coords(g)
    W = np.stack([Y * Z, -3.0 * Y * Z])
    uf = _faces_const(g, vel)
    _, _, wh = O.advect_src(g, O.make_advect_bc(), dt, 1.0, W, uf, whalf=True)
    nu = [v * dt / dx for v in vel]
    FX, FY, FZ = _face_coords(g, 0)
    exact = FY * FZ - nu[1] / 2 * dx * FZ - nu[2] / 2 * dx * FY + nu[1] * nu[2] * dx * dx / 3.0
    m = _inner(g, 0)
    np.testing.assert_allclose(wh[0][0][m], exact[m], rtol=0, atol=1e-13)
    np.testing.assert_allclose(wh[0][1][m], -3.0 * exact[m], rtol=0, atol=3e-13)


def _mc(r):
    return np.maximum(0.0, np.minimum(np.minimum(2 * r, 0.5 * (1 + r)), 2.0))


@pytest.mark.parametrize("u", [0.8, -0.55])
def test_muscl_hancock_1d_and_tvd(u):
    n = 64
    g = O.make_grid([n, 8], periodic=(1, 1, 0))
    dt = 0.9 * g.dx / abs(u)   # CFL 0.9
    rng = np.random.default_rng(5)
    line = np.where(np.arange(n) < n // 3, 1.0, 0.0) + 0.3 * np.sin(np.arange(n) * 0.4) + 0.2 * rng.random(n)
    W = np.zeros(O.cell_shape(g, 2))
    W[0, 0, 1:-1, 1:-1] = line[None, :]
    W[1, 0, 1:-1, 1:-1] = -2.0 * line[None, :] ** 2
    uf = [np.full(O.face_shape(g, 0, 1), u), np.zeros(O.face_shape(g, 1, 1))]
    out, flux, wh = O.advect_src(g, O.make_advect_bc(), dt, 1.0, W, uf, whalf=True)
    nu = u * dt / g.dx
    for c in range(2):
        w = W[c, 0, 1, 1:-1]
        dwr = np.roll(w, -1) - w                  # W[i+1]-W[i]
        dwl = w - np.roll(w, 1)                   # W[i]-W[i-1]
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.where(dwr != 0, dwl / dwr, 0.0)
        s = _mc(r) * dwr                          # limited slope of cell i
        if u > 0:
            face = w + 0.5 * (1 - nu) * s         # face i+1/2 from cell i
            face = np.roll(face, 1)               # face index i (between i-1 and i)
        else:
            face = w - 0.5 * (1 + nu) * s         # face i-1/2 from cell i
        got = wh[0][c, 0, 0, :n]
        np.testing.assert_allclose(got, face, rtol=0, atol=1e-14)
        # one step of the update is TVD
        new = w + dt * out[c, 0, 1, 1:-1]
        tv0 = np.abs(np.diff(np.append(w, w[0]))).sum()
        tv1 = np.abs(np.diff(np.append(new, new[0]))).sum()
        assert tv1 <= tv0 + 1e-12, (tv0, tv1)


@pytest.mark.parametrize("dim", [2, 3])
def test_conservation(dim):
    n = [20, 16] if dim == 2 else [10, 8, 12]
    per = (1, 0, 0) if dim == 2 else (1, 0, 1)
    g = O.make_grid(n, periodic=per)
    rng = np.random.default_rng(11)
    W = rng.random(O.cell_shape(g, 2))
    S = rng.random(O.cell_shape(g, 2)) - 0.5
    uf = []
    for d in range(dim):
        f = rng.random(O.face_shape(g, d, 1)) - 0.5
        if per[d]:   # the two periodic end faces are one face
            sl0 = [slice(None)] * f.ndim
            sl1 = [slice(None)] * f.ndim
            ax = f.ndim - 1 - d
            sl0[ax], sl1[ax] = 0, -1
            f[tuple(sl1)] = f[tuple(sl0)]
        else:        # no normal velocity through a wall
            sl = [slice(None)] * f.ndim
            ax = f.ndim - 1 - d
            sl[ax] = 0
            f[tuple(sl)] = 0.0
            sl[ax] = -1
            f[tuple(sl)] = 0.0
        uf.append(f)
    abc = O.make_advect_bc({(d, s, c): (O.ADV_DIRICHLET, 1.3) for d in range(dim) for s in range(2) for c in range(2)})
    out, flux, _ = O.advect_src(g, abc, 0.3 * g.dx, 1.0, W, uf, S=S)
    tot = O.valid(g, out).reshape(2, -1).sum(axis=1)
    assert np.all(np.abs(tot) < 1e-10 * np.abs(O.valid(g, out)).sum()), tot


def test_wall_faces():
    n = 16
    g = O.make_grid([n, n], periodic=(1, 0, 0))
    dt = 0.3 * g.dx
    dx = g.dx
    X, Y, _ = _coords(g)
    W = _W2(0.5 + 2.0 * Y, -1.0 + 0.7 * Y)
    for v in (0.6, -0.6):
        uf = _faces_const(g, (0.0, v))
        abc = O.make_advect_bc({(1, 0, 0): (O.ADV_DIRICHLET, 3.7), (1, 1, 0): (O.ADV_NEUMANN, 0.25),
                                (1, 0, 1): (O.ADV_EXTRAP, 0.0), (1, 1, 1): (O.ADV_INFLOW_OUTFLOW, -4.0)})
        _, _, wh = O.advect_src(g, abc, dt, 1.0, W, uf, whalf=True)
        h = wh[1]
        assert np.all(h[0, 0, 0, :] == 3.7)
        np.testing.assert_array_equal(h[0, 0, n, :], h[0, 0, n - 1, :] + (1.0 * dx) * 0.25)
        # extrapolation at the low wall: W- of the first cell, W(0) when the
        # flow leaves through it... W(0) - 0.7*(v dt/2) when v < 0
        ext = -1.0 - (0.7 * v * dt / 2 if v < 0 else 0.0)
        np.testing.assert_allclose(h[1, 0, 0, :], ext, rtol=0, atol=1e-14)
        # inflow/outflow at the high wall: outflow (v > 0) extrapolates W+ of
        # the last cell, inflow takes the value
        if v > 0:
            np.testing.assert_allclose(h[1, 0, n, :], -1.0 + 0.7 * (1.0 - v * dt / 2), rtol=0, atol=1e-14)
        else:
            assert np.all(h[1, 0, n, :] == -4.0)


def test_flux_coefficient_per_component():
    """The reference's component loop multiplies the advection coefficient
    into every component once per component: F = W_half * u * coeff^2 for
    two components."""
    n = 12
    g = O.make_grid([n, n], periodic=(1, 1, 0))
    rng = np.random.default_rng(2)
    W = rng.random(O.cell_shape(g, 2))
    uf = [rng.random(O.face_shape(g, d, 1)) - 0.5 for d in range(2)]
    for f, d in zip(uf, range(2)):
        ax = f.ndim - 1 - d
        sl0 = [slice(None)] * f.ndim
        sl1 = [slice(None)] * f.ndim
        sl0[ax], sl1[ax] = 0, -1
        f[tuple(sl1)] = f[tuple(sl0)]
    coeff = 1.7
    _, flux, wh = O.advect_src(g, O.make_advect_bc(), 0.2 * g.dx, coeff, W, uf, whalf=True)
    for d in range(2):
        np.testing.assert_allclose(flux[d], wh[d] * uf[d] * coeff * coeff, rtol=1e-15, atol=0)


def test_edge_to_cell():
    g = O.make_grid([6, 5], periodic=(0, 0, 0))
    rng = np.random.default_rng(4)
    uf = [rng.random(O.face_shape(g, d, 1)) for d in range(2)]
    vc = O.edge_to_cell(g, uf)
    np.testing.assert_array_equal(O.valid(g, vc[0])[0, 0], 0.5 * (uf[0][0, 0, :, :-1] + uf[0][0, 0, :, 1:]))
    np.testing.assert_array_equal(O.valid(g, vc[1])[0, 0], 0.5 * (uf[1][0, 0, :-1, :] + uf[1][0, 0, 1:, :]))
