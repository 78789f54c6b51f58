"""Pins of the oracle's state update and velocity closure (oracle/mlref_flow.c)
against what the mathematics fixes, independent of its formulas:

* enthalpy variables: closed forms per regime, the solid concentration from
  an independent brentq root of the mush relations, and the mass balance
  chi Theta_l + (1 - chi) Theta_s = Theta in the mush;
* the pressure gradient: exact for quadratic pressure fields at every cell,
  walls included (second-order extrapolation is exact for quadratics);
* the buoyancy: its closed form on arbitrary fields;
* the projection's right-hand side: the cell-centred divergence is exact for
  linear velocity components inside, telescopes to zero over a domain with
  no-slip walls and periodic sides, and converges at second order to the
  analytic divergence;
* the whole projection: the projected velocity's divergence falls at second
  order under refinement while the unprojected one does not, with variable
  porosity, and solid cells carry no velocity.

A plausible slip (a wrong extrapolation coefficient, a missing porosity
factor, a sign or a 1/dt scale) fails one of these."""
import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import mlref as O
from oracle import steps as S

P = dict(stefan=5.0, compositionRatio=2.0, lewis=200.0)


def _grid(n, periodic=(1, 0, 0)):
    return O.make_grid(list(n), 1.0 / n[0], list(periodic))


def _coords(g):
    """Cell-centre coordinates (x, z) of the valid cells, shape (ny, nx)."""
    x = (np.arange(g.n[0]) + 0.5) * g.dx
    z = (np.arange(g.n[1]) + 0.5) * g.dx
    return np.meshgrid(x, z)


# ------------------------------------------------------- enthalpy variables

def test_enthalpy_vars_regimes():
    for St, CR, pc in [(5.0, 2.0, 1e-5), (2.0, 8.0, 0.05)]:
        p = O.default_params(stefan=St, compositionRatio=CR, waterDistributionCoeff=pc)
        # solid
        T, Cl, Cs, chi = O.enthalpy_vars_point(-1.0, -1.0, p)
        assert (T, Cl, Cs, chi) == (-1.0, 0.0, -1.0, 0.0)
        # liquid
        T, Cl, Cs, chi = O.enthalpy_vars_point(St + 1.5, -0.5, p)
        assert (T, Cl, Cs, chi) == (St + 1.5 - St, -0.5, 0.0, 1.0)
        # mush: chi from an independent root solve of the mush relations
        # Theta = Theta_l (chi + pc (1 - chi)) - CR (1 - chi), H = chi St - Theta_l
        for C, w in [(-0.8, 0.5), (-0.5 * CR, 0.3), (-0.1, 0.9)]:
            He, Hl = (CR + C) / CR * St, St - C          # eutectic and liquidus enthalpies
            H = He + w * (Hl - He)
            T, Cl, Cs, chi = O.enthalpy_vars_point(H, C, p)
            f = lambda x: (x * St - H) * (x + pc * (1 - x)) - CR * (1 - x) - C
            ref = brentq(f, 1e-9, 1 - 1e-12, xtol=1e-15)
            assert chi == pytest.approx(ref, abs=1e-12)
            assert T == pytest.approx(H - chi * St, abs=1e-10)          # H = chi St + theta
            assert Cl == pytest.approx(-T, abs=1e-14)                   # liquidus Theta_l = -theta
            assert chi * Cl + (1 - chi) * Cs == pytest.approx(C, abs=1e-12)   # mass balance


def test_enthalpy_vars_eutectic_solid_concentration():
    """On the eutectic plateau the reference takes the solid concentration
    from the mush-branch porosity (CALCULATECS), not from the eutectic
    porosity H/St that CALCULATEPOROSITY reports."""
    St, CR, pc = 5.0, 2.0, 1e-5
    p = O.default_params(stefan=St, compositionRatio=CR, waterDistributionCoeff=pc)
    H, C = 1.0, -1.0
    T, Cl, Cs, chi = O.enthalpy_vars_point(H, C, p)
    assert (T, Cl) == (0.0, 0.0)
    assert chi == pytest.approx(H / St, rel=1e-15)
    # the root in [0, 1] of the mush relations continued below H_e
    # (c = 1, theta_e = Theta_e = 0), found independently
    f = lambda x: (x * St - H) * (x + pc * (1 - x)) - CR * (1 - x) - C
    pm = brentq(f, 0.0, 1 - 1e-12, xtol=1e-15)
    assert Cs == pytest.approx(C / (1 - pm), rel=1e-12)


def test_enthalpy_vars_field_matches_points_and_closure():
    g = _grid((12, 10))
    p = O.default_params(**P)
    rng = np.random.default_rng(1)
    phi = np.zeros(O.cell_shape(g, 2))
    v = O.valid(g, phi)
    v[0] = rng.uniform(-1, 8, v[0].shape)
    v[1] = rng.uniform(-1.9, -0.2, v[1].shape)
    out = O.enthalpy_vars(g, p, phi)
    vo = O.valid(g, out)
    for j in range(10):
        for i in range(12):
            t, cl = O.closure_point(v[0, 0, j, i], v[1, 0, j, i], p)
            assert vo[0, 0, j, i] == t and vo[1, 0, j, i] == cl
            assert vo[3, 0, j, i] == O.porosity_point(v[0, 0, j, i], v[1, 0, j, i], p)


# ---------------------------------------------------- U* source and gradient

def _vars(g, rng, chi=None):
    v = np.zeros(O.cell_shape(g, 4))
    w = O.valid(g, v)
    w[0] = rng.uniform(-0.5, 1.0, w[0].shape)
    w[1] = rng.uniform(-1.0, 0.5, w[1].shape)
    w[2] = rng.uniform(-2.0, 0.0, w[2].shape)
    w[3] = rng.uniform(0.1, 1.0, w[3].shape) if chi is None else chi
    return v


@pytest.mark.parametrize("n,periodic", [((16, 12), (1, 0, 0)), ((12, 16), (0, 0, 0))])
def test_ustar_src_buoyancy_and_quadratic_pressure(n, periodic):
    g = _grid(n, periodic)
    rng = np.random.default_rng(2)
    vo, vn = _vars(g, rng), _vars(g, rng)
    X, Z = _coords(g)
    a2, a1 = 0.7, -1.3
    pi = np.zeros(O.cell_shape(g, 1))
    # quadratic in z; linear-periodic-free in x unless x is walled
    px = 0.0 if periodic[0] else 1.0
    O.valid(g, pi)[0, 0] = a2 * Z ** 2 + a1 * Z + px * (0.4 * X ** 2 - 0.2 * X)
    bT, bS = 3.0, 2000.0
    src = O.ustar_src(g, bT, bS, 1, vo, vn, pi)
    th = 0.5 * O.valid(g, vo)[0] + 0.5 * O.valid(g, vn)[0]
    thl = 0.5 * O.valid(g, vo)[1] + 0.5 * O.valid(g, vn)[1]
    chi = 0.5 * O.valid(g, vo)[3] + 0.5 * O.valid(g, vn)[3]
    gz = 2 * a2 * Z + a1                 # exact derivative at cell centres
    gx = px * (0.8 * X - 0.2)
    np.testing.assert_allclose(O.valid(g, src[1])[0, 0], chi[0] * (bT * th[0] - bS * thl[0]) - chi[0] * gz,
                               rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(O.valid(g, src[0])[0, 0], -chi[0] * gx, rtol=1e-11, atol=1e-11)


# ------------------------------------------------- projection right-hand side

def test_divergence_linear_exact_and_conservative():
    n = (16, 16)
    g = _grid(n, (1, 0, 0))
    X, Z = _coords(g)
    rng = np.random.default_rng(3)
    chi = np.full(O.valid(g, np.zeros(O.cell_shape(g, 1)))[0].shape, 0.9)
    vo = _vars(g, rng, chi)
    vn = _vars(g, rng, chi)
    pi = np.zeros(O.cell_shape(g, 1))
    ustar = [np.zeros(O.cell_shape(g, 1)) for _ in range(2)]
    O.valid(g, ustar[0])[0, 0] = np.sin(2 * np.pi * X) * (1 + Z)
    O.valid(g, ustar[1])[0, 0] = 0.3 + 2.0 * Z + np.cos(2 * np.pi * X)
    dt = 1e-3
    u, rhs = O.cc_project_rhs(g, dt, 0.05, 1, vo, vn, pi, ustar)
    div = O.valid(g, rhs)[0, 0] * dt
    # with pi = 0 the velocity is U* itself
    np.testing.assert_array_equal(O.valid(g, u[0]), O.valid(g, ustar[0]))
    # interior rows: d(u_z)/dz = 2 exactly for the linear part; the periodic
    # x part is the central difference of sin
    h = g.dx
    dux = (np.sin(2 * np.pi * (X + h)) - np.sin(2 * np.pi * (X - h))) / (2 * h) * (1 + Z)
    np.testing.assert_allclose(div[1:-1], (dux + 2.0)[1:-1], rtol=1e-11, atol=1e-10)
    # walls: the face value there is 0 (no-slip ghost), so the sum of the
    # divergence telescopes to zero
    assert abs(div.sum()) <= 1e-10 * np.abs(div).max()


def test_divergence_second_order():
    """div of u = (sin(2 pi x) z(1-z), z^2 (1-z)^2 cos(2 pi x)), which
    vanishes on the walls, converges at second order away from the walls and
    at first order in the wall rows (the no-slip ghost -u1 makes the face
    value at the wall exactly 0 but the one-sided difference there O(h))."""
    inner, wall = [], []
    for m in (32, 64, 128):
        g = _grid((m, m), (1, 0, 0))
        X, Z = _coords(g)
        one = np.ones(O.valid(g, np.zeros(O.cell_shape(g, 1)))[0].shape)
        rng = np.random.default_rng(4)
        vo, vn = _vars(g, rng, one), _vars(g, rng, one)
        ustar = [np.zeros(O.cell_shape(g, 1)) for _ in range(2)]
        O.valid(g, ustar[0])[0, 0] = np.sin(2 * np.pi * X) * Z * (1 - Z)
        O.valid(g, ustar[1])[0, 0] = Z ** 2 * (1 - Z) ** 2 * np.cos(2 * np.pi * X)
        _, rhs = O.cc_project_rhs(g, 1.0, 0.05, 1, vo, vn, np.zeros(O.cell_shape(g, 1)), ustar)
        exact = (2 * np.pi * np.cos(2 * np.pi * X) * Z * (1 - Z) +
                 (2 * Z * (1 - Z) ** 2 - 2 * Z ** 2 * (1 - Z)) * np.cos(2 * np.pi * X))
        e = np.abs(O.valid(g, rhs)[0, 0] - exact)
        inner.append(e[1:-1].max())
        wall.append(e[[0, -1]].max())
    assert inner[0] / inner[1] > 3.5 and inner[1] / inner[2] > 3.5, inner
    assert wall[0] / wall[1] > 1.8 and wall[1] / wall[2] > 1.8, wall


def test_pressure_faces_arithmetic_mean():
    g = _grid((8, 6), (0, 0, 0))
    rng = np.random.default_rng(5)
    vo, vn = _vars(g, rng), _vars(g, rng)
    b = O.pressure_faces(g, vo, vn)
    c = (0.5 * O.valid(g, vo)[3] + 0.5 * O.valid(g, vn)[3])[0]
    np.testing.assert_allclose(b[0][0, 0, :, 1:-1], -0.5 * (c[:, :-1] + c[:, 1:]), rtol=1e-15)
    np.testing.assert_array_equal(b[0][0, 0, :, 0], -c[:, 0])
    np.testing.assert_array_equal(b[0][0, 0, :, -1], -c[:, -1])
    np.testing.assert_allclose(b[1][0, 0, 1:-1, :], -0.5 * (c[:-1] + c[1:]), rtol=1e-15)
    np.testing.assert_array_equal(b[1][0, 0, 0, :], -c[0])


# ------------------------------------------------------- the whole projection

def _project(m, chi_fn, limit=0.05):
    g = _grid((m, m), (1, 0, 0))
    X, Z = _coords(g)
    chi = chi_fn(X, Z)
    rng = np.random.default_rng(6)
    vo, vn = _vars(g, rng, chi), _vars(g, rng, chi)
    ustar = [np.zeros(O.cell_shape(g, 1)) for _ in range(2)]
    O.valid(g, ustar[0])[0, 0] = np.sin(2 * np.pi * X) * Z * (1 - Z) + 0.3 * np.cos(2 * np.pi * X)
    O.valid(g, ustar[1])[0, 0] = np.sin(np.pi * Z) * (1 + 0.5 * np.sin(2 * np.pi * X))
    dt = 1e-3
    pi0 = np.zeros(O.cell_shape(g, 1))
    u, rhs = O.cc_project_rhs(g, dt, limit, 1, vo, vn, pi0, ustar)
    bp = O.pressure_faces(g, vo, vn)
    pres = np.zeros(O.cell_shape(g, 1))
    bcp = O.make_bc({(1, 0, 0): (1, 0.0), (1, 1, 0): (1, 0.0)})
    st = O.lin_mg_solve(g, 1, bcp, 0.0, None, 1.0, None, bp, pres, rhs, opts=O.default_opts(**S.P_OPTS))
    O.cc_correct(g, dt, limit, 1, vo, vn, pres, u)
    _, rhs1 = O.cc_project_rhs(g, dt, limit, 1, vo, vn, np.zeros(O.cell_shape(g, 1)), u)
    return g, O.valid(g, rhs)[0, 0] * dt, O.valid(g, rhs1)[0, 0] * dt, u, vn, st


@pytest.mark.parametrize("chi_fn", [lambda X, Z: np.ones_like(X),
                                    lambda X, Z: 0.6 + 0.3 * np.sin(2 * np.pi * X) * np.cos(np.pi * Z)])
def test_projection_removes_divergence_at_second_order(chi_fn):
    """The cell-centred projection is approximate: the divergence of the
    projected velocity is not zero but falls at second order away from the
    walls (rows next to them: first order), from an O(1) divergence before."""
    before, inner, full = [], [], []
    for m in (32, 64, 128):
        g, d0, d1, u, vn, st = _project(m, chi_fn)
        # the projection's tolerance is 1e-15 relative; the hang rule ends it
        # once the residual stops falling: the rounding floor of this singular
        # (periodic, all-Neumann) problem, 1e-14 to 4e-12 of its start with
        # the BiCGStab bottom
        assert st.final_residual <= 1e-11 * st.initial_residual
        before.append(np.abs(d0).max())
        inner.append(np.abs(d1[2:-2]).max())
        full.append(np.abs(d1).max())
    assert min(before) > 5.0
    assert full[0] < 0.05 * before[0]
    assert inner[0] / inner[1] > 3.4 and inner[1] / inner[2] > 3.4, inner
    assert full[0] / full[1] > 1.8 and full[1] / full[2] > 1.8, full


def test_projection_zeroes_solid_cells():
    def chi_fn(X, Z):
        c = np.full_like(X, 0.8)
        c[Z < 0.25] = 0.02                       # a solid bottom layer
        return c
    g, d0, d1, u, vn, st = _project(32, chi_fn)
    solid = O.valid(g, vn)[3, 0] <= 0.05
    assert solid.any()
    for d in range(2):
        assert np.all(O.valid(g, u[d])[0, 0][solid] == 0.0)
        assert np.abs(O.valid(g, u[d])[0, 0][~solid]).max() > 1e-3


# ---------------------------------------------------------- permeability laws

def test_once_rounded_cube_is_correctly_rounded():
    """The Kozeny-Carman and cubic laws use x^3 rounded once (FMA remainders);
    checked here against the correctly rounded cube from exact rational
    arithmetic, on random porosities and near-1 values."""
    from fractions import Fraction
    rng = np.random.default_rng(12)
    xs = np.concatenate([rng.uniform(0.0, 1.0, 2000), 1.0 - rng.uniform(0.0, 1e-6, 200), [0.05, 0.5, 1.0]])
    g = O.make_grid([xs.size, 1], 1.0, [0, 0, 0])
    chi = np.zeros(O.cell_shape(g, 1))
    O.valid(g, chi)[0, 0, 0, :] = xs
    cub = O.valid(g, O.darcy_ccoef(g, chi, 1.0, "cubic"))[0, 0, 0, :]   # chi / chi^3
    for x, c in zip(xs, cub):
        if x == 0.0:
            continue
        x3 = float(Fraction(x) ** 3)          # correctly rounded x^3
        assert c == x / x3


@pytest.mark.parametrize("law,fn", [
    ("kozeny_carman", lambda x: x ** 3 / (1 - x) ** 2), ("cubic", lambda x: x ** 3),
    ("log", lambda x: -x ** 2 * np.log(1 - x)), ("porosity", lambda x: x), ("pure_fluid", lambda x: 1.0 + 0 * x)])
def test_darcy_coefficient_closed_forms(law, fn):
    """c = coeff * chi / Pi(chi) for every permeability law
    (src/MushyLayerParams.cpp:1031-1090), and the Hele-Shaw harmonic mean."""
    rng = np.random.default_rng(13)
    xs = rng.uniform(0.05, 0.95, 300)
    g = O.make_grid([xs.size, 1], 1.0, [0, 0, 0])
    chi = np.zeros(O.cell_shape(g, 1))
    O.valid(g, chi)[0, 0, 0, :] = xs
    c = O.valid(g, O.darcy_ccoef(g, chi, 7.0, law))[0, 0, 0, :]
    np.testing.assert_allclose(c, 7.0 * xs / fn(xs), rtol=1e-14)
    hs = 0.3
    c = O.valid(g, O.darcy_ccoef(g, chi, 7.0, law, hs))[0, 0, 0, :]
    np.testing.assert_allclose(c, 7.0 * xs * (1 / hs + 1 / fn(xs)), rtol=1e-13)
