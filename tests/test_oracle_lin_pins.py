"""Pins of the oracle's linear variable-coefficient operators (the
Darcy-Brinkman Helmholtz form alpha*a*u - beta*c*u - beta*div(b grad u) and
the projection's Poisson form) against what the mathematics fixes:

* second-order convergence of the operator to a manufactured solution with
  smooth b and c (a wrong face index, sign or dropped term breaks it);
* symmetry of the homogeneous discrete operator (a shifted face index or a
  transposed direction breaks it);
* the multigrid solve equals a dense solve of the assembled system;
* one red/black sweep equals textbook Gauss-Seidel in red-then-black order;
* the restricted residual is the child average of the fine residual and the
  constant prolongation is injection.
"""
import numpy as np
import pytest

from oracle import mlref as O


def _grid(n, periodic):
    return O.make_grid(n, 1.0 / n[0], list(periodic) + [0] * (3 - len(periodic)))


def _centres(g):
    n = [g.n[0], g.n[1], g.n[2] if g.dim == 3 else 1]
    x = (np.arange(n[0]) + 0.5) * g.dx
    y = (np.arange(n[1]) + 0.5) * g.dx
    z = (np.arange(n[2]) + 0.5) * g.dx
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def _face_pts(g, d):
    n = [g.n[0], g.n[1], g.n[2] if g.dim == 3 else 1]
    ax = []
    for e in range(3):
        if e == d:
            ax.append(np.arange(n[e] + 1) * g.dx)
        else:
            ax.append((np.arange(n[e]) + 0.5) * g.dx)
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X, Y, Z


TWO_PI = 2 * np.pi


def _manufactured(dim):
    """u, b, c and L u for alpha*u - beta*c*u - beta*div(b grad u) (a = 1);
    x (and z) periodic, y with zero normal derivative of u at y = 0, 1."""
    def u(X, Y, Z):
        r = np.sin(TWO_PI * X) * np.cos(np.pi * Y) + 0.3 * np.cos(np.pi * Y)
        return r * (np.cos(TWO_PI * Z) + 1.5) if dim == 3 else r

    def ux(X, Y, Z):
        r = TWO_PI * np.cos(TWO_PI * X) * np.cos(np.pi * Y)
        return r * (np.cos(TWO_PI * Z) + 1.5) if dim == 3 else r

    def uy(X, Y, Z):
        r = -np.pi * (np.sin(TWO_PI * X) + 0.3) * np.sin(np.pi * Y)
        return r * (np.cos(TWO_PI * Z) + 1.5) if dim == 3 else r

    def uz(X, Y, Z):
        return (np.sin(TWO_PI * X) * np.cos(np.pi * Y) + 0.3 * np.cos(np.pi * Y)) * (-TWO_PI * np.sin(TWO_PI * Z))

    def b(X, Y, Z):
        return 1.0 + 0.5 * np.sin(TWO_PI * X) * np.sin(np.pi * Y) + (0.2 * np.cos(TWO_PI * Z) if dim == 3 else 0.0)

    def c(X, Y, Z):
        return 2.0 + np.cos(TWO_PI * X) * Y

    def div_b_grad(X, Y, Z, h=1e-5):
        # central differences of the exact flux (analytic to ~h^2 = 1e-10)
        fx = lambda x: b(x, Y, Z) * ux(x, Y, Z)
        fy = lambda y: b(X, y, Z) * uy(X, y, Z)
        s = (fx(X + h) - fx(X - h)) / (2 * h) + (fy(Y + h) - fy(Y - h)) / (2 * h)
        if dim == 3:
            fz = lambda z: b(X, Y, z) * uz(X, Y, z)
            s = s + (fz(Z + h) - fz(Z - h)) / (2 * h)
        return s

    return u, b, c, div_b_grad


@pytest.mark.parametrize("dim", [2, 3])
def test_operator_converges_second_order_to_manufactured_solution(dim):
    alpha, beta = 1.0, -0.05
    u, bf, cf, dbg = _manufactured(dim)
    errs = []
    for n0 in (16, 32):
        n = (n0,) * dim
        g = _grid(n, [1, 0, 1][:dim])
        bc = O.make_bc({(1, 0, 0): (1, 0.0), (1, 1, 0): (1, 0.0)})
        X, Y, Z = _centres(g)
        phi = np.zeros(O.cell_shape(g, 1))
        O.valid(g, phi)[0] = u(X, Y, Z)
        c = np.zeros(O.cell_shape(g, 1))
        O.valid(g, c)[0] = cf(X, Y, Z)
        b = []
        for d in range(dim):
            FX, FY, FZ = _face_pts(g, d)
            b.append(np.ascontiguousarray(bf(FX, FY, FZ)[None]))
        lhs = O.lin_apply_op(g, 1, bc, alpha, None, beta, c, b, phi)
        exact = alpha * u(X, Y, Z) - beta * cf(X, Y, Z) * u(X, Y, Z) - beta * dbg(X, Y, Z)
        errs.append(np.max(np.abs(O.valid(g, lhs)[0] - exact)))
    rate = np.log2(errs[0] / errs[1])
    assert 1.8 < rate < 2.3, (errs, rate)


def _assemble(g, ncomp, bc, alpha, a, beta, c, b):
    """Dense matrix of the homogeneous operator and its inhomogeneous offset,
    from unit vectors."""
    shp = O.cell_shape(g, ncomp)
    v = O.valid(g, np.zeros(shp))
    N = v.size
    A = np.zeros((N, N))
    for m in range(N):
        x = np.zeros(shp)
        e = np.zeros(v.shape)
        e.reshape(-1)[m] = 1.0
        O.valid(g, x)[...] = e
        A[:, m] = O.valid(g, O.lin_apply_op(g, ncomp, bc, alpha, a, beta, c, b, x, homogeneous=1)).ravel()
    f = O.valid(g, O.lin_apply_op(g, ncomp, bc, alpha, a, beta, c, b, np.zeros(shp), homogeneous=0)).ravel()
    return A, f


def _random_coefs(g, ncomp, seed, with_c=True):
    rng = np.random.default_rng(seed)
    b = [np.ascontiguousarray(-(0.5 + rng.random(O.face_shape(g, d, ncomp)))) for d in range(g.dim)]
    for d in range(g.dim):
        # a periodic direction's first and last faces are the same face
        if g.periodic[d]:
            ax = 3 - d
            sl_first = [slice(None)] * 4
            sl_last = [slice(None)] * 4
            sl_first[ax], sl_last[ax] = 0, -1
            b[d][tuple(sl_last)] = b[d][tuple(sl_first)]
    c = None
    if with_c:
        c = np.zeros(O.cell_shape(g, ncomp))
        O.valid(g, c)[...] = 0.1 + rng.random(O.valid(g, c).shape)
    return b, c


@pytest.mark.parametrize("n,periodic", [((6, 5), (0, 0)), ((4, 6), (1, 0)), ((4, 3, 4), (0, 1, 0))])
def test_homogeneous_operator_is_symmetric(n, periodic):
    g = _grid(n, periodic)
    bc = O.make_bc({(0, 0, 0): (0, 0.0), (1, 1, 0): (0, 0.0), (2, 0, 0): (0, 0.0)})
    b, c = _random_coefs(g, 1, 4)
    A, _ = _assemble(g, 1, bc, 1.0, None, -0.01, c, b)
    np.testing.assert_allclose(A, A.T, rtol=1e-13, atol=1e-12)
    # symmetric positive definite for alpha > 0, beta < 0 (b < 0, c > 0)
    assert np.all(np.linalg.eigvalsh(A) > 0)


@pytest.mark.parametrize("with_c,prolong", [(True, 0), (False, 0), (True, 1)])
def test_mg_solve_matches_dense_solve(with_c, prolong):
    g = _grid((16, 16), (1, 0))
    bc = O.make_bc({(1, 0, 0): (0, 0.4), (1, 1, 0): (1, -0.3)})
    b, c = _random_coefs(g, 1, 5, with_c)
    alpha, beta = (1.0, -0.02) if with_c else (0.3, 1.0)
    if not with_c:
        b = [-x for x in b]   # projection sign convention: beta = 1, b = -S < 0 flipped to keep SPD
        beta = -1.0
    A, f = _assemble(g, 1, bc, alpha, None, beta, c, b)
    rng = np.random.default_rng(6)
    rhs = np.zeros(O.cell_shape(g, 1))
    O.valid(g, rhs)[...] = rng.standard_normal(O.valid(g, rhs).shape)
    x_dense = np.linalg.solve(A, O.valid(g, rhs).ravel() - f)
    phi = np.zeros(O.cell_shape(g, 1))
    st = O.lin_mg_solve(g, 1, bc, alpha, None, beta, c, b, phi, rhs, prolong_linear=prolong,
                        opts=O.default_opts(nu1=2, nu2=2, bottom=40, max_iter=60, eps=1e-13, norm_thresh=1e-14))
    assert st.exit_status & (1 | 8), (st.exit_status, st.iterations, st.final_residual)
    np.testing.assert_allclose(O.valid(g, phi).ravel(), x_dense, rtol=1e-10, atol=1e-11)


def test_gsrb_is_textbook_red_black_gauss_seidel():
    g = _grid((6, 4), (1, 1))
    bc = O.make_bc({})
    b, c = _random_coefs(g, 1, 7)
    alpha, beta = 1.0, -0.05
    A, f = _assemble(g, 1, bc, alpha, None, beta, c, b)
    assert np.allclose(f, 0.0)
    rng = np.random.default_rng(8)
    phi = np.zeros(O.cell_shape(g, 1))
    rhs = np.zeros(O.cell_shape(g, 1))
    O.valid(g, phi)[...] = rng.standard_normal(O.valid(g, phi).shape)
    O.valid(g, rhs)[...] = rng.standard_normal(O.valid(g, rhs).shape)
    x = O.valid(g, phi).ravel().copy()
    r = O.valid(g, rhs).ravel()
    nx, ny = 6, 4
    for colour in (0, 1):
        for j in range(ny):
            for i in range(nx):
                if (i + j + colour) % 2 == 0:
                    m = i + nx * j
                    x[m] = x[m] + (r[m] - A[m] @ x) / A[m, m]
    out = O.lin_gsrb(g, 1, bc, alpha, None, beta, c, b, phi, rhs)
    np.testing.assert_allclose(O.valid(g, out).ravel(), x, rtol=1e-13, atol=1e-13)


def test_restrict_residual_is_child_average_and_prolong_is_injection():
    g = _grid((8, 8, 4), (0, 1, 0))
    bc = O.make_bc({(0, 0, 0): (0, 1.0), (2, 1, 0): (1, 2.0)})
    b, c = _random_coefs(g, 1, 9)
    rng = np.random.default_rng(10)
    phi = np.zeros(O.cell_shape(g, 1))
    rhs = np.zeros(O.cell_shape(g, 1))
    O.valid(g, phi)[...] = rng.standard_normal(O.valid(g, phi).shape)
    O.valid(g, rhs)[...] = rng.standard_normal(O.valid(g, rhs).shape)
    res = O.lin_residual(g, 1, bc, 1.0, None, -0.1, c, b, phi.copy(), rhs, homogeneous=1)
    resc = O.lin_restrict_residual(g, 1, bc, 1.0, None, -0.1, c, b, phi.copy(), rhs)
    v = O.valid(g, res)[0]
    avg = v.reshape(2, 2, 4, 2, 4, 2).mean(axis=(1, 3, 5))
    gc = O.coarsen(g, 2)
    np.testing.assert_allclose(O.valid(gc, resc)[0], avg, rtol=1e-13, atol=1e-14)
    corr = rng.standard_normal(O.cell_shape(gc, 1))
    fine = np.zeros(O.cell_shape(g, 1))
    O.prolong_constant(g, 1, fine, corr)
    inj = np.repeat(np.repeat(np.repeat(O.valid(gc, corr)[0], 2, 0), 2, 1), 2, 2)
    np.testing.assert_array_equal(O.valid(g, fine)[0], inj)
