"""Pins of the oracle's BiCGStab bottom solver of the linear multigrid (the
reference's bottom solver of the projection and U* solves,
srcSubcycle/Projector.cpp:3004-3023, srcSubcycle/AMRLevelMushyLayerInit.cpp:739):

* the dot product equals numpy's (to rounding; exactly for integer data);
* preCond is lambda*rhs followed by two textbook red/black Gauss-Seidel
  sweeps of the assembled matrix;
* every iterate equals scipy's independent BiCGStab
  (scipy.sparse.linalg.bicgstab, right preconditioning with the same preCond
  as a linear operator) after 1, 2, 3 and 5 iterations;
* the residual norm it reports is the true residual of the returned
  solution, and repeated calls converge to the dense solve;
* a multigrid solve with the BiCGStab bottom converges to the dense solve,
  and on a level that does not coarsen one V-cycle is one BiCGStab solve.
"""
import numpy as np
import pytest
from scipy.sparse.linalg import LinearOperator, bicgstab

from oracle import mlref as O


def _grid(n, periodic):
    return O.make_grid(n, 1.0 / n[0], list(periodic) + [0] * (3 - len(periodic)))


def _coefs(g, seed, with_c=True):
    rng = np.random.default_rng(seed)
    b = [np.ascontiguousarray(-(0.5 + rng.random(O.face_shape(g, d, 1)))) for d in range(g.dim)]
    for d in range(g.dim):
        if g.periodic[d]:   # a periodic direction's first and last faces are one face
            ax = 3 - d
            sl_first = [slice(None)] * 4
            sl_last = [slice(None)] * 4
            sl_first[ax], sl_last[ax] = 0, -1
            b[d][tuple(sl_last)] = b[d][tuple(sl_first)]
    c = None
    if with_c:
        c = np.zeros(O.cell_shape(g, 1))
        O.valid(g, c)[...] = 0.1 + rng.random(O.valid(g, c).shape)
    return b, c


def _assemble(g, bc, alpha, beta, c, b):
    shp = O.cell_shape(g, 1)
    v = O.valid(g, np.zeros(shp))
    N = v.size
    A = np.zeros((N, N))
    for m in range(N):
        x = np.zeros(shp)
        e = np.zeros(v.shape)
        e.reshape(-1)[m] = 1.0
        O.valid(g, x)[...] = e
        A[:, m] = O.valid(g, O.lin_apply_op(g, 1, bc, alpha, None, beta, c, b, x, homogeneous=1)).ravel()
    return A


def _field(g, vals):
    x = np.zeros(O.cell_shape(g, 1))
    O.valid(g, x)[...] = np.asarray(vals).reshape(O.valid(g, x).shape)
    return x


CASES = [((8, 6), (0, 0), True), ((6, 8), (1, 0), False), ((4, 4, 4), (0, 1, 0), True)]
BC = O.make_bc({(0, 0, 0): (0, 0.0), (1, 1, 0): (0, 0.0), (1, 0, 0): (1, 0.0), (2, 0, 0): (0, 0.0)})


def _problem(n, periodic, with_c, seed):
    g = _grid(n, periodic)
    b, c = _coefs(g, seed, with_c)
    alpha, beta = (1.0, -0.02) if with_c else (0.0, 1.0)
    if not with_c:
        b = [-x for x in b]   # -div(|b| grad) with beta = 1: SPD
    return g, b, c, alpha, beta


def test_dot_matches_numpy():
    g = _grid((13, 11), (0, 0))
    rng = np.random.default_rng(1)
    x = _field(g, rng.standard_normal(13 * 11))
    y = _field(g, rng.standard_normal(13 * 11))
    ref = float(np.dot(O.valid(g, x).ravel(), O.valid(g, y).ravel()))
    assert abs(O.lin_dot(g, 1, x, y) - ref) <= 1e-13 * np.sum(np.abs(O.valid(g, x) * O.valid(g, y)))
    xi = _field(g, rng.integers(-50, 50, 13 * 11).astype(float))
    yi = _field(g, rng.integers(-50, 50, 13 * 11).astype(float))
    assert O.lin_dot(g, 1, xi, yi) == float(np.dot(O.valid(g, xi).ravel(), O.valid(g, yi).ravel()))
    # ghost cells do not take part
    x2 = x.copy()
    x2[0, 0, 0, :] = 1e30
    assert O.lin_dot(g, 1, x2, y) == O.lin_dot(g, 1, x, y)


def test_precond_is_lambda_then_two_red_black_sweeps():
    g = _grid((6, 4), (1, 1))   # periodic: lambda is exactly 1/diag
    bc = O.make_bc({})
    b, c = _coefs(g, 7)
    alpha, beta = 1.0, -0.05
    A = _assemble(g, bc, alpha, beta, c, b)
    rng = np.random.default_rng(8)
    r = rng.standard_normal(24)
    x = r / np.diag(A)
    nx, ny = 6, 4
    for _ in range(2):
        for colour in (0, 1):
            for j in range(ny):
                for i in range(nx):
                    if (i + j + colour) % 2 == 0:
                        m = i + nx * j
                        x[m] = x[m] + (r[m] - A[m] @ x) / A[m, m]
    out = O.lin_precond(g, 1, bc, alpha, None, beta, c, b, _field(g, r))
    np.testing.assert_allclose(O.valid(g, out).ravel(), x, rtol=1e-12, atol=1e-13)


def _precond_operator(g, bc, alpha, beta, c, b):
    N = int(np.prod(O.valid(g, np.zeros(O.cell_shape(g, 1))).shape))

    def mv(v):
        return O.valid(g, O.lin_precond(g, 1, bc, alpha, None, beta, c, b, _field(g, np.ravel(v)))).ravel().copy()
    return LinearOperator((N, N), matvec=mv, dtype=float)


@pytest.mark.parametrize("n,periodic,with_c", CASES)
@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_iterates_match_scipy_bicgstab(n, periodic, with_c, k):
    g, b, c, alpha, beta = _problem(n, periodic, with_c, 11)
    A = _assemble(g, BC, alpha, beta, c, b)
    rng = np.random.default_rng(12)
    rhs = rng.standard_normal(A.shape[0])
    M = _precond_operator(g, BC, alpha, beta, c, b)
    # stay below the iteration that converges: there the solver stops after
    # the alpha half-step (|r| <= eps r0), where scipy completes the iteration
    _, it_conv, _ = O.lin_bicgstab(g, 1, BC, alpha, None, beta, c, b, np.zeros(O.cell_shape(g, 1)),
                                   _field(g, rhs), imax=100)
    k = min(k, it_conv - 1)
    assert k >= 1
    # scipy: x0 = 0, k iterations, no early stop
    x_sp, _ = bicgstab(A, rhs, x0=np.zeros_like(rhs), rtol=1e-300, atol=0.0, maxiter=k, M=M)
    phi = np.zeros(O.cell_shape(g, 1))
    status, it, _ = O.lin_bicgstab(g, 1, BC, alpha, None, beta, c, b, phi, _field(g, rhs), imax=k)
    assert it == k and status == 2
    x = O.valid(g, phi).ravel()
    scale = np.linalg.norm(np.linalg.solve(A, rhs))
    assert np.linalg.norm(x - x_sp) <= 1e-11 * scale, (np.linalg.norm(x - x_sp) / scale, status)


@pytest.mark.parametrize("n,periodic,with_c", CASES)
def test_reported_residual_is_true_residual_and_converges_to_dense(n, periodic, with_c):
    g, b, c, alpha, beta = _problem(n, periodic, with_c, 21)
    A = _assemble(g, BC, alpha, beta, c, b)
    rng = np.random.default_rng(22)
    rhs = rng.standard_normal(A.shape[0])
    phi = np.zeros(O.cell_shape(g, 1))
    r0 = np.linalg.norm(rhs)
    status, it, fn = O.lin_bicgstab(g, 1, BC, alpha, None, beta, c, b, phi, _field(g, rhs))
    assert status == 1 and it < A.shape[0]
    true = np.linalg.norm(rhs - A @ O.valid(g, phi).ravel())
    assert fn <= 1e-6 * r0 and abs(true - fn) <= 1e-9 * r0, (fn / r0, true / r0)
    for _ in range(4):   # each call starts from the true residual of the last
        O.lin_bicgstab(g, 1, BC, alpha, None, beta, c, b, phi, _field(g, rhs))
    x_dense = np.linalg.solve(A, rhs)
    np.testing.assert_allclose(O.valid(g, phi).ravel(), x_dense, rtol=1e-9, atol=1e-10 * np.abs(x_dense).max())


def test_zero_rhs_leaves_phi():
    g, b, c, alpha, beta = _problem((8, 6), (0, 0), True, 31)
    phi = np.zeros(O.cell_shape(g, 1))
    status, it, fn = O.lin_bicgstab(g, 1, BC, alpha, None, beta, c, b, phi, np.zeros_like(phi))
    assert it == 0 and fn == 0.0 and not phi.any()


@pytest.mark.parametrize("with_c", [True, False])
def test_mg_solve_with_bicgstab_bottom_matches_dense(with_c):
    g, b, c, alpha, beta = _problem((16, 16), (1, 0), with_c, 41)
    A = _assemble(g, BC, alpha, beta, c, b)
    rng = np.random.default_rng(42)
    rhs = rng.standard_normal(A.shape[0])
    phi = np.zeros(O.cell_shape(g, 1))
    o = O.default_opts(nu1=2, nu2=2, bottom=2, max_iter=60, eps=1e-13, norm_thresh=1e-14,
                       bottom_solver=O.BOTTOM_BICGSTAB, bottom_max_iter=20)
    st = O.lin_mg_solve(g, 1, BC, alpha, None, beta, c, b, phi, _field(g, rhs), opts=o)
    assert st.exit_status & (1 | 8)
    x_dense = np.linalg.solve(A, rhs)
    np.testing.assert_allclose(O.valid(g, phi).ravel(), x_dense, rtol=1e-10, atol=1e-11 * np.abs(x_dense).max())


def test_single_level_vcycle_is_one_bicgstab_solve():
    # 7 x 5 does not coarsen: the V-cycle is the bottom (no relaxations, BiCGStab)
    g, b, c, alpha, beta = _problem((7, 5), (0, 0), True, 51)
    rng = np.random.default_rng(52)
    rhs = _field(g, rng.standard_normal(35))
    o = O.default_opts(nu1=2, nu2=2, bottom=0, max_iter=1, eps=1e-30, hang=-1e300, norm_thresh=0.0,
                       bottom_solver=O.BOTTOM_BICGSTAB, bottom_max_iter=20)
    phi = np.zeros(O.cell_shape(g, 1))
    O.lin_mg_solve(g, 1, BC, alpha, None, beta, c, b, phi, rhs, opts=o)
    e = np.zeros(O.cell_shape(g, 1))
    O.lin_bicgstab(g, 1, BC, alpha, None, beta, c, b, e, rhs, imax=20)
    # phi = 0 + e, with the homogeneous-BC residual of phi = 0 equal to rhs
    np.testing.assert_array_equal(O.valid(g, phi), O.valid(g, e))
    # relaxation bottom of 0 sweeps leaves the correction at zero
    o.bottom_solver = O.BOTTOM_RELAX
    phi2 = np.zeros(O.cell_shape(g, 1))
    O.lin_mg_solve(g, 1, BC, alpha, None, beta, c, b, phi2, rhs, opts=o)
    assert not O.valid(g, phi2).any()
