"""The advective source of the (H, Theta) equations on the GPU, through the C
ABI, against the oracle (oracle/mlref_advect.c):

* ml_hc_advect_src: src - div(flux) and the face fluxes bit-identical on
  random states, velocities and diffusive sources, in 2D and 3D, periodic
  and walled, with every supported wall type per side and component
  (Dirichlet, extrapolation, inflow/outflow, Neumann) and an advection
  coefficient other than 1;
* ml_hc_diffusive_src (computeScalDiffusion): bit-identical to the oracle's
  operator applied with alpha = 0, beta = 1;
* ml_vc_face_velocity: bit-identical face averages, zero on walls;
* the plume inflow type is refused.
The whole Darcy-Brinkman steps with the advective source in front of the
(H, Theta) solve are compared in test_gpu_flow.py."""
import numpy as np
import pytest

from oracle import mlref as O
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu

CASES = [
    ((24, 16), (1, 0)),
    ((18, 20), (0, 0)),
    ((12, 10, 8), (1, 0, 1)),
    ((10, 9, 11), (0, 0, 0)),
]

WALL_TYPES = [O.ADV_DIRICHLET, O.ADV_EXTRAP, O.ADV_INFLOW_OUTFLOW, O.ADV_NEUMANN]


def _grid(n, periodic):
    return O.make_grid(list(n), 1.0 / n[0], list(periodic) + [0] * (3 - len(periodic)))


def _periodic_faces(g, f, d):
    """Make the two end faces of a periodic direction one face."""
    if g.periodic[d]:
        ax = f.ndim - 1 - d
        s0 = [slice(None)] * f.ndim
        s1 = [slice(None)] * f.ndim
        s0[ax], s1[ax] = 0, -1
        f[tuple(s1)] = f[tuple(s0)]
    return f


def _adv_bc(g, rng):
    spec = {}
    for d in range(g.dim):
        for s in range(2):
            for c in range(2):
                spec[(d, s, c)] = (WALL_TYPES[rng.integers(len(WALL_TYPES))], float(rng.uniform(-1, 1)))
    return spec


@pytest.mark.parametrize("n,periodic", CASES)
@pytest.mark.parametrize("coeff", [1.0, 1.7])
def test_advect_src_bitwise(ml, gpu_ctx, n, periodic, coeff):
    g = _grid(n, periodic)
    dim = len(n)
    rng = np.random.default_rng(3 + sum(n) + int(coeff * 10))
    W = rng.uniform(-1, 1, O.cell_shape(g, 2))
    S = rng.uniform(-2, 2, O.cell_shape(g, 2))
    src0 = rng.uniform(-1, 1, O.cell_shape(g, 2))
    uf = [_periodic_faces(g, rng.uniform(-3, 3, O.face_shape(g, d, 1)), d) for d in range(dim)]
    # a few faces with exactly zero velocity (the averaged Riemann state)
    for d in range(dim):
        uf[d].reshape(-1)[:: 7] = 0.0
        _periodic_faces(g, uf[d], d)
    spec = _adv_bc(g, rng)
    dt = 0.2 * g.dx / 3.0
    ref, rflux, _ = O.advect_src(g, O.make_advect_bc(spec), dt, coeff, W, uf, S=S, src=src0)
    op = ml.HCOp(gpu_ctx, list(n), g.dx, list(periodic), ml.make_params(), ml.make_bc())
    dW = ml.DeviceArray(gpu_ctx, W.shape, host=W)
    dS = ml.DeviceArray(gpu_ctx, S.shape, host=S)
    dsrc = ml.DeviceArray(gpu_ctx, src0.shape, host=src0)
    duf = [ml.DeviceArray(gpu_ctx, u.shape, host=u) for u in uf]
    dfl = [ml.DeviceArray(gpu_ctx, O.face_shape(g, d, 2)) for d in range(dim)]
    ap = ml.make_advect_params(None, coeff, spec)
    op.advect_src(ap, dt, dW, duf, dS, dsrc, flux=dfl)
    got = dsrc.get()
    assert HC.bitwise_equal(O.valid(g, got), O.valid(g, ref))
    for d in range(dim):
        assert HC.bitwise_equal(dfl[d].get(), rflux[d]), d
    # without a diffusive source and without flux outputs (own workspace)
    ref2, _, _ = O.advect_src(g, O.make_advect_bc(spec), dt, coeff, W, uf, S=None, src=src0)
    dsrc2 = ml.DeviceArray(gpu_ctx, src0.shape, host=src0)
    op.advect_src(ap, dt, dW, duf, None, dsrc2)
    assert HC.bitwise_equal(O.valid(g, dsrc2.get()), O.valid(g, ref2))
    op.close()


@pytest.mark.parametrize("n,periodic", CASES)
def test_diffusive_src_and_face_velocity_bitwise(ml, gpu_ctx, n, periodic):
    g = _grid(n, periodic)
    dim = len(n)
    p = O.default_params(stefan=4.0, compositionRatio=3.0, lewis=100.0)
    bcspec = {}
    for d in range(dim):
        if not periodic[d]:
            bcspec[(d, 0, 0)] = (0, -0.3)
            bcspec[(d, 1, 0)] = (0, 0.9)
    bc = O.make_bc(bcspec)
    phi = HC.random_state(list(n), seed=5 + sum(n))
    chi = np.ascontiguousarray(O.enthalpy_vars(g, p, phi)[3:4])
    b = O.hc_bcoef(g, p, chi)
    ref = O.hc_apply_op(g, p, bc, 0.0, None, 1.0, b, phi.copy())
    op = ml.HCOp(gpu_ctx, list(n), g.dx, list(periodic),
                 ml.make_params(stefan=4.0, compositionRatio=3.0, lewis=100.0), ml.make_bc(bcspec))
    db = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in b]
    op.set_coefs(None, db, 1.0, -1e-3)
    dphi = ml.DeviceArray(gpu_ctx, phi.shape, host=phi)
    out = ml.DeviceArray(gpu_ctx, phi.shape)
    op.diffusive_src(dphi, out)
    assert HC.bitwise_equal(O.valid(g, out.get()), O.valid(g, ref))
    op.close()
    rng = np.random.default_rng(9)
    u = []
    for _ in range(dim):
        x = np.zeros(O.cell_shape(g, 1))
        O.valid(g, x)[...] = rng.standard_normal(O.valid(g, x).shape)
        u.append(x)
    rf = O.face_velocity(g, u)
    vop = ml.VCOp(gpu_ctx, list(n), g.dx, list(periodic), ml.make_bc())
    du = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in u]
    df = [ml.DeviceArray(gpu_ctx, O.face_shape(g, d, 1)) for d in range(dim)]
    vop.face_velocity(du, df)
    for d in range(dim):
        got = df[d].get()
        assert HC.bitwise_equal(got, rf[d]), d
        if not periodic[d]:
            ax = got.ndim - 1 - d
            assert np.all(np.take(got, 0, axis=ax) == 0.0) and np.all(np.take(got, -1, axis=ax) == 0.0)
    vop.close()


def test_advect_src_rejects_plume(ml, gpu_ctx):
    n = [16, 16]
    op = ml.HCOp(gpu_ctx, n, 1.0 / 16, [1, 0], ml.make_params(), ml.make_bc())
    x = ml.DeviceArray(gpu_ctx, (2, 1, 18, 18))
    uf = [ml.DeviceArray(gpu_ctx, (1, 1, 16, 17)), ml.DeviceArray(gpu_ctx, (1, 1, 17, 16))]
    ap = ml.make_advect_params(None, 1.0, {(1, 0, 0): (ml.ADV_PLUME_INFLOW, 0.0)})
    with pytest.raises(ml.MLError, match="unsupported advection boundary"):
        op.advect_src(ap, 1e-3, x, uf, None, x)
    op.close()
