"""ml_hc_relax_noalloc — the (H, Theta) relaxation with its partner array in
the library's static device scratch (no allocation) — against ml_hc_relax
through the C ABI: bit-identical states on two calls in a row, the second on
a smaller level (the scratch keeps the first call's contents), and a clear
error for a level larger than the scratch."""
import numpy as np
import pytest

from tests import hc_cases as HC

pytestmark = pytest.mark.gpu


def _relaxed(ml, ctx, op, case, depth, iters, noalloc):
    """Level `depth` of a mushy state relaxed `iters` times (host array)."""
    n, _ = op.level_n(depth)
    shp = ml.cell_shape(n, 2)
    rng = np.random.default_rng(sum(n) + depth)
    phi0 = np.zeros(shp)
    v = ml.valid(n, phi0)
    v[0] = rng.uniform(-0.5, 7.5, v[0].shape)
    v[1] = rng.uniform(-1.9, -0.2, v[1].shape)
    rhs0 = phi0 + 1e-3 * rng.standard_normal(shp)
    phi = ml.DeviceArray(ctx, shp, host=phi0)
    rhs = ml.DeviceArray(ctx, shp, host=rhs0)
    if noalloc:
        op.relax_noalloc(depth, phi, rhs, iters)
    else:
        op.relax(depth, phi, rhs, iters)
    out = phi.get()
    phi.free()
    rhs.free()
    return out


def _op_with_coefs(ml, ctx, case):
    op = HC.gpu_setup(ml, ctx, case)
    n = case["n"]
    chi = ml.DeviceArray(ctx, (1,) + case["phi"].shape[1:])
    b = [ml.DeviceArray(ctx, ml.face_shape(n, d, 2)) for d in range(len(n))]
    phi = ml.DeviceArray(ctx, case["phi"].shape, host=case["phi"])
    op.porosity(phi, chi)
    op.assemble_bcoef(chi, b)
    op.set_coefs(None, b, 1.0, -case["dt"])
    return op


def test_noalloc_relax_matches_relax(ml, gpu_ctx, wl):
    big = wl.mushy(n=(256, 256), seed=21)
    small = wl.mushy(n=(64, 32), seed=22, St=3.0, CR=6.0)
    small["periodic"] = [0, 0]
    ops = [_op_with_coefs(ml, gpu_ctx, c) for c in (big, small)]
    # two calls in a row, the second with a smaller input; then a coarse
    # level of the first operator, and a 3-iteration call
    for k, (op, case, depth, iters) in enumerate([(ops[0], big, 0, 4), (ops[1], small, 0, 1), (ops[0], big, 2, 3)]):
        print("noalloc call %d: n %s depth %d iters %d" % (k, op.level_n(depth)[0], depth, iters), flush=True)
        a = _relaxed(ml, gpu_ctx, op, case, depth, iters, noalloc=True)
        r = _relaxed(ml, gpu_ctx, op, case, depth, iters, noalloc=False)
        assert HC.bitwise_equal(a, r), (k, float(np.max(np.abs(a - r))))
        assert np.max(np.abs(a)) > 0
    for op in ops:
        op.close()


def test_noalloc_rejects_levels_beyond_scratch(ml, gpu_ctx, wl):
    case = wl.diffusive(n=(1024, 1024))
    op = HC.gpu_setup(ml, gpu_ctx, case)
    x = ml.DeviceArray(gpu_ctx, case["phi"].shape, host=case["phi"])
    with pytest.raises(ml.MLError, match="exceeds the static scratch"):
        op.relax_noalloc(0, x, x, 1)
    # a coarse level of the same operator fits
    n1, _ = op.level_n(1)
    assert 2 * 8 * (n1[0] + 2) * (n1[1] + 2) <= 8 << 20
    x.free()
    op.close()
