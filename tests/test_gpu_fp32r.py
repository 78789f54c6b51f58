"""FP32-refined form of the (H, Theta) solve on the GPU: FP32 smoothing and
transfers inside an FP64 defect-correction loop (FP64 outer residual, FP64
closure regime decision, FP64 coarsest-level solve).

The converged state is compared with the FP64 oracle's converged
backward-Euler step (FAS V-cycles to the same 1e-10 relative residual) field
by field: H and Theta (the solved state) and the porosity derived from it.
Both solutions carry a residual of at most 1e-10 * |r0|, so they may differ
by that much through the operator's inverse; the bar is written as
TOL_STATE (relative max-norm on H and Theta) and TOL_CHI (absolute on the
porosity, which lies in [0, 1]).
"""
import numpy as np
import pytest

from oracle import mlref as O
from tests import fp32r_model as FM
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu

TOL_STATE = 1e-9
TOL_CHI = 1e-9
EPS = 1e-10


def _diffusive(wl, n, periodic=None, seed=3):
    c = wl.diffusive(n=n, seed=seed)
    if periodic is not None:
        c["periodic"] = periodic
    c["apply_bc"] = 1
    return c


def _mixed(case, seed):
    n = case["n"]
    rng = np.random.default_rng(seed)
    phi = case["phi"].copy()
    v = HC.valid(n, phi)
    z = (np.arange(n[-1]) + 0.5) / n[-1]
    zz = z.reshape((-1,) + (1,) * (len(n) - 1)) if len(n) == 3 else z.reshape(-1, 1)
    v[0] = -1.0 + 8.0 * zz + 0.3 * rng.uniform(-1, 1, size=v[0].shape)
    v[1] = -1.0 + 0.2 * rng.uniform(-1, 1, size=v[1].shape)
    case["phi"] = phi
    return case


CASES = {
    "diffusive_64": lambda wl: _diffusive(wl, (64, 64)),
    "diffusive_128": lambda wl: _diffusive(wl, (128, 128)),
    "mixed_64x32_walls": lambda wl: _mixed(_diffusive(wl, (64, 32), periodic=[0, 0]), 15),
    "mixed_48x32_walls": lambda wl: _mixed(_diffusive(wl, (48, 32), periodic=[0, 0]), 12),
    "mushy_128": lambda wl: wl.mushy(n=(128, 128)),
    "mushy_256": lambda wl: wl.mushy(n=(256, 256)),
    "mixed_3d_16": lambda wl: _mixed(_diffusive(wl, (16, 16, 16)), 13),
    "mixed_3d_32x32x16": lambda wl: _mixed(_diffusive(wl, (32, 32, 16), periodic=[1, 0, 0]), 14),
}


def _field_errors(case, p, g, out, ref):
    n = case["n"]
    eH = HC.rel_err(HC.valid(n, out)[0], HC.valid(n, ref)[0])
    eC = HC.rel_err(HC.valid(n, out)[1], HC.valid(n, ref)[1])
    chi_o = O.porosity_field(g, p, out)
    chi_r = O.porosity_field(g, p, ref)
    eX = float(np.max(np.abs(HC.valid(n, chi_o) - HC.valid(n, chi_r))))
    return eH, eC, eX


@pytest.mark.parametrize("name", sorted(CASES))
def test_refined_be_step_matches_fp64_oracle(ml, gpu_ctx, wl, name):
    case = CASES[name](wl)
    case.setdefault("apply_bc", 1)
    n = case["n"]
    g, p, bc = HC.oracle_setup(case)
    b = HC.bcoef_from_state(case)
    dt = case["dt"]
    old = case["phi"].copy()
    ref = old.copy()
    st_ref = O.hc_be_step(g, p, bc, None, b, dt, ref, old, opts=O.default_opts(max_iter=100, eps=EPS, hang=-1.0))
    assert st_ref.exit_status & (1 | 8), st_ref.exit_status

    op = HC.gpu_setup(ml, gpu_ctx, case)
    db = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in b]
    op.set_coefs(None, db, 1.0, -dt)
    dold = ml.DeviceArray(gpu_ctx, old.shape, host=old)
    dnew = ml.DeviceArray(gpu_ctx, old.shape, host=old)
    opts = ml.default_opts(max_iter=100, eps=EPS, hang=-1.0, precision=ml.PRECISION_FP32_REFINED)
    st = op.solve_be(dnew, dold, dt, opts=opts)
    out = dnew.get()
    assert st.exit_status & (1 | 8), (st.exit_status, st.iterations, st.final_residual)
    assert st.final_residual <= max(EPS * st.initial_residual, opts.norm_thresh)
    # the FP64 residual of the initial state is the oracle's, to rounding
    assert abs(st.initial_residual - st_ref.initial_residual) <= 1e-12 * st_ref.initial_residual
    eH, eC, eX = _field_errors(case, p, g, out, ref)
    # the same algorithm evaluated in FP64 on the host takes (about) as many
    # outer iterations: FP32 smoothing does not slow the defect correction
    model = old.copy()
    _, it_model, _ = FM.solve(g, p, bc, 1.0, -dt, b, model, old.copy(), max_iter=100, eps=EPS, hang=-1.0)
    print("%s: outer iterations %d (FP64 model %d, FAS V-cycles %d), err H %.3e Theta %.3e chi %.3e"
          % (name, st.iterations, it_model, st_ref.iterations, eH, eC, eX))
    assert eH <= TOL_STATE and eC <= TOL_STATE and eX <= TOL_CHI, (eH, eC, eX)
    assert st.iterations <= it_model + 2, (st.iterations, it_model)


def test_refined_uses_fewer_outer_iterations(ml, gpu_ctx, wl):
    """The Newton-type outer loop converges in fewer iterations than the
    FAS V-cycles of the plain FP64 solve on the mushy 256x256 step."""
    case = wl.mushy(n=(256, 256))
    case.setdefault("apply_bc", 1)
    b = HC.bcoef_from_state(case)
    dt = case["dt"]
    old = case["phi"]
    op = HC.gpu_setup(ml, gpu_ctx, case)
    db = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in b]
    op.set_coefs(None, db, 1.0, -dt)
    dold = ml.DeviceArray(gpu_ctx, old.shape, host=old)
    its = {}
    for prec in (ml.PRECISION_FP64, ml.PRECISION_FP32_REFINED):
        dnew = ml.DeviceArray(gpu_ctx, old.shape, host=old)
        st = op.solve_be(dnew, dold, dt, opts=ml.default_opts(max_iter=100, hang=-1.0, precision=prec))
        assert st.exit_status & (1 | 8)
        its[prec] = st.iterations
    assert its[ml.PRECISION_FP32_REFINED] < its[ml.PRECISION_FP64], its


def test_refined_rejects_acoef_and_host_driver(ml, gpu_ctx, wl):
    case = _diffusive(wl, (32, 32))
    op = HC.gpu_setup(ml, gpu_ctx, case)
    old = case["phi"]
    dold = ml.DeviceArray(gpu_ctx, old.shape, host=old)
    dnew = ml.DeviceArray(gpu_ctx, old.shape, host=old)
    opts = ml.default_opts(precision=ml.PRECISION_FP32_REFINED, flags=ml.SOLVE_HOST_DRIVER)
    with pytest.raises(ml.MLError):
        op.solve_be(dnew, dold, case["dt"], opts=opts)
