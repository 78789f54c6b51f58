"""Batched form of the FP32-refined (H, Theta) step on the GPU: independent
instances of different grid sizes, parameters and BCs share every launch.

Each instance's step must equal, bit for bit (states and solve statistics),
the single-instance sequence porosity -> bCoef -> set_coefs -> solve_be
(FP32-refined) on its own operator, which the FP32-refined tests pin against
the FP64 oracle; the smallest instances are also checked against the
oracle's FP64 backward-Euler step directly, to the FP32-refined form's
tolerances (TOL_STATE relative on H and Theta, TOL_CHI absolute on the
porosity).  The instance set covers every alignment of the tiled levels
(single-workgroup level lc = 0 .. 3), periodic and walled sides, and two
consecutive steps (the second reuses the uploaded tables and the captured
graph)."""
import numpy as np
import pytest

from oracle import mlref as O
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu

TOL_STATE = 1e-9
TOL_CHI = 1e-9
EPS = 1e-10


def _instances(wl):
    spec = [((32, 32), 2.0, 3.0, None), ((64, 64), 5.0, 2.0, None), ((128, 64), 7.5, 12.0, [0, 0]),
            ((64, 64), 1.5, 18.0, [0, 0]), ((128, 128), 9.0, 4.0, None), ((256, 128), 3.0, 8.0, None)]
    cases = []
    for i, (n, St, CR, per) in enumerate(spec):
        c = wl.mushy(n=n, seed=40 + i, St=St, CR=CR)
        if per is not None:
            c["periodic"] = per
        c["apply_bc"] = 1
        cases.append(c)
    return cases


def _single_step(ml, ctx, case, phi, opts):
    """The single-instance step the batch must reproduce."""
    op = HC.gpu_setup(ml, ctx, case)
    n = case["n"]
    shp = phi.shape
    chi = ml.DeviceArray(ctx, (1,) + shp[1:])
    b = [ml.DeviceArray(ctx, ml.face_shape(n, d, 2)) for d in range(len(n))]
    old = ml.DeviceArray(ctx, shp)
    op.porosity(phi, chi)
    op.assemble_bcoef(chi, b)
    op.set_coefs(None, b, 1.0, -case["dt"])
    ctx.check(ctx.lib.ml_buf_copy(ctx.h, old.ptr, phi.ptr, phi.nbytes), "copy")
    st = op.solve_be(phi, old, case["dt"], opts=opts)
    op.close()
    return st


@pytest.mark.parametrize("flags", ["graph", "no_graph_no_lookahead"])
def test_batch_step_matches_single_instance_steps(ml, gpu_ctx, wl, flags):
    cases = _instances(wl)
    dt = cases[0]["dt"]
    fl = 0 if flags == "graph" else (ml.SOLVE_NO_GRAPH | ml.SOLVE_NO_LOOKAHEAD)
    opts = ml.default_opts(max_iter=100, eps=EPS, hang=-1.0, precision=ml.PRECISION_FP32_REFINED, flags=fl)
    ops = [HC.gpu_setup(ml, gpu_ctx, c) for c in cases]
    batch = ml.HCBatch(gpu_ctx, ops)
    bphi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"]) for c in cases]
    sphi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"]) for c in cases]
    for step in range(2):
        print("batch step %d, %d instances, flags %s" % (step, len(cases), flags), flush=True)
        bst = batch.step(bphi, dt, opts)
        for i, c in enumerate(cases):
            sst = _single_step(ml, gpu_ctx, c, sphi[i], opts)
            a, r = bphi[i].get(), sphi[i].get()
            assert HC.bitwise_equal(a, r), (step, i, c["n"], float(np.max(np.abs(a - r))))
            assert (bst[i].exit_status, bst[i].iterations) == (sst.exit_status, sst.iterations), (step, i)
            assert bst[i].final_residual == sst.final_residual and bst[i].initial_residual == sst.initial_residual
            assert bst[i].exit_status & (1 | 8), (step, i, bst[i].exit_status)
    batch.close()
    for op in ops:
        op.close()


def test_batch_tile_shapes_give_identical_steps(ml, gpu_ctx, wl):
    """Every tile shape of the batched relaxation gives the single-instance
    steps bit for bit; the autotuner times them all and keeps the fastest."""
    cases = _instances(wl)
    dt = cases[0]["dt"]
    opts = ml.default_opts(max_iter=100, eps=EPS, hang=-1.0, precision=ml.PRECISION_FP32_REFINED)
    ops = [HC.gpu_setup(ml, gpu_ctx, c) for c in cases]
    batch = ml.HCBatch(gpu_ctx, ops)
    ref = []
    for c in cases:
        x = ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"])
        _single_step(ml, gpu_ctx, c, x, opts)
        ref.append(x.get())
    nshapes = len(batch.tile()[2])
    assert nshapes >= 2
    for v in list(range(nshapes)) + [-1]:
        batch.set_tile(v)
        bphi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"]) for c in cases]
        print("tile shape %d" % v, flush=True)
        batch.step(bphi, dt, opts)
        for i in range(len(cases)):
            assert HC.bitwise_equal(bphi[i].get(), ref[i]), (v, i)
    chosen, shape, ms = batch.tile()
    print("autotuned shape %d %s, ms per shape %s" % (chosen, shape, ms))
    assert all(t > 0 for t in ms) and ms[chosen] == min(ms)
    batch.close()
    for op in ops:
        op.close()


def test_batch_compacted_stragglers_match_single_instance_steps(ml, gpu_ctx, wl):
    """15 small instances converge well before one 512x256 instance with a
    large concentration ratio; once at most 1/8 of the batch runs, later
    iterations launch only the stragglers' tiles.  Results stay those of the
    single-instance steps, bit for bit."""
    big = wl.batch_instances(count=54, seed=5)[53]
    assert big["n"] == [512, 256]
    cases = [wl.mushy(n=(32, 32), seed=60 + i, St=2.0 + 0.5 * i, CR=1.0 + i) for i in range(15)] + [big]
    for c in cases:
        c["apply_bc"] = 1
    dt = cases[0]["dt"]
    opts = ml.default_opts(max_iter=100, eps=EPS, precision=ml.PRECISION_FP32_REFINED, flags=ml.SOLVE_NO_LOOKAHEAD)
    ops = [HC.gpu_setup(ml, gpu_ctx, c) for c in cases]
    batch = ml.HCBatch(gpu_ctx, ops)
    bphi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"]) for c in cases]
    sphi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"]) for c in cases]
    for step in range(3):
        print("compaction batch step %d" % step, flush=True)
        bst = batch.step(bphi, dt, opts)
        its = []
        for i, c in enumerate(cases):
            sst = _single_step(ml, gpu_ctx, c, sphi[i], opts)
            assert HC.bitwise_equal(bphi[i].get(), sphi[i].get()), (step, i)
            assert (bst[i].exit_status, bst[i].iterations) == (sst.exit_status, sst.iterations), (step, i)
            its.append(sst.iterations)
        print("  iterations small max %d, large %d" % (max(its[:-1]), its[-1]))
    batch.close()
    for op in ops:
        op.close()


def test_batch_step_matches_fp64_oracle(ml, gpu_ctx, wl):
    cases = _instances(wl)[:4]
    dt = cases[0]["dt"]
    opts = ml.default_opts(max_iter=100, eps=EPS, hang=-1.0, precision=ml.PRECISION_FP32_REFINED)
    ops = [HC.gpu_setup(ml, gpu_ctx, c) for c in cases]
    batch = ml.HCBatch(gpu_ctx, ops)
    bphi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"]) for c in cases]
    batch.step(bphi, dt, opts)
    for i, c in enumerate(cases):
        n = c["n"]
        g, p, bc = HC.oracle_setup(c)
        b = HC.bcoef_from_state(c)
        old = c["phi"].copy()
        ref = old.copy()
        st = O.hc_be_step(g, p, bc, None, b, dt, ref, old, opts=O.default_opts(max_iter=100, eps=EPS, hang=-1.0))
        assert st.exit_status & (1 | 8)
        out = bphi[i].get()
        eH = HC.rel_err(HC.valid(n, out)[0], HC.valid(n, ref)[0])
        eC = HC.rel_err(HC.valid(n, out)[1], HC.valid(n, ref)[1])
        chi_o = O.porosity_field(g, p, out)
        chi_r = O.porosity_field(g, p, ref)
        eX = float(np.max(np.abs(HC.valid(n, chi_o) - HC.valid(n, chi_r))))
        print("instance %d %s: err H %.3e Theta %.3e chi %.3e" % (i, n, eH, eC, eX))
        assert eH <= TOL_STATE and eC <= TOL_STATE and eX <= TOL_CHI, (i, eH, eC, eX)
    batch.close()
    for op in ops:
        op.close()


def test_batch_rejects_unsupported(ml, gpu_ctx, wl):
    c3 = wl.diffusive(n=(16, 16, 16))
    op3 = HC.gpu_setup(ml, gpu_ctx, c3)
    with pytest.raises(ml.MLError):
        ml.HCBatch(gpu_ctx, [op3])
    op3.close()
    c = wl.mushy(n=(64, 64), seed=3)
    op = HC.gpu_setup(ml, gpu_ctx, c)
    batch = ml.HCBatch(gpu_ctx, [op])
    phi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"])]
    with pytest.raises(ml.MLError):
        batch.step(phi, c["dt"], ml.default_opts(precision=ml.PRECISION_FP64))
    with pytest.raises(ml.MLError):
        batch.step(phi, c["dt"], ml.default_opts(precision=ml.PRECISION_FP32_REFINED, nu2=5))
    batch.close()
    op.close()
