"""The workload configurations at their specified sizes, on the GPU through
the C ABI:

* 64x64 diffusive solidification: its own initial state, 10 plain-FP64
  backward-Euler steps under the reference's solver options (V(1,4), relax
  bottom 40, at most 10 V-cycles, eps = hang = norm_thresh = 1e-10); every
  step's state bit-identical to the oracle's, with the same iterations and
  exit status;
* 3D mushy layer (correlation length 0.08): FP64 (H, Theta) steps at 32^3
  bit-identical to the oracle; at 128^3, one step in each form converging
  (status 1 or 8) with the FP32-refined state within 1e-9 of the FP64 one;
* the batched form on 48 instances of the seeded 512-instance draw (64^2,
  128^2, 256^2 and 512x256 grids, St, CR, Ra_C, Da drawn per instance),
  each instance bit-identical to its single-instance step, set up with one
  device allocation and one fill;
* the adjoint at 128x128 over 50 steps (dt 5e-4): one directional
  derivative of the GPU adjoint against central differences of the
  oracle's FP64 forward model (16 OpenMP threads for the oracle)."""
import importlib
import os

import numpy as np
import pytest

from oracle import mlref as O
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu


def _gpu_step(ml, ctx, op, phi, case, opts, chi, b, old):
    op.porosity(phi, chi)
    op.assemble_bcoef(chi, b)
    op.set_coefs(None, b, 1.0, -case["dt"])
    ctx.check(ctx.lib.ml_buf_copy(ctx.h, old.ptr, phi.ptr, phi.nbytes), "copy")
    return op.solve_be(phi, old, case["dt"], opts=opts)


def _work(ml, ctx, case):
    n = case["n"]
    shp = case["phi"].shape
    return (ml.DeviceArray(ctx, (1,) + shp[1:]), [ml.DeviceArray(ctx, ml.face_shape(n, d, 2)) for d in range(len(n))],
            ml.DeviceArray(ctx, shp))


def test_diffusive_64_ten_steps_reference_options(ml, gpu_ctx, wl, oracle):
    case = wl.diffusive(n=(64, 64))
    assert case["steps"] == 10 and case["dt"] == 1e-3
    g, p, bc = HC.oracle_setup(case)
    op = HC.gpu_setup(ml, gpu_ctx, case)
    phi = ml.DeviceArray(gpu_ctx, case["phi"].shape, host=case["phi"])
    chi, b, old = _work(ml, gpu_ctx, case)
    ref = case["phi"].copy()
    gopts = ml.default_opts()
    oopts = O.default_opts()
    assert (gopts.nu1, gopts.nu2, gopts.bottom, gopts.max_iter) == (1, 4, 40, 10)
    for step in range(case["steps"]):
        st = _gpu_step(ml, gpu_ctx, op, phi, case, gopts, chi, b, old)
        bo = O.hc_bcoef(g, p, O.porosity_field(g, p, ref))
        so = O.hc_be_step(g, p, bc, None, bo, case["dt"], ref, ref.copy(), opts=oopts)
        out = phi.get()
        print("step %d: GPU status %d iters %d resid %.3e | oracle status %d iters %d"
              % (step, st.exit_status, st.iterations, st.final_residual, so.exit_status, so.iterations), flush=True)
        assert (st.exit_status, st.iterations) == (so.exit_status, so.iterations), step
        assert st.final_residual == so.final_residual, step
        assert HC.bitwise_equal(HC.valid(case["n"], out), HC.valid(case["n"], ref)), step
    op.close()


def test_mushy_3d_fp64_steps_match_oracle(ml, gpu_ctx, wl, oracle):
    case = wl.mushy(n=(32, 32, 32), corr=0.08)
    g, p, bc = HC.oracle_setup(case)
    op = HC.gpu_setup(ml, gpu_ctx, case)
    phi = ml.DeviceArray(gpu_ctx, case["phi"].shape, host=case["phi"])
    chi, b, old = _work(ml, gpu_ctx, case)
    ref = case["phi"].copy()
    opts = ml.default_opts(max_iter=100, hang=-1.0)
    for step in range(3):
        st = _gpu_step(ml, gpu_ctx, op, phi, case, opts, chi, b, old)
        bo = O.hc_bcoef(g, p, O.porosity_field(g, p, ref))
        so = O.hc_be_step(g, p, bc, None, bo, case["dt"], ref, ref.copy(),
                          opts=O.default_opts(max_iter=100, hang=-1.0))
        print("3D step %d: GPU status %d iters %d | oracle status %d iters %d"
              % (step, st.exit_status, st.iterations, so.exit_status, so.iterations), flush=True)
        assert st.exit_status & (1 | 8)
        assert (st.exit_status, st.iterations) == (so.exit_status, so.iterations), step
        assert HC.bitwise_equal(HC.valid(case["n"], phi.get()), HC.valid(case["n"], ref)), step
    op.close()


def test_mushy_128_cubed_fp32_refined_matches_fp64(ml, gpu_ctx, wl):
    case = wl.mushy(n=(128, 128, 128), corr=0.08)
    n = case["n"]
    out = {}
    for prec in (ml.PRECISION_FP64, ml.PRECISION_FP32_REFINED):
        op = HC.gpu_setup(ml, gpu_ctx, case)
        phi = ml.DeviceArray(gpu_ctx, case["phi"].shape, host=case["phi"])
        chi, b, old = _work(ml, gpu_ctx, case)
        st = _gpu_step(ml, gpu_ctx, op, phi, case, ml.default_opts(max_iter=100, hang=-1.0, precision=prec),
                       chi, b, old)
        op.porosity(phi, chi)
        print("128^3 precision %d: status %d iterations %d resid %.3e"
              % (prec, st.exit_status, st.iterations, st.final_residual), flush=True)
        assert st.exit_status & (1 | 8)
        out[prec] = (HC.valid(n, phi.get()), HC.valid(n, chi.get()))
        for x in (phi, chi, old, *b):
            x.free()
        op.close()
    a, r = out[ml.PRECISION_FP32_REFINED], out[ml.PRECISION_FP64]
    eH, eC = HC.rel_err(a[0][0], r[0][0]), HC.rel_err(a[0][1], r[0][1])
    eX = float(np.max(np.abs(a[1] - r[1])))
    print("FP32-refined vs FP64 at 128^3: H %.3e Theta %.3e porosity %.3e" % (eH, eC, eX))
    assert eH <= 1e-9 and eC <= 1e-9 and eX <= 1e-9


def test_batch_of_drawn_instances(ml, gpu_ctx, wl):
    draw = wl.batch_instances(count=512, seed=5)
    cases = draw[:47] + [draw[53]]
    sizes = sorted(set(tuple(c["n"]) for c in cases))
    assert (512, 256) in sizes and (64, 64) in sizes and (256, 256) in sizes
    for c in cases:
        c["apply_bc"] = 1
    dt = cases[0]["dt"]
    opts = ml.default_opts(max_iter=100, precision=ml.PRECISION_FP32_REFINED)
    ops = [HC.gpu_setup(ml, gpu_ctx, c) for c in cases]
    batch = ml.HCBatch(gpu_ctx, ops)
    nbytes, nalloc, nfill = batch.memory()
    print("batch of %d instances %s: %d device bytes, %d allocation(s), %d fill(s)"
          % (len(cases), sizes, nbytes, nalloc, nfill), flush=True)
    assert nalloc == 1 and nfill == 1 and nbytes > 0
    bphi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"]) for c in cases]
    sphi = [ml.DeviceArray(gpu_ctx, c["phi"].shape, host=c["phi"]) for c in cases]
    work = [_work(ml, gpu_ctx, c) for c in cases]
    for step in range(2):
        bst = batch.step(bphi, dt, opts)
        for i, c in enumerate(cases):
            sst = _gpu_step(ml, gpu_ctx, ops[i], sphi[i], c, opts, *work[i])
            assert HC.bitwise_equal(bphi[i].get(), sphi[i].get()), (step, i, c["n"])
            assert (bst[i].exit_status, bst[i].iterations) == (sst.exit_status, sst.iterations), (step, i)
        print("step %d: iterations mean %.2f max %d, statuses %s" % (
            step, np.mean([s.iterations for s in bst]), max(s.iterations for s in bst),
            sorted(set(s.exit_status for s in bst))), flush=True)
    batch.close()
    for op in ops:
        op.close()


def test_adjoint_128_fifty_steps_against_oracle(ml, gpu_ctx, wl, oracle):
    torch = pytest.importorskip("torch")
    adj = importlib.import_module("mushy-layer_amd.adjoint")
    from tests.test_gpu_adjoint import _adjoint_grads, _loss_np, _oracle_forward
    case = wl.heat_flux(n=(128, 128), steps=50, dt=5e-4)
    model = adj.HeatFluxAdjoint(gpu_ctx, case, checkpoint=10)
    F_true = wl.flux_profile(128, F0=12.0, modes=((1, 0.3, 0.0), (2, 0.1, 1.0)))
    tphi, tchi, cps = model.forward(F_true, 5.5)
    cps.close()
    F = wl.flux_profile(128, F0=10.0)
    St = 5.0
    L, gF, gperm, gSt = _adjoint_grads(torch, adj, model, F, St, tphi, tchi)
    assert gperm == 0.0 and L > 0
    nthr = O.set_threads(min(16, os.cpu_count() or 1))
    try:
        v = np.cos(2 * np.pi * (np.arange(128) + 0.5) / 128) + 0.5 * np.sin(6 * np.pi * (np.arange(128) + 0.5) / 128)
        v /= np.linalg.norm(v)
        h = 1e-6 * np.max(np.abs(F))
        lp = _loss_np(case["n"], *_oracle_forward(case, F + h * v, St), tphi, tchi)
        lm = _loss_np(case["n"], *_oracle_forward(case, F - h * v, St), tphi, tchi)
    finally:
        O.set_threads(1)
    fd = (lp - lm) / (2 * h)
    ad = float(gF @ v)
    print("128^2 x 50 steps, oracle on %d threads: adjoint %.9e, central difference %.9e, rel %.2e"
          % (nthr, ad, fd, abs(ad - fd) / abs(fd)))
    assert abs(ad - fd) <= 2e-5 * abs(fd)
    model.close()
