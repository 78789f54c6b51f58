"""Batched Darcy-Brinkman steps (steps.BatchDarcyBrinkmanStepper): drawn
mushy-layer instances of uneven sizes, each with its own St, CR, Ra_C and Da,
step together -- the (H, Theta) solves with their advective sources in one
HCBatch (ml_hc_batch_step_src), the U* and projection solves in VCBatches.
Every instance's state, velocity, pressure and solve statistics must equal
its own DarcyBrinkmanStepper (FP32-refined (H, Theta) solve) bit for bit over
two steps, and the drawn Ra_C and Da must change the flow."""
import importlib

import numpy as np
import pytest

from tests import hc_cases as HC

pytestmark = pytest.mark.gpu

wl = importlib.import_module("mushy-layer_amd.workloads")
stp = importlib.import_module("mushy-layer_amd.steps")


def _flow(c):
    return wl.flow(dim=2, Pr=10.0, Da=c["Da"], Ra_C=c["Ra_C"])


def test_batched_darcy_brinkman_steps_equal_single_instance_steps(ml, gpu_ctx):
    ctx = gpu_ctx
    cases = wl.batch_instances(count=6, seed=4)   # 512x256, 128^2 and 64^2 instances
    assert {tuple(c["n"]) for c in cases} == {(64, 64), (128, 128), (512, 256)}
    hc_opts = ml.default_opts(max_iter=60, precision=ml.PRECISION_FP32_REFINED)
    batch = stp.BatchDarcyBrinkmanStepper(ml, ctx, cases, _flow, hc_opts)
    singles = [stp.DarcyBrinkmanStepper(ml, ctx, c, ml.default_opts(max_iter=60, precision=ml.PRECISION_FP32_REFINED),
                                        _flow(c)) for c in cases]
    for _ in range(2):
        out = batch.step()
        for s in singles:
            s.step()
        for i, (s, b) in enumerate(zip(singles, batch.st)):
            assert s.stats[-1] == b.stats[-1], (i, s.stats[-1], b.stats[-1])
            assert s.vstats[-1] == b.vstats[-1], (i, s.vstats[-1], b.vstats[-1])
            assert HC.bitwise_equal(s.phi.get(), b.phi.get()), i
            for d in range(2):
                assert HC.bitwise_equal(s.u[d].get(), b.u[d].get()), (i, d)
            assert HC.bitwise_equal(s.pi.get(), b.pi.get()), i
        assert all(o[0].exit_status & (1 | 8) for o in out)
    # the drawn parameters act: the instances' velocities differ in size
    umax = [max(np.abs(HC.valid(c["n"], b.u[d].get())).max() for d in range(2)) for c, b in zip(cases, batch.st)]
    assert max(umax) > 0 and len({round(float(v), 12) for v in umax}) == len(umax)
    batch.close()
    for s in singles:
        s.close()
