"""Checkpoint and restart of a Darcy-Brinkman run through the reference's
Chombo HDF5 layout (libmlio.so; writeCheckpointLevel / readCheckpointLevel,
srcSubcycle/AMRLevelMushyLayerIO.cpp:184-420): three steps, a checkpoint,
a fresh stepper restored from it, and two more steps on both -- the
restarted run equals the uninterrupted one bit for bit (H, Theta, velocity,
pressure), with the same solver iterations."""
import importlib

import numpy as np
import pytest

from oracle import steps as S
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu


def test_checkpoint_restart_is_exact(ml, gpu_ctx, wl, tmp_path):
    cio = importlib.import_module("mushy-layer_amd.chombo_io")
    try:
        cio.lib()
    except (OSError, cio.H5Error) as e:
        pytest.skip("libmlio.so not loadable here: %s" % e)
    stp = importlib.import_module("mushy-layer_amd.steps")
    case = wl.mushy(n=(64, 64), seed=21)
    flow = wl.flow(dim=2, Pr=10.0, Da=1e-3, Ra_C=200.0)
    opts = ml.default_opts(max_iter=100, hang=S.NO_HANG)
    a = stp.DarcyBrinkmanStepper(ml, gpu_ctx, case, opts, flow)
    for _ in range(3):
        a.step()
    path = str(tmp_path / "chk000003.2d.hdf5")
    a.save_checkpoint(path, 3, 3 * case["dt"])
    b = stp.DarcyBrinkmanStepper(ml, gpu_ctx, case, ml.default_opts(max_iter=100, hang=S.NO_HANG), flow)
    it, t = b.load_checkpoint(path)
    assert it == 3 and t == 3 * case["dt"]
    for k in range(2):
        a.step()
        b.step()
        gpu_ctx.sync()
        print("after restart step %d: HC %s / %s, velocity %s / %s" % (k, a.stats[-1][:2], b.stats[-1][:2],
                                                                      a.vstats[-1], b.vstats[-1]), flush=True)
        assert a.stats[-1][:2] == b.stats[-1][:2] and a.vstats[-1] == b.vstats[-1]
    n = case["n"]
    assert HC.bitwise_equal(HC.valid(n, a.phi.get()), HC.valid(n, b.phi.get()))
    fa, fb = a.flow_fields(), b.flow_fields()
    for key in fa:
        assert HC.bitwise_equal(fa[key], fb[key]), key
    assert np.abs(fa["u1"]).max() > 0
    a.close()
    b.close()
