"""Seeded workload generators (host only)."""
import numpy as np


def test_batch_instances_draw(wl):
    xs = wl.batch_instances(count=48, seed=7, sizes=((16, 16), (32, 32), (32, 16), (64, 32)))
    assert len(xs) == 48
    sizes = [tuple(x["n"]) for x in xs]
    # ratio 8:4:2:1 of the sizes, in expectation
    assert sizes.count((16, 16)) > sizes.count((64, 32))
    for x in xs:
        p = x["params"]
        assert 1.0 <= p["stefan"] <= 10.0 and 1.0 <= p["compositionRatio"] <= 20.0
        assert 0.0 <= x["Ra_C"] <= 400.0 and 1e-4 <= x["Da"] <= 1e-2
        chi = x["chi"]
        assert chi.min() >= 0.05 - 1e-12 and chi.max() <= 1.0 + 1e-12
    # the same seed gives the same instances
    ys = wl.batch_instances(count=48, seed=7, sizes=((16, 16), (32, 32), (32, 16), (64, 32)))
    assert all(np.array_equal(a["phi"], b["phi"]) for a, b in zip(xs, ys))
