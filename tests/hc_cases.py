"""Shared case builders for the (H, Theta) operator tests: the same seeded
host arrays feed the CPU oracle and the GPU library."""
import numpy as np

from oracle import mlref as O


def oracle_setup(case):
    """Oracle grid / params / bc objects for a workload dict."""
    g = O.make_grid(case["n"], case["dx"], case["periodic"] + [0] * (3 - len(case["periodic"])))
    p = O.default_params(**case["params"])
    bc = O.make_bc(case["bc"], apply=case.get("apply_bc", 1))
    return g, p, bc


def gpu_setup(ml, ctx, case):
    p = ml.make_params(**case["params"])
    bc = ml.make_bc(case["bc"], apply=case.get("apply_bc", 1))
    op = ml.HCOp(ctx, case["n"], case["dx"], case["periodic"], p, bc)
    return op


def bcoef_from_state(case):
    """Porosity of the state then the HC face coefficients (oracle)."""
    g, p, _ = oracle_setup(case)
    chi = O.porosity_field(g, p, case["phi"])
    return O.hc_bcoef(g, p, chi)


def random_state(n, seed, Hrange=(-2.0, 8.0), Crange=(-2.6, 0.4)):
    rng = np.random.default_rng(seed)
    shape = (2, n[2] + 2, n[1] + 2, n[0] + 2) if len(n) == 3 else (2, 1, n[1] + 2, n[0] + 2)
    phi = np.zeros(shape)
    phi[0] = rng.uniform(*Hrange, size=phi[0].shape)
    phi[1] = rng.uniform(*Crange, size=phi[1].shape)
    return phi


def valid(n, x):
    if len(n) == 3:
        return x[:, 1:-1, 1:-1, 1:-1]
    return x[:, :, 1:-1, 1:-1]


def rel_err(a, b):
    """max|a-b| / max|b| (0 when both are zero)."""
    den = np.max(np.abs(b))
    num = np.max(np.abs(a - b))
    if den == 0:
        return float(num)
    return float(num / den)


def bitwise_equal(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))
