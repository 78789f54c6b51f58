"""The state update and the cell-centred velocity closure of a Darcy-Brinkman
step on the GPU, through the C ABI, against the oracle (oracle/mlref_flow.c,
oracle/steps.py):

* enthalpy variables, U* source, velocity before the projection and the
  projection's right-hand side, projection faces, projection correction:
  bit-identical on random inputs (2D periodic/walled, 2D walled, 3D);
* the U* Darcy coefficient for every permeability law: bit-identical for the
  laws built from the once-rounded cube and square (Kozeny-Carman, cubic,
  porosity, pure fluid), within 4 ulp for the log law (device and host log
  may round differently);
* whole Darcy-Brinkman steps (HC solve + U* solves + projection) at 64^2 and
  32^3 (the 3D workload's correlation length 0.08, Ra_C 300, Da 5e-4): H,
  Theta, velocity and pressure bit-identical, every solve with the same
  iterations and exit status."""
import importlib

import numpy as np
import pytest

from oracle import mlref as O
from oracle import steps as S
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu

def _grid_case(n, periodic):
    g = O.make_grid(list(n), 1.0 / n[0], list(periodic) + [0] * (3 - len(periodic)))
    return g


def _rand_vars(g, rng, solid_frac=0.1):
    v = np.zeros(O.cell_shape(g, 4))
    w = O.valid(g, v)
    w[0] = rng.uniform(-0.5, 1.0, w[0].shape)
    w[1] = rng.uniform(-1.0, 0.5, w[1].shape)
    w[2] = rng.uniform(-2.0, 0.0, w[2].shape)
    w[3] = rng.uniform(0.0, 1.0, w[3].shape)
    w[3][rng.random(w[3].shape) < solid_frac] = 0.02
    return v


def _rand_cell(g, rng, scale=1.0):
    x = np.zeros(O.cell_shape(g, 1))
    O.valid(g, x)[...] = scale * rng.standard_normal(O.valid(g, x).shape)
    return x


CASES = [((24, 16), (1, 0)), ((16, 20), (0, 0)), ((12, 10, 8), (1, 1, 0)), ((8, 12, 10), (0, 1, 0))]


@pytest.mark.parametrize("n,periodic", CASES)
def test_enthalpy_vars_bitwise(ml, gpu_ctx, n, periodic):
    g = _grid_case(n, periodic)
    p = O.default_params(stefan=4.0, compositionRatio=3.0, lewis=100.0)
    phi = HC.random_state(list(n), seed=sum(n))
    ref = O.enthalpy_vars(g, p, phi)
    op = ml.HCOp(gpu_ctx, list(n), g.dx, list(periodic), ml.make_params(stefan=4.0, compositionRatio=3.0,
                                                                          lewis=100.0), ml.make_bc())
    d = ml.DeviceArray(gpu_ctx, phi.shape, host=phi)
    out = ml.DeviceArray(gpu_ctx, (4,) + phi.shape[1:])
    op.enthalpy_vars(d, out)
    got = out.get()
    assert HC.bitwise_equal(O.valid(g, got), O.valid(g, ref))
    # every regime present
    chi = O.valid(g, ref)[3]
    assert (chi == 0).any() and (chi == 1).any() and ((chi > 0) & (chi < 1)).any()
    op.close()


@pytest.mark.parametrize("n,periodic", CASES)
def test_velocity_closure_pieces_bitwise(ml, gpu_ctx, n, periodic):
    g = _grid_case(n, periodic)
    dim = len(n)
    rng = np.random.default_rng(7 + sum(n))
    vo, vn = _rand_vars(g, rng), _rand_vars(g, rng)
    pi = _rand_cell(g, rng)
    ustar = [_rand_cell(g, rng, 0.3) for _ in range(dim)]
    dt = 2e-3
    bT, bS = 5.0, 2000.0
    # oracle
    src_ref = O.ustar_src(g, bT, bS, 1, vo, vn, pi.copy())
    u_ref, rhs_ref = O.cc_project_rhs(g, dt, 0.05, 1, vo, vn, pi.copy(), ustar)
    bp_ref = O.pressure_faces(g, vo, vn)
    pi2 = _rand_cell(g, rng)
    uc_ref = [x.copy() for x in u_ref]
    O.cc_correct(g, dt, 0.05, 1, vo, vn, pi2.copy(), uc_ref)
    # GPU
    bc = ml.make_bc({(dim - 1, 0, 0): (0, 0.0), (dim - 1, 1, 0): (0, 0.0)})
    op = ml.VCOp(gpu_ctx, list(n), g.dx, list(periodic), bc)
    D = lambda x: ml.DeviceArray(gpu_ctx, x.shape, host=x)
    dvo, dvn, dpi, dpi2 = D(vo), D(vn), D(pi), D(pi2)
    dus = [D(x) for x in ustar]
    src = [ml.DeviceArray(gpu_ctx, pi.shape) for _ in range(dim)]
    u = [ml.DeviceArray(gpu_ctx, pi.shape) for _ in range(dim)]
    rhs = ml.DeviceArray(gpu_ctx, pi.shape)
    bp = [ml.DeviceArray(gpu_ctx, ml.face_shape(list(n), d, 1)) for d in range(dim)]
    fp = ml.make_flow_params(bT, bS, 0.05, 1)
    op.ustar_src(fp, dvo, dvn, dpi, src)
    op.cc_project_rhs(fp, dt, dvo, dvn, dpi, dus, u, rhs)
    op.pressure_faces(dvo, dvn, bp)
    for d in range(dim):
        assert HC.bitwise_equal(O.valid(g, src[d].get()), O.valid(g, src_ref[d])), ("src", d)
        assert HC.bitwise_equal(O.valid(g, u[d].get()), O.valid(g, u_ref[d])), ("u", d)
        assert HC.bitwise_equal(bp[d].get(), bp_ref[d]), ("faces", d)
    assert HC.bitwise_equal(O.valid(g, rhs.get()), O.valid(g, rhs_ref)), "rhs"
    op.cc_correct(fp, dt, dvo, dvn, dpi2, u)
    for d in range(dim):
        assert HC.bitwise_equal(O.valid(g, u[d].get()), O.valid(g, uc_ref[d])), ("corrected u", d)
    # solid cells carry no velocity
    solid = O.valid(g, vn)[3] <= 0.05
    assert solid.any() and all(np.all(O.valid(g, u[d].get())[0][solid] == 0) for d in range(dim))
    op.close()


@pytest.mark.parametrize("law", ["kozeny_carman", "cubic", "log", "porosity", "pure_fluid"])
def test_darcy_coefficient_laws(ml, gpu_ctx, law):
    n = [32, 16]
    g = _grid_case(n, (1, 0))
    rng = np.random.default_rng(11)
    chi = np.zeros(O.cell_shape(g, 1))
    v = rng.uniform(0.05, 0.999, O.valid(g, chi).shape)
    v.flat[:5] = [1.0, 0.05, 0.5, 0.999, 0.2]
    O.valid(g, chi)[...] = v
    ref = O.valid(g, O.darcy_ccoef(g, chi, 1e4, law))[0]
    op = ml.VCOp(gpu_ctx, n, g.dx, [1, 0], ml.make_bc())
    dc = ml.DeviceArray(gpu_ctx, chi.shape, host=chi)
    c = ml.DeviceArray(gpu_ctx, chi.shape)
    op.darcy_ccoef_law(dc, 1e4, law, c)
    got = O.valid(g, c.get())[0]
    if law == "log":
        np.testing.assert_allclose(got, ref, rtol=4 * np.finfo(float).eps, atol=0)
    else:
        assert HC.bitwise_equal(got, ref)
    if law == "kozeny_carman":
        assert got.flat[0] == 0.0       # chi = 1: infinite permeability, no Darcy drag
    op.close()


def _db_compare(ml, ctx, case, flow, steps, max_iter=100):
    stp = importlib.import_module("mushy-layer_amd.steps")
    opts = ml.default_opts(max_iter=max_iter, hang=S.NO_HANG)
    gpu = stp.DarcyBrinkmanStepper(ml, ctx, case, opts, flow)
    cpu = S.DarcyBrinkmanStep(case, flow, max_iter=max_iter)
    n = case["n"]
    g = cpu.g
    for k in range(steps):
        gpu.step()
        cpu.step()
        ctx.sync()
        a, r = gpu.phi.get(), cpu.phi
        fl = gpu.flow_fields()
        errs = {}
        for d in range(len(n)):
            errs["u%d" % d] = HC.rel_err(fl["u%d" % d], O.valid(g, cpu.u[d])[0])
        errs["p"] = HC.rel_err(fl["p"], O.valid(g, cpu.pres)[0])
        print("step %d: HC %s, velocity solves GPU %s oracle %s, errors %s"
              % (k, cpu.hc_stats[-1], gpu.vstats[-1], cpu.vstats[-1], {x: "%.2e" % y for x, y in errs.items()}),
              flush=True)
        hs_g = [(s[0], s[1]) for s in gpu.stats]
        assert hs_g == cpu.hc_stats, (k, hs_g, cpu.hc_stats)
        assert gpu.vstats == cpu.vstats, (k, gpu.vstats, cpu.vstats)
        assert HC.bitwise_equal(HC.valid(n, a), HC.valid(n, r)), k
        for d in range(len(n)):
            assert HC.bitwise_equal(fl["u%d" % d], O.valid(g, cpu.u[d])[0]), (k, d, errs)
        assert HC.bitwise_equal(fl["p"], O.valid(g, cpu.pres)[0]), (k, errs)
        umax = max(np.abs(fl["u%d" % d]).max() for d in range(len(n)))
        assert umax > 0
    gpu.close()


def test_darcy_brinkman_steps_2d(ml, gpu_ctx, wl):
    case = wl.mushy(n=(64, 64), seed=12)
    flow = wl.flow(dim=2, Pr=10.0, Da=1e-3, Ra_C=200.0)
    _db_compare(ml, gpu_ctx, case, flow, steps=3)


def test_darcy_brinkman_steps_3d(ml, gpu_ctx, wl):
    case = wl.mushy(n=(32, 32, 32), corr=0.08, seed=13)
    flow = wl.flow(dim=3, Pr=10.0, Da=5e-4, Ra_C=300.0)
    _db_compare(ml, gpu_ctx, case, flow, steps=2)
