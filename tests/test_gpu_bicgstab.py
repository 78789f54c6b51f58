"""GPU (FP64) parity of the linear multigrid with the reference's BiCGStab
bottom solver (the projection's, srcSubcycle/Projector.cpp:3004-3023, and the
U* solves', srcSubcycle/AMRLevelMushyLayerInit.cpp:739), through the C ABI:

* ml_vc_dot and ml_vc_precond bit-identical to the oracle's dot product (its
  fixed summation order) and preCond, on fine and coarse MG levels;
* whole solves with the BiCGStab bottom inside the single-workgroup V-cycle:
  same iterations, exit status, final residual and bits as the oracle, for
  coarsest levels run by one wavefront (2x2, 6x4, 6x4x2 cells) and by the
  whole workgroup (24x10 cells, and 31x20 with an odd row length);
* the reference-API host driver (host/BiCGStabSolver.H over
  HipDarcyBrinkmanOp) bit-identical to the fused engine, graph replays equal
  to kernel-by-kernel launches;
* a coarsest level above one workgroup's cells is refused (ML_ENOTSUP).
"""
import importlib

import numpy as np
import pytest

from oracle import mlref as O
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu

wlm = importlib.import_module("mushy-layer_amd.workloads")

# (n, kind): coarsest MG levels 2x2, 24x10, 6x4, 31x20, 2x2x2, 6x4x2, 2x1
CASES = {
    "db_2d_64": ((64, 64), "db"),
    "db_2d_96x40": ((96, 40), "db"),
    "proj_2d_48x32": ((48, 32), "proj"),
    "db_2d_31x20": ((31, 20), "db_walls"),
    "db_3d_16": ((16, 16, 16), "db"),
    "proj_3d_24x16x8": ((24, 16, 8), "proj"),
    "proj_2d_1024x512": ((1024, 512), "proj"),
}
U_KW = dict(nu1=2, nu2=2, bottom=2, max_iter=10, eps=1e-10, hang=1e-10, norm_thresh=1e-10, bottom_solver=1,
            bottom_max_iter=80)
P_KW = dict(nu1=2, nu2=4, bottom=4, max_iter=20, eps=1e-15, hang=1e-15, norm_thresh=1e-15, bottom_solver=1,
            bottom_max_iter=20)


def _setup(ml, ctx, name):
    n, kind = CASES[name]
    w = wlm.darcy_brinkman(n=n, seed=3, corr=0.15)
    if kind == "db_walls":
        # no-slip walls on every side (an odd periodic length is refused on a
        # level the single-workgroup kernel relaxes)
        kind = "db"
        w["periodic"] = [0] * len(n)
        w["bc_u"] = {(d, s, 0): (0, 0.0) for d in range(len(n)) for s in (0, 1)}
    if kind == "db":
        C = dict(b=w["b_u"], c=w["c"], bc=w["bc_u"], alpha=1.0, beta=-w["dt"])
    else:
        C = dict(b=w["b_p"], c=None, bc=w["bc_p"], alpha=0.0, beta=1.0)
    g = O.make_grid(w["n"], w["dx"], w["periodic"] + [0] * (3 - len(w["periodic"])))
    op = ml.VCOp(ctx, w["n"], w["dx"], w["periodic"], ml.make_bc(C["bc"]))
    db = [ml.DeviceArray(ctx, x.shape, host=x) for x in C["b"]]
    dc = ml.DeviceArray(ctx, C["c"].shape, host=C["c"]) if C["c"] is not None else None
    op.set_coefs(None, db, dc, C["alpha"], C["beta"])
    return C, w, g, O.make_bc(C["bc"]), op, (db, dc), kind


def test_odd_periodic_length_on_a_single_workgroup_level_is_refused(ml, gpu_ctx):
    with pytest.raises(RuntimeError, match="odd periodic length"):
        ml.VCOp(gpu_ctx, [31, 20], 1.0 / 31, [1, 0], ml.make_bc({}))


def _rand(g, seed, zero_mean=False):
    rng = np.random.default_rng(seed)
    x = np.zeros(O.cell_shape(g, 1))
    O.valid(g, x)[...] = rng.standard_normal(O.valid(g, x).shape)
    if zero_mean:
        O.valid(g, x)[...] -= O.valid(g, x).mean()
    return x


def _check(n, out, ref):
    e = HC.rel_err(HC.valid(n, out), HC.valid(n, ref))
    assert e <= 1e-12, e
    assert HC.bitwise_equal(np.ascontiguousarray(HC.valid(n, out)), np.ascontiguousarray(HC.valid(n, ref)))


def _level(g, depth, C):
    """Grid and averaged coefficients of MG depth `depth` (the factory's
    arithmetic averages)."""
    if depth == 0:
        return g, C["b"], C["c"]
    r = 1 << depth
    gc = O.coarsen(g, r)
    c = None
    if C["c"] is not None:
        c = O.average_cells(g, 1, C["c"], r)
    b = O.average_faces(g, 1, C["b"], r)
    return gc, b, c


@pytest.mark.parametrize("name", ["db_2d_64", "proj_2d_48x32", "db_3d_16", "db_2d_31x20"])
def test_dot_and_precond_parity(ml, gpu_ctx, name):
    C, w, g, bc, op, keep, kind = _setup(ml, gpu_ctx, name)
    for depth in sorted({0, O.lib().mlr_mg_depth(O.C.byref(g)) - 1}):
        gl, b, c = _level(g, depth, C)
        nl = [gl.n[d] for d in range(gl.dim)]
        x, y = _rand(gl, 1 + depth), _rand(gl, 2 + depth)
        dx_, dy_ = (ml.DeviceArray(gpu_ctx, v.shape, host=v) for v in (x, y))
        assert op.dot(depth, dx_, dy_) == O.lin_dot(gl, 1, x, y), depth
        ref = O.lin_precond(gl, 1, bc, C["alpha"], None, C["beta"], c, b, y)
        dz = ml.DeviceArray(gpu_ctx, y.shape, host=np.zeros_like(y))
        op.precond(depth, dz, dy_)
        _check(nl, dz.get(), ref)


@pytest.mark.parametrize("name", sorted(CASES))
def test_solve_with_bicgstab_bottom_parity(ml, gpu_ctx, name):
    """Whole solves with the reference's options (U*: BiCGStab of 80
    iterations after 2 smoothings; projection: 20 after 4): same iterations,
    exit status, final residual and bits as the oracle."""
    C, w, g, bc, op, keep, kind = _setup(ml, gpu_ctx, name)
    kw = U_KW if kind == "db" else P_KW
    rhs = _rand(g, 6, zero_mean=kind == "proj")
    phi = np.zeros(O.cell_shape(g, 1))
    st_ref = O.lin_mg_solve(g, 1, bc, C["alpha"], None, C["beta"], C["c"], C["b"], phi, rhs,
                            opts=O.default_opts(**{k: v for k, v in kw.items()}))
    dphi = ml.DeviceArray(gpu_ctx, phi.shape, host=np.zeros_like(phi))
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    st = op.solve(dphi, drhs, opts=ml.default_opts(**kw))
    print("%s: %d iterations, status %d, |r| %.3e" % (name, st.iterations, st.exit_status, st.final_residual))
    assert (st.iterations, st.exit_status) == (st_ref.iterations, st_ref.exit_status)
    assert st.final_residual == st_ref.final_residual
    _check(w["n"], dphi.get(), phi)


@pytest.mark.parametrize("name", ["db_2d_64", "db_2d_96x40", "proj_3d_24x16x8"])
def test_host_driver_and_graph_equal_engine(ml, gpu_ctx, name):
    """The reference-API host driver (one C-ABI call per LinearOp method,
    BiCGStabSolver on the coarsest HipDarcyBrinkmanOp), the graph-replayed
    engine (three solves in a row) and the kernel-by-kernel engine give the
    same statistics and bits."""
    C, w, g, bc, op, keep, kind = _setup(ml, gpu_ctx, name)
    kw = dict(U_KW if kind == "db" else P_KW)
    rhs = _rand(g, 21, zero_mean=kind == "proj")
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    out = []
    for flags in (ml.SOLVE_NO_GRAPH, ml.SOLVE_HOST_DRIVER, 0, 0, 0):
        dphi = ml.DeviceArray(gpu_ctx, rhs.shape, host=np.zeros_like(rhs))
        st = op.solve(dphi, drhs, opts=ml.default_opts(flags=flags, **kw))
        out.append((dphi.get(), (st.exit_status, st.iterations, st.initial_residual, st.final_residual)))
        gpu_ctx.sync()
    for k in range(1, len(out)):
        assert out[k][1] == out[0][1], k
        assert HC.bitwise_equal(out[k][0], out[0][0]), k


def test_bicgstab_bottom_needs_a_small_coarsest_level(ml, gpu_ctx):
    # 40 x 33 does not coarsen: its 1320-cell level is relaxed by the tiled
    # kernels, and BiCGStab (which runs inside the single-workgroup V-cycle)
    # is refused; the relaxation bottom still solves
    n = (40, 33)
    w = wlm.darcy_brinkman(n=n, seed=3, corr=0.15)
    op = ml.VCOp(gpu_ctx, w["n"], w["dx"], w["periodic"], ml.make_bc(w["bc_u"]))
    db = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in w["b_u"]]
    dc = ml.DeviceArray(gpu_ctx, w["c"].shape, host=w["c"])
    op.set_coefs(None, db, dc, 1.0, -w["dt"])
    g = O.make_grid(w["n"], w["dx"], w["periodic"] + [0])
    rhs = _rand(g, 3)
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    dphi = ml.DeviceArray(gpu_ctx, rhs.shape, host=np.zeros_like(rhs))
    with pytest.raises(RuntimeError, match="BiCGStab bottom"):
        op.solve(dphi, drhs, opts=ml.default_opts(**U_KW))
    st = op.solve(dphi, drhs, opts=ml.default_opts(**dict(U_KW, bottom_solver=0, bottom=8)))
    assert st.exit_status in (1, 2, 8)
