"""GPU (FP64) parity of the linear variable-coefficient operator — the
Darcy-Brinkman U* Helmholtz form and the projection's Poisson form — against
the CPU oracle, through the C ABI.  The kernels reproduce the oracle's
evaluation order, so every comparison is bit-for-bit; the relative bar of
1e-12 is asserted first so that a failure reports its size."""
import importlib

import numpy as np
import pytest

from oracle import mlref as O
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu

TOL = 1e-12
wlm = importlib.import_module("mushy-layer_amd.workloads")


def _case(name):
    n, kind = CASES[name]
    w = wlm.darcy_brinkman(n=n, seed=3, corr=0.15)
    if kind == "db":
        return dict(w=w, b=w["b_u"], c=w["c"], bc=w["bc_u"], alpha=1.0, beta=-w["dt"])
    return dict(w=w, b=w["b_p"], c=None, bc=w["bc_p"], alpha=0.0, beta=1.0)


CASES = {
    "db_2d_64": ((64, 64), "db"),
    "db_2d_96x40": ((96, 40), "db"),
    "proj_2d_64": ((64, 64), "proj"),
    "proj_2d_48x32": ((48, 32), "proj"),
    "db_2d_512": ((512, 512), "db"),
    "db_3d_16": ((16, 16, 16), "db"),
    "proj_3d_24x16x8": ((24, 16, 8), "proj"),
    "db_3d_64": ((64, 64, 64), "db"),
    # above 512^2 cells: the wide tiles (short halo) of large levels
    "db_2d_1024x512": ((1024, 512), "db"),
    "proj_2d_768x512": ((768, 512), "proj"),
}
SMALL = [k for k in CASES if k not in ("db_2d_512", "db_3d_64", "db_2d_1024x512", "proj_2d_768x512")]
WIDE = ["db_2d_1024x512", "proj_2d_768x512"]


def _setup(ml, ctx, name):
    C = _case(name)
    w = C["w"]
    g = O.make_grid(w["n"], w["dx"], w["periodic"] + [0] * (3 - len(w["periodic"])))
    bc_o = O.make_bc(C["bc"])
    op = ml.VCOp(ctx, w["n"], w["dx"], w["periodic"], ml.make_bc(C["bc"]))
    db = [ml.DeviceArray(ctx, x.shape, host=x) for x in C["b"]]
    dc = ml.DeviceArray(ctx, C["c"].shape, host=C["c"]) if C["c"] is not None else None
    op.set_coefs(None, db, dc, C["alpha"], C["beta"])
    return C, w, g, bc_o, op, (db, dc)


def _rand(g, seed):
    rng = np.random.default_rng(seed)
    x = np.zeros(O.cell_shape(g, 1))
    O.valid(g, x)[...] = rng.standard_normal(O.valid(g, x).shape)
    return x


def _check(n, out, ref):
    e = HC.rel_err(HC.valid(n, out), HC.valid(n, ref))
    assert e <= TOL, e
    assert HC.bitwise_equal(np.ascontiguousarray(HC.valid(n, out)), np.ascontiguousarray(HC.valid(n, ref)))


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("iters", [1, 3])
def test_vc_relax_parity(ml, gpu_ctx, name, iters):
    C, w, g, bc, op, keep = _setup(ml, gpu_ctx, name)
    n = w["n"]
    phi, rhs = _rand(g, 1), _rand(g, 2)
    ref = O.lin_gsrb(g, 1, bc, C["alpha"], None, C["beta"], C["c"], C["b"], phi.copy(), rhs, iters=iters)
    dphi = ml.DeviceArray(gpu_ctx, phi.shape, host=phi)
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    op.relax(0, dphi, drhs, iters)
    _check(n, dphi.get(), ref)


@pytest.mark.parametrize("name", SMALL)
def test_vc_residual_apply_restrict_prolong(ml, gpu_ctx, name):
    C, w, g, bc, op, keep = _setup(ml, gpu_ctx, name)
    n = w["n"]
    a, be, c, b = C["alpha"], C["beta"], C["c"], C["b"]
    phi, rhs = _rand(g, 3), _rand(g, 4)
    dphi = ml.DeviceArray(gpu_ctx, phi.shape, host=phi)
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    dout = ml.DeviceArray(gpu_ctx, phi.shape)
    for homog in (0, 1):
        op.residual(0, dout, dphi, drhs, homog)
        _check(n, dout.get(), O.lin_residual(g, 1, bc, a, None, be, c, b, phi.copy(), rhs, homogeneous=homog))
        op.apply_op(0, dout, dphi, homog)
        _check(n, dout.get(), O.lin_apply_op(g, 1, bc, a, None, be, c, b, phi.copy(), homogeneous=homog))
    gc = O.coarsen(g, 2)
    nc = [gc.n[d] for d in range(g.dim)]
    dresc = ml.DeviceArray(gpu_ctx, O.cell_shape(gc, 1))
    op.restrict_residual(0, dresc, dphi, drhs)
    _check(nc, dresc.get(), O.lin_restrict_residual(g, 1, bc, a, None, be, c, b, phi.copy(), rhs))
    corr = _rand(gc, 5)
    dcorr = ml.DeviceArray(gpu_ctx, corr.shape, host=corr)
    op.prolong_increment(0, dphi, dcorr)
    _check(n, dphi.get(), O.prolong_constant(g, 1, phi.copy(), corr))


@pytest.mark.parametrize("name", sorted(CASES))
def test_vc_solve_parity(ml, gpu_ctx, name):
    """Whole residual-correction MG solve: same iterations, exit status and
    bit-identical solution as the oracle."""
    C, w, g, bc, op, keep = _setup(ml, gpu_ctx, name)
    n = w["n"]
    a, be, c, b = C["alpha"], C["beta"], C["c"], C["b"]
    rhs = _rand(g, 6)
    if C["c"] is None:
        # projection: a compatible right-hand side (zero mean) for the
        # singular all-Neumann/periodic Poisson problem
        O.valid(g, rhs)[...] -= O.valid(g, rhs).mean()
    phi = np.zeros(O.cell_shape(g, 1))
    kw = dict(nu1=2, nu2=4, bottom=8, max_iter=6, eps=1e-10, hang=1e-10, norm_thresh=1e-10)
    st_ref = O.lin_mg_solve(g, 1, bc, a, None, be, c, b, phi, rhs, prolong_linear=0, opts=O.default_opts(**kw))
    dphi = ml.DeviceArray(gpu_ctx, phi.shape, host=np.zeros_like(phi))
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    st = op.solve(dphi, drhs, opts=ml.default_opts(**kw))
    assert (st.iterations, st.exit_status) == (st_ref.iterations, st_ref.exit_status)
    assert st.final_residual == st_ref.final_residual
    _check(n, dphi.get(), phi)


@pytest.mark.parametrize("name", ["db_2d_64", "db_2d_512", "db_3d_16"])
def test_db_be_step_parity(ml, gpu_ctx, name):
    """Darcy-Brinkman U* backward-Euler step: rhs = u_old + dt*src, solved to
    1e-10 (VelocityMultigrid: smoothing 2, at most 10 iterations)."""
    C, w, g, bc, op, keep = _setup(ml, gpu_ctx, name)
    n = w["n"]
    dt = w["dt"]
    u0, src = w["u0"], w["src"]
    rhs = np.zeros_like(u0)
    O.valid(g, rhs)[...] = O.valid(g, u0) * 1.0 + dt * O.valid(g, src)
    kw = dict(nu1=2, nu2=2, bottom=8, max_iter=10, eps=1e-10, hang=1e-10, norm_thresh=1e-10)
    ref = u0.copy()
    st_ref = O.lin_mg_solve(g, 1, bc, 1.0, None, -dt, C["c"], C["b"], ref, rhs, opts=O.default_opts(**kw))
    du0 = ml.DeviceArray(gpu_ctx, u0.shape, host=u0)
    dnew = ml.DeviceArray(gpu_ctx, u0.shape, host=u0)
    dsrc = ml.DeviceArray(gpu_ctx, src.shape, host=src)
    st = op.solve_be(dnew, du0, dt, src=dsrc, opts=ml.default_opts(**kw))
    assert (st.iterations, st.exit_status) == (st_ref.iterations, st_ref.exit_status)
    _check(n, dnew.get(), ref)


@pytest.mark.parametrize("name", ["db_2d_64", "db_2d_512", "db_3d_16"])
def test_db_be_many_matches_oracle_per_component(ml, gpu_ctx, name):
    """Concurrent U* solves (one per velocity component, sharing the
    operator): each is bit-identical to the oracle's backward-Euler solve of
    its own component, with the same iterations and exit status; a second
    call with the workspaces already created gives the same results."""
    C, w, g, bc, op, keep = _setup(ml, gpu_ctx, name)
    n = w["n"]
    dt = w["dt"]
    dim = len(n)
    kw = dict(nu1=2, nu2=2, bottom=8, max_iter=10, eps=1e-10, hang=1e-10, norm_thresh=1e-10)
    olds, srcs = [], []
    for comp in range(dim):
        u = _rand(g, 20 + comp) + (w["u0"] if comp == dim - 1 else 0.0)
        olds.append(u)
        srcs.append(w["src"] if comp == dim - 1 else _rand(g, 30 + comp))
    refs, st_refs = [], []
    for u, src in zip(olds, srcs):
        rhs = np.zeros_like(u)
        O.valid(g, rhs)[...] = O.valid(g, u) * 1.0 + dt * O.valid(g, src)
        ref = u.copy()
        st_refs.append(O.lin_mg_solve(g, 1, bc, 1.0, None, -dt, C["c"], C["b"], ref, rhs,
                                      opts=O.default_opts(**kw)))
        refs.append(ref)
    dold = [ml.DeviceArray(gpu_ctx, u.shape, host=u) for u in olds]
    dsrc = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in srcs]
    for _ in range(2):
        dnew = [ml.DeviceArray(gpu_ctx, u.shape, host=u) for u in olds]
        sts = op.solve_be_many(dnew, dold, dt, src=dsrc, opts=ml.default_opts(**kw))
        for st, st_ref, d, ref in zip(sts, st_refs, dnew, refs):
            assert (st.iterations, st.exit_status) == (st_ref.iterations, st_ref.exit_status)
            _check(n, d.get(), ref)


@pytest.mark.parametrize("name", ["db_2d_64", "proj_2d_48x32", "db_3d_16"])
def test_graph_replays_match_direct_launches(ml, gpu_ctx, name):
    """The solve replays two captured iterations as a HIP graph; three solves
    in a row, with other eager work and synchronisations between them, each
    give exactly the state of the solve launched kernel by kernel."""
    C, w, g, bc, op, keep = _setup(ml, gpu_ctx, name)
    rhs = _rand(g, 8)
    if C["c"] is None:
        O.valid(g, rhs)[...] -= O.valid(g, rhs).mean()
    kw = dict(nu1=2, nu2=4, bottom=8, max_iter=9, eps=1e-12, hang=1e-12, norm_thresh=1e-14)
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    dref = ml.DeviceArray(gpu_ctx, rhs.shape, host=np.zeros_like(rhs))
    st_ref = op.solve(dref, drhs, opts=ml.default_opts(flags=ml.SOLVE_NO_GRAPH, **kw))
    ref = dref.get()
    scratch = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    d = ml.DeviceArray(gpu_ctx, rhs.shape)
    for rep in range(3):
        d.set(np.zeros_like(rhs))   # same buffers every time: the captured graph is replayed
        st = op.solve(d, drhs, opts=ml.default_opts(**kw))
        assert (st.iterations, st.exit_status, st.final_residual) == \
            (st_ref.iterations, st_ref.exit_status, st_ref.final_residual), rep
        _check(w["n"], d.get(), ref)
        op.relax(0, scratch, drhs, 1)   # eager work between the replays
        gpu_ctx.sync()


@pytest.mark.parametrize("n", [(64, 64), (24, 16, 8)])
def test_darcy_and_projection_coefficients(ml, gpu_ctx, n):
    """cCoef = (chi/Pi) Pr/Da with Kozeny-Carman Pi, and the projection's
    -sqrt(|chi_lo chi_hi|) faces, from a device porosity field.  pow() of the
    device math library may differ from the host's in the last bit, so cCoef
    is held to 1e-14 relative; the faces are bit-identical."""
    w = wlm.darcy_brinkman(n=n, seed=3, corr=0.15)
    chi = np.zeros(O.cell_shape(O.make_grid(n), 1))
    O.valid(O.make_grid(n), chi)[0] = w["chi"].reshape(O.valid(O.make_grid(n), chi)[0].shape)
    op = ml.VCOp(gpu_ctx, n, w["dx"], w["periodic"], ml.make_bc(w["bc_p"]))
    dchi = ml.DeviceArray(gpu_ctx, chi.shape, host=chi)
    dc = ml.DeviceArray(gpu_ctx, chi.shape)
    op.darcy_ccoef(dchi, w["Pr"] / w["Da"], dc)
    c = HC.valid(n, dc.get())
    ref = HC.valid(n, w["c"])
    assert np.max(np.abs(c - ref)) <= 1e-14 * np.max(np.abs(ref))
    db = [ml.DeviceArray(gpu_ctx, x.shape) for x in w["b_p"]]
    op.porosity_faces(dchi, -1.0, db)
    for d in range(len(n)):
        assert HC.bitwise_equal(db[d].get(), w["b_p"][d])


def test_solves_follow_changed_options_and_coefficients(ml, gpu_ctx):
    """Consecutive solves on one operator with different smoothing counts and
    then a different beta: each solve uses its own settings (the device
    descriptor is resent when they change and reused when they do not) and
    matches the oracle bit for bit."""
    C, w, g, bc, op, keep = _setup(ml, gpu_ctx, "db_2d_64")
    n = w["n"]
    a, c, b = C["alpha"], C["c"], C["b"]
    rhs = _rand(g, 11)
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    runs = [(dict(nu1=2, nu2=4, bottom=8), C["beta"]), (dict(nu1=1, nu2=2, bottom=4), C["beta"]),
            (dict(nu1=1, nu2=2, bottom=4), C["beta"]), (dict(nu1=2, nu2=4, bottom=8), 2.0 * C["beta"])]
    for k, (sm, be) in enumerate(runs):
        kw = dict(max_iter=6, eps=1e-10, hang=1e-10, norm_thresh=1e-10, **sm)
        if be != C["beta"]:
            op.set_coefs(None, keep[0], keep[1], a, be)
        phi = np.zeros(O.cell_shape(g, 1))
        st_ref = O.lin_mg_solve(g, 1, bc, a, None, be, c, b, phi, rhs, opts=O.default_opts(**kw))
        dphi = ml.DeviceArray(gpu_ctx, phi.shape, host=np.zeros_like(phi))
        st = op.solve(dphi, drhs, opts=ml.default_opts(**kw))
        assert (st.iterations, st.exit_status) == (st_ref.iterations, st_ref.exit_status), k
        _check(n, dphi.get(), phi)


@pytest.mark.parametrize("name", ["db_2d_64", "proj_2d_48x32", "db_3d_16", "proj_3d_24x16x8", "db_2d_1024x512"])
def test_host_driver_equals_engine(ml, gpu_ctx, name):
    """ML_SOLVE_HOST_DRIVER runs the linear MG through the reference-API
    operator objects (host/HipVCOps.H, HipDarcyBrinkmanOp per MG depth) and
    the host driver (host/LinearMultiGrid.H), one C-ABI call per operator
    method: the same iterations, statistics and bits as the fused engine,
    for the plain solve and the backward-Euler form of two components."""
    C, w, g, bc, op, keep = _setup(ml, gpu_ctx, name)
    rhs = _rand(g, 21)
    if C["c"] is None:
        O.valid(g, rhs)[...] -= O.valid(g, rhs).mean()
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    out = {}
    for flags in (0, ml.SOLVE_HOST_DRIVER):
        o = ml.default_opts(nu1=2, nu2=2, bottom=8, max_iter=12, eps=1e-12, hang=1e-14, flags=flags)
        dphi = ml.DeviceArray(gpu_ctx, rhs.shape)
        st = op.solve(dphi, drhs, opts=o)
        out[flags] = (dphi.get(), (st.exit_status, st.iterations, st.initial_residual, st.final_residual))
    e, h = out[0], out[ml.SOLVE_HOST_DRIVER]
    print("%s: engine %s, host driver %s" % (name, e[1], h[1]))
    assert e[1] == h[1]
    assert HC.bitwise_equal(e[0], h[0])
    if C["c"] is None:
        return
    # backward-Euler solves of two components (the U* form)
    old = [_rand(g, 30 + i) for i in range(2)]
    src = [_rand(g, 40 + i) for i in range(2)]
    res = {}
    for flags in (0, ml.SOLVE_HOST_DRIVER):
        o = ml.default_opts(nu1=2, nu2=2, bottom=8, max_iter=10, eps=1e-10, hang=1e-10, flags=flags)
        dn = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in old]
        do = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in old]
        ds = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in src]
        sts = op.solve_be_many(dn, do, C["w"]["dt"], src=ds, opts=o)
        res[flags] = ([x.get() for x in dn], [(s.exit_status, s.iterations) for s in sts])
    assert res[0][1] == res[ml.SOLVE_HOST_DRIVER][1]
    assert all(HC.bitwise_equal(a, b) for a, b in zip(res[0][0], res[ml.SOLVE_HOST_DRIVER][0]))


# Large levels relax by streaming along the last direction (2D above 512^2
# cells, 3D above 96^3): wall and periodic sides in every direction, ragged
# block and segment counts, and several launches per call.
STREAM = {
    "2d_700x600_walls": ((700, 600), (0, 0), {(0, 0, 0): (0, 0.0), (0, 1, 0): (1, 0.0), (1, 0, 0): (1, 0.0),
                                              (1, 1, 0): (0, 0.0)}),
    "2d_530x520_periodic": ((530, 520), (1, 1), {}),
    "2d_1030x261_mixed": ((1030, 261), (1, 0), {(1, 0, 0): (0, 0.0), (1, 1, 0): (0, 0.0)}),
    "3d_100x98x97_walls_xz": ((100, 98, 97), (0, 1, 0), {(0, 0, 0): (0, 0.0), (0, 1, 0): (0, 0.0),
                                                         (2, 0, 0): (1, 0.0), (2, 1, 0): (0, 0.0)}),
    "3d_100x100x100_periodic_z": ((100, 100, 100), (1, 0, 1), {(1, 0, 0): (0, 0.0), (1, 1, 0): (1, 0.0)}),
}


@pytest.mark.parametrize("name", sorted(STREAM))
def test_vc_relax_stream_parity(ml, gpu_ctx, name):
    n, per, spec = STREAM[name]
    dim = len(n)
    g = O.make_grid(list(n), 1.0 / n[0], list(per) + [0] * (3 - dim))
    rng = np.random.default_rng(sum(n))
    b = [0.5 + rng.random(O.face_shape(g, d, 1)) for d in range(dim)]
    c = _rand(g, 7)
    c = np.abs(c) * 50.0
    bc_o = O.make_bc(spec)
    op = ml.VCOp(gpu_ctx, list(n), g.dx, list(per), ml.make_bc(spec))
    db = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in b]
    dc = ml.DeviceArray(gpu_ctx, c.shape, host=c)
    alpha, beta = 1.0, -2e-3
    op.set_coefs(None, db, dc, alpha, beta)
    phi, rhs = _rand(g, 11), _rand(g, 12)
    dphi = ml.DeviceArray(gpu_ctx, phi.shape, host=phi)
    drhs = ml.DeviceArray(gpu_ctx, rhs.shape, host=rhs)
    ref = phi.copy()
    for iters in (1, 2):   # 2 then 4 colour passes: one and two streaming launches
        ref = O.lin_gsrb(g, 1, bc_o, alpha, None, beta, c, b, ref, rhs, iters=iters)
        op.relax(0, dphi, drhs, iters)
        _check(list(n), dphi.get(), ref)
    op.close()
