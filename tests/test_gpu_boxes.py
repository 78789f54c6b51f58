"""Multi-box levels (a DisjointBoxLayout with one FAB per box) through the C
ABI: the 256^2 mushy layer cut into four 128^2 boxes (the reference's
max_grid_size 128), an uneven 2D layout and a 3D layout of eight boxes.
The boxes' FABs gather into the domain array bit for bit (cells and face
coefficients), the backward-Euler step on the gathered level equals the
one-box step bit for bit, and the scatter returns every box its valid cells
and its exchange ghosts (neighbouring boxes and periodic images)."""
import numpy as np
import pytest

from oracle import mlref as O
from tests import hc_cases as HC

pytestmark = pytest.mark.gpu


def _fab(x, n, lo, hi):
    """Box [lo, hi] of a domain cell array, grown by one (ghosts from the
    domain array, periodic images taken by the caller's padding)."""
    if len(n) == 2:
        return np.ascontiguousarray(x[:, :, lo[1]:hi[1] + 3, lo[0]:hi[0] + 3])
    return np.ascontiguousarray(x[:, lo[2]:hi[2] + 3, lo[1]:hi[1] + 3, lo[0]:hi[0] + 3])


def _face_fab(f, n, d, lo, hi):
    sl = [slice(lo[e], hi[e] + 1 + (e == d)) for e in range(len(n))]
    if len(n) == 2:
        return np.ascontiguousarray(f[:, :, sl[1], sl[0]])
    return np.ascontiguousarray(f[:, sl[2], sl[1], sl[0]])


def _wrap_ghosts(x, n, periodic):
    """The domain array with its ghost layer set to the periodic images (the
    values a scattered box's ghosts take across periodic sides)."""
    y = x.copy()
    dim = len(n)
    for d in range(dim):
        if not periodic[d]:
            continue
        ax = 3 - d
        idx = [slice(None)] * 4
        src = [slice(None)] * 4
        idx[ax], src[ax] = 0, n[d]
        y[tuple(idx)] = y[tuple(src)]
        idx[ax], src[ax] = n[d] + 1, 1
        y[tuple(idx)] = y[tuple(src)]
    return y


LAYOUTS = {
    "2d_256_four_128": ((256, 256), [((0, 0), (127, 127)), ((128, 0), (255, 127)),
                                     ((0, 128), (127, 255)), ((128, 128), (255, 255))]),
    "2d_96x64_uneven": ((96, 64), [((0, 0), (31, 63)), ((32, 0), (95, 15)), ((32, 16), (95, 63))]),
    "3d_32_eight": ((32, 32, 32), [((x, y, z), (x + 15, y + 15, z + 15)) for z in (0, 16) for y in (0, 16)
                                   for x in (0, 16)]),
}


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_boxes_gather_solve_scatter(ml, gpu_ctx, wl, name):
    n, boxes = LAYOUTS[name]
    n = list(n)
    dim = len(n)
    case = wl.mushy(n=n, seed=31, corr=0.1 if dim == 2 else 0.08)
    per = case["periodic"]
    phi = case["phi"]
    g, p, bc = HC.oracle_setup(case)
    # the one-box step
    chi = O.porosity_field(g, p, phi)
    b = O.hc_bcoef(g, p, chi)
    op = HC.gpu_setup(ml, gpu_ctx, case)
    opts = ml.default_opts(max_iter=100, hang=-1.0)
    db = [ml.DeviceArray(gpu_ctx, x.shape, host=x) for x in b]
    op.set_coefs(None, db, 1.0, -case["dt"])
    one = ml.DeviceArray(gpu_ctx, phi.shape, host=phi)
    old = ml.DeviceArray(gpu_ctx, phi.shape, host=phi)
    st1 = op.solve_be(one, old, case["dt"], opts=opts)
    ref = one.get()
    # the boxes: FABs cut from the state, faces from the coefficients
    padded = _wrap_ghosts(phi, n, per)
    fabs = [ml.DeviceArray(gpu_ctx, f.shape, host=f) for f in (_fab(padded, n, lo, hi) for lo, hi in boxes)]
    level = ml.DeviceArray(gpu_ctx, phi.shape)
    ml.level_gather(gpu_ctx, n, per, 2, boxes, fabs, level)
    assert HC.bitwise_equal(HC.valid(n, level.get()), HC.valid(n, phi))
    lb = []
    for d in range(dim):
        ff = [ml.DeviceArray(gpu_ctx, f.shape, host=f) for f in (_face_fab(b[d], n, d, lo, hi) for lo, hi in boxes)]
        out = ml.DeviceArray(gpu_ctx, b[d].shape)
        ml.level_gather_faces(gpu_ctx, n, 2, boxes, d, ff, out)
        assert HC.bitwise_equal(out.get(), b[d]), d
        lb.append(out)
    # the step on the gathered level, with the gathered coefficients
    op2 = HC.gpu_setup(ml, gpu_ctx, case)
    op2.set_coefs(None, lb, 1.0, -case["dt"])
    old2 = ml.DeviceArray(gpu_ctx, phi.shape)
    gpu_ctx.check(gpu_ctx.lib.ml_buf_copy(gpu_ctx.h, old2.ptr, level.ptr, level.nbytes), "copy")
    st2 = op2.solve_be(level, old2, case["dt"], opts=opts)
    assert (st1.exit_status, st1.iterations) == (st2.exit_status, st2.iterations)
    assert HC.bitwise_equal(HC.valid(n, level.get()), HC.valid(n, ref))
    # scatter: every box's valid cells and exchange ghosts
    for f in fabs:
        gpu_ctx.check(gpu_ctx.lib.ml_vec_set(gpu_ctx.h, f.nbytes // 8, f.ptr, -7.0), "ml_vec_set")
    ml.level_scatter(gpu_ctx, n, per, 2, boxes, level, fabs)
    refp = _wrap_ghosts(ref, n, per)
    for (lo, hi), f in zip(boxes, fabs):
        got = f.get()
        exp = _fab(refp, n, lo, hi)
        # cells of the grown box inside the domain or across a periodic side
        mask = np.ones(exp.shape[1:], dtype=bool)
        for d in range(dim):
            ax = dim - 1 - d if dim == 3 else 2 - d
            if per[d]:
                continue
            m = np.ones(exp.shape[1 + ax], dtype=bool)
            if lo[d] == 0:
                m[0] = False
            if hi[d] == n[d] - 1:
                m[-1] = False
            shape = [1] * len(exp.shape[1:])
            shape[ax] = -1
            mask &= m.reshape(shape)
        # corners of the grown box are not exchange ghosts of a 5/7-point
        # stencil but the scatter fills them too when both sides are inside
        for c in range(2):
            assert HC.bitwise_equal(got[c][mask], exp[c][mask]), (lo, hi, c)
            assert np.all(got[c][~mask] == -7.0), (lo, hi, c)
    for x in (op, op2):
        x.close()


def test_boxes_rejected(ml, gpu_ctx):
    n = [32, 32]
    x = ml.DeviceArray(gpu_ctx, (2, 1, 34, 34))
    f = ml.DeviceArray(gpu_ctx, (2, 1, 34, 34))
    with pytest.raises(ml.MLError, match="overlap"):
        ml.level_gather(gpu_ctx, n, [1, 0], 2, [((0, 0), (31, 31)), ((0, 0), (15, 15))], [f, f], x)
    with pytest.raises(ml.MLError, match="cover"):
        ml.level_gather(gpu_ctx, n, [1, 0], 2, [((0, 0), (15, 31))], [f], x)
    with pytest.raises(ml.MLError, match="outside"):
        ml.level_gather(gpu_ctx, n, [1, 0], 2, [((0, 0), (32, 31))], [f], x)
