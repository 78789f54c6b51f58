"""Chombo HDF5 checkpoint and plot files (the reference's on-disk format,
srcSubcycle/AMRLevelMushyLayerIO.cpp:131-656) through libmlio.so.

`H5File` wraps the C ABI of include/mlio.h.  On top of it:

  read_level(path, level)     header, boxes and every stored variable of one
                              level (per-box FABs, one ghost layer), e.g. the
                              reference's own checkpoints and plot files;
  domain_array(level, name)   a variable of a level whose boxes cover its
                              problem domain, as one array in this package's
                              layout (ncomp, [nz+2,] ny+2, nx+2);
  write_checkpoint(...)       a single-level checkpoint as writeCheckpointHeader
                              / writeCheckpointLevel lay it out;
  write_plotfile(...)         a single-level plot file as writePlotHeader /
                              writePlotLevel lay it out.
Arrays cross the boundary as host numpy float64 arrays; device state is
copied to the host by the caller (DeviceArray.get()).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB = os.path.join(_HERE, "libmlio.so")
_LIB = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
H5_OK, H5_EINVAL, H5_EIO, H5_ENOTFOUND = 0, -1, -2, -3


class H5Error(RuntimeError):
    pass


class Level(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3), ("periodic", C.c_int * 3), ("ref_ratio", C.c_int),
                ("tag_buffer_size", C.c_int), ("timestep_failed", C.c_int), ("dx", C.c_double), ("dt", C.c_double),
                ("time", C.c_double)]


class PlotParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("compositionRatio", "heatConductivityRatio", "specificHeatRatio",
                                          "waterDistributionCoeff", "stefan", "lewis", "darcy")] + \
               [("timescale", C.c_char_p), ("velocity_scale", C.c_char_p)]


_SIG = {
    "ml_h5_error": (C.c_char_p, []),
    "ml_h5_create": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(_vp)]),
    "ml_h5_open": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "ml_h5_close": (C.c_int, [_vp]),
    "ml_h5_dim": (C.c_int, [_vp]),
    "ml_h5_put_int": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_int]),
    "ml_h5_put_real": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_double]),
    "ml_h5_put_string": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_char_p]),
    "ml_h5_put_box": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _ip, _ip]),
    "ml_h5_get_int": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _ip]),
    "ml_h5_get_real": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _dp]),
    "ml_h5_get_string": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int]),
    "ml_h5_get_box": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _ip, _ip]),
    "ml_h5_put_boxes": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip]),
    "ml_h5_get_boxes": (C.c_int, [_vp, C.c_int, C.c_int, _ip, _ip, _ip]),
    "ml_h5_put_fabs": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, _ip, _ip, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(_dp)]),
    "ml_h5_fab_info": (C.c_int, [_vp, C.c_int, C.c_char_p, _ip, _ip]),
    "ml_h5_get_fabs": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, _ip, _ip, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(_dp)]),
    "ml_h5_put_faces": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, _ip, _ip, C.c_int, C.POINTER(_dp)]),
    "ml_h5_get_faces": (C.c_int, [_vp, C.c_int, C.c_char_p, C.c_int, _ip, _ip, C.c_int, C.POINTER(_dp)]),
    "ml_h5_write_checkpoint_header": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, _ip, _ip, C.c_int,
                                                C.POINTER(C.c_char_p)]),
    "ml_h5_write_checkpoint_level_header": (C.c_int, [_vp, C.c_int, C.POINTER(Level)]),
    "ml_h5_write_plot_header": (C.c_int, [_vp, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_char_p),
                                          C.POINTER(PlotParams)]),
    "ml_h5_write_plot_level_header": (C.c_int, [_vp, C.c_int, C.POINTER(Level)]),
    "ml_h5_read_level_header": (C.c_int, [_vp, C.c_int, C.POINTER(Level)]),
    "ml_h5_describe": (C.c_int, [_vp, C.c_char_p, C.c_int]),
}


def lib():
    """The loaded libmlio.so; raises when it was not built (no HDF5)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(IO_LIB):
            raise H5Error("libmlio.so is not built (python mushy-layer_amd/build.py; needs HDF5)")
        L = C.CDLL(IO_LIB)
        for name, (res, args) in _SIG.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc, what):
    if rc != H5_OK:
        raise H5Error("%s: %s (code %d)" % (what, lib().ml_h5_error().decode(), rc))


def _b(s):
    return s.encode() if isinstance(s, str) else s


def _iarr(v):
    return (C.c_int * len(v))(*v)


def _boxes_arrays(boxes):
    lo, hi = [], []
    for b_lo, b_hi in boxes:
        lo += list(b_lo) + [0] * (3 - len(b_lo))
        hi += list(b_hi) + [0] * (3 - len(b_hi))
    return _iarr(lo or [0]), _iarr(hi or [0])


def _ptrs(arrs):
    out = (_dp * max(1, len(arrs)))()
    for i, a in enumerate(arrs):
        assert a.dtype == np.float64 and a.flags.c_contiguous
        out[i] = a.ctypes.data_as(_dp)
    return out


def fab_shape(dim, lo, hi, ncomp, ghost=1):
    """Shape of a box's FAB (box grown by ghost, component outermost)."""
    ext = [hi[d] - lo[d] + 1 + 2 * ghost for d in range(dim)]
    return (ncomp,) + tuple(reversed(ext)) if dim == 3 else (ncomp, 1, ext[1], ext[0])


def face_fab_shape(dim, lo, hi, d, ncomp):
    ext = [hi[e] - lo[e] + 1 + (e == d) for e in range(dim)]
    return (ncomp,) + tuple(reversed(ext)) if dim == 3 else (ncomp, 1, ext[1], ext[0])


class H5File:
    """An open Chombo HDF5 file (ml_h5)."""

    def __init__(self, path, mode="r", dim=2):
        self.h = _vp()
        if mode == "w":
            _check(lib().ml_h5_create(_b(path), dim, C.byref(self.h)), "create " + path)
        else:
            _check(lib().ml_h5_open(_b(path), C.byref(self.h)), "open " + path)
        self.dim = lib().ml_h5_dim(self.h)

    def close(self):
        if self.h:
            _check(lib().ml_h5_close(self.h), "close")
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # header attributes
    def put(self, group, name, value):
        if isinstance(value, str):
            _check(lib().ml_h5_put_string(self.h, _b(group), _b(name), _b(value)), name)
        elif isinstance(value, (bool, int, np.integer)):
            _check(lib().ml_h5_put_int(self.h, _b(group), _b(name), int(value)), name)
        else:
            _check(lib().ml_h5_put_real(self.h, _b(group), _b(name), float(value)), name)

    def put_box(self, group, name, lo, hi):
        _check(lib().ml_h5_put_box(self.h, _b(group), _b(name), _iarr(list(lo) + [0]), _iarr(list(hi) + [0])), name)

    def get_int(self, group, name, default=None):
        v = C.c_int()
        rc = lib().ml_h5_get_int(self.h, _b(group), _b(name), C.byref(v))
        if rc == H5_ENOTFOUND and default is not None:
            return default
        _check(rc, name)
        return v.value

    def get_real(self, group, name, default=None):
        v = C.c_double()
        rc = lib().ml_h5_get_real(self.h, _b(group), _b(name), C.byref(v))
        if rc == H5_ENOTFOUND and default is not None:
            return default
        _check(rc, name)
        return v.value

    def get_string(self, group, name, default=None):
        buf = C.create_string_buffer(4096)
        rc = lib().ml_h5_get_string(self.h, _b(group), _b(name), buf, 4096)
        if rc == H5_ENOTFOUND and default is not None:
            return default
        _check(rc, name)
        return buf.value.decode()

    def get_box(self, group, name):
        lo, hi = (C.c_int * 3)(), (C.c_int * 3)()
        _check(lib().ml_h5_get_box(self.h, _b(group), _b(name), lo, hi), name)
        return [lo[d] for d in range(self.dim)], [hi[d] for d in range(self.dim)]

    # boxes and level data
    def put_boxes(self, level, boxes):
        lo, hi = _boxes_arrays(boxes)
        _check(lib().ml_h5_put_boxes(self.h, level, len(boxes), lo, hi), "boxes")

    def boxes(self, level):
        n = C.c_int()
        _check(lib().ml_h5_get_boxes(self.h, level, 0, C.byref(n), None, None), "boxes")
        lo, hi = (C.c_int * (3 * max(1, n.value)))(), (C.c_int * (3 * max(1, n.value)))()
        _check(lib().ml_h5_get_boxes(self.h, level, n.value, C.byref(n), lo, hi), "boxes")
        return [([lo[3 * b + d] for d in range(self.dim)], [hi[3 * b + d] for d in range(self.dim)])
                for b in range(n.value)]

    def put_fabs(self, level, name, boxes, fabs, ghost=1, out_ghost=0):
        lo, hi = _boxes_arrays(boxes)
        ncomp = fabs[0].shape[0]
        _check(lib().ml_h5_put_fabs(self.h, level, _b(name), len(boxes), lo, hi, ncomp, ghost, out_ghost,
                                    _ptrs(fabs)), "put " + name)

    def fab_info(self, level, name):
        nc, og = C.c_int(), C.c_int()
        _check(lib().ml_h5_fab_info(self.h, level, _b(name), C.byref(nc), C.byref(og)), name)
        return nc.value, og.value

    def get_fabs(self, level, name, boxes, c0=0, ncomp=None, ghost=1):
        if ncomp is None:
            ncomp = self.fab_info(level, name)[0] - c0
        fabs = [np.zeros(fab_shape(self.dim, lo, hi, ncomp, ghost)) for lo, hi in boxes]
        blo, bhi = _boxes_arrays(boxes)
        _check(lib().ml_h5_get_fabs(self.h, level, _b(name), len(boxes), blo, bhi, c0, ncomp, ghost, _ptrs(fabs)),
               "read " + name)
        return fabs

    def put_faces(self, level, name, boxes, faces):
        """faces[b][d]: face FAB of box b, direction d."""
        lo, hi = _boxes_arrays(boxes)
        flat = [faces[b][d] for b in range(len(boxes)) for d in range(self.dim)]
        _check(lib().ml_h5_put_faces(self.h, level, _b(name), len(boxes), lo, hi, flat[0].shape[0], _ptrs(flat)),
               "put " + name)

    def get_faces(self, level, name, boxes, ncomp=None):
        if ncomp is None:
            ncomp = self.fab_info(level, name)[0]
        faces = [[np.zeros(face_fab_shape(self.dim, lo, hi, d, ncomp)) for d in range(self.dim)] for lo, hi in boxes]
        flat = [faces[b][d] for b in range(len(boxes)) for d in range(self.dim)]
        blo, bhi = _boxes_arrays(boxes)
        _check(lib().ml_h5_get_faces(self.h, level, _b(name), len(boxes), blo, bhi, ncomp, _ptrs(flat)),
               "read " + name)
        return faces

    def level_header(self, level):
        L = Level()
        _check(lib().ml_h5_read_level_header(self.h, level, C.byref(L)), "level %d header" % level)
        d = self.dim
        return dict(lo=list(L.lo)[:d], hi=list(L.hi)[:d], periodic=list(L.periodic)[:d], ref_ratio=L.ref_ratio,
                    tag_buffer_size=L.tag_buffer_size, timestep_failed=L.timestep_failed, dx=L.dx, dt=L.dt,
                    time=L.time)

    def describe(self):
        n = lib().ml_h5_describe(self.h, None, 0)
        _check(0 if n >= 0 else n, "describe")
        buf = C.create_string_buffer(n + 1)
        lib().ml_h5_describe(self.h, buf, n + 1)
        return buf.value.decode().splitlines()


def _level_struct(h):
    L = Level()
    for d in range(len(h["lo"])):
        L.lo[d], L.hi[d] = h["lo"][d], h["hi"][d]
        L.periodic[d] = int(h.get("periodic", [0] * 3)[d])
    L.ref_ratio = h.get("ref_ratio", 2)
    L.tag_buffer_size = h.get("tag_buffer_size", 0)
    L.timestep_failed = h.get("timestep_failed", 0)
    L.dx, L.dt, L.time = h["dx"], h["dt"], h["time"]
    return L


def _names(names):
    arr = (C.c_char_p * max(1, len(names)))()
    for i, s in enumerate(names):
        arr[i] = _b(s)
    return arr


def read_level(path, level=0):
    """Header, boxes and variables of one level: {"header", "boxes",
    "root": {num_levels, iteration, time, components}, "fabs": {name: [FAB per
    box]}, "faces": {name: [[face FAB per direction] per box]}}."""
    with H5File(path) as f:
        root = {"num_levels": f.get_int("/", "num_levels"), "iteration": f.get_int("/", "iteration", -1),
                "time": f.get_real("/", "time", 0.0)}
        nc = f.get_int("/", "num_components", 0)
        root["components"] = [f.get_string("/", "component_%d" % i) for i in range(nc)]
        h = f.level_header(level)
        boxes = f.boxes(level)
        lines = f.describe()
        prefix = "/level_%d/" % level
        names = sorted({ln[len(prefix):].split(":datatype=0")[0] for ln in lines
                        if ln.startswith(prefix) and ":datatype=0 dataset" in ln})
        fabs, faces = {}, {}
        for nm in names:
            kind = f.get_string("level_%d/%s_attributes" % (level, nm), "objectType", "FArrayBox")
            if kind == "FArrayBox":
                fabs[nm] = f.get_fabs(level, nm, boxes)
            else:
                faces[nm] = f.get_faces(level, nm, boxes)
        return {"root": root, "header": h, "boxes": boxes, "fabs": fabs, "faces": faces, "dim": f.dim}


def domain_array(lev, name, comps=None):
    """A variable of a level whose boxes cover its problem domain, gathered into
    one array of this package's layout (ghost cells zero)."""
    h, dim = lev["header"], lev["dim"]
    lo, hi = h["lo"], h["hi"]
    n = [hi[d] - lo[d] + 1 for d in range(dim)]
    fabs = lev["fabs"][name]
    ncomp = fabs[0].shape[0]
    out = np.zeros(fab_shape(dim, [0] * dim, [x - 1 for x in n], ncomp, 1))
    covered = 0
    for (blo, bhi), fab in zip(lev["boxes"], fabs):
        s = [slice(blo[d] - lo[d] + 1, bhi[d] - lo[d] + 2) for d in range(dim)]
        v = fab[(slice(None),) + ((slice(1, -1),) * 3 if dim == 3 else (slice(None), slice(1, -1), slice(1, -1)))]
        if dim == 3:
            out[:, s[2], s[1], s[0]] = v
        else:
            out[:, :, s[1], s[0]] = v
        covered += int(np.prod([bhi[d] - blo[d] + 1 for d in range(dim)]))
    if covered != int(np.prod(n)):
        raise H5Error("%s: the level's boxes do not cover its domain" % name)
    return out if comps is None else out[comps]


def _single_box(n):
    return [([0] * len(n), [x - 1 for x in n])]


def _fab_of(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def write_checkpoint(path, n, dx, dt, time, iteration, periodic, scalars, vectors, faces=None,
                     tag_buffer_size=0, ref_ratio=2):
    """Single-level checkpoint in the reference's layout
    (writeCheckpointHeader / writeCheckpointLevel): scalars {name: 1-component
    array}, vectors {name: dim-component array} (components named "x<name>",
    "y<name>"[, "z<name>"] in the header), faces {name: [face array per
    direction]} (e.g. "advVel"); arrays in this package's layout."""
    dim = len(n)
    names = list(scalars) + ["%s%s" % ("xyz"[d], v) for v in vectors for d in range(dim)]
    boxes = _single_box(n)
    with H5File(path, "w", dim) as f:
        _check(lib().ml_h5_write_checkpoint_header(f.h, 1, int(iteration), float(time), None, None, len(names),
                                                   _names(names)), "checkpoint header")
        h = dict(lo=[0] * dim, hi=[x - 1 for x in n], periodic=list(periodic), dx=dx, dt=dt, time=time,
                 tag_buffer_size=tag_buffer_size, ref_ratio=ref_ratio)
        _check(lib().ml_h5_write_checkpoint_level_header(f.h, 0, C.byref(_level_struct(h))), "level header")
        f.put_boxes(0, boxes)
        for nm, a in list(scalars.items()) + list(vectors.items()):
            f.put_fabs(0, nm, boxes, [_fab_of(a)], ghost=1, out_ghost=0)
        for nm, fl in (faces or {}).items():
            f.put_faces(0, nm, boxes, [[_fab_of(x) for x in fl]])


def write_plotfile(path, n, dx, dt, time, iteration, fields, params, timestep_failed=0, ref_ratio=2):
    """Single-level plot file in the reference's layout (writePlotHeader /
    writePlotLevel): fields is an ordered {component name: 1-component
    array}; params the MushyLayerParams values written to the header (C, k,
    cp, pc, S, L, D, timescale, velocity_scale)."""
    dim = len(n)
    names = list(fields)
    boxes = _single_box(n)
    data = np.concatenate([_fab_of(fields[k]) for k in names], axis=0)
    P = PlotParams()
    P.compositionRatio = params.get("C", 0.0)
    P.heatConductivityRatio = params.get("k", 1.0)
    P.specificHeatRatio = params.get("cp", 1.0)
    P.waterDistributionCoeff = params.get("pc", 0.0)
    P.stefan = params.get("S", 0.0)
    P.lewis = params.get("L", 0.0)
    P.darcy = params.get("D", 0.0)
    ts, vs = _b(params.get("timescale", "Diffusive")), _b(params.get("velocity_scale", "Advective"))
    P.timescale, P.velocity_scale = ts, vs
    with H5File(path, "w", dim) as f:
        _check(lib().ml_h5_write_plot_header(f.h, 1, int(iteration), float(time), len(names), _names(names),
                                             C.byref(P)), "plot header")
        h = dict(lo=[0] * dim, hi=[x - 1 for x in n], dx=dx, dt=dt, time=time, timestep_failed=timestep_failed,
                 ref_ratio=ref_ratio)
        _check(lib().ml_h5_write_plot_level_header(f.h, 0, C.byref(_level_struct(h))), "level header")
        f.put_boxes(0, boxes)
        f.put_fabs(0, "data", boxes, [np.ascontiguousarray(data)], ghost=1, out_ghost=0)
