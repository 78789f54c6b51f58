"""Python binding of libmlhip.so — the MI355X implementation of the
single-level implicit mushy-layer solve (SOFTBALL's AMRNonLinearMultiCompOp
FAS multigrid and the variable-coefficient Helmholtz/Poisson operators).

The package directory name contains a hyphen; import it with
``importlib.import_module("mushy-layer_amd")``.

The compute path is the HIP library only: if ``libmlhip.so`` is missing or
the GPU is absent, the calls raise — there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmlhip.so")

_LIB = None


class MLError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "stefan", "compositionRatio", "waterDistributionCoeff", "specificHeatRatio",
        "heatConductivityRatio", "thetaEutectic", "ThetaEutectic", "heatDiffusionCoeff",
        "saltDiffusionCoeff")]


class BC(C.Structure):
    _fields_ = [("type", C.c_int * 2 * 2 * 3), ("val", C.c_double * 2 * 2 * 3), ("apply", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("nu1", C.c_int), ("nu2", C.c_int), ("bottom", C.c_int), ("max_iter", C.c_int),
                ("eps", C.c_double), ("hang", C.c_double), ("norm_thresh", C.c_double),
                ("precision", C.c_int), ("flags", C.c_int),
                ("bottom_solver", C.c_int), ("bottom_max_iter", C.c_int)]


BOTTOM_RELAX, BOTTOM_BICGSTAB = 0, 1   # SolveOpts.bottom_solver (linear operator solves)


SOLVE_HOST_DRIVER = 1
SOLVE_NO_LOOKAHEAD = 2   # one iteration enqueued at a time (per-launch measurements)
SOLVE_NO_GRAPH = 4   # launch every kernel from the host (no captured iterations)
PRECISION_FP64, PRECISION_FP32_REFINED = 0, 1
KT_VC_RELAX = 1   # ml_ctx_timing_start kernel family: tiled relaxation of the linear VC operator
KT_HC_RELAX = 2   # ... of the (H, Theta) operator with the closure
KT_BATCH_RELAX = 3   # batched form: FP32 relaxation of every instance's level


class FlowParams(C.Structure):
    _fields_ = [("buoyancyTCoeff", C.c_double), ("buoyancySCoeff", C.c_double), ("porosityLimit", C.c_double),
                ("scalePressure", C.c_int)]


class AdvectParams(C.Structure):
    _fields_ = [("advectionCoeff", C.c_double), ("type", C.c_int * 2 * 2 * 3), ("val", C.c_double * 2 * 2 * 3)]


# advection boundary types, AdvectIBC's numbering (BCutil/AdvectIBC.H:137-144)
ADV_DIRICHLET, ADV_EXTRAP, ADV_INFLOW_OUTFLOW, ADV_PLUME_INFLOW, ADV_NEUMANN = 0, 1, 2, 3, 4


def make_advect_params(hc_bc=None, coeff=1.0, spec=None):
    """Advection BCs of (theta, Theta_l): from the HC operator's BC as
    PhysBCUtil::scalarTraceTSl_IBC converts them (Dirichlet -> Dirichlet,
    Neumann -> Neumann, same values; BCutil/PhysBCUtil.cpp:3380-3434,
    4059-4087), else extrapolation; spec {(dir, side, comp): (type, value)}
    overrides.  coeff: m_advectionCoeff."""
    a = AdvectParams()
    a.advectionCoeff = coeff
    for d in range(3):
        for s in range(2):
            for c in range(2):
                t, v = ADV_EXTRAP, 0.0
                if hc_bc is not None:
                    t = ADV_DIRICHLET if hc_bc.type[d][s][c] == 0 else ADV_NEUMANN
                    v = hc_bc.val[d][s][c]
                a.type[d][s][c] = t
                a.val[d][s][c] = v
    for (d, s, c), (t, v) in (spec or {}).items():
        a.type[d][s][c] = t
        a.val[d][s][c] = v
    return a


PERM_PURE_FLUID, PERM_CUBIC, PERM_KOZENY_CARMAN, PERM_LOG, PERM_POROSITY = 0, 1, 2, 3, 4
PERM_LAWS = {"pure_fluid": PERM_PURE_FLUID, "cubic": PERM_CUBIC, "kozeny_carman": PERM_KOZENY_CARMAN,
             "log": PERM_LOG, "porosity": PERM_POROSITY}


def make_flow_params(buoyancyTCoeff=0.0, buoyancySCoeff=0.0, porosityLimit=0.05, scalePressure=1):
    f = FlowParams()
    f.buoyancyTCoeff = buoyancyTCoeff
    f.buoyancySCoeff = buoyancySCoeff
    f.porosityLimit = porosityLimit
    f.scalePressure = scalePressure
    return f


class SolveStats(C.Structure):
    _fields_ = [("exit_status", C.c_int), ("iterations", C.c_int),
                ("initial_residual", C.c_double), ("final_residual", C.c_double)]


_vp = C.c_void_p
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)

# name -> (restype, argtypes); mirrors include/mlhip.h
SIGNATURES = {
    "ml_ctx_create": (C.c_int, [C.c_int, C.c_uint, C.POINTER(_vp)]),
    "ml_ctx_destroy": (C.c_int, [_vp]),
    "ml_last_error": (C.c_char_p, [_vp]),
    "ml_ctx_synchronize": (C.c_int, [_vp]),
    "ml_ctx_stream": (_vp, [_vp]),
    "ml_ctx_event_record": (C.c_int, [_vp, C.c_int]),
    "ml_ctx_event_elapsed": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "ml_ctx_timing_start": (C.c_int, [_vp, C.c_int, C.c_int]),
    "ml_ctx_timing_stop": (C.c_int, [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                      C.POINTER(C.c_longlong)]),
    "ml_buf_alloc": (C.c_int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    "ml_buf_free": (C.c_int, [_vp, _vp]),
    "ml_buf_upload": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "ml_buf_download": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "ml_buf_copy": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "ml_cell_count": (C.c_size_t, [C.c_int, _ip]),
    "ml_face_count": (C.c_size_t, [C.c_int, _ip, C.c_int]),
    "ml_mg_depth": (C.c_int, [C.c_int, _ip]),
    "ml_default_opts": (None, [C.POINTER(SolveOpts)]),
    "ml_hc_op_create": (C.c_int, [_vp, C.c_int, _ip, _ip, C.c_double, _ip, C.POINTER(Params),
                                  C.POINTER(BC), C.POINTER(_vp)]),
    "ml_hc_op_destroy": (C.c_int, [_vp]),
    "ml_hc_level_info": (C.c_int, [_vp, C.c_int, _ip, _dp]),
    "ml_hc_num_levels": (C.c_int, [_vp]),
    "ml_hc_set_coefs": (C.c_int, [_vp, _vp, C.POINTER(_vp), C.c_double, C.c_double]),
    "ml_hc_set_alpha_beta": (C.c_int, [_vp, C.c_double, C.c_double]),
    "ml_hc_diagonal_scale": (C.c_int, [_vp, C.c_int, _vp]),
    "ml_hc_compute_diffused_var": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
    "ml_hc_relax": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
    "ml_hc_relax_noalloc": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
    "ml_hc_jacobi": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
    "ml_hc_precond": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "ml_hc_get_flux": (C.c_int, [_vp, C.c_int, _vp, C.POINTER(_vp)]),
    "ml_hc_residual": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int]),
    "ml_hc_apply_op": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
    "ml_hc_restrict_residual": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "ml_hc_restrict_r": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "ml_hc_prolong_increment": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "ml_hc_solve": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "ml_hc_solve_be": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.POINTER(SolveOpts),
                                 C.POINTER(SolveStats)]),
    "ml_vc_op_create": (C.c_int, [_vp, C.c_int, _ip, _ip, C.c_double, _ip, C.POINTER(BC), C.POINTER(_vp)]),
    "ml_vc_op_destroy": (C.c_int, [_vp]),
    "ml_vc_num_levels": (C.c_int, [_vp]),
    "ml_vc_set_coefs": (C.c_int, [_vp, _vp, C.POINTER(_vp), _vp, C.c_double, C.c_double]),
    "ml_vc_set_alpha_beta": (C.c_int, [_vp, C.c_double, C.c_double]),
    "ml_vc_relax": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
    "ml_vc_residual": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_int]),
    "ml_vc_apply_op": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int]),
    "ml_vc_precond": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "ml_vc_dot": (C.c_int, [_vp, C.c_int, _vp, _vp, C.POINTER(C.c_double)]),
    "ml_vc_restrict_residual": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp]),
    "ml_vc_prolong_increment": (C.c_int, [_vp, C.c_int, _vp, _vp]),
    "ml_vc_solve": (C.c_int, [_vp, _vp, _vp, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "ml_vc_solve_be": (C.c_int, [_vp, _vp, _vp, _vp, C.c_double, C.POINTER(SolveOpts),
                                 C.POINTER(SolveStats)]),
    "ml_vc_solve_be_many": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.c_double,
                                      C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "ml_vc_darcy_ccoef": (C.c_int, [_vp, _vp, C.c_double, _vp]),
    "ml_vc_porosity_faces": (C.c_int, [_vp, _vp, C.c_double, C.POINTER(_vp)]),
    "ml_hc_batch_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp)]),
    "ml_hc_batch_destroy": (C.c_int, [_vp]),
    "ml_hc_batch_size": (C.c_int, [_vp]),
    "ml_hc_batch_memory": (C.c_int, [_vp, C.POINTER(C.c_size_t), _ip, _ip]),
    "ml_hc_batch_set_tile": (C.c_int, [_vp, C.c_int]),
    "ml_hc_batch_tile": (C.c_int, [_vp, _ip, _ip, _ip, _ip, C.POINTER(C.c_float), C.c_int]),
    "ml_hc_batch_step": (C.c_int, [_vp, C.POINTER(_vp), C.c_double, C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "ml_hc_batch_step_src": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.c_double, C.POINTER(SolveOpts),
                                       C.POINTER(SolveStats)]),
    "ml_vc_batch_create": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.c_int, C.POINTER(_vp)]),
    "ml_vc_set_coefs_many": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "ml_hc_set_coefs_many": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "ml_hc_db_advect_src_many": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.c_double,
                                           C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp),
                                           C.POINTER(_vp), C.POINTER(_vp)]),
    "ml_hc_db_coefs_many": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "ml_vc_db_ustar_many": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(C.c_double), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "ml_vc_db_project_many": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.c_double, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "ml_vc_db_correct_many": (C.c_int, [_vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp), C.c_double, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "ml_vc_batch_destroy": (C.c_int, [_vp]),
    "ml_vc_batch_solve": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "ml_vc_batch_solve_be": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.c_double,
                                       C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "ml_hc_flux_source": (C.c_int, [_vp, _vp, _vp]),
    "ml_hc_adjoint_step": (C.c_int, [_vp, _vp, _vp, C.c_double, _vp, _vp, _vp, _dp, C.POINTER(SolveOpts),
                                     C.POINTER(SolveStats)]),
    "ml_hc_porosity_vjp": (C.c_int, [_vp, _vp, _vp, _vp, _dp]),
    "ml_hc_porosity": (C.c_int, [_vp, _vp, _vp]),
    "ml_hc_assemble_bcoef": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "ml_hc_enthalpy_vars": (C.c_int, [_vp, _vp, _vp]),
    "ml_vc_ustar_src": (C.c_int, [_vp, C.POINTER(FlowParams), _vp, _vp, _vp, C.POINTER(_vp)]),
    "ml_vc_cc_project_rhs": (C.c_int, [_vp, C.POINTER(FlowParams), C.c_double, _vp, _vp, _vp, C.POINTER(_vp),
                                       C.POINTER(_vp), _vp]),
    "ml_vc_pressure_faces": (C.c_int, [_vp, _vp, _vp, C.POINTER(_vp)]),
    "ml_vc_cc_correct": (C.c_int, [_vp, C.POINTER(FlowParams), C.c_double, _vp, _vp, _vp, C.POINTER(_vp)]),
    "ml_vc_darcy_ccoef_law": (C.c_int, [_vp, _vp, C.c_double, C.c_int, C.c_double, _vp]),
    "ml_hc_diffusive_src": (C.c_int, [_vp, _vp, _vp]),
    "ml_hc_advect_src": (C.c_int, [_vp, C.POINTER(AdvectParams), C.c_double, _vp, C.POINTER(_vp), _vp, _vp,
                                   C.POINTER(_vp)]),
    "ml_vc_face_velocity": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "ml_level_gather": (C.c_int, [_vp, C.c_int, _ip, _ip, C.c_int, C.c_int, _ip, _ip, C.POINTER(_vp), _vp]),
    "ml_level_scatter": (C.c_int, [_vp, C.c_int, _ip, _ip, C.c_int, C.c_int, _ip, _ip, _vp, C.POINTER(_vp)]),
    "ml_level_gather_faces": (C.c_int, [_vp, C.c_int, _ip, C.c_int, C.c_int, _ip, _ip, C.c_int, C.POINTER(_vp),
                                        _vp]),
    "ml_norm_max": (C.c_int, [_vp, C.c_int, _ip, C.c_int, _vp, _dp]),
    "ml_vec_axby": (C.c_int, [_vp, C.c_size_t, _vp, _vp, _vp, C.c_double, C.c_double]),
    "ml_vec_incr": (C.c_int, [_vp, C.c_size_t, _vp, _vp, C.c_double]),
    "ml_vec_mult": (C.c_int, [_vp, C.c_size_t, _vp, _vp]),
    "ml_vec_set": (C.c_int, [_vp, C.c_size_t, _vp, C.c_double]),
}


def lib():
    """Load libmlhip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise MLError("libmlhip.so not built: run python mushy-layer_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _iarr(v, n=3, fill=0):
    a = (C.c_int * n)()
    for i in range(n):
        a[i] = int(v[i]) if i < len(v) else fill
    return a


def default_opts(**kw):
    o = SolveOpts()
    lib().ml_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def make_params(**kw):
    p = Params()
    p.stefan = kw.get("stefan", 5.0)
    p.compositionRatio = kw.get("compositionRatio", 2.0)
    p.waterDistributionCoeff = kw.get("waterDistributionCoeff", 1e-5)
    p.specificHeatRatio = kw.get("specificHeatRatio", 1.0)
    p.heatConductivityRatio = kw.get("heatConductivityRatio", 1.0)
    p.thetaEutectic = kw.get("thetaEutectic", 0.0)
    p.ThetaEutectic = kw.get("ThetaEutectic", 0.0)
    p.heatDiffusionCoeff = kw.get("heatDiffusionCoeff", 1.0)
    p.saltDiffusionCoeff = kw.get("saltDiffusionCoeff", 1.0 / kw.get("lewis", 200.0))
    return p


def make_bc(spec=None, apply=1):
    """spec: {(dir, side, comp): (type, value)}; unspecified entries Neumann 0."""
    b = BC()
    for d in range(3):
        for s in range(2):
            for c in range(2):
                b.type[d][s][c] = 1
                b.val[d][s][c] = 0.0
    for (d, s, c), (t, v) in (spec or {}).items():
        b.type[d][s][c] = t
        b.val[d][s][c] = v
    b.apply = apply
    return b


class Context:
    """One device context (a HIP stream) of libmlhip.so."""

    def __init__(self, device=0):
        self.lib = lib()
        h = _vp()
        rc = self.lib.ml_ctx_create(device, 0, C.byref(h))
        if rc != 0:
            raise MLError("ml_ctx_create failed (%d): is a GPU visible?" % rc)
        self.h = h
        self._bufs = []

    def check(self, rc, what):
        if rc != 0:
            msg = self.lib.ml_last_error(self.h)
            raise MLError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def alloc(self, nbytes):
        p = _vp()
        self.check(self.lib.ml_buf_alloc(self.h, nbytes, C.byref(p)), "ml_buf_alloc")
        return p.value

    def free(self, p):
        self.check(self.lib.ml_buf_free(self.h, p), "ml_buf_free")

    def upload(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        self.check(self.lib.ml_buf_upload(self.h, dptr, arr.ctypes.data, arr.nbytes), "ml_buf_upload")

    def download(self, dptr, arr):
        assert arr.flags.c_contiguous
        self.check(self.lib.ml_buf_download(self.h, arr.ctypes.data, dptr, arr.nbytes), "ml_buf_download")
        return arr

    def sync(self):
        self.check(self.lib.ml_ctx_synchronize(self.h), "ml_ctx_synchronize")

    def record(self, slot):
        """Record HIP event `slot` on the library stream."""
        self.check(self.lib.ml_ctx_event_record(self.h, slot), "ml_ctx_event_record")

    def elapsed_ms(self, a, b):
        ms = C.c_float()
        self.check(self.lib.ml_ctx_event_elapsed(self.h, a, b, C.byref(ms)), "ml_ctx_event_elapsed")
        return ms.value

    def timing_start(self, kernel, max_launches):
        """Bracket the launches of a kernel family (KT_*) with HIP events."""
        self.check(self.lib.ml_ctx_timing_start(self.h, kernel, max_launches), "ml_ctx_timing_start")

    def timing_stop(self):
        """(timed launches, summed ms, summed algorithmic bytes, untimed
        launches) since timing_start."""
        n, ms, b, missed = C.c_longlong(), C.c_double(), C.c_double(), C.c_longlong()
        self.check(self.lib.ml_ctx_timing_stop(self.h, C.byref(n), C.byref(ms), C.byref(b), C.byref(missed)),
                   "ml_ctx_timing_stop")
        return n.value, ms.value, b.value, missed.value

    def close(self):
        if self.h:
            self.lib.ml_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A device buffer holding a numpy-shaped float64 array."""

    def __init__(self, ctx, shape, dtype=np.float64, host=None):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = ctx.alloc(self.nbytes)
        if host is not None:
            self.set(host)
        else:
            ctx.check(ctx.lib.ml_vec_set(ctx.h, self.nbytes // 8, self.ptr, 0.0), "ml_vec_set")

    def set(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype).reshape(self.shape)
        self.ctx.upload(self.ptr, host)

    def get(self):
        out = np.empty(self.shape, self.dtype)
        return self.ctx.download(self.ptr, out)

    def comp_ptr(self, c):
        """Device address of component c of a cell-layout array."""
        return self.ptr + c * (self.nbytes // self.shape[0])

    def free(self):
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = None


def _ptrs(xs):
    arr = (_vp * max(1, len(xs)))()
    for i, x in enumerate(xs):
        arr[i] = x.ptr if x is not None else None
    return arr


def _doubles(v):
    return (C.c_double * max(1, len(v)))(*v)


def vc_set_coefs_many(ctx, ops, a, b, c, alpha, beta):
    """VCOp.set_coefs of many operators in a few launches (ml_vc_set_coefs_many):
    a, c: lists of DeviceArray or None (or None for all); b: list of lists of
    face arrays; alpha, beta: scalars or lists."""
    n = len(ops)
    alpha = alpha if isinstance(alpha, (list, tuple)) else [alpha] * n
    beta = beta if isinstance(beta, (list, tuple)) else [beta] * n
    hs = (_vp * max(1, n))(*[op.h for op in ops])
    ctx.check(ctx.lib.ml_vc_set_coefs_many(ctx.h, n, hs, _ptrs(a) if a is not None else None,
                                           _ptrs([x for bs in b for x in bs]),
                                           _ptrs(c) if c is not None else None, _doubles(alpha), _doubles(beta)),
              "ml_vc_set_coefs_many")


def hc_set_coefs_many(ctx, ops, a, b, alpha, beta):
    """HCOp.set_coefs of many operators in a few launches (ml_hc_set_coefs_many)."""
    n = len(ops)
    alpha = alpha if isinstance(alpha, (list, tuple)) else [alpha] * n
    beta = beta if isinstance(beta, (list, tuple)) else [beta] * n
    hs = (_vp * max(1, n))(*[op.h for op in ops])
    ctx.check(ctx.lib.ml_hc_set_coefs_many(ctx.h, n, hs, _ptrs(a) if a is not None else None,
                                           _ptrs([x for bs in b for x in bs]), _doubles(alpha), _doubles(beta)),
              "ml_hc_set_coefs_many")


def hc_db_advect_src_many(ctx, ops, aps, dt, phi, W, u, dsrc, u_face, src):
    """The advective source of a Darcy-Brinkman step for many operators
    (ml_hc_db_advect_src_many): per operator i the diffusive source of phi[i]
    into dsrc[i], face velocities of u[i] (list of per-direction arrays) into
    u_face[i], src[i] = 0 and the advective source of W[i] (DeviceArray or
    device address).  aps: AdvectParams per operator."""
    n = len(ops)
    hs = (_vp * max(1, n))(*[op.h for op in ops])
    ap = (_vp * max(1, n))(*[C.cast(C.pointer(a), _vp) for a in aps])
    w = (_vp * max(1, n))(*[x.ptr if hasattr(x, "ptr") else x for x in W])
    ctx.check(ctx.lib.ml_hc_db_advect_src_many(ctx.h, n, hs, ap, dt, _ptrs(phi), w, _ptrs([x for us in u for x in us]),
                                               _ptrs(dsrc), _ptrs([x for fs in u_face for x in fs]), _ptrs(src)),
              "ml_hc_db_advect_src_many")


def _hs(ops):
    return (_vp * max(1, len(ops)))(*[op.h for op in ops])


def _structs(xs):
    return (_vp * max(1, len(xs)))(*[C.cast(C.pointer(x), _vp) for x in xs])


def _flat(xss):
    return _ptrs([x for xs in xss for x in xs])


def hc_db_coefs_many(ctx, ops, phi, vars_, b):
    """Enthalpy variables of phi[i] into vars_[i] and the HC face coefficients
    b[i] of their porosity, for many operators (ml_hc_db_coefs_many)."""
    ctx.check(ctx.lib.ml_hc_db_coefs_many(ctx.h, len(ops), _hs(ops), _ptrs(phi), _ptrs(vars_), _flat(b)),
              "ml_hc_db_coefs_many")


def vc_db_ustar_many(ctx, hops, uops, fps, darcy, laws, hele_shaw, phi, vold, vnew, pi, c, src, u, uold):
    """Enthalpy variables at t^{n+1}, Darcy coefficients, U* sources and the
    saved velocities of many Darcy-Brinkman instances (ml_vc_db_ustar_many)."""
    n = len(uops)
    law = (C.c_int * max(1, n))(*[PERM_LAWS.get(x, x) for x in laws])
    ctx.check(ctx.lib.ml_vc_db_ustar_many(ctx.h, n, _hs(hops), _hs(uops), _structs(fps), _doubles(darcy), law,
                                          _doubles(hele_shaw), _ptrs(phi), _ptrs(vold), _ptrs(vnew), _ptrs(pi),
                                          _ptrs(c), _flat(src), _flat(u), _flat(uold)), "ml_vc_db_ustar_many")


def vc_db_project_many(ctx, uops, fps, dt, vold, vnew, pi, ustar, u, rhs, bp):
    """Projection right-hand sides, pressure faces and pi = 0 for many
    instances (ml_vc_db_project_many)."""
    ctx.check(ctx.lib.ml_vc_db_project_many(ctx.h, len(uops), _hs(uops), _structs(fps), dt, _ptrs(vold), _ptrs(vnew),
                                            _ptrs(pi), _flat(ustar), _flat(u), _ptrs(rhs), _flat(bp)),
              "ml_vc_db_project_many")


def vc_db_correct_many(ctx, uops, fps, dt, vold, vnew, pi, u):
    """Projection corrections of many instances (ml_vc_db_correct_many)."""
    ctx.check(ctx.lib.ml_vc_db_correct_many(ctx.h, len(uops), _hs(uops), _structs(fps), dt, _ptrs(vold), _ptrs(vnew),
                                            _ptrs(pi), _flat(u)), "ml_vc_db_correct_many")


def cell_shape(n, ncomp):
    if len(n) == 3:
        return (ncomp, n[2] + 2, n[1] + 2, n[0] + 2)
    return (ncomp, 1, n[1] + 2, n[0] + 2)


def face_shape(n, d, ncomp):
    m = list(n) + [1] * (3 - len(n))
    m[d] += 1
    return (ncomp, m[2], m[1], m[0])


def valid(n, x):
    """Valid-cell view (ghost ring removed) of a host array in cell layout."""
    if len(n) == 3:
        return x[:, 1:-1, 1:-1, 1:-1]
    return x[:, :, 1:-1, 1:-1]


def _boxes(boxes):
    nb = len(boxes)
    lo = (C.c_int * (3 * nb))()
    hi = (C.c_int * (3 * nb))()
    for b, (l, h) in enumerate(boxes):
        for d in range(3):
            lo[3 * b + d] = l[d] if d < len(l) else 0
            hi[3 * b + d] = h[d] if d < len(h) else 0
    return nb, lo, hi


def level_gather(ctx, n, periodic, ncomp, boxes, fabs, level):
    """The domain array `level` from the boxes' FABs; boxes: [(lo, hi), ...]
    inclusive valid-cell bounds (a DisjointBoxLayout)."""
    nb, lo, hi = _boxes(boxes)
    ctx.check(ctx.lib.ml_level_gather(ctx.h, len(n), _iarr(n), _iarr(periodic), ncomp, nb, lo, hi,
                                      (_vp * nb)(*[f.ptr for f in fabs]), level.ptr), "ml_level_gather")


def level_scatter(ctx, n, periodic, ncomp, boxes, level, fabs):
    """The boxes' FABs (valid cells and exchange ghosts) from `level`."""
    nb, lo, hi = _boxes(boxes)
    ctx.check(ctx.lib.ml_level_scatter(ctx.h, len(n), _iarr(n), _iarr(periodic), ncomp, nb, lo, hi, level.ptr,
                                       (_vp * nb)(*[f.ptr for f in fabs])), "ml_level_scatter")


def level_gather_faces(ctx, n, ncomp, boxes, d, fabs, level_faces):
    nb, lo, hi = _boxes(boxes)
    ctx.check(ctx.lib.ml_level_gather_faces(ctx.h, len(n), _iarr(n), ncomp, nb, lo, hi, d,
                                            (_vp * nb)(*[f.ptr for f in fabs]), level_faces.ptr),
              "ml_level_gather_faces")


class HCOp:
    """AMRNonLinearMultiCompOp (+ its factory's MG hierarchy) on the GPU."""

    def __init__(self, ctx, n, dx, periodic, params, bc):
        self.ctx = ctx
        self.n = list(n)
        self.dim = len(n)
        self.dx = dx
        h = _vp()
        lo = _iarr([0] * 3)
        hi = _iarr([x - 1 for x in n])
        ctx.check(ctx.lib.ml_hc_op_create(ctx.h, self.dim, lo, hi, dx, _iarr(periodic), C.byref(params),
                                          C.byref(bc), C.byref(h)), "ml_hc_op_create")
        self.h = h
        self.bc = bc
        self.nlev = ctx.lib.ml_hc_num_levels(h)

    def level_n(self, depth):
        n = (C.c_int * 3)()
        dx = C.c_double()
        self.ctx.check(self.ctx.lib.ml_hc_level_info(self.h, depth, n, C.byref(dx)), "ml_hc_level_info")
        return [n[d] for d in range(self.dim)], dx.value

    def set_coefs(self, a, b, alpha, beta):
        arr = (_vp * 3)()
        for d in range(self.dim):
            arr[d] = b[d].ptr
        self.ctx.check(self.ctx.lib.ml_hc_set_coefs(self.h, a.ptr if a is not None else None, arr, alpha, beta),
                       "ml_hc_set_coefs")

    def set_alpha_beta(self, alpha, beta):
        self.ctx.check(self.ctx.lib.ml_hc_set_alpha_beta(self.h, alpha, beta), "ml_hc_set_alpha_beta")

    def relax(self, depth, phi, rhs, iters):
        self.ctx.check(self.ctx.lib.ml_hc_relax(self.h, depth, phi.ptr, rhs.ptr, iters), "ml_hc_relax")

    def jacobi(self, depth, phi, rhs, iters):
        self.ctx.check(self.ctx.lib.ml_hc_jacobi(self.h, depth, phi.ptr, rhs.ptr, iters), "ml_hc_jacobi")

    def precond(self, depth, phi, rhs):
        self.ctx.check(self.ctx.lib.ml_hc_precond(self.h, depth, phi.ptr, rhs.ptr), "ml_hc_precond")

    def get_flux(self, depth, phi, flux):
        arr = (_vp * 3)()
        for d in range(self.dim):
            arr[d] = flux[d].ptr
        self.ctx.check(self.ctx.lib.ml_hc_get_flux(self.h, depth, phi.ptr, arr), "ml_hc_get_flux")

    def relax_noalloc(self, depth, phi, rhs, iters):
        """relax() with the partner array in the library's static scratch
        (no device allocation)."""
        self.ctx.check(self.ctx.lib.ml_hc_relax_noalloc(self.h, depth, phi.ptr, rhs.ptr, iters),
                       "ml_hc_relax_noalloc")

    def residual(self, depth, res, phi, rhs, homogeneous=0):
        self.ctx.check(self.ctx.lib.ml_hc_residual(self.h, depth, res.ptr, phi.ptr, rhs.ptr, homogeneous),
                       "ml_hc_residual")

    def apply_op(self, depth, lhs, phi, homogeneous=0):
        self.ctx.check(self.ctx.lib.ml_hc_apply_op(self.h, depth, lhs.ptr, phi.ptr, homogeneous),
                       "ml_hc_apply_op")

    def diffused_var(self, depth, phi, out, homogeneous=0):
        self.ctx.check(self.ctx.lib.ml_hc_compute_diffused_var(self.h, depth, phi.ptr, out.ptr, homogeneous),
                       "ml_hc_compute_diffused_var")

    def restrict_residual(self, depth, resc, phi, rhs):
        self.ctx.check(self.ctx.lib.ml_hc_restrict_residual(self.h, depth, resc.ptr, phi.ptr, rhs.ptr),
                       "ml_hc_restrict_residual")

    def restrict_r(self, depth, phic, phif):
        self.ctx.check(self.ctx.lib.ml_hc_restrict_r(self.h, depth, phic.ptr, phif.ptr), "ml_hc_restrict_r")

    def prolong_increment(self, depth, phif, corrc):
        self.ctx.check(self.ctx.lib.ml_hc_prolong_increment(self.h, depth, phif.ptr, corrc.ptr),
                       "ml_hc_prolong_increment")

    def enthalpy_vars(self, phi, out):
        """out (4 components) = (theta, Theta_l, Theta_s, porosity) of phi."""
        self.ctx.check(self.ctx.lib.ml_hc_enthalpy_vars(self.h, phi.ptr, out.ptr), "ml_hc_enthalpy_vars")

    def porosity(self, phi, chi):
        self.ctx.check(self.ctx.lib.ml_hc_porosity(self.h, phi.ptr, chi.ptr), "ml_hc_porosity")

    def diffusive_src(self, phi, out):
        """out = the operator with alpha = 0, beta = 1 applied to phi
        (computeScalDiffusion)."""
        self.ctx.check(self.ctx.lib.ml_hc_diffusive_src(self.h, phi.ptr, out.ptr), "ml_hc_diffusive_src")

    def advect_src(self, ap, dt, W, u_face, S, src, flux=None):
        """src -= div(W_half u_face coeff) (computeScalarAdvectiveSrcHC).  W: a
        DeviceArray or device address of the 2-component (theta, Theta_l)
        field; S: the diffusive source or None; flux: optional face outputs."""
        uf = (_vp * 3)()
        for d, x in enumerate(u_face):
            uf[d] = x.ptr
        fl = None
        if flux is not None:
            fl = (_vp * 3)()
            for d, x in enumerate(flux):
                fl[d] = x.ptr
        w = W.ptr if hasattr(W, "ptr") else W
        self.ctx.check(self.ctx.lib.ml_hc_advect_src(self.h, C.byref(ap), dt, w, uf, S.ptr if S is not None else None,
                                                     src.ptr, fl), "ml_hc_advect_src")

    def assemble_bcoef(self, chi, b):
        """b (face arrays) from the cell porosity chi (a DeviceArray or a
        device address, e.g. component 3 of the enthalpy variables)."""
        arr = (_vp * 3)()
        for d in range(self.dim):
            arr[d] = b[d].ptr
        p = chi.ptr if hasattr(chi, "ptr") else chi
        self.ctx.check(self.ctx.lib.ml_hc_assemble_bcoef(self.h, p, arr), "ml_hc_assemble_bcoef")

    def solve(self, phi, rhs, opts=None):
        st = SolveStats()
        o = opts or default_opts()
        self.ctx.check(self.ctx.lib.ml_hc_solve(self.h, phi.ptr, rhs.ptr, C.byref(o), C.byref(st)), "ml_hc_solve")
        return st

    def flux_source(self, flux, src):
        """src = the bottom heat flux's source (theta component of the bottom
        cells), from the current level-0 bCoef."""
        self.ctx.check(self.ctx.lib.ml_hc_flux_source(self.h, flux.ptr, src.ptr), "ml_hc_flux_source")

    def adjoint_step(self, phi_new, phi_old, dt, g_new, g_old, g_flux=None, opts=None):
        """Adjoint of the backward-Euler step phi_old -> phi_new: g_old, adds to
        g_flux; returns (dL/dSt of this step, SolveStats of the J^T solve)."""
        st = SolveStats()
        gst = C.c_double(0.0)
        o = opts or default_opts(eps=1e-13, max_iter=500)
        self.ctx.check(self.ctx.lib.ml_hc_adjoint_step(self.h, phi_new.ptr, phi_old.ptr, dt, g_new.ptr, g_old.ptr,
                                                       g_flux.ptr if g_flux is not None else None, C.byref(gst),
                                                       C.byref(o), C.byref(st)), "ml_hc_adjoint_step")
        return gst.value, st

    def porosity_vjp(self, phi, g_chi, g_phi):
        """g_phi += g_chi dchi/dphi; returns sum g_chi dchi/dSt."""
        gst = C.c_double(0.0)
        self.ctx.check(self.ctx.lib.ml_hc_porosity_vjp(self.h, phi.ptr, g_chi.ptr, g_phi.ptr, C.byref(gst)),
                       "ml_hc_porosity_vjp")
        return gst.value

    def solve_be(self, phi_new, phi_old, dt, src=None, opts=None):
        st = SolveStats()
        o = opts or default_opts()
        self.ctx.check(self.ctx.lib.ml_hc_solve_be(self.h, phi_new.ptr, phi_old.ptr,
                                                   src.ptr if src is not None else None, dt, C.byref(o),
                                                   C.byref(st)), "ml_hc_solve_be")
        return st

    def close(self):
        if self.h:
            self.ctx.lib.ml_hc_op_destroy(self.h)
            self.h = None


class HCBatch:
    """The batched form: independent (H, Theta) instances (HCOp each, own
    grid, parameters, BCs) that advance one FP32-refined implicit step
    together, every launch covering all of them.  Instance i's step equals
    porosity -> assemble_bcoef -> set_coefs(None, b, 1, -dt) ->
    solve_be(phi_i, copy, dt, FP32-refined) on its HCOp, bit for bit."""

    def __init__(self, ctx, ops):
        self.ctx = ctx
        self.ops = list(ops)
        n = len(self.ops)
        h = _vp()
        ctx.check(ctx.lib.ml_hc_batch_create(ctx.h, n, (_vp * n)(*[op.h for op in self.ops]), C.byref(h)),
                  "ml_hc_batch_create")
        self.h = h

    def step(self, phis, dt, opts=None, srcs=None):
        """Advance every state (list of DeviceArray, one per instance) in
        place, with a source per instance when `srcs` is given; returns the
        per-instance SolveStats."""
        n = len(self.ops)
        st = (SolveStats * n)()
        o = opts or default_opts(precision=PRECISION_FP32_REFINED)
        if srcs is None:
            self.ctx.check(self.ctx.lib.ml_hc_batch_step(self.h, (_vp * n)(*[x.ptr for x in phis]), dt, C.byref(o),
                                                         st), "ml_hc_batch_step")
        else:
            self.ctx.check(self.ctx.lib.ml_hc_batch_step_src(self.h, (_vp * n)(*[x.ptr for x in phis]),
                                                             (_vp * n)(*[x.ptr if x is not None else None
                                                                         for x in srcs]), dt, C.byref(o), st),
                           "ml_hc_batch_step_src")
        return list(st)

    def memory(self):
        """(device bytes, device allocations, fill launches) of the batch's set-up."""
        nb, na, nf = C.c_size_t(), C.c_int(), C.c_int()
        self.ctx.check(self.ctx.lib.ml_hc_batch_memory(self.h, C.byref(nb), C.byref(na), C.byref(nf)),
                       "ml_hc_batch_memory")
        return nb.value, na.value, nf.value

    def set_tile(self, variant):
        """Fix the relaxation tile shape (-1: tune again at the next step)."""
        self.ctx.check(self.ctx.lib.ml_hc_batch_set_tile(self.h, variant), "ml_hc_batch_set_tile")

    def tile(self):
        """(variant, (tile_x, tile_y, threads), tuning ms per shape)."""
        v, tx, ty, nt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        ms = (C.c_float * 16)()
        n = self.ctx.lib.ml_hc_batch_tile(self.h, C.byref(v), C.byref(tx), C.byref(ty), C.byref(nt), ms, 16)
        self.ctx.check(n if n < 0 else 0, "ml_hc_batch_tile")
        return v.value, (tx.value, ty.value, nt.value), [ms[i] for i in range(n)]

    def close(self):
        if self.h:
            self.ctx.lib.ml_hc_batch_destroy(self.h)
            self.h = None


class VCBatch:
    """Batched linear solves: `per_op` solves on each of many VCOps (the U*
    components or the projection of many instances) iterate together, every
    launch covering all of them; unit u is solve u % per_op of ops[u //
    per_op], and gets exactly the result and statistics of solve_many /
    solve_be_many on its operator."""

    def __init__(self, ctx, ops, per_op=1):
        self.ctx = ctx
        self.ops = list(ops)
        self.per_op = per_op
        n = len(self.ops)
        h = _vp()
        ctx.check(ctx.lib.ml_vc_batch_create(ctx.h, n, (_vp * n)(*[op.h for op in self.ops]), per_op, C.byref(h)),
                  "ml_vc_batch_create")
        self.h = h

    def solve(self, phis, rhss, opts=None):
        n = len(phis)
        st = (SolveStats * n)()
        o = opts or default_opts()
        self.ctx.check(self.ctx.lib.ml_vc_batch_solve(self.h, (_vp * n)(*[x.ptr for x in phis]),
                                                      (_vp * n)(*[x.ptr for x in rhss]), C.byref(o), st),
                       "ml_vc_batch_solve")
        return list(st)

    def solve_be(self, news, olds, dt, srcs=None, opts=None):
        n = len(news)
        st = (SolveStats * n)()
        o = opts or default_opts()
        src = (_vp * n)(*[x.ptr if x is not None else None for x in srcs]) if srcs is not None else None
        self.ctx.check(self.ctx.lib.ml_vc_batch_solve_be(self.h, (_vp * n)(*[x.ptr for x in news]),
                                                         (_vp * n)(*[x.ptr for x in olds]), src, dt, C.byref(o), st),
                       "ml_vc_batch_solve_be")
        return list(st)

    def close(self):
        if self.h:
            self.ctx.lib.ml_vc_batch_destroy(self.h)
            self.h = None


class VCOp:
    """DarcyBrinkmanOp / AMRProjectionOp (+ the factory's MG hierarchy) on the
    GPU: L u = alpha*a*u - beta*c*u - beta*div(b grad u), one component."""

    def __init__(self, ctx, n, dx, periodic, bc):
        self.ctx = ctx
        self.n = list(n)
        self.dim = len(n)
        self.dx = dx
        h = _vp()
        lo = _iarr([0] * 3)
        hi = _iarr([x - 1 for x in n])
        ctx.check(ctx.lib.ml_vc_op_create(ctx.h, self.dim, lo, hi, dx, _iarr(periodic), C.byref(bc), C.byref(h)),
                  "ml_vc_op_create")
        self.h = h
        self.nlev = ctx.lib.ml_vc_num_levels(h)

    def set_coefs(self, a, b, c, alpha, beta):
        arr = (_vp * 3)()
        for d in range(self.dim):
            arr[d] = b[d].ptr
        self.ctx.check(self.ctx.lib.ml_vc_set_coefs(self.h, a.ptr if a is not None else None, arr,
                                                    c.ptr if c is not None else None, alpha, beta),
                       "ml_vc_set_coefs")

    def set_alpha_beta(self, alpha, beta):
        self.ctx.check(self.ctx.lib.ml_vc_set_alpha_beta(self.h, alpha, beta), "ml_vc_set_alpha_beta")

    def relax(self, depth, phi, rhs, iters):
        self.ctx.check(self.ctx.lib.ml_vc_relax(self.h, depth, phi.ptr, rhs.ptr, iters), "ml_vc_relax")

    def residual(self, depth, res, phi, rhs, homogeneous=0):
        self.ctx.check(self.ctx.lib.ml_vc_residual(self.h, depth, res.ptr, phi.ptr, rhs.ptr, homogeneous),
                       "ml_vc_residual")

    def apply_op(self, depth, lhs, phi, homogeneous=0):
        self.ctx.check(self.ctx.lib.ml_vc_apply_op(self.h, depth, lhs.ptr, phi.ptr, homogeneous), "ml_vc_apply_op")

    def precond(self, depth, phi, rhs):
        """preCond: phi = lambda*rhs, then two red/black iterations."""
        self.ctx.check(self.ctx.lib.ml_vc_precond(self.h, depth, phi.ptr, rhs.ptr), "ml_vc_precond")

    def dot(self, depth, x, y):
        """(x, y) over the valid cells, in the BiCGStab bottom's order."""
        out = C.c_double(0.0)
        self.ctx.check(self.ctx.lib.ml_vc_dot(self.h, depth, x.ptr, y.ptr, C.byref(out)), "ml_vc_dot")
        return out.value

    def restrict_residual(self, depth, resc, phi, rhs):
        self.ctx.check(self.ctx.lib.ml_vc_restrict_residual(self.h, depth, resc.ptr, phi.ptr, rhs.ptr),
                       "ml_vc_restrict_residual")

    def prolong_increment(self, depth, phif, corrc):
        self.ctx.check(self.ctx.lib.ml_vc_prolong_increment(self.h, depth, phif.ptr, corrc.ptr),
                       "ml_vc_prolong_increment")

    def darcy_ccoef(self, chi, coeff, c):
        self.ctx.check(self.ctx.lib.ml_vc_darcy_ccoef(self.h, chi.ptr, coeff, c.ptr), "ml_vc_darcy_ccoef")

    def darcy_ccoef_law(self, chi_ptr, coeff, law, c, hele_shaw=0.0):
        """c = (chi/Pi(chi))*coeff for a permeability law (PERM_* or its name);
        chi_ptr: a DeviceArray or a device address."""
        law = PERM_LAWS.get(law, law)
        p = chi_ptr.ptr if hasattr(chi_ptr, "ptr") else chi_ptr
        self.ctx.check(self.ctx.lib.ml_vc_darcy_ccoef_law(self.h, p, coeff, law, hele_shaw, c.ptr),
                       "ml_vc_darcy_ccoef_law")

    @staticmethod
    def _arr(xs):
        arr = (_vp * 3)()
        for d, x in enumerate(xs):
            arr[d] = x.ptr
        return arr

    def ustar_src(self, fp, vold, vnew, pi, src):
        self.ctx.check(self.ctx.lib.ml_vc_ustar_src(self.h, C.byref(fp), vold.ptr, vnew.ptr, pi.ptr, self._arr(src)),
                       "ml_vc_ustar_src")

    def cc_project_rhs(self, fp, dt, vold, vnew, pi, ustar, u, rhs):
        self.ctx.check(self.ctx.lib.ml_vc_cc_project_rhs(self.h, C.byref(fp), dt, vold.ptr, vnew.ptr, pi.ptr,
                                                         self._arr(ustar), self._arr(u), rhs.ptr),
                       "ml_vc_cc_project_rhs")

    def pressure_faces(self, vold, vnew, b):
        self.ctx.check(self.ctx.lib.ml_vc_pressure_faces(self.h, vold.ptr, vnew.ptr, self._arr(b)),
                       "ml_vc_pressure_faces")

    def face_velocity(self, u, u_face):
        """u_face[d] = CellToEdge of the no-slip cell velocity (0 on walls)."""
        self.ctx.check(self.ctx.lib.ml_vc_face_velocity(self.h, self._arr(u), self._arr(u_face)),
                       "ml_vc_face_velocity")

    def cc_correct(self, fp, dt, vold, vnew, pi, u):
        self.ctx.check(self.ctx.lib.ml_vc_cc_correct(self.h, C.byref(fp), dt, vold.ptr, vnew.ptr, pi.ptr,
                                                     self._arr(u)), "ml_vc_cc_correct")

    def porosity_faces(self, chi, scale, b):
        arr = (_vp * 3)()
        for d in range(self.dim):
            arr[d] = b[d].ptr
        self.ctx.check(self.ctx.lib.ml_vc_porosity_faces(self.h, chi.ptr, scale, arr), "ml_vc_porosity_faces")

    def solve(self, phi, rhs, opts=None):
        st = SolveStats()
        o = opts or default_opts()
        self.ctx.check(self.ctx.lib.ml_vc_solve(self.h, phi.ptr, rhs.ptr, C.byref(o), C.byref(st)), "ml_vc_solve")
        return st

    def solve_be(self, phi_new, phi_old, dt, src=None, opts=None):
        st = SolveStats()
        o = opts or default_opts()
        self.ctx.check(self.ctx.lib.ml_vc_solve_be(self.h, phi_new.ptr, phi_old.ptr,
                                                   src.ptr if src is not None else None, dt, C.byref(o),
                                                   C.byref(st)), "ml_vc_solve_be")
        return st

    def solve_be_many(self, phi_new, phi_old, dt, src=None, opts=None):
        """Independent backward-Euler solves (lists of DeviceArray), run
        concurrently; returns their SolveStats."""
        n = len(phi_new)
        st = (SolveStats * n)()
        o = opts or default_opts()
        arr = lambda xs: (_vp * n)(*[x.ptr for x in xs])
        self.ctx.check(self.ctx.lib.ml_vc_solve_be_many(self.h, n, arr(phi_new), arr(phi_old),
                                                        arr(src) if src is not None else None, dt, C.byref(o), st),
                       "ml_vc_solve_be_many")
        return list(st)

    def close(self):
        if self.h:
            self.ctx.lib.ml_vc_op_destroy(self.h)
            self.h = None

