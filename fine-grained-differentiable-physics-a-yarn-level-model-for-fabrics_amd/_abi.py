"""ctypes binding of include/yarnstep.h (libyarnstep.so)."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_HANDLES = 16

YC_OK, YC_ERR_NONFINITE, YC_ERR_NOT_CONVERGED, YC_ERR_HIP, YC_ERR_ARG = 0, 1, 2, 3, 4


class YcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("yarnstep error %d: %s" % (code, msg))
        self.code = code


class NonFinite(YcError):
    pass


class NotConverged(YcError):
    pass


class YcDesc(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int), ("length", ctypes.c_int),
        ("hight", ctypes.c_double), ("L", ctypes.c_double), ("R", ctypes.c_double),
        ("n_yarns", ctypes.c_int),
        ("rho", ctypes.c_double * 2), ("Y", ctypes.c_double * 2), ("B", ctypes.c_double * 2),
        ("kc", ctypes.c_double), ("kf", ctypes.c_double), ("df", ctypes.c_double),
        ("mu", ctypes.c_double), ("S", ctypes.c_double), ("handle_k", ctypes.c_double),
        ("pattern", ctypes.c_int), ("pose", ctypes.c_int),
        ("G", ctypes.c_double * 3),
        ("wind_vel", ctypes.c_double * 3), ("wind_density", ctypes.c_double), ("wind_drag", ctypes.c_double),
        ("n_handles", ctypes.c_int),
        ("handles", ctypes.c_int * (2 * MAX_HANDLES)),
    ]


class YcBatchOpts(ctypes.Structure):
    _fields_ = [("graph", ctypes.c_int), ("group_order", ctypes.c_int), ("node_block", ctypes.c_int)]


class YcStepInfo(ctypes.Structure):
    _fields_ = [("cg_iters", ctypes.c_int), ("cg_relres", ctypes.c_double), ("status", ctypes.c_int)]


# every symbol include/yarnstep.h declares
EXPORTS = ["yc_create", "yc_destroy", "yc_last_error", "yc_sizes", "yc_set_state", "yc_get_state",
           "yc_set_ext_force", "yc_assemble", "yc_export_csr", "yc_export_vec", "yc_export_crossing",
           "yc_export_normals", "yc_solve", "yc_update", "yc_step", "yc_run", "yc_time_spmv", "yc_spmv_bytes",
           "yc_mlp_param_count", "yc_mlp_create", "yc_mlp_destroy", "yc_mlp_set_params", "yc_mlp_get_params",
           "yc_mlp_set_norm", "yc_mlp_set_data", "yc_mlp_forward", "yc_mlp_grad", "yc_mlp_train", "yc_features",
           "yc_targets", "yc_step_warm", "yc_run_warm", "yc_mlp_time_forward", "yc_set_cg_variant",
           "yc_tangent_begin", "yc_step_tangent", "yc_tangent_get", "yc_tangent_end",
           "yc_time_solve", "yc_cg_iteration_bytes",
           "yc_batch_create", "yc_batch_destroy", "yc_batch_set_opts", "yc_batch_get_opts", "yc_batch_run",
           "yc_batch_info", "yc_time_assemble", "yc_tangent_set_ext", "yc_set_preconditioner",
           "yc_set_state_device", "yc_get_state_device", "yc_export_precond", "yc_export_matrix",
           "yc_load_system", "yc_update_vel", "yc_run_guess", "yc_mlp_forward_noalloc", "yc_warm_guess", "yc_time_warm_guess",
           "yc_adjoint_begin", "yc_adjoint_step", "yc_adjoint_run", "yc_adjoint_set_targets", "yc_adjoint_backward", "yc_adjoint_end",
           "yc_batch_adjoint_run", "yc_batch_adjoint_backward"]


def lib_path():
    return os.path.join(_HERE, "libyarnstep.so")


_lib = None


def load_library():
    """Load libyarnstep.so from the package directory; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    p = lib_path()
    if not os.path.exists(p):
        raise YcError(YC_ERR_HIP, "libyarnstep.so is not built (run __graft_entry__.build())")
    # One HIP runtime per process: torch ships its own libamdhip64 (soname libamdhip64.so.7,
    # requested by torch as "libamdhip64.so").  Loaded after this library, torch would map a
    # second runtime beside the system one this library requests; loaded first, its runtime
    # is the one this library binds to (same soname).  Two runtimes in one process have
    # failed to find the device in torch after many contexts had come and gone.
    try:
        import torch  # noqa: F401
    except Exception:   # absent or unusable: the library loads its own runtime
        pass
    L = ctypes.CDLL(p)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    vp = ctypes.c_void_p
    L.yc_create.argtypes = [ctypes.POINTER(YcDesc), ctypes.POINTER(vp)]
    L.yc_destroy.argtypes = [vp]
    L.yc_destroy.restype = None
    L.yc_last_error.restype = ctypes.c_char_p
    L.yc_sizes.argtypes = [vp, ip, ip, ip, ip]
    L.yc_set_state.argtypes = [vp] + [dp] * 6
    L.yc_get_state.argtypes = [vp] + [dp] * 6
    L.yc_set_ext_force.argtypes = [vp, dp]
    L.yc_assemble.argtypes = [vp, ctypes.c_double]
    L.yc_export_csr.argtypes = [vp, ip, ip, dp, ip]
    L.yc_export_vec.argtypes = [vp, ctypes.c_int, dp]
    L.yc_export_matrix.argtypes = [vp, ctypes.c_int, ip, ip, dp, ip]
    L.yc_load_system.argtypes = [vp, ip, ip, dp, dp]
    L.yc_update_vel.argtypes = [vp, dp, dp, ctypes.c_double]
    L.yc_export_crossing.argtypes = [vp, dp, dp, dp]
    L.yc_export_normals.argtypes = [vp, dp]
    L.yc_solve.argtypes = [vp, dp, ctypes.c_double, ctypes.c_int, ip, dp]
    L.yc_update.argtypes = [vp, ctypes.c_double]
    L.yc_step.argtypes = [vp, ctypes.c_double, dp, ctypes.c_double, ctypes.c_int, ctypes.POINTER(YcStepInfo)]
    L.yc_run.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)]
    L.yc_time_spmv.argtypes = [vp, ctypes.c_int, dp]
    L.yc_run_guess.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_longlong)]
    L.yc_set_cg_variant.argtypes = [vp, ctypes.c_int]
    L.yc_set_preconditioner.argtypes = [vp, ctypes.c_int]
    L.yc_set_state_device.argtypes = [vp] + [ctypes.c_void_p] * 7
    L.yc_export_precond.argtypes = [vp, dp]
    L.yc_get_state_device.argtypes = [vp] + [ctypes.c_void_p] * 7
    L.yc_tangent_begin.argtypes = [vp, ctypes.c_int, dp, dp]
    L.yc_step_tangent.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(YcStepInfo)]
    L.yc_tangent_get.argtypes = [vp, ctypes.c_int] + [dp] * 6
    L.yc_tangent_end.argtypes = [vp]
    L.yc_tangent_set_ext.argtypes = [vp, dp]
    L.yc_adjoint_begin.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.yc_adjoint_step.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(YcStepInfo)]
    L.yc_adjoint_run.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                 ctypes.POINTER(ctypes.c_longlong)]
    L.yc_adjoint_set_targets.argtypes = [vp, dp, dp, ctypes.c_double]
    L.yc_adjoint_backward.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp, dp, dp, dp,
                                      ctypes.POINTER(ctypes.c_longlong)]
    L.yc_adjoint_end.argtypes = [vp]
    L.yc_batch_adjoint_run.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_int)]
    L.yc_batch_adjoint_backward.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp,
                                            ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_int)]
    L.yc_time_solve.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, ip]
    L.yc_time_assemble.argtypes = [vp, ctypes.c_double, ctypes.c_int, dp]
    L.yc_cg_iteration_bytes.argtypes = [vp]
    L.yc_cg_iteration_bytes.restype = ctypes.c_double
    L.yc_spmv_bytes.argtypes = [vp]
    L.yc_spmv_bytes.restype = ctypes.c_double
    fp = ctypes.POINTER(ctypes.c_float)
    L.yc_mlp_param_count.argtypes = []
    L.yc_mlp_create.argtypes = [ctypes.POINTER(vp)]
    L.yc_mlp_destroy.argtypes = [vp]
    L.yc_mlp_destroy.restype = None
    L.yc_mlp_set_params.argtypes = [vp, fp]
    L.yc_mlp_get_params.argtypes = [vp, fp]
    L.yc_mlp_set_norm.argtypes = [vp, fp, fp, fp]
    L.yc_mlp_set_data.argtypes = [vp, fp, fp, fp, ctypes.c_int]
    L.yc_mlp_forward.argtypes = [vp, fp, ctypes.c_int, fp]
    L.yc_mlp_forward_noalloc.argtypes = [vp, fp, ctypes.c_int, fp]
    L.yc_mlp_grad.argtypes = [vp, ip, ctypes.c_int, fp, fp]
    L.yc_mlp_train.argtypes = [vp, ip, ctypes.c_int, ctypes.c_int, ctypes.c_float, fp]
    L.yc_mlp_time_forward.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.yc_features.argtypes = [vp, fp]
    L.yc_warm_guess.argtypes = [vp, vp]
    L.yc_time_warm_guess.argtypes = [vp, vp, ctypes.c_int, dp, dp]
    L.yc_targets.argtypes = [vp, fp]
    L.yc_step_warm.argtypes = [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.POINTER(YcStepInfo)]
    L.yc_run_warm.argtypes = [vp, vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_longlong)]
    L.yc_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    L.yc_batch_destroy.argtypes = [vp]
    L.yc_batch_destroy.restype = None
    L.yc_batch_set_opts.argtypes = [vp, ctypes.POINTER(YcBatchOpts)]
    L.yc_batch_get_opts.argtypes = [vp, ctypes.POINTER(YcBatchOpts)]
    L.yc_batch_run.argtypes = [vp, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                               ctypes.POINTER(ctypes.c_longlong), ip]
    L.yc_batch_info.argtypes = [vp, ip, ip, ip, ip]
    _lib = L
    return L


def check(rc):
    if rc == YC_OK:
        return rc
    msg = load_library().yc_last_error().decode()
    if rc == YC_ERR_NONFINITE:
        raise NonFinite(rc, msg)
    if rc == YC_ERR_NOT_CONVERGED:
        raise NotConverged(rc, msg)
    raise YcError(rc, msg)


def make_desc(spec):
    d = YcDesc()
    d.width, d.length = int(spec["width"]), int(spec["length"])
    d.hight, d.L, d.R = float(spec["hight"]), float(spec["L"]), float(spec["R"])
    yarns = spec["yarns"]
    d.n_yarns = len(yarns)
    for i, y in enumerate(yarns):
        d.rho[i], d.Y[i], d.B[i] = float(y["rho"]), float(y["Y"]), float(y["B"])
    for k in ("kc", "kf", "df", "mu", "S", "handle_k"):
        setattr(d, k, float(spec[k]))
    d.pattern = {"plain": 0, "twill": 1, "satin": 2}[spec["pattern"]]
    d.pose = {"upright": 0, "flat": 1}[spec["pose"]]
    for i in range(3):
        d.G[i] = float(spec["G"][i])
        d.wind_vel[i] = float(spec["wind_vel"][i])
    d.wind_density, d.wind_drag = float(spec["wind_density"]), float(spec["wind_drag"])
    hs = spec["handles"]
    if len(hs) > MAX_HANDLES:
        raise ValueError("at most %d handles" % MAX_HANDLES)
    d.n_handles = len(hs)
    for i, (u, v) in enumerate(hs):
        d.handles[2 * i], d.handles[2 * i + 1] = int(u), int(v)
    return d
