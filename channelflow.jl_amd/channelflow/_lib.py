"""ctypes binding of libcfhip.so (C ABI in include/channelflow_hip.h).

The library is the only compute path: loading fails loudly when it is
missing, and there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "lib", "libcfhip.so")

CF_OK = 0
CF_ERR_ARGUMENT = 1
CF_ERR_DIMENSION = 2
CF_ERR_DEVICE = 3
CF_ERR_UNSUPPORTED = 4

FORMS = {"rotating": 0, "convective": 1, "divergence": 2}


class cf_config(C.Structure):
    _fields_ = [
        ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
        ("Lx", C.c_double), ("Lz", C.c_double),
        ("nu", C.c_double), ("dt", C.c_double),
        ("baseflow", C.POINTER(C.c_double)),
        ("form", C.c_int), ("constraint", C.c_int),
        ("constraint_values", C.c_double * 2),
        ("device", C.c_int),
        ("precision", C.c_int),
    ]


PRECISIONS = {"fp64": 0, "mixed": 1}


_lib = None

# symbols declared in include/channelflow_hip.h
SYMBOLS = ["cf_problem_create", "cf_problem_destroy", "cf_last_error", "cf_problem_sizes", "cf_step", "cf_advance",
           "cf_step_timed",
           "cf_last_gradient", "cf_forward", "cf_inverse", "cf_nonlinear", "cf_stokes_solve", "cf_project",
           "cf_pressure", "cf_dot", "cf_dissipation", "cf_power_input", "cf_plane_statistics", "cf_advance_to", "cf_advance_to_noalloc", "cf_gaussian_noise", "cf_random_state",
           "cf_ensemble_create", "cf_ensemble_advance", "cf_ensemble_info", "cf_ensemble_destroy", "cf_set_forcing", "cf_set_closure", "cf_set_nonlinear",
           "cf_velocity_gradient", "cf_box_filter", "cf_mlp_features", "cf_mlp_infer", "cf_mlp_work_floats",
           "cf_mlp_grad", "cf_mlp_adam", "cf_problem_tune", "cf_problem_tiles", "cf_problem_get_tiles", "cf_problem_set_tiles",
           "cf_forward3", "cf_inverse3"]

# int (*cf_forcing_fn)(void* user, double tstage, const double* U0, U1, U2, double* F0, F1, F2, void* stream)
FORCING_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                         C.c_void_p, C.c_void_p, C.c_void_p)


def load(path=LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
    L = C.CDLL(path)
    vp, d, i = C.c_void_p, C.c_double, C.c_int
    dp = C.POINTER(C.c_double)
    L.cf_last_error.restype = C.c_char_p
    L.cf_problem_create.argtypes = [C.POINTER(cf_config), C.POINTER(vp)]
    L.cf_problem_destroy.argtypes = [vp]
    L.cf_problem_sizes.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.cf_step.argtypes = [vp, vp, vp, vp, vp, d, i, dp, dp, vp]
    L.cf_advance.argtypes = [vp, vp, vp, vp, vp, d, i, i, dp, dp, vp]
    L.cf_step_timed.argtypes = [vp, vp, vp, vp, vp, d, i, dp, dp, dp, vp]
    L.cf_last_gradient.argtypes = [vp, dp, vp]
    L.cf_forward.argtypes = [vp, vp, vp, vp]
    L.cf_inverse.argtypes = [vp, vp, vp, vp]
    L.cf_nonlinear.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.cf_stokes_solve.argtypes = [vp, d, d, vp, vp, vp, vp, vp, vp, vp, i, dp, dp, vp]
    L.cf_project.argtypes = [vp, vp, vp, vp, vp]
    L.cf_pressure.argtypes = [vp, vp, vp, vp, vp, vp]
    L.cf_dot.argtypes = [vp, vp, vp, vp, vp, vp, vp, dp, vp]
    L.cf_dissipation.argtypes = [vp, vp, vp, vp, d, dp, vp]
    L.cf_power_input.argtypes = [vp, vp, vp, d, dp, dp, vp]
    L.cf_plane_statistics.argtypes = [vp, vp, vp, vp, dp, vp]
    L.cf_advance_to.argtypes = [vp, vp, vp, vp, vp, d, d, i, dp, dp, vp]
    L.cf_advance_to_noalloc.argtypes = [vp, vp, vp, vp, vp, d, d, i, dp, dp, vp]
    L.cf_gaussian_noise.argtypes = [vp, d, C.c_ulonglong, vp, vp]
    L.cf_random_state.argtypes = [vp, d, C.c_ulonglong, vp, vp, vp, vp, vp]
    L.cf_ensemble_create.argtypes = [C.POINTER(vp), C.POINTER(vp), i, i, C.POINTER(vp)]
    L.cf_ensemble_advance.argtypes = [vp, i, vp]
    L.cf_ensemble_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.cf_ensemble_destroy.argtypes = [vp]
    L.cf_set_forcing.argtypes = [vp, FORCING_FN, vp, vp, vp, vp]
    L.cf_set_closure.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i]
    L.cf_set_nonlinear.argtypes = [vp, FORCING_FN, vp, vp, vp, vp]
    L.cf_velocity_gradient.argtypes = [vp, vp, vp, vp, vp, vp]
    L.cf_box_filter.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, d, d, vp]
    L.cf_mlp_features.argtypes = [vp, vp, i, i, i, vp, vp, vp]
    L.cf_mlp_infer.argtypes = [vp, vp, vp, vp, i, i, i, vp, vp]
    L.cf_mlp_work_floats.argtypes = [C.c_longlong]
    L.cf_mlp_grad.argtypes = [vp, vp, vp, vp, vp, vp, C.c_longlong, vp, vp, vp]
    L.cf_mlp_adam.argtypes = [vp, vp, vp, vp, vp, C.c_float, i, vp]
    L.cf_problem_tune.argtypes = [vp, vp]
    L.cf_problem_tiles.argtypes = [vp, C.POINTER(C.c_int)]
    L.cf_problem_get_tiles.argtypes = [vp, C.POINTER(C.c_int)]
    L.cf_problem_set_tiles.argtypes = [vp, C.POINTER(C.c_int)]
    L.cf_forward3.argtypes = [vp, vp, vp, vp, vp, vp]
    L.cf_inverse3.argtypes = [vp, vp, vp, vp, vp, vp]
    for name in SYMBOLS:
        getattr(L, name).restype = C.c_int if name != "cf_last_error" else C.c_char_p
    L.cf_mlp_work_floats.restype = C.c_longlong
    _lib = L
    return L


class ArgumentError(ValueError):
    """Mirror of Julia's ArgumentError."""


class DimensionMismatch(ValueError):
    """Mirror of Julia's DimensionMismatch."""


class DeviceError(RuntimeError):
    pass


def check(rc):
    if rc == CF_OK:
        return
    msg = load().cf_last_error().decode()
    if rc == CF_ERR_ARGUMENT:
        raise ArgumentError(msg)
    if rc == CF_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if rc == CF_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise DeviceError(msg)
