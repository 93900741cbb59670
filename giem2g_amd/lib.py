"""ctypes binding of libgiem2g_amd.so (the C ABI in include/giem2g_amd.h).

The library is required: importing a compute entry point without the built
library raises, there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libgiem2g_amd.so")

# Symbols declared in include/giem2g_amd.h (checked by tests/test_abi.py).
REFERENCE_SYMBOLS = [
    "giem2g_calc_data_sizes", "giem2g_prepare_ie_kernel", "giem2g_calc_ie_kernel",
    "giem2g_calc_fft_of_ie_kernel", "giem2g_set_anomaly_conductivity",
    "giem2g_apply_ie_operator", "giem2g_set_logger", "giem2g_print_stats",
    "giem2g_delete_ie_operator",
]
EXTENDED_SYMBOLS = [
    "giem2g_last_error", "giem2g_set_component_order", "giem2g_op_create", "giem2g_op_destroy",
    "giem2g_op_set_background", "giem2g_op_build_tensor", "giem2g_op_tensor_stage",
    "giem2g_op_get_tensor", "giem2g_op_put_tensor", "giem2g_op_set_sigma",
    "giem2g_op_set_sigma_dev", "giem2g_op_apply_dev", "giem2g_op_apply_dev_f32",
    "giem2g_op_apply_pair_dev", "giem2g_op_apply_batch_dev", "giem2g_op_apply_batch_dev_noalloc",
    "giem2g_op_apply_host", "giem2g_op_stream", "giem2g_op_sync", "giem2g_op_stats",
    "giem2g_op_time_apply", "giem2g_plane_wave", "giem2g_solve", "giem2g_solve_dev",
    "giem2g_last_history", "giem2g_unet_create", "giem2g_unet_destroy", "giem2g_unet_nparams",
    "giem2g_unet_set_params", "giem2g_unet_get_params", "giem2g_unet_get_grads",
    "giem2g_unet_forward_dev", "giem2g_unet_backward_dev", "giem2g_unet_train_step",
    "giem2g_op_set_learned", "giem2g_op_learned_features", "giem2g_op_learned_apply",
    "giem2g_op_scale", "giem2g_op_set_contraction", "giem2g_solve_pair_dev", "giem2g_unet_adam",
    "giem2g_op_apply_adjoint_dev", "giem2g_rc_create", "giem2g_rc_destroy", "giem2g_rc_calc_kernel",
    "giem2g_rc_calc_fft_of_kernel", "giem2g_rc_get_tensor", "giem2g_rc_put_tensor",
    "giem2g_rc_receivers",
    "giem2g_rc_set_anomaly_conductivity", "giem2g_rc_set_anomaly_conductivity_dev",
    "giem2g_rc_recalculate", "giem2g_rc_recalculate_dev", "giem2g_rc_plane_wave", "giem2g_rc_stats",
    "giem2g_op_set_geometry", "giem2g_op_set_profiling", "giem2g_op_counters",
    "giem2g_op_reset_counters",
]


class Background(C.Structure):
    _fields_ = [("Nl", C.c_int), ("csigb", C.c_void_p), ("thickness", C.c_void_p)]


class Anomaly(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int), ("dx", C.c_double),
                ("dy", C.c_double), ("z", C.c_void_p), ("dz", C.c_void_p)]


class Data(C.Structure):
    _fields_ = [("tensor_size", C.c_longlong), ("ie_kernel_buffer_length", C.c_longlong),
                ("fft_buffers_length", C.c_longlong), ("comm64", C.c_longlong),
                ("giem2g_tensor", C.c_void_p), ("dz", C.c_void_p), ("sqsigb", C.c_void_p),
                ("csigb", C.c_void_p), ("kernel_buffer", C.c_void_p),
                ("fft_buffer_in", C.c_void_p), ("fft_buffer_out", C.c_void_p)]


class SolverParams(C.Structure):
    _fields_ = [("outer_m", C.c_int), ("inner_m", C.c_int), ("max_restart", C.c_int),
                ("tol", C.c_double), ("inner_precision", C.c_int), ("precond", C.c_int),
                ("verbose", C.c_int), ("learned_patience", C.c_int), ("inner_tol", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("outer_iterations", C.c_int), ("status", C.c_int),
                ("arnoldi_residual", C.c_double), ("true_residual", C.c_double),
                ("matvecs_fp64", C.c_longlong), ("matvecs_fp32", C.c_longlong),
                ("precond_calls", C.c_longlong), ("precond_fallbacks", C.c_longlong),
                ("seconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


LOGGER = C.CFUNCTYPE(None, C.c_char_p)

_lib = None


def available():
    return os.path.exists(LIB_PATH)


def load():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libgiem2g_amd.so is not built (run python -m giem2g_amd.build); "
                           "there is no CPU fallback")
    # PyTorch ships its own libamdhip64.so.7.  Loading torch first makes this
    # library bind to that same HIP runtime (one runtime per process), so HBM
    # tensors and streams can be shared with torch.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    sig = {
        "giem2g_last_error": (C.c_char_p, []),
        "giem2g_set_component_order": (i, [vp, i]),
        "giem2g_op_create": (i, [i, i, i, d, d, vp, C.POINTER(vp)]),
        "giem2g_op_destroy": (i, [vp]),
        "giem2g_op_set_background": (i, [vp, i, vp, vp, d]),
        "giem2g_op_build_tensor": (i, [vp]),
        "giem2g_op_tensor_stage": (i, [vp, i]),
        "giem2g_op_get_tensor": (i, [vp, vp]),
        "giem2g_op_put_tensor": (i, [vp, vp, i]),
        "giem2g_op_set_sigma": (i, [vp, vp]),
        "giem2g_op_set_sigma_dev": (i, [vp, vp]),
        "giem2g_op_apply_dev": (i, [vp, vp, vp]),
        "giem2g_op_apply_dev_f32": (i, [vp, vp, vp]),
        "giem2g_op_apply_pair_dev": (i, [vp, i, vp, vp, vp, vp]),
        "giem2g_op_apply_batch_dev": (i, [vp, i, i, vp, vp]),
        "giem2g_op_apply_batch_dev_noalloc": (i, [vp, i, i, vp, vp]),
        "giem2g_op_apply_host": (i, [vp, vp, vp]),
        "giem2g_op_stream": (i, [vp, C.POINTER(vp)]),
        "giem2g_op_sync": (i, [vp]),
        "giem2g_op_stats": (i, [vp, vp]),
        "giem2g_op_time_apply": (i, [vp, i, vp, vp, i, vp]),
        "giem2g_plane_wave": (i, [vp, i, vp]),
        "giem2g_solve": (i, [vp, vp, vp, C.POINTER(SolverParams), vp, C.POINTER(SolveInfo)]),
        "giem2g_solve_dev": (i, [vp, vp, vp, C.POINTER(SolverParams), vp, C.POINTER(SolveInfo)]),
        "giem2g_last_history": (i, [vp, vp, i, C.POINTER(i)]),
        "giem2g_unet_create": (i, [C.POINTER(vp)]),
        "giem2g_unet_destroy": (i, [vp]),
        "giem2g_unet_nparams": (i, [C.POINTER(C.c_longlong)]),
        "giem2g_unet_set_params": (i, [vp, vp]),
        "giem2g_unet_get_params": (i, [vp, vp]),
        "giem2g_unet_get_grads": (i, [vp, vp]),
        "giem2g_unet_forward_dev": (i, [vp, vp, vp, i, i, i, i, i]),
        "giem2g_unet_backward_dev": (i, [vp, vp]),
        "giem2g_unet_train_step": (i, [vp, vp, vp, i, i, i, i, C.c_float, C.POINTER(d)]),
        "giem2g_op_set_learned": (i, [vp, vp, d]),
        "giem2g_op_learned_features": (i, [vp, vp, vp, vp, vp]),
        "giem2g_op_learned_apply": (i, [vp, vp, vp]),
        "giem2g_op_scale": (i, [vp, i, vp, vp]),
        "giem2g_op_set_contraction": (i, [vp, i]),
        "giem2g_unet_adam": (i, [vp, C.c_float]),
        "giem2g_op_apply_adjoint_dev": (i, [vp, vp, vp]),
        "giem2g_solve_pair_dev": (i, [vp, vp, vp, vp, vp, C.POINTER(SolverParams), vp, vp,
                                      C.POINTER(SolveInfo), C.POINTER(C.c_longlong)]),
        "giem2g_rc_create": (i, [vp, i, vp, vp, vp, C.POINTER(vp)]),
        "giem2g_rc_destroy": (i, [vp]),
        "giem2g_rc_calc_kernel": (i, [vp]),
        "giem2g_rc_calc_fft_of_kernel": (i, [vp]),
        "giem2g_rc_get_tensor": (i, [vp, vp]),
        "giem2g_rc_put_tensor": (i, [vp, vp, i]),
        "giem2g_rc_receivers": (i, [vp, vp, vp, vp]),
        "giem2g_rc_set_anomaly_conductivity": (i, [vp, vp]),
        "giem2g_rc_set_anomaly_conductivity_dev": (i, [vp, vp]),
        "giem2g_rc_recalculate": (i, [vp, vp, vp, vp]),
        "giem2g_rc_recalculate_dev": (i, [vp, vp, vp, vp]),
        "giem2g_rc_plane_wave": (i, [vp, i, vp]),
        "giem2g_rc_stats": (i, [vp, vp]),
        "giem2g_op_set_geometry": (i, [vp, d, d, vp]),
        "giem2g_op_set_profiling": (i, [vp, i]),
        "giem2g_op_counters": (i, [vp, vp, i]),
        "giem2g_op_reset_counters": (i, [vp]),
        "giem2g_calc_data_sizes": (None, [Anomaly, C.POINTER(Data)]),
        "giem2g_prepare_ie_kernel": (None, [Anomaly, Data]),
        "giem2g_calc_ie_kernel": (None, [vp, Background, Anomaly, d]),
        "giem2g_calc_fft_of_ie_kernel": (None, [vp]),
        "giem2g_set_anomaly_conductivity": (None, [vp, vp]),
        "giem2g_apply_ie_operator": (None, [vp, vp, vp]),
        "giem2g_set_logger": (None, [LOGGER]),
        "giem2g_print_stats": (None, [vp]),
        "giem2g_delete_ie_operator": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


class G2GError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        msg = load().giem2g_last_error()
        raise G2GError("giem2g error %d: %s" % (rc, msg.decode() if msg else "?"))
