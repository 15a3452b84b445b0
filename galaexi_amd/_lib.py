"""ctypes binding of include/galx.h (libgalx.so, built in-tree for gfx950).

There is no fallback: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgalx.so")

GALX_FORM_WEAK, GALX_FORM_SPLIT_PI = 0, 1
GALX_RIEMANN_LF, GALX_RIEMANN_ROE, GALX_RIEMANN_ROEEF = 0, 1, 3
GALX_PREC_FP64, GALX_PREC_MIXED = 0, 1
GALX_ADJ_STAGES, GALX_ADJ_STEPS = 0, 1
GALX_E_ARG = -1

# every symbol declared in include/galx.h
EXPORTS = (
    "galx_version", "galx_supported_degrees", "galx_last_error", "galx_create", "galx_destroy", "galx_set_mesh",
    "galx_set_state", "galx_get_state", "galx_set_state_device", "galx_get_state_device", "galx_rhs", "galx_get_ut",
    "galx_get_ut_device", "galx_lserk_step", "galx_lserk_steps", "galx_calc_dt", "galx_prolong_to_face",
    "galx_surf_int", "galx_timers_enable", "galx_timer_read", "galx_timers_reset", "galx_memory_bytes",
    "galx_rhs_vjp", "galx_rhs_vjp_device", "galx_lserk_forward_ckpt", "galx_lserk_adjoint",
    "galx_lserk_adjoint_device", "galx_lserk_stage", "galx_get_gradients", "galx_get_gradients_device",
    "galx_step_vjp", "galx_step_vjp_device", "galx_set_source", "galx_set_adjoint_storage",
    "galx_get_lifted_gradients", "galx_get_lifted_gradients_device", "galx_surf_int_noalloc",
)


class GalxConfig(C.Structure):
    _fields_ = [("N", C.c_int), ("form", C.c_int), ("riemann", C.c_int), ("precision", C.c_int),
                ("kappa", C.c_double), ("R", C.c_double), ("Pr", C.c_double), ("mu0", C.c_double),
                ("CFLScale", C.c_double), ("DFLScale", C.c_double), ("device", C.c_int)]


class GalxError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GalxError(f"{LIB_PATH} is missing: build the HIP extension first (python -m galaexi_amd.build)")
    # One HIP runtime per process: torch ships its own libamdhip64/libhsa-runtime64 (same sonames as ROCm's),
    # and a second runtime instance loaded after the first cannot open the device.  Loading torch first makes
    # libgalx bind to torch's instance (sonames match), so device pointers and streams are shared with torch.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, ip, dp, i, d = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int, C.c_double
    sig = {
        "galx_version": ([], C.c_char_p),
        "galx_supported_degrees": ([i], C.c_uint),
        "galx_last_error": ([], C.c_char_p),
        "galx_create": ([C.POINTER(GalxConfig), C.POINTER(vp)], i),
        "galx_destroy": ([vp], i),
        "galx_set_mesh": ([vp, i, i, ip, ip, ip, dp, dp, dp, dp, dp, dp, dp, dp], i),
        "galx_set_state": ([vp, dp], i),
        "galx_get_state": ([vp, dp], i),
        "galx_set_state_device": ([vp, vp], i),
        "galx_get_state_device": ([vp, vp], i),
        "galx_set_source": ([vp, i, dp, dp], i),
        "galx_rhs": ([vp, d], i),
        "galx_get_ut": ([vp, dp], i),
        "galx_get_ut_device": ([vp, vp], i),
        "galx_lserk_step": ([vp, d, d], i),
        "galx_lserk_steps": ([vp, d, d, i], i),
        "galx_calc_dt": ([vp, dp, ip], i),
        "galx_prolong_to_face": ([i, i, i, i, i, ip, ip, dp, dp, dp], i),
        "galx_surf_int": ([i, i, i, i, i, ip, ip, dp, dp, d, d, dp], i),
        "galx_surf_int_noalloc": ([i, i, i, i, i, ip, ip, dp, dp, d, d, dp], i),
        "galx_timers_enable": ([vp, i], i),
        "galx_timer_read": ([vp, C.c_char_p, dp, C.POINTER(C.c_long)], i),
        "galx_timers_reset": ([vp], i),
        "galx_memory_bytes": ([vp, C.POINTER(C.c_longlong)], i),
        "galx_rhs_vjp": ([vp, dp, dp], i),
        "galx_rhs_vjp_device": ([vp, vp, vp], i),
        "galx_lserk_forward_ckpt": ([vp, d, d, i, i], i),
        "galx_set_adjoint_storage": ([vp, i], i),
        "galx_lserk_adjoint": ([vp, dp], i),
        "galx_lserk_adjoint_device": ([vp, vp], i),
        "galx_step_vjp": ([vp, C.c_double, C.c_double, dp], i),
        "galx_step_vjp_device": ([vp, C.c_double, C.c_double, vp], i),
        "galx_lserk_stage": ([vp, d, d, i], i),
        "galx_get_gradients": ([vp, dp, dp, dp], i),
        "galx_get_gradients_device": ([vp, vp, vp, vp], i),
        "galx_get_lifted_gradients": ([vp, i, dp, dp, dp], i),
        "galx_get_lifted_gradients_device": ([vp, i, vp, vp, vp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc, what="galx call"):
    if rc != 0:
        msg = lib().galx_last_error().decode(errors="replace")
        raise GalxError(f"{what} failed with {rc}: {msg}")
    return rc
