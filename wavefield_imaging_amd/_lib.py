"""ctypes binding of libawefd (include/awefd.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C wavefield_imaging_amd/csrc``) into ``wavefield_imaging_amd/libawefd.so``.
There is no fallback: if the library is missing, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libawefd.so")

CART2D, CART3D, SPHER2D, SPHER3D = 0, 1, 2, 3
KIND_NAMES = {"cart2d": CART2D, "cart3d": CART3D, "spher2d": SPHER2D, "spher3d": SPHER3D}

# every symbol include/awefd.h declares
EXPORTS = (
    "awefd_params_default", "awefd_create", "awefd_padded_dims", "awefd_set_model", "awefd_set_model_dev",
    "awefd_set_sources", "awefd_set_receivers", "awefd_set_wavelet", "awefd_run",
    "awefd_run_dev", "awefd_reset", "awefd_get_field", "awefd_debug_points",
    "awefd_profile", "awefd_kernel_stats", "awefd_device", "awefd_last_error",
    "awefd_destroy", "awefd_get_traces", "awefd_batch_create", "awefd_batch_set_tile",
    "awefd_batch_tune", "awefd_batch_run", "awefd_batch_destroy", "awefd_run_ckpt",
    "awefd_adjoint_grad", "awefd_adjoint_release", "awefd_batch_run_ckpt", "awefd_batch_adjoint_grad",
    "awefd_traces_dev", "awefd_batch_set_depth", "awefd_batch_tune_depth", "awefd_run_noalloc",
)


class AwefdError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("n", C.c_int * 3), ("o", C.c_float * 3), ("d", C.c_float * 3),
        ("nt", C.c_int), ("dt", C.c_float), ("nb", C.c_int), ("fdorder", C.c_int),
        ("optfd", C.c_int), ("dabc", C.c_int), ("fsrf", C.c_int), ("expl", C.c_int),
        ("cden", C.c_int), ("sse", C.c_int), ("jdata", C.c_int), ("device", C.c_int),
        ("adj", C.c_int), ("kt", C.c_int), ("hybrid", C.c_int), ("sinc", C.c_int),
    ]


SNAP_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int))

_lib = None


def load():
    """Load libawefd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AwefdError("libawefd.so not built (run __graft_entry__.build()): %s" % LIB_PATH)
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7, but torch asks for it as libamdhip64.so).  Loaded
    # first, it also serves libawefd's libamdhip64.so.7; loaded after
    # libawefd, it would be a second runtime that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    sig = {
        "awefd_params_default": (C.c_int, [C.c_int, P(Params)]),
        "awefd_create": (C.c_int, [P(Params), P(C.c_void_p)]),
        "awefd_padded_dims": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int)]),
        "awefd_set_model": (C.c_int, [C.c_void_p, P(C.c_float), P(C.c_float)]),
        "awefd_set_model_dev": (C.c_int, [C.c_void_p, C.c_void_p]),
        "awefd_set_sources": (C.c_int, [C.c_void_p, C.c_int, P(C.c_float)]),
        "awefd_set_receivers": (C.c_int, [C.c_void_p, C.c_int, P(C.c_float)]),
        "awefd_set_wavelet": (C.c_int, [C.c_void_p, C.c_long, P(C.c_float)]),
        "awefd_run": (C.c_int, [C.c_void_p, P(C.c_float), C.c_int, SNAP_FN, C.c_void_p]),
        "awefd_run_dev": (C.c_int, [C.c_void_p, C.c_void_p]),
        "awefd_reset": (C.c_int, [C.c_void_p]),
        "awefd_get_field": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "awefd_debug_points": (C.c_int, [C.c_void_p, C.c_int, P(C.c_int), P(C.c_float)]),
        "awefd_profile": (C.c_int, [C.c_void_p, C.c_int]),
        "awefd_kernel_stats": (C.c_int, [C.c_void_p, P(C.c_long), P(C.c_double)]),
        "awefd_device": (C.c_int, [C.c_void_p]),
        "awefd_last_error": (C.c_char_p, []),
        "awefd_destroy": (None, [C.c_void_p]),
        "awefd_get_traces": (C.c_int, [C.c_void_p, P(C.c_float)]),
        "awefd_batch_create": (C.c_int, [C.c_int, P(C.c_void_p), P(C.c_void_p)]),
        "awefd_batch_set_tile": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "awefd_batch_tune": (C.c_int, [C.c_void_p, C.c_int, P(C.c_int), P(C.c_int), P(C.c_double)]),
        "awefd_batch_run": (C.c_int, [C.c_void_p]),
        "awefd_batch_destroy": (None, [C.c_void_p]),
        "awefd_run_ckpt": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
        "awefd_adjoint_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
        "awefd_adjoint_release": (None, [C.c_void_p]),
        "awefd_batch_run_ckpt": (C.c_int, [C.c_void_p, C.c_int]),
        "awefd_traces_dev": (C.c_int, [C.c_void_p, P(C.c_void_p)]),
        "awefd_batch_set_depth": (C.c_int, [C.c_void_p, C.c_int]),
        "awefd_batch_tune_depth": (C.c_int, [C.c_void_p, C.c_int, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_double)]),
        "awefd_batch_adjoint_grad": (C.c_int, [C.c_void_p, P(C.c_void_p), C.c_void_p]),
        "awefd_run_noalloc": (C.c_int, [C.c_void_p, P(C.c_float)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().awefd_last_error().decode(errors="replace")
        raise AwefdError("libawefd error %d: %s" % (rc, msg))


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)
