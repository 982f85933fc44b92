"""ctypes binding of the C ABI in ``include/elastic_iso_3d.h``.

The library ``elasticisolib_3d_amd/lib/libelastic_iso_3d.so`` is built in-tree
by ``python -m elasticisolib_3d_amd.build``. Loading it never falls back to
anything else: a missing library is an error.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "lib", "libelastic_iso_3d.so")

ELA_OK = 0
ELA_MEM_HOST = 0
ELA_MEM_DEVICE = 1
ELA_FP32 = 4
ELA_FP64 = 8

# Every symbol the header declares (checked by the CPU test-suite).
EXPORTS = (
    "ela_last_error", "ela_version", "ela_create", "ela_destroy", "ela_set_stream",
    "ela_synchronize", "ela_set_medium", "ela_scale_medium", "ela_nl_forward",
    "ela_nl_adjoint", "ela_born_forward", "ela_born_adjoint", "ela_last_timing",
    "ela_interp_linear", "ela_interp_weights", "ela_time_step_kernels", "ela_set_tiling", "ela_get_tiling",
    "ela_set_tile_order", "ela_set_born_storage", "ela_born_storage_used",
    "ela_born_stagger", "ela_born_unstagger", "ela_medium_gradient",
    "ela_gradient_forward", "ela_gradient_backward", "ela_set_gradient_checkpoints", "ela_gradient_plan_used",
    "ela_set_windows", "ela_set_graph", "ela_stream_probe", "ela_set_elastic", "ela_set_born_mode",
    "ela_nl_forward_noalloc", "ela_get_medium", "ela_set_gradient_derivatives", "ela_gradient_derivatives_used",
    "ela_time_gradient_kernels", "ela_set_tiling_transposed", "ela_get_tiling_transposed",
)


class ElaConfig(ctypes.Structure):
    _fields_ = [
        ("nz", ctypes.c_int64), ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("nts", ctypes.c_int64), ("sub", ctypes.c_int64),
        ("dz", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("dts", ctypes.c_double), ("min_pad", ctypes.c_int64), ("alpha_cos", ctypes.c_double),
        ("precision", ctypes.c_int32), ("device", ctypes.c_int32),
    ]


class ElaPoints(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_void_p * 7), ("n", ctypes.c_int64 * 7)]


class ElaTraces(ctypes.Structure):
    _fields_ = [("tr", ctypes.c_void_p * 9)]


class ElaInterpAxes(ctypes.Structure):
    _fields_ = [
        ("nz", ctypes.c_int64), ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("oz", ctypes.c_double), ("ox", ctypes.c_double), ("oy", ctypes.c_double),
        ("dz", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
        ("fat", ctypes.c_int64), ("z_pad_minus", ctypes.c_int64), ("z_pad_plus", ctypes.c_int64),
        ("x_pad_minus", ctypes.c_int64), ("x_pad_plus", ctypes.c_int64), ("y_pad", ctypes.c_int64),
        ("tolerance", ctypes.c_double),
    ]


class ElaError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status."""


_lib = None


def _share_hip_runtime() -> None:
    """PyTorch-ROCm ships its own libamdhip64 with the same SONAME as the system
    one. Loading torch first makes this library bind to that same runtime, so
    that one process has one HIP/HSA runtime (two runtimes cannot both own the
    GPU)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load() -> ctypes.CDLL:
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    _share_hip_runtime()
    if not os.path.exists(LIB_PATH):
        raise ElaError(
            f"{LIB_PATH} not found: build it with `python -m elasticisolib_3d_amd.build` "
            "(the propagator has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    lib.ela_last_error.restype = ctypes.c_char_p
    lib.ela_version.restype = ctypes.c_char_p
    lib.ela_create.argtypes = [ctypes.POINTER(ElaConfig), ctypes.POINTER(vp)]
    lib.ela_destroy.argtypes = [vp]
    lib.ela_set_stream.argtypes = [vp, vp]
    lib.ela_synchronize.argtypes = [vp]
    lib.ela_set_medium.argtypes = [vp, vp, i32]
    lib.ela_scale_medium.argtypes = [i64, i64, i64, ctypes.c_double, vp, vp, vp, vp]
    for name in ("ela_born_stagger", "ela_born_unstagger", "ela_medium_gradient"):
        getattr(lib, name).argtypes = [i64, i64, i64, ctypes.c_double, vp, vp, vp]
    P, Tr = ctypes.POINTER(ElaPoints), ctypes.POINTER(ElaTraces)
    lib.ela_nl_forward.argtypes = [vp, P, Tr, P, Tr, i32]
    lib.ela_nl_forward_noalloc.argtypes = [vp, P, Tr, P, Tr]
    lib.ela_set_born_mode.argtypes = [vp, i32]
    lib.ela_nl_adjoint.argtypes = [vp, P, Tr, P, Tr, i32]
    lib.ela_born_forward.argtypes = [vp, P, Tr, vp, P, Tr, i32]
    lib.ela_born_adjoint.argtypes = [vp, P, Tr, vp, P, Tr, i32]
    lib.ela_interp_linear.argtypes = [ctypes.POINTER(ElaInterpAxes), i64, vp, vp, vp, vp, vp, vp, ctypes.POINTER(i64)]
    lib.ela_interp_weights.argtypes = [ctypes.POINTER(ElaInterpAxes), i64, vp, vp, vp, i32, i32, i32,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, vp, vp, vp,
                                       ctypes.POINTER(i64)]
    lib.ela_set_tiling.argtypes = [vp, i32, i32]
    lib.ela_set_tile_order.argtypes = [vp, i32, i32]
    lib.ela_set_born_storage.argtypes = [vp, i64, i64]
    lib.ela_set_windows.argtypes = [vp, i32, vp, vp, vp]
    lib.ela_set_graph.argtypes = [vp, i32]
    lib.ela_stream_probe.argtypes = [vp, i32, i32, i64, i32, ctypes.POINTER(ctypes.c_double)]
    lib.ela_set_elastic.argtypes = [vp, vp, ctypes.c_double, i32]
    lib.ela_get_medium.argtypes = [vp, vp, i32]
    lib.ela_gradient_forward.argtypes = [vp, P, Tr, P, Tr, i32]
    lib.ela_gradient_backward.argtypes = [vp, P, Tr, P, Tr, vp, Tr, i32]
    lib.ela_set_gradient_checkpoints.argtypes = [vp, i64, i64, i32, i64]
    lib.ela_set_gradient_derivatives.argtypes = [vp, i32]
    lib.ela_gradient_derivatives_used.argtypes = [vp, ctypes.POINTER(i32)]
    lib.ela_time_gradient_kernels.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_double)]
    lib.ela_set_tiling_transposed.argtypes = [vp, i32]
    lib.ela_get_tiling_transposed.argtypes = [vp, ctypes.POINTER(i32)]
    lib.ela_gradient_plan_used.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i32),
                                           ctypes.POINTER(i64)]
    lib.ela_born_storage_used.argtypes = [vp, ctypes.POINTER(i64)]
    lib.ela_get_tiling.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.ela_time_step_kernels.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.ela_last_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i64), ctypes.POINTER(i64)]
    for name in EXPORTS:
        if name not in ("ela_last_error", "ela_version"):
            getattr(lib, name).restype = i32
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != ELA_OK:
        msg = load().ela_last_error().decode(errors="replace")
        raise ElaError(f"{what} failed (status {rc}): {msg}")
