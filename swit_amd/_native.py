"""ctypes binding of the C ABI in include/swit_fd.h (libswit_fd.so, gfx950).

The library is built in-tree by ``swit_amd.build`` (hipcc, --offload-arch=gfx950).
There is no fallback: if the library is missing, loading fails loudly.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libswit_fd.so")

# Error codes and flags, include/swit_fd.h
SWIT_OK = 0
SWIT_EINVAL = 1
SWIT_ENOMEM = 2
SWIT_EHIP = 3
SWIT_ESTATE = 4
SWIT_PATH_PLAIN = 0x0
SWIT_PATH_BATCHED = 0x1
SWIT_STRIPS_F16 = 0x10
SWIT_STRIPS_BF16 = 0x20
SWIT_PLAN_EXACT = 0x40
SWIT_PLAN_EXACT_CKPT = 0x80
SWIT_PLAN_NO_STRIPS = 0x100

# Every entry point include/swit_fd.h declares.
EXPORTS = (
    "swit_plan_create",
    "swit_plan_set_model",
    "swit_forward",
    "swit_adjoint",
    "swit_plan_destroy",
    "swit_last_error",
    "swit_plan_scratch_bytes",
    "swit_plan_exact_generation",
    "swit_plan_strips",
    "swit_plan_autotune",
    "swit_forward_exact",
    "swit_adjoint_exact",
    "swit_misfit",
    "swit_group_create",
    "swit_group_forward",
    "swit_group_adjoint",
    "swit_group_destroy",
    "swit_group_gradient",
    "swit_plan_set_option",
    "swit_plan_kernel_time",
    "swit_grad_precond_work",
    "swit_grad_precond",
    "swit_grad_precond_noalloc",
    "swit_process_work",
    "swit_process",
)

SWIT_MISFIT_WAVEFORM = 0
SWIT_PRECOND_SCRATCH_BYTES = 64 << 20
SWIT_MISFIT_GLOBAL_CORRELATION = 1
SWIT_MISFIT_ENVELOPE = 2
SWIT_MISFIT_TRAVELTIME = 3
SWIT_OPT_GRAPH = 1
SWIT_OPT_PROFILE = 2
SWIT_OPT_FWD_GROUP = 4
SWIT_OPT_FWD_COLGROUP = 8
SWIT_OPT_BWD_COLGROUP = 16
SWIT_OPT_BWD_STEPS = 32
SWIT_OPT_EXACT_FUSED = 64
SWIT_OPT_BWD_TILE_X = 128
SWIT_OPT_STREAMS = 256
SWIT_OPT_BWD_BALANCE = 512


class SwitPrecond(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nz", ctypes.c_int), ("marine", ctypes.c_int), ("grad_mute", ctypes.c_int),
                ("grad_thred", ctypes.c_float), ("grad_smooth", ctypes.c_int), ("vpmax", ctypes.c_double)]


SWIT_PROCESS_MAX_TAPS = 16
SWIT_NORM = {"none": 0, "l1": 1, "l2": 2, "max": 3}


class SwitProcessParams(ctypes.Structure):
    _fields_ = [("ngathers", ctypes.c_int), ("ntr", ctypes.c_int), ("nt", ctypes.c_int), ("dt", ctypes.c_double),
                ("mute_offset_short", ctypes.c_int), ("mute_offset_long", ctypes.c_int),
                ("mute_offset_short_dis", ctypes.c_double), ("mute_offset_long_dis", ctypes.c_double),
                ("mute_late_arrival", ctypes.c_int), ("mute_late_window", ctypes.c_double),
                ("filter_taps", ctypes.c_int), ("b", ctypes.c_double * SWIT_PROCESS_MAX_TAPS),
                ("a", ctypes.c_double * SWIT_PROCESS_MAX_TAPS), ("zi", ctypes.c_double * SWIT_PROCESS_MAX_TAPS),
                ("detrend_taper", ctypes.c_int), ("norm_trace", ctypes.c_int), ("norm_event", ctypes.c_int)]


class SwitGrid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("npml", ctypes.c_int32),
        ("nt", ctypes.c_int32),
        ("free_surface", ctypes.c_int32),
        ("dx", ctypes.c_float),
        ("dt", ctypes.c_float),
    ]


class SwitShots(ctypes.Structure):
    _fields_ = [
        ("ns", ctypes.c_int32),
        ("ngmax", ctypes.c_int32),
        ("ng", ctypes.POINTER(ctypes.c_int32)),
        ("src_xz", ctypes.POINTER(ctypes.c_float)),
        ("rec_xz", ctypes.POINTER(ctypes.c_float)),
    ]


class NativeLibraryMissing(ImportError):
    pass


class SwitError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"swit error {code}: {message}")
        self.code = code


_lib = None


def load(path: str | None = None):
    """Load libswit_fd.so and declare its signatures (no GPU call is made)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise NativeLibraryMissing(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(p)
    vp = ctypes.c_void_p
    L.swit_plan_create.argtypes = [ctypes.POINTER(SwitGrid), ctypes.POINTER(SwitShots), ctypes.c_int,
                                   ctypes.POINTER(vp)]
    L.swit_plan_create.restype = ctypes.c_int
    L.swit_plan_set_model.argtypes = [vp, vp, vp, vp]
    L.swit_plan_set_model.restype = ctypes.c_int
    L.swit_forward.argtypes = [vp, vp, vp, ctypes.c_int, vp]
    L.swit_forward.restype = ctypes.c_int
    L.swit_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.swit_adjoint.restype = ctypes.c_int
    L.swit_plan_destroy.argtypes = [vp]
    L.swit_plan_destroy.restype = ctypes.c_int
    L.swit_last_error.argtypes = []
    L.swit_last_error.restype = ctypes.c_char_p
    L.swit_forward_exact.argtypes = [vp, vp, vp, vp]
    L.swit_forward_exact.restype = ctypes.c_int
    L.swit_adjoint_exact.argtypes = [vp, vp, vp, vp]
    L.swit_adjoint_exact.restype = ctypes.c_int
    L.swit_plan_autotune.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), vp]
    L.swit_plan_autotune.restype = ctypes.c_int
    L.swit_misfit.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, vp, vp, vp, vp, vp, vp]
    L.swit_misfit.restype = ctypes.c_int
    L.swit_grad_precond_work.argtypes = [ctypes.POINTER(SwitPrecond), ctypes.POINTER(ctypes.c_size_t)]
    L.swit_grad_precond_work.restype = ctypes.c_int
    L.swit_grad_precond.argtypes = [ctypes.POINTER(SwitPrecond), vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp]
    L.swit_grad_precond.restype = ctypes.c_int
    L.swit_grad_precond_noalloc.argtypes = [ctypes.POINTER(SwitPrecond), vp, vp, vp, vp, vp, vp]
    L.swit_grad_precond_noalloc.restype = ctypes.c_int
    L.swit_process_work.argtypes = [ctypes.POINTER(SwitProcessParams), ctypes.POINTER(ctypes.c_size_t)]
    L.swit_process_work.restype = ctypes.c_int
    L.swit_process.argtypes = [ctypes.POINTER(SwitProcessParams), vp, vp, vp, vp, vp, vp, ctypes.c_size_t, vp]
    L.swit_process.restype = ctypes.c_int
    pp = ctypes.POINTER(vp)
    L.swit_group_create.argtypes = [pp, ctypes.c_int, pp]
    L.swit_group_create.restype = ctypes.c_int
    L.swit_group_forward.argtypes = [vp, pp, pp, ctypes.c_int, vp]
    L.swit_group_forward.restype = ctypes.c_int
    L.swit_group_adjoint.argtypes = [vp, pp, pp, pp, pp, pp, vp]
    L.swit_group_adjoint.restype = ctypes.c_int
    L.swit_group_gradient.argtypes = [vp, ctypes.c_int, pp, pp, pp, pp, pp, pp, pp, pp, pp, vp]
    L.swit_group_gradient.restype = ctypes.c_int
    L.swit_group_destroy.argtypes = [vp]
    L.swit_group_destroy.restype = ctypes.c_int
    L.swit_plan_strips.argtypes = [vp, vp, vp]
    L.swit_plan_strips.restype = ctypes.c_int
    L.swit_plan_scratch_bytes.argtypes = [vp]
    L.swit_plan_scratch_bytes.restype = ctypes.c_int64
    L.swit_plan_exact_generation.argtypes = [vp]
    L.swit_plan_exact_generation.restype = ctypes.c_int64
    L.swit_plan_set_option.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.swit_plan_set_option.restype = ctypes.c_int
    L.swit_plan_kernel_time.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int64)]
    L.swit_plan_kernel_time.restype = ctypes.c_int
    if path is None:
        _lib = L
    return L


def check(rc: int):
    if rc != SWIT_OK:
        msg = load().swit_last_error().decode(errors="replace")
        raise SwitError(rc, msg)
