"""ctypes binding of libssam.so (the C ABI declared in include/ssam.h).

The library is loaded from this package directory only; if it is missing the
import fails loudly (there is no CPU fallback for any entry point).
"""
from __future__ import annotations

import ctypes
import pathlib

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libssam.so"
HEADER = HERE.parent / "include" / "ssam.h"

OK, E_INVALID, E_UNSUPPORTED, E_HIP, E_NOMEM = 0, 1, 2, 3, 4
F32, F64, BF16, F16 = 0, 1, 2, 3
STAR, BOX, DENSE, FACE_EDGE = 0, 1, 2, 3
ZERO, CLAMP, HOLD, DIRICHLET0 = 0, 1, 2, 3
FLAG_COMPENSATED = 1
FLAG_WORK_FP32 = 2


class SsamError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {strerror(status)} (status {status})")


class Stencil2DDesc(ctypes.Structure):
    _fields_ = [
        ("storage", ctypes.c_int32),
        ("shape", ctypes.c_int32),
        ("fw", ctypes.c_int32),
        ("fh", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("boundary", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("coeffs", ctypes.POINTER(ctypes.c_double)),
    ]


class Stencil3DDesc(ctypes.Structure):
    _fields_ = [
        ("storage", ctypes.c_int32),
        ("shape", ctypes.c_int32),
        ("radius", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("boundary", ctypes.c_int32),
        ("flags", ctypes.c_uint32),
        ("coeffs", ctypes.POINTER(ctypes.c_double)),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("storage", ctypes.c_int32),
        ("shape", ctypes.c_int32),
        ("radius", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("nx", ctypes.c_int64),
        ("ny", ctypes.c_int64),
        ("nz", ctypes.c_int64),
        ("pitch", ctypes.c_int64),
        ("src", ctypes.c_void_p),
        ("dst", ctypes.c_void_p),
        ("work", ctypes.c_void_p),
        ("coeffs", ctypes.POINTER(ctypes.c_double)),
    ]


BATCH_EAGER = 1

_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_f = ctypes.c_float
_d = ctypes.c_double

_PROTOS = {
    "ssam_version": (ctypes.c_char_p, []),
    "ssam_strerror": (ctypes.c_char_p, [_i]),
    "ssam_stencil2d": (_i, [ctypes.POINTER(Stencil2DDesc), _vp, _vp, _i64, _i64, _i64, _vp, _vp]),
    "ssam_j2d5pt_f32": (_i, [_vp, _vp, _i, _i, _f, _f, _f, _f, _f, _vp]),
    "ssam_j2d5pt_f64": (_i, [_vp, _vp, _i, _i, _d, _d, _d, _d, _d, _vp]),
    "ssam_j2dstar_f32": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "ssam_j2dstar_f64": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "ssam_j2dbox_f32": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "ssam_j2dbox_f64": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "ssam_j2d5pt_temporal_f32": (_i, [_vp, _vp, _i, _i, _f, _f, _f, _f, _f, _i, _vp, _vp]),
    "ssam_j2d5pt_temporal_f64": (_i, [_vp, _vp, _i, _i, _d, _d, _d, _d, _d, _i, _vp, _vp]),
    "ssam_j2d9pt_temporal_f32": (_i, [_vp, _vp, _i, _i] + [_f] * 9 + [_i, _vp, _vp]),
    "ssam_j2d9pt_temporal_f64": (_i, [_vp, _vp, _i, _i] + [_d] * 9 + [_i, _vp, _vp]),
    "ssam_conv2d_f32": (_i, [_vp, _vp, _i, _i, _i, _vp, _i, _i, _vp]),
    "ssam_conv2d_f64": (_i, [_vp, _vp, _i, _i, _i, _vp, _i, _i, _vp]),
    "ssam_stencil3d": (_i, [ctypes.POINTER(Stencil3DDesc), _vp, _vp, _i64, _i64, _i64, _vp, _vp]),
    "ssam_j3d7pt_f32": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "ssam_j3d7pt_f64": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "ssam_j3d13pt_f32": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "ssam_j3d13pt_f64": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "ssam_poisson_f32": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "ssam_poisson_f64": (_i, [_vp, _vp, _i, _i, _vp, _vp]),
    "ssam_jacobi19": (_i, [_i, _vp, _vp, _i64, _i64, _i64, _i, _vp, ctypes.POINTER(_d), _vp]),
    "ssam_jacobi19_noalloc": (_i, [_i, _vp, _vp, _i64, _i64, _i64, _i, _vp, _vp, _vp]),
    "ssam_tune_stencil2d": (_i, [ctypes.POINTER(Stencil2DDesc), _vp, _vp, _i64, _i64, _i64, _vp, _vp,
                                 ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "ssam_tune_clear": (None, []),
    "ssam_tune_save": (_i, [ctypes.c_char_p]),
    "ssam_tune_load": (_i, [ctypes.c_char_p]),
    "ssam_batch_plan_create": (_i, [ctypes.POINTER(BatchItem), _i, ctypes.POINTER(ctypes.c_void_p)]),
    "ssam_batch_run": (_i, [ctypes.c_void_p, _vp, ctypes.c_uint32]),
    "ssam_batch_plan_info": (_i, [ctypes.c_void_p, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                  ctypes.POINTER(_i64)]),
    "ssam_batch_plan_destroy": (None, [ctypes.c_void_p]),
    "ssam_batch_plan_tune": (_i, [ctypes.c_void_p, _vp]),
    "ssam_microbench": (_i, [_i, ctypes.POINTER(ctypes.c_double), _vp]),
    "ssam_batch_plan_tiling": (_i, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                    ctypes.POINTER(ctypes.c_int32)]),
    "ssam_fill_uniform": (_i, [_vp, _i, _i64, _u64, _u64, _d, _d, _vp]),
    "ssam_stream_copy": (_i, [_vp, _vp, _i64, _vp]),
}

_LIB = None


def lib() -> ctypes.CDLL:
    """The loaded libssam.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m ssam_amd.build` "
                "(the SSAM entry points have no fallback implementation)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def strerror(status: int) -> str:
    return lib().ssam_strerror(status).decode()


def check(status: int, what: str) -> None:
    if status != OK:
        raise SsamError(status, what)


def declared_symbols() -> list[str]:
    """Every function name include/ssam.h declares."""
    import re
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(ssam_\w+)\s*\(", text, re.M)))
