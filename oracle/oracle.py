"""ctypes binding of the CPU oracle (oracle/gold.c -> oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Each function restates a reference
gold loop (cited in gold.c).  `ref_conv2d` calls the reference's own
`Convolution<T>` (ssai-2dconv/common.h:186-216) compiled into oracle/_ref when that
build exists.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "libref_gold.so"

_lib = None
_ref = None
_vp = ctypes.c_void_p


def build(reference_dir: str | None = None) -> None:
    """make -C oracle (and the _ref build when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    ref = pathlib.Path(reference_dir or "/root/reference")
    if (ref / "ssai-2dconv" / "common.h").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REFERENCE_DIR={ref}"], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        L.orc_splitmix64_at.restype = ctypes.c_uint64
        L.orc_splitmix64_at.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        for n in ("orc_uniform_f32", "orc_uniform_f64"):
            getattr(L, n).argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64,
                                      ctypes.c_double, ctypes.c_double, _vp]
        for n in ("orc_conv2d_f32", "orc_conv2d_f64"):
            getattr(L, n).argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                      ctypes.c_int]
        for n in ("orc_conv2d_iter_f32", "orc_conv2d_iter_f64"):
            getattr(L, n).argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int]
        for n in ("orc_stencil3d_taps_f32", "orc_stencil3d_taps_f64"):
            getattr(L, n).argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _vp, _vp]
        for n in ("orc_poisson_ref_f32", "orc_poisson_ref_f64"):
            getattr(L, n).argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for n in ("orc_jacobi19_f32", "orc_jacobi19_f64"):
            getattr(L, n).argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            getattr(L, n).restype = ctypes.c_double
        L.orc_err_f64.argtypes = [_vp, _vp, _vp, ctypes.c_int64, _vp]
        L.orc_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def ref_lib():
    """The reference's own gold (oracle/_ref), or None where it was not built."""
    global _ref
    if _ref is None and REF_PATH.exists():
        R = ctypes.CDLL(str(REF_PATH))
        for n in ("ref_conv2d_f32", "ref_conv2d_f64"):
            getattr(R, n).argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                      ctypes.c_int]
        _ref = R
    return _ref


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype) -> str:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return "f32"
    if dt == np.float64:
        return "f64"
    raise TypeError(f"oracle supports float32/float64, got {dt}")


def max_threads() -> int:
    return lib().orc_max_threads()


def splitmix64_at(seed: int, i: int) -> int:
    return lib().orc_splitmix64_at(seed, i)


def uniform(seed: int, n: int, a: float = 0.0, b: float = 1.0, dtype=np.float32,
            offset: int = 0) -> np.ndarray:
    out = np.empty(n, dtype=dtype)
    getattr(lib(), f"orc_uniform_{_sfx(dtype)}")(seed, offset, n, a, b, _p(out))
    return out


def conv2d(src: np.ndarray, filt: np.ndarray) -> np.ndarray:
    """Reference 2D gold restated: zero-padded correlation (common.h:186-216)."""
    src = np.ascontiguousarray(src)
    f = np.ascontiguousarray(filt, dtype=src.dtype)
    h, w = src.shape
    out = np.empty_like(src)
    getattr(lib(), f"orc_conv2d_{_sfx(src.dtype)}")(_p(src), _p(out), w, h, _p(f), f.shape[1],
                                                     f.shape[0])
    return out


def conv2d_iter(src: np.ndarray, filt: np.ndarray, steps: int) -> np.ndarray:
    """`steps` zero-padded sweeps (stencil2d_temp_5pt.cu:169-175)."""
    img = np.array(src, copy=True, order="C")
    tmp = np.empty_like(img)
    f = np.ascontiguousarray(filt, dtype=img.dtype)
    h, w = img.shape
    getattr(lib(), f"orc_conv2d_iter_{_sfx(img.dtype)}")(_p(img), _p(tmp), w, h, _p(f),
                                                          f.shape[1], f.shape[0], steps)
    return img


def ref_conv2d(src: np.ndarray, filt: np.ndarray) -> np.ndarray | None:
    """The reference's Convolution<T> itself (None when oracle/_ref is absent)."""
    R = ref_lib()
    if R is None:
        return None
    src = np.ascontiguousarray(src)
    f = np.ascontiguousarray(filt, dtype=src.dtype)
    h, w = src.shape
    out = np.empty_like(src)
    getattr(R, f"ref_conv2d_{_sfx(src.dtype)}")(_p(src), _p(out), w, h, _p(f), f.shape[1],
                                                 f.shape[0])
    return out


def stencil3d_taps(inp: np.ndarray, taps, coef, radius: int, out: np.ndarray | None = None):
    """Interior-only 3D gold (tmp3d7pt-v2.cu:278-298 / tmp3d13pt.cu:301-327 form):
    out = sum_t coef[t] * in[p + tap_t], left to right in T; boundary untouched
    (zero when `out` is None)."""
    inp = np.ascontiguousarray(inp)
    nz, ny, nx = inp.shape
    if out is None:
        out = np.zeros_like(inp)
    t = np.ascontiguousarray(np.asarray(taps, dtype=np.int32).reshape(-1, 3))
    c = np.ascontiguousarray(np.asarray(coef, dtype=np.float64))
    getattr(lib(), f"orc_stencil3d_taps_{_sfx(inp.dtype)}")(_p(inp), _p(out), nx, ny, nz, radius,
                                                             len(c), _p(t), _p(c))
    return out


def poisson_ref(inp: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
    """poisson_gold (tmpPoisson.cu:369-386), interior only."""
    inp = np.ascontiguousarray(inp)
    nz, ny, nx = inp.shape
    if out is None:
        out = np.zeros_like(inp)
    getattr(lib(), f"orc_poisson_ref_{_sfx(inp.dtype)}")(_p(inp), _p(out), nx, ny, nz)
    return out


def jacobi19(inp: np.ndarray, f: np.ndarray | None = None) -> tuple[np.ndarray, float]:
    """One 19-point Jacobi sweep with zero Dirichlet boundary; returns (out, sum r^2)."""
    inp = np.ascontiguousarray(inp)
    nz, ny, nx = inp.shape
    out = np.empty_like(inp)
    fp = _p(np.ascontiguousarray(f, dtype=inp.dtype)) if f is not None else None
    r2 = getattr(lib(), f"orc_jacobi19_{_sfx(inp.dtype)}")(_p(inp), _p(out), fp, nx, ny, nz)
    return out, r2


def err(got: np.ndarray, ref: np.ndarray, scale: np.ndarray | None = None) -> dict:
    """max|d|, rms|d|, max |d|/scale (robust relative), sum|d|, exact matches, count."""
    g = np.ascontiguousarray(got, dtype=np.float64).ravel()
    r = np.ascontiguousarray(ref, dtype=np.float64).ravel()
    s = None if scale is None else np.ascontiguousarray(scale, dtype=np.float64).ravel()
    o = np.zeros(6)
    lib().orc_err_f64(_p(g), _p(r), _p(s) if s is not None else None, g.size, _p(o))
    return dict(max_abs=o[0], rms=o[1], max_rel=o[2], sum_abs=o[3], match=int(o[4]),
                count=int(o[5]))
