"""Host-side operators over libssam.so for torch device tensors.

Mirrors the reference harness operators (ssai-2dstencil/stencil2D_*.cu `Test*`,
ssai-2dconv/conv2D.cu `Test_conv2D`, ssai-3dstencil/tmp3d*.cu `_3d*`): the caller
owns the device buffers, src and dst never alias, and every call goes through the
C ABI of include/ssam.h.  Errors raise `SsamError` with the library's status; there
is no CPU fallback.  torch is used only for device memory and streams.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from ._lib import check

_DTYPES = {torch.float32: L.F32, torch.float64: L.F64, torch.bfloat16: L.BF16,
           torch.float16: L.F16}


def dtype_code(t: torch.dtype) -> int:
    try:
        return _DTYPES[t]
    except KeyError:
        raise TypeError(f"unsupported storage dtype {t}") from None


def _stream(stream: Optional[torch.cuda.Stream]) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _dev(t: Optional[torch.Tensor], name: str) -> ctypes.c_void_p:
    if t is None:
        return ctypes.c_void_p(0)
    if not t.is_cuda:
        raise ValueError(f"{name} must be a device tensor (got {t.device})")
    return ctypes.c_void_p(t.data_ptr())


def _check_2d(src: torch.Tensor, dst: torch.Tensor) -> tuple[int, int, int]:
    if src.dim() != 2 or dst.dim() != 2 or src.shape != dst.shape:
        raise ValueError("src and dst must be 2-D tensors of one shape")
    if src.dtype != dst.dtype:
        raise ValueError("src and dst must share the storage dtype")
    if src.stride(1) != 1 or dst.stride(1) != 1 or src.stride(0) != dst.stride(0):
        raise ValueError("rows must be contiguous with one common pitch")
    h, w = src.shape
    return w, h, src.stride(0)


def _work_flags(src: torch.Tensor, work: Optional[torch.Tensor], flags: int) -> int:
    """BF16 / FP16 grids with a float32 `work` of 2 * height * width values: the
    running state of multi-launch runs stays FP32 there (SSAM_FLAG_WORK_FP32)."""
    if work is not None and src.element_size() == 2 and work.dtype == torch.float32:
        h, w = src.shape
        if work.numel() < 2 * h * w or not work.is_contiguous():
            raise ValueError("a float32 work area for BF16/FP16 grids holds 2 * height * width values")
        flags |= L.FLAG_WORK_FP32
    return flags


def stencil2d(src: torch.Tensor, dst: torch.Tensor, coeffs, shape: int = L.BOX, steps: int = 1,
              boundary: int = L.ZERO, work: Optional[torch.Tensor] = None, flags: int = 0,
              stream: Optional[torch.cuda.Stream] = None) -> None:
    """`steps` sweeps of the fh x fw correlation `coeffs` (row-major, taps weight
    in(x + n - fw/2, y + m - fh/2)) over src -> dst (ssam_stencil2d).  BF16 / FP16
    runs of more than one fused launch need a float32 `work` of 2 * h * w values."""
    w, h, pitch = _check_2d(src, dst)
    flags = _work_flags(src, work, flags)
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    if c.ndim != 2:
        raise ValueError("coeffs must be a 2-D (fh, fw) array")
    d = L.Stencil2DDesc()
    d.storage = dtype_code(src.dtype)
    d.shape = shape
    d.fh, d.fw = c.shape
    d.steps = steps
    d.boundary = boundary
    d.flags = flags
    d.coeffs = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = L.lib().ssam_stencil2d(ctypes.byref(d), _dev(src, "src"), _dev(dst, "dst"), w, h, pitch,
                                _dev(work, "work"), _stream(stream))
    check(rc, "ssam_stencil2d")


def tune2d(src: torch.Tensor, dst: torch.Tensor, coeffs, shape: int = L.BOX, steps: int = 1,
           boundary: int = L.ZERO, work: Optional[torch.Tensor] = None, flags: int = 0,
           stream: Optional[torch.cuda.Stream] = None) -> tuple[int, int]:
    """Autotune the tiling of this request (ssam_tune_stencil2d): returns
    (waves_per_simd, fused_steps_per_launch); later stencil2d calls with the same
    (storage, shape, radius, steps, width, height) use it.  dst/work are overwritten."""
    w, h, pitch = _check_2d(src, dst)
    flags = _work_flags(src, work, flags)
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    d = L.Stencil2DDesc()
    d.storage = dtype_code(src.dtype)
    d.shape = shape
    d.fh, d.fw = c.shape
    d.steps = steps
    d.boundary = boundary
    d.flags = flags
    d.coeffs = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    wps, tch = ctypes.c_int(0), ctypes.c_int(0)
    check(L.lib().ssam_tune_stencil2d(ctypes.byref(d), _dev(src, "src"), _dev(dst, "dst"), w, h,
                                      pitch, _dev(work, "work"), _stream(stream), ctypes.byref(wps),
                                      ctypes.byref(tch)), "ssam_tune_stencil2d")
    return wps.value, tch.value


def tune_save(path: str) -> None:
    check(L.lib().ssam_tune_save(str(path).encode()), "ssam_tune_save")


def tune_load(path: str) -> None:
    check(L.lib().ssam_tune_load(str(path).encode()), "ssam_tune_load")


def tune_clear() -> None:
    L.lib().ssam_tune_clear()


def _sfx(t: torch.Tensor) -> str:
    if t.dtype == torch.float32:
        return "f32"
    if t.dtype == torch.float64:
        return "f64"
    raise TypeError("the reference-kernel entry points take float32 or float64")


def j2d5pt(src, dst, fc, fn, fs, fw, fe, stream=None) -> None:
    """stencil2d_v1::j2d5pt (stencil2D_5pt-v1.cu:8-9) via ssam_j2d5pt_*."""
    w, h, pitch = _check_2d(src, dst)
    if pitch != w:
        raise ValueError("j2d5pt takes unpitched grids (pitch == width)")
    fn_ = getattr(L.lib(), f"ssam_j2d5pt_{_sfx(src)}")
    check(fn_(_dev(src, "src"), _dev(dst, "dst"), w, h, fc, fn, fs, fw, fe, _stream(stream)),
          "ssam_j2d5pt")


def j2dstar(src, dst, radius: int, coefs: Sequence[float], stream=None) -> None:
    """j2d9pt..j2d25pt: coefs = fc, fn0, fs0, fw0, fe0, fn1, ... (4r+1 values)."""
    w, h, pitch = _check_2d(src, dst)
    if pitch != w:
        raise ValueError("j2dstar takes unpitched grids (pitch == width)")
    ct = np.ascontiguousarray(np.asarray(coefs, dtype=np.float32 if src.dtype == torch.float32
                                         else np.float64))
    if ct.size != 4 * radius + 1:
        raise ValueError("a star of radius r takes 4r+1 coefficients")
    fn_ = getattr(L.lib(), f"ssam_j2dstar_{_sfx(src)}")
    check(fn_(_dev(src, "src"), _dev(dst, "dst"), w, h, radius, ct.ctypes.data_as(ctypes.c_void_p),
              _stream(stream)), "ssam_j2dstar")


def j2dbox(src, dst, radius: int, wts, stream=None) -> None:
    """j2d64pt / j2d81pt / j2d121pt with run-time taps w[(2r+1)^2]."""
    w, h, pitch = _check_2d(src, dst)
    if pitch != w:
        raise ValueError("j2dbox takes unpitched grids (pitch == width)")
    ct = np.ascontiguousarray(np.asarray(wts, dtype=np.float32 if src.dtype == torch.float32
                                         else np.float64)).ravel()
    if ct.size != (2 * radius + 1) ** 2:
        raise ValueError("a box of radius r takes (2r+1)^2 taps")
    fn_ = getattr(L.lib(), f"ssam_j2dbox_{_sfx(src)}")
    check(fn_(_dev(src, "src"), _dev(dst, "dst"), w, h, radius, ct.ctypes.data_as(ctypes.c_void_p),
              _stream(stream)), "ssam_j2dbox")


def j2d5pt_temporal(src, dst, fc, fn, fs, fw, fe, steps: int, work=None, stream=None) -> None:
    """stencil2d_tmp::j2d5pt (stencil2d_temp_5pt.cu:9-10): `steps` fused sweeps."""
    w, h, pitch = _check_2d(src, dst)
    fn_ = getattr(L.lib(), f"ssam_j2d5pt_temporal_{_sfx(src)}")
    check(fn_(_dev(src, "src"), _dev(dst, "dst"), w, h, fc, fn, fs, fw, fe, steps,
              _dev(work, "work"), _stream(stream)), "ssam_j2d5pt_temporal")


def j2d9pt_temporal(src, dst, c9: Sequence[float], steps: int, work=None, stream=None) -> None:
    """stencil2d_tmp::j2d9pt (stencil2d_temp_9pt.cu:12-14); c9 = fc, fn0, fs0, fw0,
    fe0, fn1, fs1, fw1, fe1."""
    w, h, pitch = _check_2d(src, dst)
    fn_ = getattr(L.lib(), f"ssam_j2d9pt_temporal_{_sfx(src)}")
    check(fn_(_dev(src, "src"), _dev(dst, "dst"), w, h, *[float(v) for v in c9], steps,
              _dev(work, "work"), _stream(stream)), "ssam_j2d9pt_temporal")


def conv2d(src, dst, weight: torch.Tensor, stream=None) -> None:
    """CONV2D::kernel_convolution2D (conv2D.cu:9-12): pitched rows, DEVICE weights
    (fh, fw).  Square filters are read by the kernel in place (asynchronous, so the
    weights must stay alive until the stream has run the call); other shapes are
    read back, which synchronises the stream."""
    w, h, pitch = _check_2d(src, dst)
    if weight.dim() != 2 or weight.dtype != src.dtype or not weight.is_contiguous():
        raise ValueError("weight must be a contiguous 2-D device tensor of the grid dtype")
    fh, fw = weight.shape
    fn_ = getattr(L.lib(), f"ssam_conv2d_{_sfx(src)}")
    check(fn_(_dev(src, "src"), _dev(dst, "dst"), w, pitch, h, _dev(weight, "weight"),
              fw, fh, _stream(stream)), "ssam_conv2d")


def _check_3d(src, dst) -> tuple[int, int, int]:
    if src.dim() != 3 or src.shape != dst.shape or not src.is_contiguous() or not dst.is_contiguous():
        raise ValueError("src and dst must be contiguous 3-D tensors (nz, ny, nx) of one shape")
    if src.dtype != dst.dtype:
        raise ValueError("src and dst must share the dtype")
    nz, ny, nx = src.shape
    return nx, ny, nz


def _work3d_flags(src: torch.Tensor, work: Optional[torch.Tensor]) -> int:
    """BF16 / FP16 grids with a float32 `work` of 2 * nx * ny * nz values: the
    running state between launches stays FP32 there (SSAM_FLAG_WORK_FP32)."""
    if work is not None and src.element_size() == 2 and work.dtype == torch.float32:
        if work.numel() < 2 * src.numel() or not work.is_contiguous():
            raise ValueError("a float32 work area for BF16/FP16 grids holds 2 * nx * ny * nz values")
        return L.FLAG_WORK_FP32
    return 0


def stencil3d(src, dst, coeffs, shape: int = L.STAR, radius: int = 1, steps: int = 1,
              work=None, stream=None) -> None:
    """ssam_stencil3d: `steps` sweeps, zero Dirichlet output boundary.  BF16 / FP16
    grids of more than one launch need a float32 `work` of 2 * nx * ny * nz values
    (their running state stays FP32)."""
    nx, ny, nz = _check_3d(src, dst)
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    d = L.Stencil3DDesc()
    d.storage = dtype_code(src.dtype)
    d.shape = shape
    d.radius = radius
    d.steps = steps
    d.boundary = L.DIRICHLET0
    d.flags = _work3d_flags(src, work)
    d.coeffs = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    check(L.lib().ssam_stencil3d(ctypes.byref(d), _dev(src, "src"), _dev(dst, "dst"), nx, ny, nz,
                                 _dev(work, "work"), _stream(stream)), "ssam_stencil3d")


def ref3d(name: str, src, dst, iters: int = 1, work=None, stream=None) -> None:
    """Reference 3D operators with reference coefficients: name in
    {'j3d7pt', 'j3d13pt', 'poisson'}; cubic grids."""
    nx, ny, nz = _check_3d(src, dst)
    if not nx == ny == nz:
        raise ValueError("the reference 3D operators take cubic grids")
    fn_ = getattr(L.lib(), f"ssam_{name}_{_sfx(src)}")
    check(fn_(_dev(src, "src"), _dev(dst, "dst"), nx, iters, _dev(work, "work"), _stream(stream)),
          f"ssam_{name}")


def jacobi19(src, dst, steps: int, work=None, residual: bool = True, stream=None) -> Optional[float]:
    """19-point Jacobi sweeps (zero Dirichlet); returns the FP64 residual L2 norm of
    the last sweep's input (synchronises the stream) when `residual`.  BF16 / FP16
    grids: `work` is float32 with 2 * nx * ny * nz values (the FP32 running state)."""
    nx, ny, nz = _check_3d(src, dst)
    if src.element_size() == 2 and work is not None:
        _work3d_flags(src, work)
        if work.dtype != torch.float32:
            raise ValueError("BF16/FP16 Jacobi runs keep their state in a float32 work area")
    r = ctypes.c_double(0.0)
    check(L.lib().ssam_jacobi19(dtype_code(src.dtype), _dev(src, "src"), _dev(dst, "dst"), nx, ny,
                                nz, steps, _dev(work, "work"),
                                ctypes.byref(r) if residual else None, _stream(stream)),
          "ssam_jacobi19")
    return r.value if residual else None


def fill_uniform(t: torch.Tensor, seed: int, a: float = 0.0, b: float = 1.0, offset: int = 0,
                 stream=None) -> torch.Tensor:
    """Seeded U[a,b) fill on the device (splitmix64, same values as the CPU oracle)."""
    if not t.is_contiguous():
        raise ValueError("fill_uniform needs a contiguous tensor")
    check(L.lib().ssam_fill_uniform(_dev(t, "t"), dtype_code(t.dtype), t.numel(), seed, offset,
                                    a, b, _stream(stream)), "ssam_fill_uniform")
    return t


def stream_copy(src: torch.Tensor, dst: torch.Tensor, stream=None) -> None:
    nbytes = src.numel() * src.element_size()
    check(L.lib().ssam_stream_copy(_dev(src, "src"), _dev(dst, "dst"), nbytes, _stream(stream)),
          "ssam_stream_copy")


class BatchPlan:
    """Many grids in one plan (ssam_batch_plan_create / ssam_batch_run).

    items: dicts with keys ndim (2/3), src, dst, coeffs, shape, radius, and for 3D
    steps and work (BF16 / FP16 3D items of more than one pass: float32 with
    2 * nx * ny * nz values, their FP32 running state).  2D tensors are (h, w) with
    contiguous rows; 3D tensors are contiguous (nz, ny, nx).  The plan binds the
    device pointers; keep the tensors alive while it exists."""

    def __init__(self, items):
        self._keep = []
        arr = (L.BatchItem * len(items))()
        for k, it in enumerate(items):
            src, dst = it["src"], it["dst"]
            b = arr[k]
            b.ndim = it["ndim"]
            b.storage = dtype_code(src.dtype)
            b.shape = it["shape"]
            b.radius = it["radius"]
            b.steps = it.get("steps", 1)
            if b.ndim == 2:
                w, h, pitch = _check_2d(src, dst)
                b.nx, b.ny, b.nz, b.pitch = w, h, 1, pitch
            else:
                nx, ny, nz = _check_3d(src, dst)
                b.nx, b.ny, b.nz, b.pitch = nx, ny, nz, 0
            b.src = _dev(src, "src").value
            b.dst = _dev(dst, "dst").value
            b.work = _dev(it.get("work"), "work").value
            c = np.ascontiguousarray(np.asarray(it["coeffs"], dtype=np.float64).ravel())
            self._keep.append(c)
            b.coeffs = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h = ctypes.c_void_p()
        check(L.lib().ssam_batch_plan_create(arr, len(items), ctypes.byref(h)), "ssam_batch_plan_create")
        self._h = h

    def run(self, eager: bool = False, stream=None) -> None:
        check(L.lib().ssam_batch_run(self._h, _stream(stream), L.BATCH_EAGER if eager else 0),
              "ssam_batch_run")

    def tune(self, stream=None) -> dict:
        """Autotune the tiling on the plan's buffers (ssam_batch_plan_tune; dst and
        work are overwritten) and return the chosen tiling."""
        check(L.lib().ssam_batch_plan_tune(self._h, _stream(stream)), "ssam_batch_plan_tune")
        return self.tiling()

    def tiling(self) -> dict:
        seg = (ctypes.c_int32 * 4)()
        aff = (ctypes.c_int32 * 9)()
        var, zseg = ctypes.c_int32(), ctypes.c_int32()
        check(L.lib().ssam_batch_plan_tiling(self._h, seg, ctypes.byref(var), ctypes.byref(zseg), aff),
              "ssam_batch_plan_tiling")
        return dict(seg_rows_2d=dict(zip(("star1", "box1", "box2", "box5"), list(seg))),
                    variant_3d=var.value, zseg_3d=zseg.value, affinity_2d=aff[0], affinity_3d=aff[1],
                    alternate_3d=aff[2], split_2d=aff[3], mfma_2d=aff[4], wps_2d=aff[5], wps_3d=aff[6],
                    grid_3d=aff[7], grid_2d=aff[8])

    def info(self) -> tuple[int, int, int]:
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(L.lib().ssam_batch_plan_info(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
              "ssam_batch_plan_info")
        return a.value, b.value, c.value

    def close(self) -> None:
        if self._h:
            L.lib().ssam_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MICROBENCH = ("fma_latency_cycles", "dpp_add_latency_cycles", "bpermute_latency_cycles",
              "lds_roundtrip_latency_cycles", "fma_per_cycle_per_cu", "dpp_add_per_cycle_per_cu",
              "bpermute_per_cycle_per_cu", "lds_roundtrip_per_cycle_per_cu", "mfma_bf16_16x16x32_per_cycle_per_cu",
              "fma_f32_tflops", "fma_f64_tflops", "mfma_f32_16x16x4_tflops", "mfma_f32_4x4x1_tflops",
              "mfma_f64_16x16x4_tflops", "fmac_vgpr_tflops", "fmac_sgpr_tflops", "fmac_16taps_tflops",
              "fmac_dpp_tflops", "fmac_vgpr_dpp_tflops", "fmac_vgpr_add_tflops")


def microbench(stream: Optional[torch.cuda.Stream] = None) -> dict:
    """All CDNA4 micro-benchmarks of ssam_microbench, by name."""
    out = {}
    for i, name in enumerate(MICROBENCH):
        v = ctypes.c_double()
        check(L.lib().ssam_microbench(i, ctypes.byref(v), _stream(stream)), "ssam_microbench")
        out[name] = v.value
    return out
