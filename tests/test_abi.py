"""C ABI checks that need no GPU: libssam.so loads, exports every function
include/ssam.h declares, and rejects malformed requests with status codes before
any HIP call (the reference harness has no such checks; its failure modes are
asserts and exits, ssai-2dconv/cudaLib.cuh:42-51)."""
import ctypes
import subprocess

import numpy as np
import pytest

from ssam_amd import _lib as L

FAKE_A = ctypes.c_void_p(0x10000000)
FAKE_B = ctypes.c_void_p(0x20000000)
FAKE_W = ctypes.c_void_p(0x30000000)


def test_library_exports_every_declared_symbol():
    syms = L.declared_symbols()
    assert len(syms) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", str(L.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    lib = L.lib()
    for s in syms:
        assert hasattr(lib, s)


def test_version_and_strerror():
    lib = L.lib()
    assert b"gfx950" in lib.ssam_version()
    assert L.strerror(L.OK) == "success"
    assert L.strerror(L.E_INVALID) == "invalid argument"
    assert L.strerror(99) == "unknown status"


def _desc(storage=L.F32, shape=L.STAR, k=3, steps=1, boundary=L.ZERO, coeffs=None):
    c = np.ones((k, k)) if coeffs is None else coeffs
    d = L.Stencil2DDesc()
    d.storage, d.shape, d.fw, d.fh = storage, shape, c.shape[1], c.shape[0]
    d.steps, d.boundary, d.flags = steps, boundary, 0
    d.coeffs = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return d, c


def _s2d(d, src=FAKE_A, dst=FAKE_B, w=64, h=64, pitch=64, work=None):
    return L.lib().ssam_stencil2d(ctypes.byref(d) if d is not None else None, src, dst, w, h,
                                  pitch, work, None)


def test_stencil2d_rejects_malformed_requests():
    d, c = _desc()
    assert _s2d(None) == L.E_INVALID
    assert _s2d(d, src=None) == L.E_INVALID
    assert _s2d(d, dst=None) == L.E_INVALID
    assert _s2d(d, w=0) == L.E_INVALID
    assert _s2d(d, h=-1) == L.E_INVALID
    assert _s2d(d, pitch=10) == L.E_INVALID                 # pitch < width
    assert _s2d(d, dst=FAKE_A) == L.E_INVALID               # src == dst
    assert _s2d(d, dst=ctypes.c_void_p(0x10000000 + 64)) == L.E_INVALID  # overlap
    for bad in [dict(storage=7), dict(shape=9), dict(steps=0), dict(boundary=5)]:
        dd, _ = _desc(**bad)
        assert _s2d(dd) == L.E_INVALID, bad
    dd, _ = _desc(k=4)                                      # even star
    assert _s2d(dd) == L.E_INVALID
    dd, _ = _desc(boundary=L.HOLD, steps=2)                 # HOLD is single-step
    assert _s2d(dd) == L.E_INVALID
    dd, _ = _desc(steps=3)                                  # 2+1 launches need work
    assert _s2d(dd) == L.E_INVALID
    dd, _ = _desc(steps=3)
    assert _s2d(dd, work=FAKE_A) == L.E_INVALID             # work aliases src
    dd, _ = _desc(shape=L.BOX, k=31)                        # radius 15 box
    assert _s2d(dd) == L.E_UNSUPPORTED
    dd, _ = _desc(shape=L.STAR, k=15)                       # radius 7 star
    assert _s2d(dd) == L.E_UNSUPPORTED
    dd, _ = _desc(shape=L.DENSE, coeffs=np.ones((3, 31)))  # 31 columns: past 29 x 29
    assert _s2d(dd) == L.E_UNSUPPORTED


def test_stencil2d_flags():
    dd, _ = _desc(shape=L.BOX)
    dd.flags = 4                                            # unknown flag bit
    assert _s2d(dd) == L.E_INVALID
    # BF16 / FP16 runs of several fused launches keep their state in FP32: without
    # the flag that declares an FP32 work area they are rejected before any launch
    for sto in (L.BF16, L.F16):
        dd, _ = _desc(storage=sto, steps=16)
        assert _s2d(dd, work=FAKE_W) == L.E_INVALID
    dd, _ = _desc(shape=L.STAR)                             # compensated star
    dd.flags = L.FLAG_COMPENSATED
    assert _s2d(dd) == L.E_UNSUPPORTED
    dd, _ = _desc(shape=L.BOX, steps=2)                     # compensated fused steps
    dd.flags = L.FLAG_COMPENSATED
    assert _s2d(dd, work=FAKE_W) == L.E_UNSUPPORTED
    dd, _ = _desc(shape=L.BOX, k=21)                        # compensated radius 10
    dd.flags = L.FLAG_COMPENSATED
    assert _s2d(dd) == L.E_UNSUPPORTED


def test_reference_entries_validate():
    lib = L.lib()
    assert lib.ssam_j2dstar_f32(FAKE_A, FAKE_B, 64, 64, 1, None, None) == L.E_INVALID
    c = (ctypes.c_float * 33)()
    assert lib.ssam_j2dstar_f32(FAKE_A, FAKE_B, 64, 64, 8, c, None) == L.E_UNSUPPORTED
    assert lib.ssam_j2dbox_f64(FAKE_A, FAKE_B, 64, 64, 15, c, None) == L.E_UNSUPPORTED
    assert lib.ssam_j2d5pt_f32(FAKE_A, FAKE_A, 64, 64, 1, 1, 1, 1, 1, None) == L.E_INVALID
    assert lib.ssam_conv2d_f32(FAKE_A, FAKE_B, 64, 64, 64, None, 3, 3, None) == L.E_INVALID
    # the reference dispatches filter sizes 2..28 (conv2D.cu:186-212); 30 is past that
    assert lib.ssam_conv2d_f32(FAKE_A, FAKE_B, 64, 64, 64, FAKE_W, 30, 30, None) == L.E_UNSUPPORTED
    assert lib.ssam_conv2d_f64(FAKE_A, FAKE_B, 64, 64, 64, FAKE_W, 31, 3, None) == L.E_UNSUPPORTED


def test_stencil3d_validates():
    lib = L.lib()
    c = np.ones(19)
    d = L.Stencil3DDesc()
    d.storage, d.shape, d.radius, d.steps, d.boundary, d.flags = L.F32, L.STAR, 1, 1, L.DIRICHLET0, 0
    d.coeffs = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    f = lambda **kw: lib.ssam_stencil3d(ctypes.byref(d), kw.get("src", FAKE_A),
                                        kw.get("dst", FAKE_B), 16, 16, 16, kw.get("work"), None)
    assert lib.ssam_stencil3d(None, FAKE_A, FAKE_B, 16, 16, 16, None, None) == L.E_INVALID
    assert f(dst=FAKE_A) == L.E_INVALID
    d.steps = 2
    assert f() == L.E_INVALID                                # needs work
    d.steps = 1
    d.radius = 3
    assert f() == L.E_UNSUPPORTED
    d.radius = 1
    d.storage = 7
    assert f() == L.E_INVALID
    # BF16 / FP16 multi-sweep runs keep their state in FP32: a work area not
    # declared as two FP32 grids is refused (the state is never rounded)
    for sto in (L.BF16, L.F16):
        d.storage, d.steps = sto, 3
        assert f(work=FAKE_W) == L.E_INVALID
        assert f() == L.E_INVALID
    d.flags = 4
    assert f(work=FAKE_W) == L.E_INVALID                     # unknown flag
    d.storage, d.steps, d.flags = L.F32, 1, 0
    d.boundary = L.ZERO
    assert f() == L.E_UNSUPPORTED
    assert lib.ssam_jacobi19(L.F32, FAKE_A, FAKE_B, 0, 4, 4, 1, None, None, None) == L.E_INVALID
    assert lib.ssam_j3d7pt_f32(FAKE_A, FAKE_B, 16, 2, None, None) == L.E_INVALID


def test_utilities_validate():
    lib = L.lib()
    assert lib.ssam_fill_uniform(None, L.F32, 10, 0, 0, 0.0, 1.0, None) == L.E_INVALID
    assert lib.ssam_fill_uniform(FAKE_A, 9, 10, 0, 0, 0.0, 1.0, None) == L.E_INVALID
    assert lib.ssam_fill_uniform(FAKE_A, L.F32, 0, 0, 0, 0.0, 1.0, None) == L.OK
    assert lib.ssam_stream_copy(FAKE_A, FAKE_B, 10, None) == L.E_INVALID
    assert lib.ssam_stream_copy(ctypes.c_void_p(0x10000004), FAKE_B, 16, None) == L.E_INVALID
    assert lib.ssam_stream_copy(FAKE_A, FAKE_B, 0, None) == L.OK


def test_tuner_cache_io(tmp_path):
    lib = L.lib()
    lib.ssam_tune_clear()
    assert lib.ssam_tune_load(str(tmp_path / "missing.txt").encode()) == L.E_INVALID
    assert lib.ssam_tune_load(None) == L.E_INVALID
    p = tmp_path / "tune.txt"
    assert lib.ssam_tune_save(str(p).encode()) == L.OK
    assert p.read_text().startswith("# ssam autotune v2")
    assert lib.ssam_tune_load(str(p).encode()) == L.E_INVALID   # header only: no entry
    p.write_text("0 0 1 1 8192 8192 8 0 0.125\n1 1 5 1 64 64 4 0 0.5\nbad line\n")
    assert lib.ssam_tune_load(str(p).encode()) == L.OK
    q = tmp_path / "again.txt"
    assert lib.ssam_tune_save(str(q).encode()) == L.OK
    lines = [l for l in q.read_text().splitlines() if not l.startswith("#")]
    assert len(lines) == 2
    lib.ssam_tune_clear()
    d, c = _desc()
    assert lib.ssam_tune_stencil2d(None, FAKE_A, FAKE_B, 64, 64, 64, None, None, None, None) == L.E_INVALID
    assert lib.ssam_tune_stencil2d(ctypes.byref(d), FAKE_A, FAKE_A, 64, 64, 64, None, None, None, None) == L.E_INVALID


def test_batch_plan_validates():
    lib = L.lib()
    h = ctypes.c_void_p()
    assert lib.ssam_batch_plan_create(None, 1, ctypes.byref(h)) == L.E_INVALID
    arr = (L.BatchItem * 1)()
    assert lib.ssam_batch_plan_create(arr, 0, ctypes.byref(h)) == L.E_INVALID
    c = np.ones(9)
    b = arr[0]
    b.ndim, b.storage, b.shape, b.radius, b.steps = 2, L.F32, L.STAR, 1, 1
    b.nx, b.ny, b.nz, b.pitch = 64, 64, 1, 64
    b.src, b.dst = 0x10000000, 0x10000000   # aliasing
    b.coeffs = c.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.ssam_batch_plan_create(arr, 1, ctypes.byref(h)) == L.E_INVALID
    b.dst = 0x20000000
    b.ndim = 4
    assert lib.ssam_batch_plan_create(arr, 1, ctypes.byref(h)) == L.E_INVALID
    b.ndim, b.pitch = 2, 10                  # pitch < width
    assert lib.ssam_batch_plan_create(arr, 1, ctypes.byref(h)) == L.E_INVALID
    b.ndim, b.nz, b.storage, b.steps, b.pitch = 3, 8, 9, 1, 0
    assert lib.ssam_batch_plan_create(arr, 1, ctypes.byref(h)) == L.E_INVALID
    b.storage, b.steps = L.BF16, 4           # 3D multi-step without work (FP32 state)
    assert lib.ssam_batch_plan_create(arr, 1, ctypes.byref(h)) == L.E_INVALID
    b.storage, b.steps = L.F32, 4            # 3D multi-step without work
    assert lib.ssam_batch_plan_create(arr, 1, ctypes.byref(h)) == L.E_INVALID
    assert lib.ssam_batch_run(None, None, 0) == L.E_INVALID
    assert lib.ssam_batch_plan_tune(None, None) == L.E_INVALID
    assert lib.ssam_batch_plan_tiling(None, None, None, None, None) == L.E_INVALID


def test_microbench_validates():
    v = ctypes.c_double()
    assert L.lib().ssam_microbench(-1, ctypes.byref(v), None) == L.E_INVALID
    assert L.lib().ssam_microbench(20, ctypes.byref(v), None) == L.E_INVALID
    assert L.lib().ssam_microbench(0, None, None) == L.E_INVALID


def test_jacobi19_noalloc_validates_without_a_device():
    """The allocation-free Jacobi entry rejects bad requests before any device
    work: null pointers, unknown storage, and grids whose residual partials would
    outgrow its fixed array (131072 tiles)."""
    lib = L.lib()
    res = ctypes.c_void_p(0x20000000)
    assert lib.ssam_jacobi19_noalloc(L.F32, None, FAKE_B, 64, 64, 64, 1, None, res, None) == L.E_INVALID
    assert lib.ssam_jacobi19_noalloc(9, FAKE_A, FAKE_B, 64, 64, 64, 1, None, res, None) == L.E_INVALID
    assert lib.ssam_jacobi19_noalloc(L.BF16, FAKE_A, FAKE_B, 64, 64, 64, 2, None, None, None) == L.E_INVALID
    big = ctypes.c_void_p(0x900000000)
    assert lib.ssam_jacobi19_noalloc(L.F32, FAKE_A, big, 1 << 16, 1 << 14, 8, 1, None, res, None) == \
        L.E_UNSUPPORTED
