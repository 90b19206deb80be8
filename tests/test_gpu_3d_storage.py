"""BF16 / FP16 storage for the 3D stencils, against the FP64 oracle.

Grids live in HBM as BF16 / FP16; every sweep accumulates in FP32, and a run of
several launches keeps its running state in FP32 (two FP32 grids in `work`), so
the result carries the rounding of the input and of the output only.  Bar:
|gpu - gold64| <= tol * s per point with s the gold of |coef| over |x|, tol 1e-2
(BF16) and 1e-3 (FP16), at every step count; the gold runs in FP64 on the
unrounded input (tmp3d7pt-v2.cu:278-298 / tmp3d13pt.cu:301-327 forms).  Results
that decay below the storage type's smallest normal number (2^-14 for FP16) are
held to that number as their scale: no 2-byte format carries 1e-3 there.
"""
import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import _lib as L
from ssam_amd import api
from ssam_amd import workloads as wl

pytestmark = pytest.mark.gpu

STORAGE = [(torch.bfloat16, 1e-2), (torch.float16, 1e-3)]
SHAPES = [("star", 1), ("star", 2), ("face_edge", 1)]


def _code(shape):
    return L.FACE_EDGE if shape == "face_edge" else L.STAR


def _gold(x, taps, coef, r, steps):
    cur, s = x.astype(np.float64), np.abs(x).astype(np.float64)
    for _ in range(steps):
        cur = oracle.stencil3d_taps(cur, taps, coef, r)
        s = oracle.stencil3d_taps(s, taps, np.abs(np.asarray(coef)), r)
    return cur, s


MIN_NORMAL = {torch.bfloat16: 2.0 ** -126, torch.float16: 2.0 ** -14}


def _check(got, want, scale, tol, what, storage=None):
    assert np.isfinite(got).all(), f"{what}: unwritten or non-finite output"
    if storage is not None:
        scale = np.maximum(scale, MIN_NORMAL[storage])
    e = oracle.err(got, want, scale)
    assert e["max_rel"] <= tol, f"{what}: {e}"
    return e


def _run(x, storage, coef, shape, r, steps, t="2", monkeypatch=None):
    src = torch.from_numpy(x).to(storage).cuda()
    dst = torch.full_like(src, float("nan"))
    work = torch.full((2 * x.size,), float("nan"), dtype=torch.float32, device="cuda") if steps > 1 else None
    if monkeypatch is not None:
        monkeypatch.setenv("SSAM_3D_T", t)
    api.stencil3d(src, dst, coef, shape=_code(shape), radius=r, steps=steps, work=work)
    torch.cuda.synchronize()
    return dst.double().cpu().numpy()


@pytest.mark.parametrize("storage,tol", STORAGE)
@pytest.mark.parametrize("shape,r", SHAPES)
@pytest.mark.parametrize("dims", [(50, 33, 27), (64, 17, 12), (7, 6, 5), (130, 20, 40), (1, 1, 1)])
def test_single_sweep_vs_fp64(storage, tol, shape, r, dims):
    nx, ny, nz = dims
    taps = wl.taps3d(shape, r)
    coef = wl.uniform(r + 60, len(taps), -1, 1)
    x = oracle.uniform(31, nx * ny * nz, -1, 1, np.float64).reshape(nz, ny, nx)
    got = _run(x, storage, coef, shape, r, 1)
    want, sc = _gold(x, taps, coef, r, 1)
    _check(got, want, sc, tol, f"{storage} {shape} r={r} {dims}", storage)


@pytest.mark.parametrize("storage,tol", STORAGE)
@pytest.mark.parametrize("shape,r", SHAPES)
@pytest.mark.parametrize("steps", [2, 3, 10])
@pytest.mark.parametrize("dims", [(50, 33, 27), (130, 20, 40), (7, 6, 5)])
def test_multi_sweep_state_stays_fp32(storage, tol, shape, r, steps, dims, monkeypatch):
    """Several sweeps: one, two or three per launch give the same result bit for
    bit (the FP32 state between launches equals the FP32 value the fused launch
    keeps in registers), within the bar at every step count, and strictly closer
    to the gold than the same sweeps run one launch at a time through the storage
    type (state rounded every sweep)."""
    nx, ny, nz = dims
    taps = wl.taps3d(shape, r)
    coef = wl.uniform(r + 70, len(taps), 0, 1)
    coef = coef / coef.sum()
    x = oracle.uniform(steps + 41, nx * ny * nz, -1, 1, np.float64).reshape(nz, ny, nx)
    fused = _run(x, storage, coef, shape, r, steps, "3", monkeypatch)
    fused2 = _run(x, storage, coef, shape, r, steps, "2", monkeypatch)
    single = _run(x, storage, coef, shape, r, steps, "1", monkeypatch)
    assert np.array_equal(fused, single)
    assert np.array_equal(fused2, single)
    want, sc = _gold(x, taps, coef, r, steps)
    e = _check(fused, want, sc, tol, f"{storage} {shape} r={r} x{steps} {dims}", storage)
    # the same sweeps with the state rounded to the storage type between them
    cur = torch.from_numpy(x).to(storage).cuda()
    for _ in range(steps):
        nxt = torch.empty_like(cur)
        api.stencil3d(cur, nxt, coef, shape=_code(shape), radius=r)
        cur = nxt
    torch.cuda.synchronize()
    if nx * ny * nz >= 1000:
        rounded = oracle.err(cur.double().cpu().numpy(), want, sc)
        assert e["rms"] < rounded["rms"], (e, rounded)


@pytest.mark.parametrize("storage,tol", STORAGE)
def test_jacobi19_fp32_state_and_fp64_residual(storage, tol):
    n, steps = 40, 6
    x = oracle.uniform(23, n ** 3, 0, 1, np.float64).reshape(n, n, n)
    src = torch.from_numpy(x).to(storage).cuda()
    dst = torch.full_like(src, float("nan"))
    work = torch.empty(2 * n ** 3, dtype=torch.float32, device="cuda")
    res = api.jacobi19(src, dst, steps, work)
    cur, r2 = x, 0.0
    for _ in range(steps):
        cur, r2 = oracle.jacobi19(cur)
    _, sc = _gold(x, wl.taps3d("face_edge", 1), wl.JACOBI19, 1, steps)
    _check(dst.double().cpu().numpy(), cur, sc, tol, f"jacobi19 {storage}", storage)
    # the residual is formed in FP64 from the FP32 state of the last sweep's input
    assert abs(res - np.sqrt(r2)) <= 5 * tol * np.sqrt(r2), (res, np.sqrt(r2))


@pytest.mark.parametrize("variant", ["5", "6"])
@pytest.mark.parametrize("storage,tol", STORAGE)
def test_batch_3d_items_in_2byte_storage(storage, tol, variant, monkeypatch):
    """BF16 / FP16 3D items in a batch plan: the two- and three-sweep tile shapes,
    FP32 state between passes in each item's work (two FP32 grids), graph replays
    and an eager run against the FP64 gold."""
    monkeypatch.setenv("SSAM_BATCH3D_VARIANT", variant)
    specs = [(70, 33, 20, "star", 1), (64, 64, 64, "face_edge", 1), (257, 9, 11, "star", 10),
             (40, 40, 40, "face_edge", 3), (33, 17, 9, "star", 2), (50, 21, 30, "star", 7)]
    items, refs = [], []
    for k, (nx, ny, nz, shape, steps) in enumerate(specs):
        x = oracle.uniform(500 + k, nx * ny * nz, -1, 1, np.float64).reshape(nz, ny, nx)
        src = torch.from_numpy(x).to(storage).cuda()
        dst = torch.full_like(src, float("nan"))
        work = torch.empty(2 * nx * ny * nz, dtype=torch.float32, device="cuda")
        taps = wl.taps3d(shape, 1)
        coef = wl.uniform(600 + k, len(taps), 0, 1)
        coef = coef / coef.sum()
        items.append(dict(ndim=3, src=src, dst=dst, work=work, coeffs=coef, steps=steps,
                          shape=_code(shape), radius=1))
        refs.append((dst, _gold(x, taps, coef, 1, steps), f"{shape} {(nx, ny, nz)} x{steps}"))
    plan = api.BatchPlan(items)
    for rep in range(3):
        for dst, _, _ in refs:
            dst.fill_(float("nan"))
        plan.run(eager=(rep == 1))
        torch.cuda.synchronize()
        for dst, (want, sc), what in refs:
            _check(dst.double().cpu().numpy(), want, sc, tol, f"{storage} {what} rep {rep}", storage)
    plan.close()


def test_2byte_chain_needs_fp32_work():
    x = np.zeros((8, 8, 8))
    src = torch.from_numpy(x).to(torch.bfloat16).cuda()
    dst = torch.empty_like(src)
    with pytest.raises(L.SsamError) as e:
        api.stencil3d(src, dst, wl.REF3D_7PT, steps=3, work=torch.empty_like(src))
    assert e.value.status == L.E_INVALID
