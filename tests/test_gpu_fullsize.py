"""Parity at the sizes the bench reports, through the C ABI.

The oracle cannot redo whole 512^3 or 8192^2 runs in test time, so each result is
checked where the oracle can reach it exactly: the input rows (2D) or planes (3D)
a band needs are regenerated on the host from the device fill's splitmix64
stream, run through the oracle, and compared on rows / planes far enough from the
band's artificial edges (r * steps) to be exact gold.  The bar is the usual one:
|gpu - gold| <= tol * sum |w x| per point (1e-5 FP32, 1e-12 FP64 per sweep; 1e-2
BF16 and 1e-3 FP16 against the FP64 gold).
"""
import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import _lib as L
from ssam_amd import api
from ssam_amd import workloads as wl

pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-5, np.float64: 1e-12}
NPDT = {torch.float32: np.float32, torch.float64: np.float64}


def _err_ok(got, want, scale, tol, what):
    e = oracle.err(got, want, scale)
    assert np.isfinite(got).all(), what
    assert e["max_rel"] <= tol, f"{what}: {e}"


def _bands2d(got, seed, a, b, f, steps, tol, rows, npdt=np.float64, band=16):
    h, w = got.shape
    rt = (f.shape[0] // 2) * steps
    for y0 in rows:
        y0 = max(0, min(y0, h - band))
        lo, hi = max(0, y0 - rt), min(h, y0 + band + rt)
        x = oracle.uniform(seed, (hi - lo) * w, a, b, npdt, offset=lo * w).reshape(hi - lo, w).astype(np.float64)
        want = oracle.conv2d_iter(x, f, steps) if steps > 1 else oracle.conv2d(x, f)
        sc = np.abs(x)
        for _ in range(steps):
            sc = oracle.conv2d(sc, np.abs(f))
        k = y0 - lo
        if lo > 0 and k < rt:
            continue
        g = got[y0:y0 + band].double().cpu().numpy()
        _err_ok(g, want[k:k + band], sc[k:k + band], tol, f"rows {y0}..{y0 + band}")


def _slabs3d(got, x_of, n_planes, taps, coef, r, steps, tol, z0s, depth=4):
    for z0 in z0s:
        z0 = max(0, min(z0, n_planes - depth))
        halo = r * steps
        lo, hi = max(0, z0 - halo), min(n_planes, z0 + depth + halo)
        x = x_of(lo, hi)
        cur, sc = x, np.abs(x)
        for _ in range(steps):
            cur = oracle.stencil3d_taps(cur, taps, coef, r)
            sc = oracle.stencil3d_taps(sc, taps, np.abs(coef), r)
        k = z0 - lo
        g = got[z0:z0 + depth].double().cpu().numpy()
        _err_ok(g, cur[k:k + depth], sc[k:k + depth], tol, f"planes {z0}..{z0 + depth}")


def _planes(seed, nx, ny, npdt):
    return lambda lo, hi: oracle.uniform(seed, (hi - lo) * nx * ny, 0.0, 1.0, npdt,
                                         offset=lo * nx * ny).reshape(hi - lo, ny, nx).astype(np.float64)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
@pytest.mark.parametrize("name,shape,r,coef", [("7pt", L.STAR, 1, wl.REF3D_7PT), ("13pt", L.STAR, 2, wl.REF3D_13PT),
                                               ("poisson19", L.FACE_EDGE, 1, wl.REF3D_POISSON)])
def test_3d_512_single_sweep(dt, name, shape, r, coef):
    n = 512
    src = torch.empty((n, n, n), dtype=dt, device="cuda")
    api.fill_uniform(src, 2, 0.0, 1.0)
    dst = torch.full_like(src, float("nan"))
    api.stencil3d(src, dst, coef, shape=shape, radius=r)
    torch.cuda.synchronize()
    taps = wl.taps3d("face_edge" if shape == L.FACE_EDGE else "star", r)
    _slabs3d(dst, _planes(2, n, n, NPDT[dt]), n, taps, coef, r, 1, TOL[NPDT[dt]], [0, 255, 508])
    del src, dst


def test_3d_512_fp64_two_sweeps_at_the_1gib_edge():
    # 512^3 FP64 is exactly 2^30 bytes: the two-sweep kernel addresses it through
    # 32-bit buffer offsets, so its last planes sit at the top of that range
    n = 512
    src = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
    api.fill_uniform(src, 8, 0.0, 1.0)
    dst = torch.full_like(src, float("nan"))
    work = torch.empty_like(src)
    api.stencil3d(src, dst, wl.REF3D_7PT, shape=L.STAR, radius=1, steps=2, work=work)
    torch.cuda.synchronize()
    _slabs3d(dst, _planes(8, n, n, np.float64), n, wl.taps3d("star", 1), wl.REF3D_7PT, 1, 2,
             2 * TOL[np.float64], [0, 254, 508])
    del src, dst, work


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_jacobi19_512_residual(dt):
    # one Jacobi-19 sweep of the bench's grid and its FP64 residual, against the
    # oracle over the whole 512^3 grid
    n = 512
    src = torch.empty((n, n, n), dtype=dt, device="cuda")
    api.fill_uniform(src, 2, 0.0, 1.0)
    dst = torch.full_like(src, float("nan"))
    res = api.jacobi19(src, dst, 1, None, residual=True)
    x = oracle.uniform(2, n ** 3, 0.0, 1.0, NPDT[dt]).reshape(n, n, n)
    want, r2 = oracle.jacobi19(x)
    rel = 1e-4 if dt == torch.float32 else 1e-10
    assert abs(res - np.sqrt(r2)) <= rel * np.sqrt(r2), (res, np.sqrt(r2))
    for z0 in (0, 255, 508):
        g = dst[z0:z0 + 4].double().cpu().numpy()
        assert np.max(np.abs(g - want[z0:z0 + 4].astype(np.float64))) <= TOL[NPDT[dt]], z0
    del src, dst, x, want


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_jacobi19_100_steps_residual(dt):
    # the 100-step run of the bench at a size the oracle iterates in seconds
    n = 128
    steps = 100
    x = oracle.uniform(9, n ** 3, 0.0, 1.0, NPDT[dt]).reshape(n, n, n)
    src = torch.from_numpy(x).cuda()
    dst = torch.full_like(src, float("nan"))
    work = torch.empty_like(src)
    res = api.jacobi19(src, dst, steps, work, residual=True)
    cur, r2 = x, 0.0
    for _ in range(steps):
        cur, r2 = oracle.jacobi19(cur)
    # weights 1/12 and 1/24 summing to 1 on values in [0, 1): |error| <= steps * tol
    assert np.max(np.abs(dst.double().cpu().numpy() - cur.astype(np.float64))) <= steps * TOL[NPDT[dt]]
    rel = 1e-4 if dt == torch.float32 else 1e-10
    assert abs(res - np.sqrt(r2)) <= rel * np.sqrt(r2), (res, np.sqrt(r2))


@pytest.mark.parametrize("name,shape,r,steps", [("box9", L.BOX, 1, 1), ("box121", L.BOX, 5, 1),
                                                ("star21", L.STAR, 5, 1), ("tmp5_T8", L.STAR, 1, 8),
                                                ("tmp9_T8", L.STAR, 2, 8)])
def test_fp64_suite_8192(name, shape, r, steps):
    n = 8192
    src = torch.empty((n, n), dtype=torch.float64, device="cuda")
    api.fill_uniform(src, 0, -1.0, 1.0)
    dst = torch.full_like(src, float("nan"))
    work = torch.empty_like(src)  # FP64 radius 2 fuses 4 steps per launch
    f = wl.normalised_taps("star" if shape == L.STAR else "box", r, 1000 + 10 * r + (shape == L.STAR))
    api.stencil2d(src, dst, f, shape=shape, steps=steps, work=work)
    torch.cuda.synchronize()
    _bands2d(dst, 0, -1.0, 1.0, f, steps, TOL[np.float64] * steps, [0, 2040, 4093, 8176])
    del src, dst, work


@pytest.mark.parametrize("storage,tol", [(torch.bfloat16, 1e-2), (torch.float16, 1e-3)])
@pytest.mark.parametrize("k", [3, 9, 15, 19])
def test_conv_8192_storage_precision(storage, tol, k):
    n = 8192
    x64 = torch.empty((n, n), dtype=torch.float64, device="cuda")
    api.fill_uniform(x64, 3, 0.0, 1.0)
    src = x64.to(storage)
    del x64
    dst = torch.full_like(src, float("nan"))
    f = wl.conv_taps(k, seed=50 + k)
    api.stencil2d(src, dst, f, shape=L.DENSE)
    torch.cuda.synchronize()
    _bands2d(dst, 3, 0.0, 1.0, f, 1, tol, [0, 3000, 8176])
    del src, dst


@pytest.mark.parametrize("storage,tol", [(torch.float32, 1e-5), (torch.bfloat16, 1e-2)])
def test_config4_seed4_plan_every_item(storage, tol):
    """The bench's uneven batch as planned there (seed 4: 192 2D grids, 64 3D grids x
    10 sweeps), tuned and replayed as a graph; every item checked on row bands (2D:
    first, middle, last 8 rows) or plane slabs (3D: first and middle 4 planes)."""
    g2, g3 = wl.config4_batch(seed=4)
    shapes = {("star", 1): L.STAR, ("box", 2): L.BOX, ("box", 5): L.BOX}
    items, keep = [], []
    for k, g in enumerate(g2):
        src = torch.empty((g["h"], g["w"]), dtype=torch.float32, device="cuda")
        api.fill_uniform(src, 1000 + k, -1.0, 1.0)
        src = src.to(storage)
        dst = torch.full_like(src, float("nan"))
        keep += [src, dst]
        items.append(dict(ndim=2, src=src, dst=dst, shape=shapes[(g["shape"], g["r"])], radius=g["r"],
                          coeffs=wl.normalised_taps(g["shape"], g["r"], 60 + g["r"])))
    for k, g in enumerate(g3):
        src = torch.empty((g["nz"], g["ny"], g["nx"]), dtype=torch.float32, device="cuda")
        api.fill_uniform(src, 5000 + k, 0.0, 1.0)
        dst, work = torch.full_like(src, float("nan")), torch.empty_like(src)
        keep += [src, dst, work]
        items.append(dict(ndim=3, src=src, dst=dst, work=work, shape=L.STAR, radius=1, steps=g["steps"],
                          coeffs=wl.REF3D_7PT))
    plan = api.BatchPlan(items)
    plan.tune()
    for it in items:
        it["dst"].fill_(float("nan"))
    plan.run()
    plan.run()  # a replay of the captured graph
    torch.cuda.synchronize()
    for k, g in enumerate(g2):
        f = wl.normalised_taps(g["shape"], g["r"], 60 + g["r"])
        _bands2d(items[k]["dst"], 1000 + k, -1.0, 1.0, f, 1, tol, [0, g["h"] // 2, g["h"] - 8],
                 npdt=np.float32, band=min(8, g["h"]))
    taps = wl.taps3d("star", 1)
    for k, g in enumerate(g3):
        _slabs3d(items[len(g2) + k]["dst"], _planes(5000 + k, g["nx"], g["ny"], np.float32), g["nz"], taps,
                 wl.REF3D_7PT, 1, g["steps"], 1e-5 * g["steps"], [0, g["nz"] // 2])
    plan.close()
