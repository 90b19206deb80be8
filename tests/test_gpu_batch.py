"""Uneven batch through ssam_batch_plan_create / ssam_batch_run: one persistent
work-queue launch for the 2D items, persistent per-step launches for 3D items,
per-item fallback for other shapes, hipGraph replay.  Every replay is compared
with the oracle (eager runs and synchronisations in between)."""
import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import _lib as L
from ssam_amd import api
from ssam_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _mk2d(h, w, seed, pad=0, dtype=torch.float32):
    x = oracle.uniform(seed, h * w, -1, 1, np.float32).reshape(h, w)
    t = torch.zeros((h, w + pad), dtype=dtype, device="cuda")
    t[:, :w] = torch.from_numpy(x).to(dtype).cuda()
    return t[:, :w], x


def _check2d(got, x, f, tol):
    want = oracle.conv2d(x.astype(np.float64), f.astype(np.float64))
    sc = oracle.conv2d(np.abs(x).astype(np.float64), np.abs(f).astype(np.float64))
    e = oracle.err(got, want, sc)
    assert e["max_rel"] <= tol, e


def _mixed_items(storage):
    specs2d = [(37, 251, "star", 1), (300, 129, "box", 2), (64, 64, "box", 5), (1, 7, "star", 1),
               (250, 997, "box", 1), (513, 300, "box", 5), (45, 600, "star", 3), (129, 130, "box", 2),
               (96, 200, "box", 2)]
    items, refs = [], []
    for k, (h, w, shape, r) in enumerate(specs2d):
        src, x = _mk2d(h, w, 100 + k, pad=(k % 3) * 4, dtype=storage)
        dst, _ = _mk2d(h, w, 0, pad=(k % 3) * 4, dtype=storage)  # same row pitch as src
        f = wl.normalised_taps(shape, r, 7 + k)
        items.append(dict(ndim=2, src=src, dst=dst, coeffs=f,
                          shape=L.STAR if shape == "star" else L.BOX, radius=r))
        refs.append(("2d", dst, src.to(torch.float64).cpu().numpy(), f))
    for k, (nx, ny, nz, shape, r) in enumerate([(70, 33, 20, "star", 1), (64, 64, 64, "face_edge", 1),
                                                 (257, 9, 11, "star", 1), (40, 40, 40, "star", 2)]):
        x = oracle.uniform(200 + k, nx * ny * nz, 0, 1, np.float32).reshape(nz, ny, nx)
        src = torch.from_numpy(x).cuda()
        dst = torch.full_like(src, float("nan"))
        work = torch.empty_like(src)
        taps = wl.taps3d(shape, r)
        coef = wl.uniform(300 + k, len(taps), -0.5, 0.5)
        items.append(dict(ndim=3, src=src, dst=dst, work=work, coeffs=coef, steps=3,
                          shape=L.FACE_EDGE if shape == "face_edge" else L.STAR, radius=r))
        refs.append(("3d", dst, x, (taps, coef, r)))
    return items, refs


def _check_refs(refs, tol):
    for kind, dst, x, f in refs:
        got = dst.to(torch.float64).cpu().numpy()
        if kind == "2d":
            _check2d(got, x, f, tol)
        else:
            taps, coef, r = f
            cur = x.astype(np.float64)
            s = np.abs(cur)
            for _ in range(3):
                cur = oracle.stencil3d_taps(cur, taps, coef, r)
                s = oracle.stencil3d_taps(s, taps, np.abs(coef), r)
            e = oracle.err(got, cur, s)
            assert e["max_rel"] <= 3e-5, e


@pytest.mark.parametrize("split,mfma,grid", [("1", "3", "0"), ("0", "3", "0"), ("1", "0", "1"), ("0", "0", "1"),
                                             ("1", "3", "1"), ("1", "1", "1"), ("0", "2", "1")])
@pytest.mark.parametrize("storage,tol", [(torch.float32, 1e-5), (torch.bfloat16, 1e-2), (torch.float16, 1e-3)])
def test_batch_mixed_2d_3d(storage, tol, split, mfma, grid, monkeypatch):
    """The 2D launch forms (one launch per kind / one for all kinds; persistent /
    one tile per wave of an ordinary grid), with the 5x5 and 11x11 boxes routed to
    the matrix cores or not (bit 0 / bit 1 of SSAM_BATCH2D_MFMA; BF16/FP16 rows of
    a multiple of 8 elements), against the oracle on every replay."""
    monkeypatch.setenv("SSAM_BATCH2D_SPLIT", split)
    monkeypatch.setenv("SSAM_BATCH2D_MFMA", mfma)
    monkeypatch.setenv("SSAM_BATCH2D_GRID", grid)
    monkeypatch.setenv("SSAM_BATCH3D_GRID", grid)
    # the two-sweep 3D shape (5) always runs as a grid; the persistent form takes a
    # one-sweep shape
    monkeypatch.setenv("SSAM_BATCH3D_VARIANT", "5" if grid == "1" else "1")
    items, refs = _mixed_items(storage)
    plan = api.BatchPlan(items)
    t2, t3, pts = plan.info()
    assert t2 > 0 and t3 > 0 and pts > 0
    t = plan.tiling()
    assert (t["split_2d"], t["mfma_2d"], t["grid_2d"], t["grid_3d"]) == (int(split), int(mfma), int(grid), int(grid))
    for rep in range(3):
        for _, dst, _, _ in refs:
            dst.fill_(float("nan"))
        plan.run(eager=(rep == 1))
        torch.cuda.synchronize()
        _check_refs(refs, tol)
    plan.close()


@pytest.mark.parametrize("seg", ["8", "16", "300"])
def test_batch_2d_tile_heights(seg, monkeypatch):
    """Short and long 2D tiles (many runs per item across the XCD queues, one tile
    per strip); every output written once, against the oracle."""
    monkeypatch.setenv("SSAM_BATCH2D_SEG", seg)
    items, refs = _mixed_items(torch.float32)
    plan = api.BatchPlan(items)
    assert plan.tiling()["seg_rows_2d"]["box5"] == int(seg)
    for _, dst, _, _ in refs:
        dst.fill_(float("nan"))
    plan.run()
    torch.cuda.synchronize()
    _check_refs(refs, 1e-5)
    plan.close()


def test_batch_tune_keeps_results():
    """ssam_batch_plan_tune picks a tiling from its candidates; runs after it (graph
    recaptured, eager) still match the oracle."""
    items, refs = _mixed_items(torch.float32)
    plan = api.BatchPlan(items)
    t = plan.tune()
    assert set(t["seg_rows_2d"].values()) <= {32, 64, 128, 256}
    assert 0 <= t["variant_3d"] <= 6 and t["zseg_3d"] in (16, 32, 64)
    assert t["affinity_2d"] in (0, 1) and t["affinity_3d"] in (0, 1)
    assert t["split_2d"] in (0, 1) and t["mfma_2d"] in (0, 1, 2, 3)
    assert t["wps_2d"] in (0, 1, 2, 4) and t["wps_3d"] in (1, 2, 4) and t["grid_3d"] in (0, 1)
    assert t["grid_2d"] in (0, 1)
    for rep in range(2):
        for _, dst, _, _ in refs:
            dst.fill_(float("nan"))
        plan.run(eager=(rep == 1))
        torch.cuda.synchronize()
        _check_refs(refs, 1e-5)
    plan.close()


def test_config4_generator_shapes():
    g2, g3 = wl.config4_batch()
    assert len(g2) == 192 and len(g3) == 64
    for g in g2:
        assert 128 <= g["w"] <= 4096 and 128 <= g["h"] <= 4096
        assert max(g["w"], g["h"]) <= 4 * min(g["w"], g["h"])
    assert {g["r"] for g in g2} == {1, 2, 5}
    assert all(64 <= g[k] <= 256 for g in g3 for k in ("nx", "ny", "nz"))


_SIZES3 = [(70, 33, 20), (1, 1, 1), (2, 3, 2), (3, 1, 5), (63, 5, 4), (62, 2, 3), (125, 9, 7),
           (124, 8, 9), (200, 17, 33), (64, 64, 12)]


@pytest.mark.parametrize("variant,affinity", [(v, v % 2) for v in range(7)])
@pytest.mark.parametrize("storage", [torch.float32, torch.float64])
def test_batch3d_tile_shapes(variant, affinity, storage, monkeypatch):
    """Every 3D tile shape of the batched launch (columns / rows per lane, prefetch
    depth; one sweep per pass, the per-grid pair, two sweeps per pass, whose odd
    step count ends in a one-sweep pass, and three, whose counts split into
    passes of three, two and one) on tiny, odd-width and multi-strip grids, short
    z segments, six step counts (groups), persistent grids of 1, 2 and all waves
    per SIMD and ordinary one-tile-per-wave grids, graph replay; against the
    oracle."""
    monkeypatch.setenv("SSAM_BATCH3D_WPS", str(variant % 3))
    monkeypatch.setenv("SSAM_BATCH3D_GRID", str((variant // 2) % 2))
    monkeypatch.setenv("SSAM_BATCH3D_VARIANT", str(variant))
    monkeypatch.setenv("SSAM_BATCH3D_AFFINITY", str(affinity))
    monkeypatch.setenv("SSAM_BATCH3D_ALTERNATE", str(variant // 4))
    monkeypatch.setenv("SSAM_BATCH3D_ZSEG", "5")
    items, refs = [], []
    for k, (nx, ny, nz) in enumerate(_SIZES3):
        shape = "face_edge" if k % 3 == 1 else "star"
        steps = (3, 2, 7, 1, 5, 4)[k % 6]
        x = oracle.uniform(500 + k, nx * ny * nz, 0, 1, np.float64).reshape(nz, ny, nx)
        src = torch.from_numpy(x).to(storage).cuda()
        dst = torch.full_like(src, float("nan"))
        work = torch.empty_like(src)
        taps = wl.taps3d(shape, 1)
        coef = wl.uniform(600 + k, len(taps), -0.5, 0.5)
        items.append(dict(ndim=3, src=src, dst=dst, work=work, coeffs=coef, steps=steps,
                          shape=L.FACE_EDGE if shape == "face_edge" else L.STAR, radius=1))
        refs.append((dst, src.to(torch.float64).cpu().numpy(), taps, coef, steps))
    plan = api.BatchPlan(items)
    tol = 3e-5 if storage == torch.float32 else 1e-12
    for rep in range(2):
        for dst, *_ in refs:
            dst.fill_(float("nan"))
        plan.run(eager=(rep == 0))
        torch.cuda.synchronize()
        for dst, x, taps, coef, steps in refs:
            cur, s = x, np.abs(x)
            for _ in range(steps):
                cur = oracle.stencil3d_taps(cur, taps, coef, 1)
                s = oracle.stencil3d_taps(s, taps, np.abs(coef), 1)
            e = oracle.err(dst.to(torch.float64).cpu().numpy(), cur, s)
            assert e["max_rel"] <= tol, (variant, x.shape, e)
    plan.close()
