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


@pytest.mark.parametrize("storage,tol", [(torch.float32, 1e-5), (torch.bfloat16, 1e-2)])
def test_batch_mixed_2d_3d(storage, tol):
    specs2d = [(37, 251, "star", 1), (300, 129, "box", 2), (64, 64, "box", 5), (1, 7, "star", 1),
               (250, 997, "box", 1), (513, 300, "box", 5), (45, 600, "star", 3), (129, 130, "box", 2)]
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
    plan = api.BatchPlan(items)
    t2, t3, pts = plan.info()
    assert t2 > 0 and t3 > 0 and pts > 0
    for rep in range(3):
        for _, dst, _, _ in refs:
            dst.fill_(float("nan"))
        plan.run(eager=(rep == 1))
        torch.cuda.synchronize()
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
    plan.close()


def test_config4_generator_shapes():
    g2, g3 = wl.config4_batch()
    assert len(g2) == 192 and len(g3) == 64
    for g in g2:
        assert 128 <= g["w"] <= 4096 and 128 <= g["h"] <= 4096
        assert max(g["w"], g["h"]) <= 4 * min(g["w"], g["h"])
    assert {g["r"] for g in g2} == {1, 2, 5}
    assert all(64 <= g[k] <= 256 for g in g3 for k in ("nx", "ny", "nz"))
