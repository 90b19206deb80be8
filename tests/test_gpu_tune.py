"""The autotuner (ssam_tune_stencil2d) times tilings on the caller's buffers and
caches the choice; results under the tuned tiling keep the parity bar."""
import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import _lib as L
from ssam_amd import api
from ssam_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,r,steps", [(L.STAR, 1, 1), (L.BOX, 3, 1), (L.STAR, 2, 8), (L.STAR, 1, 8)])
def test_tuned_parity_and_cache(tmp_path, shape, r, steps):
    api.tune_clear()
    h, w = 300, 700
    x = oracle.uniform(17, h * w, -1, 1, np.float32).reshape(h, w)
    f = wl.normalised_taps("star" if shape == L.STAR else "box", r, 9).astype(np.float32)
    src = torch.from_numpy(x).cuda()
    dst = torch.empty_like(src)
    work = torch.empty_like(src)
    wps, tch = api.tune2d(src, dst, f, shape=shape, steps=steps, work=work)
    assert wps in (2, 4, 8, 12, 24) and 1 <= tch <= 8
    dst.fill_(float("nan"))
    api.stencil2d(src, dst, f, shape=shape, steps=steps, work=work)
    torch.cuda.synchronize()
    want = oracle.conv2d_iter(x, f, steps)
    s = np.abs(x).astype(np.float64)
    for _ in range(steps):
        s = oracle.conv2d(s, np.abs(f).astype(np.float64))
    e = oracle.err(dst.cpu().numpy(), want, s)
    assert e["max_rel"] <= 1e-5 * steps, e
    p = tmp_path / "t.txt"
    api.tune_save(p)
    api.tune_clear()
    api.tune_load(p)
    assert f"{wps} {tch}" in p.read_text()
