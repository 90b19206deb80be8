"""Matrix-core (MFMA banded-GEMM) form of the 2D box / dense filters for BF16 and
FP16 storage, through the C ABI.  SSAM_2D_FORM=2 forces the form, =1 the VALU
register-cache kernels; both must meet the storage-precision bar against the FP64
gold (1e-2 BF16, 1e-3 FP16, per point relative to sum |w x|), and small-integer
data with exact taps must match the gold exactly.
"""
import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import _lib as L
from ssam_amd import api
from ssam_amd import workloads as wl

pytestmark = pytest.mark.gpu
TOL = {torch.bfloat16: 1e-2, torch.float16: 1e-3}


def _run(x, f, storage, pad=0, shape=L.DENSE):
    h, w = x.shape
    src = torch.zeros((h, w + pad), dtype=storage, device="cuda")
    src[:, :w] = torch.from_numpy(x).to(storage).cuda()
    dst = torch.full((h, w + pad), float("nan"), dtype=storage, device="cuda")
    api.stencil2d(src[:, :w], dst[:, :w], f, shape=shape)
    torch.cuda.synchronize()
    return dst[:, :w].to(torch.float64).cpu().numpy()


@pytest.fixture(params=["2", "1"], ids=["mfma", "valu"])
def form(request, monkeypatch):
    monkeypatch.setenv("SSAM_2D_FORM", request.param)
    return request.param


@pytest.mark.parametrize("r", range(1, 10))
@pytest.mark.parametrize("storage", [torch.bfloat16, torch.float16])
# pitches that are multiples of 8 elements (16-byte rows) and odd ones (rows at
# 2-byte offsets: unaligned 16-byte row loads)
@pytest.mark.parametrize("h,w,pad", [(97, 413, 3), (16, 64, 8), (130, 250, 6), (1, 9, 7), (300, 1024, 0),
                                     (50, 101, 0), (33, 77, 2), (40, 130, 5)])
def test_seeded_box(form, r, storage, h, w, pad):
    x = oracle.uniform(70 + r, h * w, -1, 1, np.float64).reshape(h, w)
    xs = torch.from_numpy(x).to(storage).to(torch.float64).numpy()  # the stored input
    f = wl.normalised_taps("box", r, 80 + r)
    got = _run(x, f, storage, pad)
    want = oracle.conv2d(xs, f)
    e = oracle.err(got, want, oracle.conv2d(np.abs(xs), np.abs(f)))
    assert e["max_rel"] <= TOL[storage], (form, e)


@pytest.mark.parametrize("k", [3, 5, 7, 9, 15, 17, 19])
@pytest.mark.parametrize("storage", [torch.bfloat16, torch.float16])
def test_conv_taps(form, k, storage):
    """The convolution benchmark's taps (U[-1,1) / K^2) on a U[0,1) image."""
    h, w = 200, 520
    x = oracle.uniform(5, h * w, 0, 1, np.float64).reshape(h, w)
    xs = torch.from_numpy(x).to(storage).to(torch.float64).numpy()
    f = wl.conv_taps(k, 17)
    got = _run(x, f, storage)
    want = oracle.conv2d(xs, f)
    e = oracle.err(got, want, oracle.conv2d(np.abs(xs), np.abs(f)))
    assert e["max_rel"] <= TOL[storage], (form, e)


@pytest.mark.parametrize("r", [1, 2, 3, 5, 8, 9])
@pytest.mark.parametrize("storage", [torch.bfloat16, torch.float16])
def test_small_integers_exact(form, r, storage):
    """Inputs and taps in {0, 1}: every product and sum is exact, so the result
    equals the gold bit for bit (taps exact in the storage type: one MFMA per row)."""
    rng = np.random.default_rng(r)
    h, w = 77, 333
    x = rng.integers(0, 2, (h, w)).astype(np.float64)
    f = rng.integers(0, 2, (2 * r + 1, 2 * r + 1)).astype(np.float64)
    got = _run(x, f, storage, shape=L.BOX)
    assert np.array_equal(got, oracle.conv2d(x, f))


def test_impulse_orientation(form):
    """A unit impulse reproduces the taps flipped about the centre (correlation)."""
    r = 5
    x = np.zeros((64, 96))
    x[30, 40] = 1.0
    f = np.arange((2 * r + 1) ** 2, dtype=np.float64).reshape(2 * r + 1, 2 * r + 1) % 7 / 8.0
    got = _run(x, f, torch.bfloat16, shape=L.BOX)
    win = got[30 - r:30 + r + 1, 40 - r:40 + r + 1]
    assert np.array_equal(win, f[::-1, ::-1])
    got[30 - r:30 + r + 1, 40 - r:40 + r + 1] = 0
    assert not got.any()
