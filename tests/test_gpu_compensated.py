"""SSAM_FLAG_COMPENSATED: box / dense filters with Kahan-compensated accumulation
of the per-filter-row partial sums, through the C ABI.

Bar: the same parity bar as the plain kernels (bit-exact on integer data, 1e-5
FP32, 1e-2 / 1e-3 vs the FP64 gold for BF16 / FP16 storage), and an FP32
accumulation error vs the FP64 gold no larger than the plain kernel's.
"""
import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import _lib as L
from ssam_amd import api
from ssam_amd import workloads as wl

pytestmark = pytest.mark.gpu
COMP = L.FLAG_COMPENSATED


def _run(x, f, storage, flags, pad=0):
    h, w = x.shape
    src = torch.zeros((h, w + pad), dtype=storage, device="cuda")
    src[:, :w] = torch.from_numpy(x).to(storage).cuda()
    dst = torch.full((h, w + pad), float("nan"), dtype=storage, device="cuda")
    api.stencil2d(src[:, :w], dst[:, :w], f, shape=L.DENSE, flags=flags)
    torch.cuda.synchronize()
    return dst[:, :w].to(torch.float64).cpu().numpy()


@pytest.mark.parametrize("k", [3, 9, 15, 19])
def test_integer_exact(k):
    rng = np.random.default_rng(k)
    x = rng.integers(0, 256, (67, 301)).astype(np.float32)
    f = rng.integers(-10, 11, (k, k)).astype(np.float32)
    got = _run(x, f.astype(np.float64), torch.float32, COMP)
    want = oracle.conv2d(x, f)
    assert np.array_equal(got, want.astype(np.float64))


@pytest.mark.parametrize("r", range(1, 10))
@pytest.mark.parametrize("storage,tol", [(torch.float32, 1e-5), (torch.bfloat16, 1e-2),
                                         (torch.float16, 1e-3)])
def test_seeded_box(r, storage, tol):
    h, w = 97, 413
    x = oracle.uniform(40 + r, h * w, -1, 1, np.float64).reshape(h, w)
    f = wl.normalised_taps("box", r, 50 + r)
    got = _run(x, f, storage, COMP, pad=4 * (r % 2))
    if storage == torch.float32:
        want = oracle.conv2d(x.astype(np.float32), f.astype(np.float32)).astype(np.float64)
    else:
        want = oracle.conv2d(x, f)
    s = oracle.conv2d(np.abs(x), np.abs(f))
    e = oracle.err(got, want, s)
    assert e["max_rel"] <= tol, e


@pytest.mark.parametrize("k", [9, 15, 19])
def test_fp32_error_not_above_plain(k):
    """The benchmark's convolution (image U[0,1), taps U[-1,1)/K^2): the
    compensated FP32 result is at least as close to the FP64 gold as the plain one."""
    h, w = 256, 1024
    x = oracle.uniform(3, h * w, 0, 1, np.float64).reshape(h, w).astype(np.float32)
    f = wl.conv_taps(k, 11).astype(np.float32).astype(np.float64)
    want = oracle.conv2d(x.astype(np.float64), f)
    plain = oracle.err(_run(x, f, torch.float32, 0), want)
    got = _run(x, f, torch.float32, COMP)
    comp = oracle.err(got, want)
    print(f"K={k} rms plain {plain['rms']:.3e} comp {comp['rms']:.3e}")
    assert comp["rms"] <= plain["rms"] and comp["max_abs"] <= 2 * plain["max_abs"]
    # and absolutely: within 1e-6 of sum |w x| at every point (the uncompensated
    # sums are those of one filter row, K terms; the output rounding adds 2^-24)
    scale = oracle.conv2d(np.abs(x).astype(np.float64), np.abs(f))
    assert oracle.err(got, want, scale)["max_rel"] <= 1e-6


def test_fp64_unchanged():
    h, w = 50, 200
    x = oracle.uniform(9, h * w, -1, 1, np.float64).reshape(h, w)
    f = wl.normalised_taps("box", 4, 9)
    assert np.array_equal(_run(x, f, torch.float64, COMP), _run(x, f, torch.float64, 0))
