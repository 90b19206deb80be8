"""The LDS-tap form of the FP32 box / dense-filter kernel against the oracle.

FP32 boxes and dense filters of radius >= 3 take this form by default; the
register-cache form stays selectable (SSAM_2D_FORM=1) and the tuner times both.
Each test runs with the form forced, so both kernels of every radius are checked:
integer KATs from the reference gold (stencil2D_{64,81,121}pt.cu tables,
conv2D fsize 2..28) bit for bit, seeded real data within 1e-5 of the per-point
scale, the clamp and hold boundaries, and an 8192^2 box121 in row bands.
"""
import numpy as np
import pytest
import torch

from oracle import oracle
from ssam_amd import _lib as L
from ssam_amd import api
from ssam_amd import workloads as wl

pytestmark = pytest.mark.gpu

FORMS = {"valu": "1", "lds": "3"}


@pytest.fixture(params=sorted(FORMS))
def form(request, monkeypatch):
    monkeypatch.setenv("SSAM_2D_FORM", FORMS[request.param])
    return request.param


def dev(x, pitch_pad=0):
    t = torch.from_numpy(np.ascontiguousarray(x))
    if pitch_pad:
        h, w = x.shape
        big = torch.zeros((h, w + pitch_pad), dtype=t.dtype)
        big[:, :w] = t
        return big.cuda()[:, :w]
    return t.cuda()


def run(x, f, shape, boundary=L.ZERO, pitch_pad=0):
    src = dev(x, pitch_pad)
    dst = dev(np.full(x.shape, np.nan, dtype=x.dtype), pitch_pad)
    api.stencil2d(src, dst, f, shape=shape, boundary=boundary)
    torch.cuda.synchronize()
    return dst.cpu().numpy()


def scale_of(x, f):
    return oracle.conv2d(np.abs(np.asarray(x, np.float64)), np.abs(np.asarray(f, np.float64)))


def check(got, want, sc, tol, what):
    assert np.isfinite(got).all(), f"{what}: unwritten or non-finite output"
    e = oracle.err(got, want, sc)
    assert e["max_rel"] <= tol, f"{what}: {e}"


@pytest.mark.parametrize("k", [3, 5, 7, 9, 15, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28])
def test_kat_conv_integer_exact(golden_dir, form, k):
    d = np.load(golden_dir / f"kat_conv{k}_int_float32.npz")
    got = run(d["src"], d["filt"], L.DENSE)
    assert np.array_equal(got, d["out"]), f"{form} K={k}"
    # pitched rows and device weights read in place (the reference conv2D entry)
    src = dev(d["src"], pitch_pad=5)
    dst = dev(np.zeros_like(d["src"]), pitch_pad=5)
    api.conv2d(src, dst, torch.from_numpy(d["filt"]).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(dst.cpu().numpy(), d["out"]), f"{form} K={k} conv2d entry"


@pytest.mark.parametrize("name", ["box64", "box81", "box121"])
def test_kat_box_reference_tables(golden_dir, form, name):
    d = np.load(golden_dir / f"kat_{name}_float32.npz")
    got = run(d["src"], d["filt"], L.BOX)
    check(got, d["out"], scale_of(d["src"], d["filt"]), 1e-5, f"{form} {name}")


SIZES = [(1, 1), (3, 7), (5, 64), (64, 5), (37, 251), (61, 240), (47, 249), (130, 497), (300, 520)]


@pytest.mark.parametrize("r", range(2, 15))
def test_parity_seeded_boxes(form, r):
    f = wl.normalised_taps("box", r, 300 + r).astype(np.float32)
    for i, (h, w) in enumerate(SIZES):
        x = oracle.uniform(2000 + i, h * w, -1, 1, np.float32).reshape(h, w)
        got = run(x, f, L.BOX, pitch_pad=(3 if i % 3 == 1 else 0))
        check(got, oracle.conv2d(x, f), scale_of(x, f), 1e-5, f"{form} r={r} {(h, w)}")


@pytest.mark.parametrize("k", [4, 6, 10, 16, 22, 28])
def test_parity_seeded_even_filters(form, k):
    f = (wl.conv_taps(k, 60 + k)).astype(np.float32)
    for i, (h, w) in enumerate(SIZES[3:]):
        x = oracle.uniform(2100 + i, h * w, 0, 1, np.float32).reshape(h, w)
        got = run(x, f, L.DENSE, pitch_pad=(2 if i % 2 else 0))
        check(got, oracle.conv2d(x, f), scale_of(x, f), 1e-5, f"{form} K={k} {(h, w)}")


@pytest.mark.parametrize("boundary", [L.CLAMP, L.HOLD])
@pytest.mark.parametrize("r", [3, 5, 9])
def test_parity_boundaries(form, boundary, r):
    f = wl.normalised_taps("box", r, 17).astype(np.float32)
    for (h, w) in [(29, 67), (130, 300), (4, 9)]:
        x = oracle.uniform(h + w + r, h * w, 0, 1, np.float32).reshape(h, w)
        got = run(x, f, L.BOX, boundary=boundary)
        if boundary == L.CLAMP:
            xp = np.pad(x, r, mode="edge")
            want = oracle.conv2d(xp, f)[r:-r, r:-r]
            sc = scale_of(xp, f)[r:-r, r:-r]
        else:
            want = oracle.conv2d(x, f)
            want[:r, :], want[-r:, :], want[:, :r], want[:, -r:] = x[:r, :], x[-r:, :], x[:, :r], x[:, -r:]
            sc = np.maximum(scale_of(x, f), np.abs(x))
        check(got, want, sc, 1e-5, f"{form} boundary={boundary} r={r} {(h, w)}")


def test_forms_agree_bit_for_bit():
    # radius <= 9: both forms run 4 columns per lane and accumulate every partial
    # in the same order, so their results are identical
    x = oracle.uniform(77, 300 * 700, -1, 1, np.float32).reshape(300, 700)
    for r in (3, 5, 9):
        f = wl.normalised_taps("box", r, 400 + r).astype(np.float32)
        out = {}
        for name, code in FORMS.items():
            with pytest.MonkeyPatch.context() as mp:
                mp.setenv("SSAM_2D_FORM", code)
                out[name] = run(x, f, L.BOX)
        assert np.array_equal(out["valu"], out["lds"]), f"r={r}"


def test_full_size_box121_bands():
    n = 8192
    src = torch.empty((n, n), dtype=torch.float32, device="cuda")
    api.fill_uniform(src, 0, -1.0, 1.0)
    f = wl.normalised_taps("box", 5, 3).astype(np.float32)
    dst = torch.full_like(src, float("nan"))
    api.stencil2d(src, dst, f, shape=L.BOX)  # default form: LDS taps
    torch.cuda.synchronize()
    x = src.cpu().numpy()
    got = dst.cpu().numpy()
    assert np.isfinite(got).all()
    for y0 in (0, 2040, 4089, 8176):
        lo, hi = max(0, y0 - 5), min(n, y0 + 16 + 5)
        slab = x[lo:hi]
        want = oracle.conv2d(slab, f)
        a = y0 - lo
        if lo == 0 or a >= 5:
            b = a + min(16, n - y0)
            check(got[y0:y0 + (b - a)], want[a:b], scale_of(slab, f)[a:b], 1e-5, f"band {y0}")
