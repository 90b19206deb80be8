"""Pins of the CPU oracle (oracle/gold.c) -- no GPU.

Each restated gold loop is checked against what fixes it independently:
  * the reference's own gold: KAT fixtures written from Convolution<T>
    (ssai-2dconv/common.h:186-216) compiled as-is, and oracle/_ref live when built;
  * closed forms: impulse response (orientation, no flip), constant fields;
  * brute force in numpy on tiny grids (an independent formulation with padding);
  * the reference coefficient tables (fixture coeffs.json) for the 3D operators;
  * splitmix64's published test vector for the generator.
"""
import json

import numpy as np
import pytest

from oracle import oracle
from ssam_amd import workloads as wl


def test_splitmix64_published_vector():
    # Vigna's splitmix64.c, seed 1234567: first five outputs
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423,
            4593380528125082431, 16408922859458223821]
    assert [oracle.splitmix64_at(1234567, i) for i in range(5)] == want
    assert [int(v) for v in wl.splitmix64(1234567, 5)] == want


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_uniform_matches_host_generator(dtype):
    a = oracle.uniform(7, 1000, -1.0, 1.0, dtype, offset=123)
    b = wl.uniform(7, 1000, -1.0, 1.0, offset=123).astype(dtype)
    assert np.array_equal(a, b)
    assert a.min() >= -1.0 and a.max() < 1.0


def _kats(golden_dir, pattern):
    files = sorted(golden_dir.glob(pattern))
    assert files, pattern
    return files


def test_conv_matches_reference_gold_fixtures(golden_dir):
    """Bit-exact against the reference's own Convolution<T> outputs."""
    for f in _kats(golden_dir, "kat_*.npz"):
        d = np.load(f)
        if "steps" in d:
            got = oracle.conv2d_iter(d["src"], d["filt"], int(d["steps"]))
        else:
            got = oracle.conv2d(d["src"], d["filt"])
        assert np.array_equal(got, d["out"]), f.name


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fh,fw", [(3, 3), (2, 2), (5, 3), (4, 7), (9, 9), (1, 1), (20, 20), (23, 23),
                                   (28, 28)])
def test_conv_matches_reference_live(dtype, fh, fw):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref (the reference gold) not built here")
    x = oracle.uniform(3, 31 * 45, -1, 1, dtype).reshape(31, 45)
    f = oracle.uniform(4, fh * fw, -1, 1, dtype).reshape(fh, fw)
    assert np.array_equal(oracle.conv2d(x, f), oracle.ref_conv2d(x, f))


def _brute(x, f):
    fh, fw = f.shape
    h, w = x.shape
    p = np.zeros((h + fh, w + fw))
    p[fh // 2:fh // 2 + h, fw // 2:fw // 2 + w] = x
    out = np.zeros((h, w))
    for m in range(fh):
        for n in range(fw):
            out += f[m, n] * p[m:m + h, n:n + w]
    return out


@pytest.mark.parametrize("fh,fw", [(3, 3), (5, 5), (2, 4), (7, 3)])
def test_conv_brute_force_fp64(fh, fw):
    x = oracle.uniform(5, 13 * 16, -1, 1, np.float64).reshape(13, 16)
    f = oracle.uniform(6, fh * fw, -1, 1, np.float64).reshape(fh, fw)
    np.testing.assert_allclose(oracle.conv2d(x, f), _brute(x, f), rtol=0, atol=1e-13)


@pytest.mark.parametrize("fh,fw", [(3, 3), (5, 5), (4, 2)])
def test_conv_impulse_response_orientation(fh, fw):
    f = np.arange(1, fh * fw + 1, dtype=np.float64).reshape(fh, fw)
    x = np.zeros((20, 20))
    y0, x0 = 9, 11
    x[y0, x0] = 1.0
    out = oracle.conv2d(x, f)
    for m in range(fh):
        for n in range(fw):
            assert out[y0 - m + fh // 2, x0 - n + fw // 2] == f[m, n]
    assert out.sum() == f.sum()


def test_conv_constant_field():
    f = wl.normalised_taps("box", 2, 9)
    x = np.full((30, 40), 3.0)
    out = oracle.conv2d(x, f)
    np.testing.assert_allclose(out[2:-2, 2:-2], 3.0, rtol=1e-14)
    assert out[0, 0] < 3.0  # zero padding at the border


def test_temporal_equals_repeated_sweeps():
    f = np.array(json.loads((oracle.HERE.parent / "tests/golden/coeffs.json").read_text())["temporal5"])
    x = oracle.uniform(8, 25 * 33, 0, 1, np.float64).reshape(25, 33)
    cur = x
    for _ in range(4):
        cur = oracle.conv2d(cur, f)
    assert np.array_equal(oracle.conv2d_iter(x, f, 4), cur)


def _macro_taps(table, base):
    """Reference macro names _F<z><y><x> -> (dz, dy, dx, coef)."""
    out = []
    for name, v in table.items():
        d = name[-3:]
        out.append((int(d[0]) - base, int(d[1]) - base, int(d[2]) - base, v))
    return out


@pytest.mark.parametrize("which,shape,r,coefs,base", [
    ("j3d7pt", "star", 1, wl.REF3D_7PT, 1),
    ("j3d13pt", "star", 2, wl.REF3D_13PT, 2),
    ("poisson", "face_edge", 1, wl.REF3D_POISSON, 1),
])
def test_3d_tables_match_reference_macros(golden_dir, which, shape, r, coefs, base):
    """The descriptor-order coefficient tables equal the reference's macros."""
    table = json.loads((golden_dir / "coeffs.json").read_text())[which]
    want = {(dz, dy, dx): v for dz, dy, dx, v in _macro_taps(table, base)}
    got = dict(zip(wl.taps3d(shape, r), coefs))
    assert got == want


@pytest.mark.parametrize("shape,r,coefs", [("star", 1, wl.REF3D_7PT), ("star", 2, wl.REF3D_13PT),
                                           ("face_edge", 1, wl.JACOBI19)])
def test_3d_taps_impulse_response(shape, r, coefs):
    n = 11
    x = np.zeros((n, n, n))
    c0 = (5, 5, 5)
    x[c0] = 1.0
    taps = wl.taps3d(shape, r)
    out = oracle.stencil3d_taps(x, taps, coefs, r)
    for (dz, dy, dx), v in zip(taps, coefs):
        assert out[c0[0] - dz, c0[1] - dy, c0[2] - dx] == v
    assert np.isclose(out.sum(), sum(coefs))


def test_poisson_ref_impulse_and_constant(golden_dir):
    table = json.loads((golden_dir / "coeffs.json").read_text())["poisson"]
    n = 9
    x = np.zeros((n, n, n))
    x[4, 4, 4] = 1.0
    out = oracle.poisson_ref(x)
    for dz, dy, dx, v in _macro_taps(table, 1):
        assert np.isclose(out[4 - dz, 4 - dy, 4 - dx], v, rtol=1e-15)
    c = np.full((n, n, n), 2.0)
    o = oracle.poisson_ref(c)
    np.testing.assert_allclose(o[1:-1, 1:-1, 1:-1], 2.0 * sum(table.values()), rtol=1e-14)
    assert np.all(o[0] == 0) and np.all(o[:, :, -1] == 0)


def test_jacobi19_invariants():
    n = 12
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    lin = (0.5 * x - 0.25 * y + 0.125 * z + 1.0).astype(np.float64)
    out, r2 = oracle.jacobi19(lin)
    # linear fields are discrete-harmonic: fixed point in the interior, zero residual
    np.testing.assert_allclose(out[1:-1, 1:-1, 1:-1], lin[1:-1, 1:-1, 1:-1], rtol=1e-14)
    assert r2 < 1e-24
    assert np.all(out[0] == 0) and np.all(out[-1] == 0)
    # maximum principle and the stencil's weights
    rnd = oracle.uniform(2, n ** 3, 0, 1, np.float64).reshape(n, n, n)
    o2, r2b = oracle.jacobi19(rnd)
    assert o2.max() <= rnd.max() and r2b > 0
    imp = np.zeros((n, n, n))
    imp[6, 6, 6] = 1.0
    oi, _ = oracle.jacobi19(imp)
    assert oi[6, 6, 5] == 1 / 12 and oi[5, 6, 6] == 1 / 12 and oi[5, 5, 6] == 1 / 24
    assert oi[6, 6, 6] == 0 and oi[5, 5, 5] == 0
    # residual definition: r = (2F + E - 24u)/6 at an isolated impulse centre
    _, r2c = oracle.jacobi19(imp)
    # centre r = -4; 6 faces 2/6; 12 edges 1/6
    assert np.isclose(r2c, 16 + 6 * (2 / 6) ** 2 + 12 * (1 / 6) ** 2, rtol=1e-14)


def test_err_metrics():
    a = np.array([1.0, 2.0, 3.0])
    b = np.array([1.0, 2.5, 2.0])
    e = oracle.err(a, b, np.array([1.0, 1.0, 2.0]))
    assert e["max_abs"] == 1.0 and e["match"] == 1 and e["count"] == 3
    assert e["max_rel"] == 0.5 and np.isclose(e["rms"], np.sqrt((0.25 + 1) / 3))


def test_err_flags_nan():
    """An unwritten output (NaN) fails every tolerance, whatever the scale."""
    a = np.array([1.0, np.nan, 3.0])
    b = np.array([1.0, 2.0, 3.0])
    e = oracle.err(a, b, np.array([1.0, 1.0, 1.0]))
    assert e["max_rel"] == np.inf and e["max_abs"] == np.inf and e["match"] == 2
    e = oracle.err(b, a)
    assert e["max_rel"] == np.inf
