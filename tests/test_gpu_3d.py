"""3D parity on the GPU through the C ABI.

The reference 3D golds (tmp3d7pt-v2.cu:278-298, tmp3d13pt.cu:301-327,
tmpPoisson.cu:369-386) write only the interior and leave a zero boundary; the
GPU path writes that boundary as 0, so whole grids are compared.  Tolerance:
|gpu - gold| <= tol * sum_i |c_i x_i|, tol 1e-5 per sweep (FP32) / 1e-12 (FP64).
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


def _scale(x, taps, coef, r, steps):
    s = np.abs(x).astype(np.float64)
    for _ in range(steps):
        s = oracle.stencil3d_taps(s, taps, np.abs(coef), r)
    return s


def _gold(x, taps, coef, r, steps):
    cur = x
    for _ in range(steps):
        cur = oracle.stencil3d_taps(cur, taps, coef, r)
    return cur


def _check(got, want, scale, tol, what):
    e = oracle.err(got, want, scale)
    assert np.isfinite(got).all(), what
    assert e["max_rel"] <= tol, f"{what}: {e}"


@pytest.mark.parametrize("name,shape,r,coef", [("j3d7pt", "star", 1, wl.REF3D_7PT),
                                                ("j3d13pt", "star", 2, wl.REF3D_13PT)])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n,iters", [(40, 1), (37, 2), (64, 3), (9, 1)])
def test_reference_star_operators(name, shape, r, coef, dt, n, iters):
    x = (oracle.uniform(n, n ** 3, 0, 1, dt) + dt(0.02121)).reshape(n, n, n)
    src = torch.from_numpy(x).cuda()
    dst = torch.full_like(src, float("nan"))
    work = torch.empty_like(src) if iters > 1 else None
    api.ref3d(name, src, dst, iters, work)
    torch.cuda.synchronize()
    taps = wl.taps3d(shape, r)
    want = _gold(x, taps, coef, r, iters)
    _check(dst.cpu().numpy(), want, _scale(x, taps, coef, r, iters), TOL[dt] * iters,
           f"{name} n={n} it={iters}")


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n,iters", [(33, 1), (48, 2)])
def test_reference_poisson(dt, n, iters):
    x = (oracle.uniform(3, n ** 3, 0, 1, dt) + dt(0.02121)).reshape(n, n, n)
    src = torch.from_numpy(x).cuda()
    dst = torch.full_like(src, float("nan"))
    work = torch.empty_like(src) if iters > 1 else None
    api.ref3d("poisson", src, dst, iters, work)
    torch.cuda.synchronize()
    want = x
    for _ in range(iters):
        want = oracle.poisson_ref(want)
    taps = wl.taps3d("face_edge", 1)
    sc = _scale(x, taps, wl.REF3D_POISSON, 1, iters)
    _check(dst.cpu().numpy(), want, sc, TOL[dt] * iters, f"poisson n={n}")


@pytest.mark.parametrize("shape,r", [("star", 1), ("star", 2), ("face_edge", 1)])
@pytest.mark.parametrize("dims", [(50, 33, 27), (64, 17, 12), (260, 9, 10), (7, 6, 5)])
def test_generic_nonuniform_grids(shape, r, dims):
    nx, ny, nz = dims
    dt = np.float64
    taps = wl.taps3d(shape, r)
    coef = wl.uniform(r + 5, len(taps), -1, 1)
    x = oracle.uniform(11, nx * ny * nz, -1, 1, dt).reshape(nz, ny, nx)
    src = torch.from_numpy(x).cuda()
    dst = torch.full_like(src, float("nan"))
    api.stencil3d(src, dst, coef, shape=L.FACE_EDGE if shape == "face_edge" else L.STAR, radius=r)
    torch.cuda.synchronize()
    want = _gold(x, taps, coef, r, 1)
    _check(dst.cpu().numpy(), want, _scale(x, taps, coef, r, 1), 1e-12, f"{shape} {dims}")


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_jacobi19_and_residual(dt):
    n = 48
    x = oracle.uniform(21, n ** 3, 0, 1, dt).reshape(n, n, n)
    steps = 5
    src = torch.from_numpy(x).cuda()
    dst = torch.empty_like(src)
    work = torch.empty_like(src)
    res = api.jacobi19(src, dst, steps, work)
    cur = x
    r2 = 0.0
    for _ in range(steps):
        cur, r2 = oracle.jacobi19(cur)
    taps = wl.taps3d("face_edge", 1)
    sc = _scale(x, taps, wl.JACOBI19, 1, steps)
    _check(dst.cpu().numpy(), cur, sc, TOL[dt] * steps, "jacobi19")
    # residual of the last sweep's input, FP64 on the GPU
    rel = 1e-4 if dt == np.float32 else 1e-10
    assert abs(res - np.sqrt(r2)) <= rel * np.sqrt(r2), (res, np.sqrt(r2))


@pytest.mark.parametrize("shape", ["star", "face_edge"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(50, 33, 27), (260, 9, 10), (7, 6, 5), (130, 20, 40), (1, 1, 1)])
@pytest.mark.parametrize("steps", [2, 3, 4, 5, 7])
def test_two_sweeps_per_launch_bit_exact(shape, dt, dims, steps, monkeypatch):
    """Multi-sweep runs launch two sweeps at a time (temporal blocking) for
    radius-1 stencils, three for the 7-point star; the result equals one sweep
    per launch bit for bit (same arithmetic per point) and the oracle within the
    per-sweep tolerance."""
    nx, ny, nz = dims
    taps = wl.taps3d(shape, 1)
    coef = wl.uniform(steps + 40, len(taps), -0.5, 0.5)
    x = oracle.uniform(17, nx * ny * nz, -1, 1, dt).reshape(nz, ny, nx)
    src = torch.from_numpy(x).cuda()
    code = L.FACE_EDGE if shape == "face_edge" else L.STAR
    outs = []
    for t in ("3", "2", "1"):
        monkeypatch.setenv("SSAM_3D_T", t)
        dst = torch.full_like(src, float("nan"))
        work = torch.full_like(src, float("nan"))
        api.stencil3d(src, dst, coef, shape=code, radius=1, steps=steps, work=work)
        torch.cuda.synchronize()
        outs.append(dst.cpu().numpy())
    assert np.array_equal(outs[0], outs[2])
    assert np.array_equal(outs[1], outs[2])
    want = _gold(x, taps, coef, 1, steps)
    _check(outs[0], want, _scale(x, taps, coef, 1, steps), TOL[dt] * steps, f"{shape} {dims} x{steps}")


@pytest.mark.parametrize("zseg", ["1", "2", "5", "64"])
@pytest.mark.parametrize("steps", [3, 6])
def test_three_sweeps_short_z_segments(zseg, steps, monkeypatch):
    """Three sweeps per launch on z-segments from 1 plane up (each tile runs 4
    fill ticks before its first output plane): bit-equal to one sweep per launch."""
    nx, ny, nz = 70, 21, 30
    taps = wl.taps3d("star", 1)
    coef = wl.uniform(steps + 90, len(taps), -0.5, 0.5)
    x = oracle.uniform(23, nx * ny * nz, -1, 1, np.float32).reshape(nz, ny, nx)
    src = torch.from_numpy(x).cuda()
    monkeypatch.setenv("SSAM_ZSEG", zseg)
    outs = []
    for t in ("3", "1"):
        monkeypatch.setenv("SSAM_3D_T", t)
        dst = torch.full_like(src, float("nan"))
        work = torch.full_like(src, float("nan"))
        api.stencil3d(src, dst, coef, shape=L.STAR, radius=1, steps=steps, work=work)
        torch.cuda.synchronize()
        outs.append(dst.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])


def test_jacobi19_residual_concurrent_threads():
    """Residual requests from several host threads, each on its own stream and
    grid size (the partials scratch grows between them): every call returns the
    norm a call alone returns."""
    import concurrent.futures as cf

    sizes = [24, 40, 56, 72]
    grids = []
    for k, n in enumerate(sizes):
        x = oracle.uniform(300 + k, n ** 3, 0, 1, np.float64).reshape(n, n, n)
        src = torch.from_numpy(x).cuda()
        grids.append((src, torch.empty_like(src), torch.empty_like(src)))
    alone = [api.jacobi19(s, d, 3, w) for s, d, w in grids]

    def run(k):
        s, d, w = grids[k]
        st = torch.cuda.Stream()
        with torch.cuda.stream(st):
            return [api.jacobi19(s, d, 3, w, stream=st) for _ in range(4)]

    with cf.ThreadPoolExecutor(max_workers=len(sizes)) as ex:
        got = list(ex.map(run, range(len(sizes))))
    for k in range(len(sizes)):
        assert got[k] == [alone[k]] * 4, (sizes[k], got[k], alone[k])


@pytest.mark.parametrize("form", ["1", "2"])
@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("dims,zseg", [((50, 33, 27), "0"), ((260, 9, 10), "3"), ((7, 6, 5), "0"),
                                        ((130, 20, 40), "5"), ((1, 1, 1), "0"), ((256, 12, 33), "7")])
def test_radius2_plane_ring_and_z_partial_sums(form, dt, dims, zseg, monkeypatch):
    """The 13-point star through the plane-ring kernel (SSAM_3D_FORM=1) and the
    z-partial-sum kernel (2, the default), against the oracle, on ragged grids and
    short z-segments (tiles that start and end inside the grid)."""
    nx, ny, nz = dims
    monkeypatch.setenv("SSAM_3D_FORM", form)
    monkeypatch.setenv("SSAM_ZSEG", zseg)
    taps = wl.taps3d("star", 2)
    coef = wl.uniform(77, len(taps), -1, 1)
    x = oracle.uniform(12, nx * ny * nz, -1, 1, dt).reshape(nz, ny, nx)
    src = torch.from_numpy(x).cuda()
    dst = torch.full_like(src, float("nan"))
    api.stencil3d(src, dst, coef, shape=L.STAR, radius=2)
    torch.cuda.synchronize()
    want = _gold(x, taps, coef, 2, 1)
    _check(dst.cpu().numpy(), want, _scale(x, taps, coef, 2, 1), TOL[dt], f"form {form} {dims} zseg {zseg}")
