"""Synthetic workloads: seeded coefficient sets and the stencil shapes of the
benchmark configurations, plus the reference's own coefficient tables.

Seeded values use splitmix64 (element i = (i+1)-th output of the stream started at
`seed`), value = a + (b-a) * (u >> 40) * 2^-24 -- the same stream the device fill
kernel (ssam_fill_uniform) and the CPU oracle produce.
"""
from __future__ import annotations

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, a: float = 0.0, b: float = 1.0, offset: int = 0) -> np.ndarray:
    u = (splitmix64(seed, n, offset) >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)
    return a + (b - a) * u


def star_mask(r: int) -> np.ndarray:
    k = 2 * r + 1
    m = np.zeros((k, k), dtype=bool)
    m[r, :] = True
    m[:, r] = True
    return m


def normalised_taps(shape: str, r: int, seed: int) -> np.ndarray:
    """Taps U[0,1) from `seed`, zero outside the shape, normalised to sum 1 (FP64)."""
    k = 2 * r + 1
    w = uniform(seed, k * k).reshape(k, k)
    if shape == "star":
        w = np.where(star_mask(r), w, 0.0)
    return w / w.sum()


def conv_taps(k: int, seed: int) -> np.ndarray:
    """Convolution taps U[-1,1) from `seed`, divided by the tap count."""
    return uniform(seed, k * k, -1.0, 1.0).reshape(k, k) / (k * k)


# ---------------------------------------------------------------- reference tables
def ref_5pt() -> np.ndarray:
    """stencil2d_temp_5pt.cu:126-142 (fn .1, fw -.4, fc .2, fe .6, fs -.3)."""
    return np.array([[0, 0.1, 0], [-0.4, 0.2, 0.6], [0, -0.3, 0]])


def ref_star_int(r: int) -> np.ndarray:
    """Integer star tables of stencil2D_{9,13,17,21,25}pt.cu (1..4r+1 laid out as in
    the reference harness), loaded from the committed fixture."""
    import json
    import pathlib
    p = pathlib.Path(__file__).resolve().parent.parent / "tests" / "golden" / "coeffs.json"
    return np.array(json.loads(p.read_text())["star_int"][str(r)], dtype=np.float64)


# 3D reference coefficients in the descriptor tap order (include/ssam.h)
REF3D_7PT = [0.3, 0.5, 0.6, 0.2, 0.4, 0.1, 0.7]                      # tmp3d7pt-v2.cu:9-15
REF3D_13PT = [0.3, 1.1, 1.2, 0.6, 0.8, 0.2, 0.4, 1.0, 1.3, 0.7, 0.9, 0.1, 0.5]  # tmp3d13pt.cu:9-21
REF3D_POISSON = [2.666] + [-0.166] * 6 + [-0.0833] * 4 + [0.0833] * 8  # tmpPoisson.cu:9-27
JACOBI19 = [0.0] + [1.0 / 12] * 6 + [1.0 / 24] * 12


def taps3d(shape: str, r: int) -> list[tuple[int, int, int]]:
    """(dz, dy, dx) of each tap in the descriptor order."""
    if shape == "face_edge":
        t = [(0, 0, 0), (0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0), (1, 0, 0)]
        for e in range(12):
            s0 = 1 if e & 1 else -1
            s1 = 1 if e & 2 else -1
            if e < 4:
                t.append((s1, s0, 0))
            elif e < 8:
                t.append((s1, 0, s0))
            else:
                t.append((0, s1, s0))
        return t
    t = [(0, 0, 0)]
    for k in range(1, r + 1):
        t += [(0, 0, -k), (0, 0, k), (0, -k, 0), (0, k, 0), (-k, 0, 0), (k, 0, 0)]
    return t


def config4_batch(seed: int = 4, n2d: int = 192, n3d: int = 64):
    """The uneven batch: n2d 2D grids with sides log-uniform in [128, 4096] (aspect
    ratio at most 4:1, redrawn otherwise), each a 5-point star, 5x5 box or 11x11 box
    at random; n3d 3D grids with sides log-uniform in [64, 256] running the 7-point
    star for 10 steps.  Returns (list2d, list3d) of dicts with sizes and shapes; all
    from one splitmix64 stream."""
    import math
    u = iter(uniform(seed, 100000))
    lo, hi = math.log(128), math.log(4096)
    g2 = []
    while len(g2) < n2d:
        w = int(round(math.exp(lo + (hi - lo) * next(u))))
        h = int(round(math.exp(lo + (hi - lo) * next(u))))
        if max(w, h) > 4 * min(w, h):
            continue
        kind = min(2, int(3 * next(u)))
        shape, r = [("star", 1), ("box", 2), ("box", 5)][kind]
        g2.append(dict(w=w, h=h, shape=shape, r=r))
    lo3, hi3 = math.log(64), math.log(256)
    g3 = []
    for _ in range(n3d):
        nx, ny, nz = (int(round(math.exp(lo3 + (hi3 - lo3) * next(u)))) for _ in range(3))
        g3.append(dict(nx=nx, ny=ny, nz=nz, shape="star", r=1, steps=10))
    return g2, g3
