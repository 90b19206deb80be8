"""Write the golden fixtures under tests/golden/ (run in the build container, where
the reference tree is present; the fixtures themselves are plain data and travel).

  coeffs.json  -- the reference's coefficient tables, read as data from its
                  harness sources: the filter[][] arrays of
                  ssai-2dstencil/stencil2D_{5pt-v1,9,13,17,21,25,64,81,121}pt.cu and
                  stencil2d_temp_{5,9}pt.cu, and the first (#if 1) definitions of
                  the 3D tap macros of ssai-3dstencil/tmp3d7pt-v2.cu,
                  tmp3d13pt.cu, tmpPoisson.cu.
  kat_*.npz    -- known-answer vectors: seeded small inputs and the output of the
                  reference's OWN gold, Convolution<T> (ssai-2dconv/common.h:186-216)
                  compiled from where it lies into oracle/_ref, iterated as the
                  temporal harness does (stencil2d_temp_5pt.cu:172-175).

    python tests/golden/make_fixtures.py [--reference /root/reference]
"""
from __future__ import annotations

import argparse
import json
import pathlib
import re
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import oracle  # noqa: E402
from ssam_amd import workloads as wl  # noqa: E402

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def _strip_if0(text: str) -> str:
    """Keep the `#if 1` branch of `#if 1 ... #else ... #endif` blocks."""
    out, skip = [], False
    depth_else = False
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("#if 1"):
            depth_else = True
            continue
        if s.startswith("#else") and depth_else:
            skip = True
            continue
        if s.startswith("#endif") and depth_else:
            skip = False
            depth_else = False
            continue
        if not skip:
            out.append(line)
    return "\n".join(out)


def filter_table(path: pathlib.Path) -> list[list[float]]:
    text = _strip_if0(path.read_text())
    i = text.index("DataType filter[FILTER_HEIGHT][FILTER_WIDTH] =")
    j = text.index("};", i)
    body = text[i:j]
    body = body[body.index("{") + 1:]
    rows = re.findall(r"\{([^{}]*)\}", body)
    tab = [[float(v) for v in re.findall(NUM, r)] for r in rows]
    return [r for r in tab if r]


def macros3d(path: pathlib.Path, prefix: str) -> dict[str, float]:
    text = _strip_if0(path.read_text())
    out = {}
    for name, val in re.findall(r"#define\s+(" + prefix + r"\d+)\s+\(\(T\)(" + NUM + r")\)", text):
        out.setdefault(name, float(val))
    return out


def write_coeffs(ref: pathlib.Path) -> dict:
    s2 = ref / "ssai-2dstencil"
    s3 = ref / "ssai-3dstencil"
    d = {
        "source": "reference harness filter tables and 3D tap macros (see make_fixtures.py)",
        "star_int": {str(r): filter_table(s2 / f"stencil2D_{n}pt.cu")
                     for r, n in [(2, 9), (3, 13), (4, 17), (5, 21), (6, 25)]},
        "box64": filter_table(s2 / "stencil2D_64pt.cu"),
        "box81": filter_table(s2 / "stencil2D_81pt.cu"),
        "box121": filter_table(s2 / "stencil2D_121pt.cu"),
        "temporal5": filter_table(s2 / "stencil2d_temp_5pt.cu"),
        "temporal9": filter_table(s2 / "stencil2d_temp_9pt.cu"),
        "j3d7pt": macros3d(s3 / "tmp3d7pt-v2.cu", "_F3d7PT_"),
        "j3d13pt": macros3d(s3 / "tmp3d13pt.cu", "_F3d13PT_"),
        "poisson": macros3d(s3 / "tmpPoisson.cu", "_F"),
    }
    # the 5pt-v1 harness draws its taps from rand(); keep its fixed template
    d["star_int"]["1"] = filter_table(s2 / "stencil2D_5pt-v1.cu")
    (HERE / "coeffs.json").write_text(json.dumps(d, indent=1))
    return d


def write_kats(coeffs: dict) -> list[str]:
    if oracle.ref_lib() is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref` first")
    written = []

    def save(name, **arrays):
        np.savez_compressed(HERE / f"{name}.npz", **arrays)
        written.append(name)

    h, w = 53, 71
    # integer-exact stars (reference integer tables, inputs 0..255): order-independent
    img = np.floor(wl.uniform(11, h * w, 0.0, 256.0)).reshape(h, w)
    for r in range(1, 7):
        f = np.array(coeffs["star_int"][str(r)], dtype=np.float64) if r > 1 else \
            np.array([[0, 1, 0], [4, 1, 2], [0, 3, 0]], dtype=np.float64)
        for dt in (np.float32, np.float64):
            out = oracle.ref_conv2d(img.astype(dt), f.astype(dt))
            save(f"kat_star_r{r}_int_{np.dtype(dt).name}", src=img.astype(dt), filt=f.astype(dt),
                 out=out)
    # reference box tables on seeded floats (rounding follows the reference gold)
    x = wl.uniform(12, h * w, -1.0, 1.0).reshape(h, w)
    for name in ("box64", "box81", "box121"):
        f = np.array(coeffs[name], dtype=np.float64)
        for dt in (np.float32, np.float64):
            save(f"kat_{name}_{np.dtype(dt).name}", src=x.astype(dt), filt=f.astype(dt),
                 out=oracle.ref_conv2d(x.astype(dt), f.astype(dt)))
    # integer convolution taps in [-10, 10] (conv2D.cu:115-119 draws rand()%21-10), over
    # the 8192^2 convolution workload's filter sizes and the reference's range up to 28
    # (conv2D.cu:186-212); 255 * 10 * 28^2 < 2^24, so every FP32 sum is exact
    for k in (2, 3, 5, 7, 9, 15, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28):
        f = np.floor(wl.uniform(100 + k, k * k, -10.0, 11.0)).reshape(k, k)
        for dt in (np.float32, np.float64):
            save(f"kat_conv{k}_int_{np.dtype(dt).name}", src=img.astype(dt), filt=f.astype(dt),
                 out=oracle.ref_conv2d(img.astype(dt), f.astype(dt)))
    # temporal: reference coefficients, iterated reference gold
    xs = (wl.uniform(13, h * w) + 0.02121).reshape(h, w)
    for name, steps in (("temporal5", 5), ("temporal5", 8), ("temporal9", 3), ("temporal9", 8)):
        f = np.array(coeffs[name], dtype=np.float64)
        for dt in (np.float32, np.float64):
            cur = xs.astype(dt)
            for _ in range(steps):
                cur = oracle.ref_conv2d(cur, f.astype(dt))
            save(f"kat_{name}_T{steps}_{np.dtype(dt).name}", src=xs.astype(dt),
                 filt=f.astype(dt), steps=np.array(steps), out=cur)
    return written


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    a = ap.parse_args()
    ref = pathlib.Path(a.reference)
    oracle.build(str(ref))
    c = write_coeffs(ref)
    names = write_kats(c)
    print(f"coeffs.json + {len(names)} KAT files")


if __name__ == "__main__":
    main()
