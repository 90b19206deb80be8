"""Build ssam_amd/libssam.so from ssam_amd/csrc with hipcc for gfx950.

Each translation unit compiles in its own hipcc process (in parallel), then the
objects link into one shared library in-tree, next to this file.  The library is
self-contained C ABI (include/ssam.h); nothing here imports torch.

    python -m ssam_amd.build [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OBJ = HERE / "_obj"
LIB = HERE / "libssam.so"
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # no packed-FP32 VALU forms in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = [
    "api.cpp",
    "util.hip",
    "stencil2d_f32.hip",
    "stencil2d_f64.hip",
    "stencil2d_bf16.hip",
    "stencil2d_f16.hip",
    "stencil2d_box_f32.hip",
    "stencil2d_box_f64.hip",
    "stencil2d_box_bf16.hip",
    "stencil2d_box_f16.hip",
    "stencil2d_even_f32.hip",
    "stencil2d_lds_box_f32.hip",
    "stencil2d_lds_even_f32.hip",
    "stencil2d_even_f64.hip",
    "stencil2d_kahan_f32.hip",
    "stencil2d_kahan_bf16.hip",
    "stencil2d_kahan_f16.hip",
    "stencil3d.hip",
    "stencil3d_f32f64.hip",
    "stencil3d_bf16.hip",
    "stencil3d_f16.hip",
    "batch.hip",
    "batch2d.hip",
    "batch3d.hip",
    "batch3d_t2.hip",
    "batch3d_t2_bf16.hip",
    "batch3d_t2_f16.hip",
    "batch3d_t3.hip",
    "batch3d_t3_bf16.hip",
    "batch3d_t3_f16.hip",
    "microbench.hip",
]


def _compile(src: str) -> tuple[str, int, str]:
    s = CSRC / src
    o = OBJ / (s.stem + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC, *CFLAGS, *lang, "-c", str(s), "-o", str(o)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    return src, p.returncode, out


def build(jobs: int | None = None, verbose: bool = False) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, out in ex.map(_compile, SOURCES):
            if out.strip() and (verbose or rc):
                print(f"[{src}]\n{out}", file=sys.stderr)
            if rc:
                failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed for: {', '.join(failed)}")
    objs = [str(OBJ / (pathlib.Path(s).stem + ".o")) for s in SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode:
        raise RuntimeError(f"link failed:\n{p.stdout}{p.stderr}")
    os.replace(tmp, LIB)
    return LIB


DRIVERS = {
    "ssai-2dstencil": "ssai_2dstencil.cpp",
    "ssai-3dstencil": "ssai_3dstencil.cpp",
    "ssai-2dconv": "ssai_2dconv.cpp",
}
BIN = HERE / "bin"


def build_drivers() -> list[pathlib.Path]:
    """The reference-command-line drivers, linked against libssam.so."""
    BIN.mkdir(exist_ok=True)
    out = []
    for exe, src in DRIVERS.items():
        target = BIN / exe
        cmd = [HIPCC, "-O3", "-std=c++17", "-fopenmp", str(HERE / "drivers" / src), "-o",
               str(target), f"-L{HERE}", "-lssam", "-Wl,-rpath,$ORIGIN/.."]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode:
            raise RuntimeError(f"driver {exe} failed:\n{p.stdout}{p.stderr}")
        out.append(target)
    return out


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
    for p in build_drivers():
        print(p)


if __name__ == "__main__":
    main()
