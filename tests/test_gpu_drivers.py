"""The three C++ drivers (ssam_amd/bin), which keep the reference's command
lines and output lines and call the kernels through the C ABI: each runs on a
small grid, reads its input image from ../data/Lena<N>x<N>.raw (the reference's
raw 8-bit loader, common.h:148-178) when present, checks itself against its CPU
loop, and the conv driver appends its log.conv2D.csv row (conv2D.cu:165-169)."""
import json
import pathlib
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BIN = pathlib.Path(__file__).resolve().parent.parent / "ssam_amd" / "bin"
_C = json.loads((pathlib.Path(__file__).resolve().parent / "golden" / "coeffs.json").read_text())
TABLES = {"64": (np.array(_C["box64"]), 1), "81": (np.array(_C["box81"]), 1),
          "121": (np.array(_C["box121"]), 1), "tmp_5": (np.array(_C["temporal5"]), 5),
          "tmp_9": (np.array(_C["temporal9"]), 3)}


@pytest.fixture()
def rundir(tmp_path):
    run = tmp_path / "run"
    run.mkdir()
    (tmp_path / "data").mkdir()
    rng = np.random.default_rng(0)
    for n in (64, 96):
        rng.integers(0, 256, n * n, dtype=np.uint8).tofile(tmp_path / "data" / f"Lena{n}x{n}.raw")
    return run


def _run(rundir, *args, expect_rc=0):
    exe = BIN / args[0]
    assert exe.exists(), f"{exe} not built"
    p = subprocess.run([str(exe), *map(str, args[1:])], cwd=rundir, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == expect_rc, p.stdout + p.stderr
    return p.stdout


@pytest.mark.parametrize("pts", ["5", "9", "13", "17", "21", "25", "64", "81", "121", "tmp_5", "tmp_9"])
@pytest.mark.parametrize("dtype", ["float", "double"])
def test_2dstencil_driver(rundir, pts, dtype):
    out = _run(rundir, "ssai-2dstencil", 64, pts, dtype)
    assert f"call_function : stencil_{pts}pt_{dtype} :" in out
    assert "Load success : ../data/Lena64x64.raw" in out
    assert re.search(r"^64x64 , \d+x\d+ , proc_count=0, cache=\d+, ", out, re.M), out
    dif = float(re.search(r"verify dif =([-+0-9.eE]+)", out).group(1))
    if pts in ("5", "9", "13", "17", "21", "25"):
        assert dif == 0.0, out  # integer tables on 8-bit images: exact
    else:
        # sum over the interior of |gpu - cpu| within tol * (sum |w|)^T * 255 per point
        w, steps = TABLES[pts]
        m = (w.shape[0] // 2) * steps
        npts = (64 - 2 * m - 1) ** 2
        tol = 1e-5 if dtype == "float" else 1e-12
        assert dif <= npts * tol * np.abs(w).sum() ** steps * 255, out


def test_2dstencil_generates_data_without_image(rundir):
    out = _run(rundir, "ssai-2dstencil", 80, "5", "float")
    assert "Load failed : ../data/Lena80x80.raw, generate random data" in out
    assert float(re.search(r"verify dif =([-+0-9.eE]+)", out).group(1)) == 0.0


@pytest.mark.parametrize("fn", ["7pt", "13pt", "poisson"])
@pytest.mark.parametrize("dtype,n", [("float", 48), ("double", 48), ("float", 200), ("double", 130)])
def test_3dstencil_driver(rundir, fn, dtype, n):
    out = _run(rundir, "ssai-3dstencil", fn, dtype, n)
    assert "use MI355X GPU" in out and "[Test] Max Error" in out
    rms = float(re.search(r"RMS Error : ([-+0-9.eE]+)", out).group(1))
    assert rms <= 1e-5, out


# the reference's run.sh loops fsize 2..21 and conv2D.cu:186-212 dispatches up to 28
@pytest.mark.parametrize("fsize", [2, 3, 5, 9, 19, 20, 21, 28])
@pytest.mark.parametrize("dtype", ["float", "double"])
def test_2dconv_driver_and_csv(rundir, fsize, dtype):
    out = _run(rundir, "ssai-2dconv", dtype, 96, fsize)
    assert "Load success : ../data/Lena96x96.raw" in out
    # 8-bit image and integer taps in [-10, 10]: every sum is exact (255*10*28^2 < 2^24)
    assert float(re.search(r"verify dif =([-+0-9.eE]+)", out).group(1)) == 0.0, out
    rows = (rundir / "log.conv2D.csv").read_text().strip().splitlines()
    assert re.fullmatch(rf"96x96, 0_{fsize}, 256, {fsize}x{fsize}, [0-9.]+", rows[-1]), rows[-1]


def test_drivers_reject_bad_arguments(rundir):
    p = subprocess.run([str(BIN / "ssai-2dstencil")], cwd=rundir, capture_output=True, text=True, timeout=60)
    assert "error, do not call functions" in p.stdout
    p = subprocess.run([str(BIN / "ssai-3dstencil"), "9pt", "float", "32"], cwd=rundir, capture_output=True,
                       text=True, timeout=60)
    assert "invalid function name!" in p.stdout
