#!/usr/bin/env python3
"""Benchmark: SSAM 2D stencil suite on one MI355X.

Workload (one "step"): one sweep of every 2D stencil of the 8192 x 8192 FP32
configuration -- boxes of radius 1..5 (9/25/49/81/121 points), stars of radius
3/4/5 (13/17/21 points), and the temporally blocked 5-point and 9-point stars
fused over T = 8 steps -- on a grid of U[-1,1) values from seed 0 with taps
U[0,1) normalised to sum 1, inputs resident in HBM.  value = total point-updates /
time of the step, in Gpoints/s.

Also measured in the same run (outside the timed steps): per-kernel times with
HIP events on the launch stream and each kernel's binding roof (HBM or FP32
compute), the stream-copy bandwidth, the same suite in FP64 and in BF16 / FP16
storage, errors of every form against the FP64 oracle on row bands, configs 0 and
2-4, and the CPU baseline (the oracle on a bounded slab, all host threads).

    python bench.py [--gpus 1] [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBPS = 8000.0
# MI355X published dense peaks (TFLOP/s): FP32 (vector and matrix), FP64, BF16/FP16 matrix
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
BF16_MFMA_PEAK_TFLOPS = 2500.0
N = 8192
BANDS = (0, 2040, 4093, 8176)  # first rows of the 16-row bands checked against the oracle

# (name, shape, radius, fused steps)
SUITE = [
    ("box9", "box", 1, 1), ("box25", "box", 2, 1), ("box49", "box", 3, 1),
    ("box81", "box", 4, 1), ("box121", "box", 5, 1),
    ("star13", "star", 3, 1), ("star17", "star", 4, 1), ("star21", "star", 5, 1),
    ("tmp5_T8", "star", 1, 8), ("tmp9_T8", "star", 2, 8),
]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip FP64/BF16 side measurements")
    ap.add_argument("--only", default="", help="comma list of suite kernel names (profiling)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (heuristic tiling)")
    ap.add_argument("--tune-file", default="",
                    help="reuse the suite's tuning from this file when it exists, else tune and save it "
                         "there (profiled runs then time the same kernels as the run that tuned)")
    ap.add_argument("--ingest-trace", metavar="TRACE_DIR",
                    help="no GPU: summarise a rocprofv3 --kernel-trace run of this bench (the "
                         "timed dispatches of each suite kernel) into profiles/ and exit")
    ap.add_argument("--ingest-pmc", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="no GPU: turn two rocprofv3 --pmc runs of this bench (FETCH_SIZE, then "
                         "WRITE_SIZE) into per-kernel HBM bytes under profiles/ and exit")
    ap.add_argument("--side", choices=("config0", "config2", "config3", "config4"),
                    help="run only this side measurement and print its JSON (for profiling)")
    ap.add_argument("--ingest-sq", nargs="+", metavar="ARG",
                    help="NAME COMMAND DIR [DIR ...]: summarise rocprofv3 --pmc runs of SQ / GRBM "
                         "counters (one directory per pass) into profiles/NAME.md and exit")
    ap.add_argument("--ingest-3d-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="no GPU: turn two rocprofv3 --pmc runs of --side config2 (FETCH_SIZE, then "
                         "WRITE_SIZE) into HBM bytes per sweep of each FP32 / FP64 3D kernel under "
                         "profiles/ and exit")
    ap.add_argument("--ingest-stats", nargs=3, metavar=("STATS_DIR", "NAME", "COMMAND"),
                    help="no GPU: summarise a rocprofv3 --kernel-trace --stats run (its "
                         "kernel_stats.csv) into profiles/NAME.md, quoting COMMAND, and exit")
    return ap.parse_args()


def ingest_stats(stats_dir: str, name: str, command: str) -> None:
    """Per-kernel calls, average and total duration of a rocprofv3 --stats run."""
    import csv

    f = next(pathlib.Path(stats_dir).rglob("*kernel_stats.csv"))
    rows = list(csv.DictReader(open(f)))
    total = sum(float(r["TotalDurationNs"]) for r in rows) or 1.0
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"# {name}", "", "Command (repository root, on one MI355X):", "", "```", command, "```", "",
           "| kernel | calls | avg us | min us | max us | share of GPU time |",
           "|---|---|---|---|---|---|"]
    for r in rows[:24]:
        out.append(f"| `{r['Name'][:150]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['MinNs']) / 1e3:.1f} | {float(r['MaxNs']) / 1e3:.1f} | "
                   f"{100 * float(r['TotalDurationNs']) / total:.1f}% |")
    dst = pathlib.Path(__file__).resolve().parent / "profiles" / f"{name}.md"
    dst.write_text("\n".join(out) + "\n")
    print(dst)


def ingest_3d_traffic(fetch_dir: str, write_dir: str) -> None:
    """HBM bytes per dispatch of the FP32 / FP64 3D kernels of a --side config2 run
    (512^3), from FETCH_SIZE and WRITE_SIZE (KiB per dispatch, scaled x2 and x1:
    the factors that make them read the bytes of the stream copy), as multiples of
    the grid's bytes, and per sweep (the launch's sweeps: ssam3d_t3_kernel three,
    ssam3d_t2_kernel two, others one)."""
    import collections
    import csv

    def load(d, counter):
        per = collections.defaultdict(list)
        for r in csv.DictReader(open(next(pathlib.Path(d).rglob("*counter_collection.csv")))):
            if "ssam3d" in r["Kernel_Name"] and r["Counter_Name"] == counter:
                per[r["Kernel_Name"].split("(")[0].replace("void ssam::", "")].append(float(r["Counter_Value"]))
        return per

    fe, wr = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    rows = []
    for k, v in fe.items():
        targ = k.split("<", 1)[1]
        if not targ.startswith(("float,", "double,")) or not targ.rstrip(">").endswith(("float", "double")):
            continue
        grid = 512 ** 3 * (8 if targ.startswith("double") else 4)
        sweeps = 3 if "t3_kernel" in k else (2 if "t2_kernel" in k else 1)
        f = sum(v) / len(v) * 1024 * 2 / grid
        w = sum(wr[k]) / len(wr[k]) * 1024 / grid if wr.get(k) else float("nan")
        rows.append(f"| `{k}` | {sweeps} | {len(v)} | {f:.2f} | {w:.2f} | {(f + w) / sweeps:.2f} |")
    cmd = "\n".join(f"timeout -s KILL 300 rocprofv3 --pmc {c} --output-format csv -d {d} -o run -- "
                     "python3 bench.py --side config2" for c, d in (("FETCH_SIZE", "F"), ("WRITE_SIZE", "W")))
    out = ["# stencil3d_512_multi_sweep_traffic", "",
           "HBM traffic per dispatch of the FP32 / FP64 3D kernels on 512³ (config 2 side run): one sweep per",
           "launch (`ssam3d_kernel`, `ssam3d_zp_kernel`), two (`ssam3d_t2_kernel`) and three (`ssam3d_t3_kernel`,",
           "7-point star).", "", "Commands (repository root, on one MI355X):", "", "```", cmd,
           "python3 bench.py --ingest-3d-traffic F W", "```", "",
           "FETCH_SIZE is scaled ×2 and WRITE_SIZE ×1, the factors that make each counter read the bytes of the",
           "1 GiB stream copy (the suite's PMC calibration). Grid: 512 MiB FP32, 1 GiB FP64. Averages over the",
           "dispatches of each kernel; template arguments: storage, shape (0 star, 1 face+edge), radius or",
           "columns per lane, ..., dst storage.", "",
           "| kernel | sweeps / launch | dispatches | FETCH (× grid) | WRITE (× grid) | HBM bytes per sweep (× grid) |",
           "|---|---|---|---|---|---|", *rows]
    dst = ROOT / "profiles" / "stencil3d_512_multi_sweep_traffic.md"
    dst.write_text("\n".join(out) + "\n")
    print(dst)


def ingest_sq(name: str, command: str, dirs) -> None:
    """Per-kernel SQ / GRBM counters of rocprofv3 --pmc runs (one directory per
    counter pass): instruction mix, issue rate and occupancy of every SSAM kernel
    dispatched after the bench's 16-byte marker copy (all dispatches when the run
    has no marker).  SQ_WAVE_CYCLES, SQ_WAIT_* and SQ_ACTIVE_* count quad-cycles;
    GRBM_GUI_ACTIVE sums the 8 XCDs' clocks."""
    import csv
    import collections

    per = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in dirs:
        f = next(pathlib.Path(d).rglob("*counter_collection.csv"))
        rows = list(csv.DictReader(open(f)))
        mark = next((int(r["Dispatch_Id"]) for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"]))
                     if "stream_copy" in r["Kernel_Name"] and int(r["Grid_Size"]) <= 256), 0)
        disp = collections.defaultdict(dict)
        for r in rows:
            if int(r["Dispatch_Id"]) <= mark or "ssam" not in r["Kernel_Name"] or "stream_copy" in r["Kernel_Name"] \
                    or "fill_uniform" in r["Kernel_Name"]:
                continue
            disp[(r["Kernel_Name"], r["Dispatch_Id"])][r["Counter_Name"]] = float(r["Counter_Value"])
        for (k, _), c in disp.items():
            for cn, v in c.items():
                per[k.split("(")[0].replace("void ", "")][cn].append(v)
    cus = 256
    md = [f"# {name}", "", "Commands (one MI355X, repository root; one counter pass per run):", "", "```",
          command, "```", "",
          "Derived per kernel (averages over its dispatches): VALU/clk/CU = SQ_INSTS_VALU / CUs / "
          "(GRBM_GUI_ACTIVE / 8); waves/SIMD = SQ_WAVE_CYCLES * 4 / (GRBM_GUI_ACTIVE / 8) / CUs / 4; "
          "wait = SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES; FMA share = SQ_INSTS_VALU_FMA_F32 / SQ_INSTS_VALU; "
          "DPP-fused adds appear in SQ_INSTS_VALU_ADD_F32.", "",
          "| kernel | dispatches | VALU/clk/CU | waves/SIMD | wait frac | FMA share | ADD_F32 share | SALU/VALU | "
          "SMEM/VALU | LDS instr/VALU | VMEM rd/VALU |", "|---|---|---|---|---|---|---|---|---|---|---|"]

    def avg(c, k):
        v = c.get(k)
        return sum(v) / len(v) if v else None

    rows = []
    for k, c in per.items():
        g = avg(c, "GRBM_GUI_ACTIVE")
        valu = avg(c, "SQ_INSTS_VALU")
        if not g or not valu:
            continue
        clk = g / 8
        wc = avg(c, "SQ_WAVE_CYCLES")
        fmt = lambda x: "" if x is None else f"{x:.3f}"
        ratio = lambda a: None if avg(c, a) is None else avg(c, a) / valu
        rows.append((valu, f"| `{k[:120]}` | {len(c['SQ_INSTS_VALU'])} | {valu / cus / clk:.2f} | "
                           f"{fmt(None if wc is None else wc * 4 / clk / cus / 4)} | "
                           f"{fmt(None if wc is None or avg(c, 'SQ_WAIT_INST_ANY') is None else avg(c, 'SQ_WAIT_INST_ANY') / wc)} | "
                           f"{fmt(ratio('SQ_INSTS_VALU_FMA_F32'))} | {fmt(ratio('SQ_INSTS_VALU_ADD_F32'))} | "
                           f"{fmt(ratio('SQ_INSTS_SALU'))} | {fmt(ratio('SQ_INSTS_SMEM'))} | "
                           f"{fmt(ratio('SQ_INSTS_LDS'))} | {fmt(ratio('SQ_INSTS_VMEM_RD'))} |"))
    md += [r for _, r in sorted(rows, reverse=True)]
    dst = ROOT / "profiles" / f"{name}.md"
    dst.write_text("\n".join(md) + "\n")
    print("\n".join(md))


def suite_kernel(k: str):
    """(suite name, steps per launch) of an FP32 suite kernel's full name, or None:
    the register-cache kernel ssam2d_kernel<float, R, SHAPE, V, T, ...> or the
    LDS-tap form ssam2d_lds_kernel<float, R, SHAPE, ...> (one step)."""
    import re

    m = re.search(r"ssam2d_kernel<float, (\d+), (\d+), \d+, (\d+),", k)
    if m:
        r, shape, t = (int(v) for v in m.groups())
    else:
        m = re.search(r"ssam2d_lds_kernel<float, (\d+), (\d+),", k)
        if not m:
            return None
        r, shape, t = int(m.group(1)), int(m.group(2)), 1
    for n, sh, rr, tt in SUITE:
        if rr == r and (1 if sh == "box" else 0) == shape:
            return n, t
    return None


def ingest_trace(trace_dir: str, steps: int, warmup: int) -> None:
    """Per-kernel duration of the suite's warm-up and timed dispatches from a
    rocprofv3 kernel trace of `bench.py --steps S --warmup W --no-extras
    --no-cpu-baseline` (the dispatches between the 16-byte marker copy and the
    bandwidth copy)."""
    import csv

    f = next(pathlib.Path(trace_dir).rglob("*kernel_trace.csv"))
    recs = []
    for r in csv.DictReader(open(f)):
        grid = int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0)
        recs.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"], grid,
                     r.get("VGPR_Count", ""), r.get("Accum_VGPR_Count", "")))
    recs.sort()
    # the timed phase: after the 16-byte marker copy, up to the bandwidth copy
    t0 = next(x[0] for x in recs if "stream_copy" in x[2] and x[3] <= 256)
    t1 = next((x[0] for x in recs if "stream_copy" in x[2] and x[3] > 256 and x[0] > t0), recs[-1][1] + 1)
    runs = {}
    for st, en, k, grid, vg, ag in recs:
        if t0 < st < t1 or "stream_copy" in k:
            runs.setdefault(k, []).append(((en - st) / 1000.0, vg, ag, grid))
    rows = []
    for k, v in runs.items():
        d = [x[0] for x in v]
        rows.append((sum(d) / len(d), k, len(d), min(d), max(d), v[-1][1], v[-1][2], v[-1][3]))
    rows.sort(reverse=True)
    md = ["# rocprofv3 kernel trace: 2D SSAM suite, 8192x8192 FP32 (bench workload)", "",
          "Commands (one MI355X, repo root; the first run tunes and saves the choices in T, the "
          "profiled run reloads them, so it launches the kernels the timed bench launches):", "",
          f"    python3 bench.py --no-extras --no-cpu-baseline --tune-file <T> --steps 3 --warmup 1",
          f"    rocprofv3 --kernel-trace --stats --output-format csv -d <D> -o run -- python3 bench.py "
          f"--steps {steps} --warmup {warmup} --no-extras --no-cpu-baseline --tune-file <T>",
          f"    python3 bench.py --ingest-trace <D> --steps {steps} --warmup {warmup}", "",
          f"Rows average the {warmup} warm-up + {steps} timed steps' dispatches (after the bench's "
          "16-byte marker copy, i.e. after the autotuner's candidates) and the stream copies; "
          "durations in microseconds.", "",
          "| kernel | dispatches | avg us | min us | max us | arch VGPR | acc VGPR | grid threads |",
          "|---|---|---|---|---|---|---|---|"]
    for avg, k, n, mn, mx, vg, ag, gr in rows:
        md.append(f"| `{k.split('(')[0]}` | {n} | {avg:.1f} | {mn:.1f} | {mx:.1f} | {vg} | {ag} | {gr} |")
    md += ["", "Template parameters: ssam2d_kernel<storage, radius, shape (0 star, 1 box), columns per "
           "lane V, fused steps T, prefetch rows PF, ring form, ticks per block U, min waves/EU>."]
    (ROOT / "profiles" / "suite_8192_f32_kernel_trace.md").write_text("\n".join(md) + "\n")
    print("\n".join(md))


def ingest_pmc(fetch_dir: str, write_dir: str) -> None:
    """Per-launch HBM traffic of each suite kernel from rocprofv3 counter CSVs.

    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what they read for
    the 1 GiB stream copy of the same run (1 GiB read, 1 GiB written), so the
    figures are bytes.  Only the dispatches after the bench's 16-byte marker copy
    count (the autotuner's trial launches come before it)."""
    import csv
    import re

    def load(d):
        f = next(pathlib.Path(d).rglob("*counter_collection.csv"))
        rows = sorted(csv.DictReader(open(f)), key=lambda r: int(r["Dispatch_Id"]))
        # dispatches after the 16-byte marker copy are the timed suite launches
        mark = next((int(r["Dispatch_Id"]) for r in rows
                     if "stream_copy" in r["Kernel_Name"] and int(r["Grid_Size"]) <= 256), 0)
        copies, timed = [], {}
        for r in rows:
            if "stream_copy" in r["Kernel_Name"] and int(r["Grid_Size"]) > 256:
                copies.append(float(r["Counter_Value"]))
            elif int(r["Dispatch_Id"]) > mark and "ssam2d_" in r["Kernel_Name"]:
                timed.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        return copies, timed

    (fc, fetch), (wc, write) = load(fetch_dir), load(write_dir)
    gib_kib = float(1 << 20)
    fs = gib_kib / (sum(fc) / len(fc))
    ws = gib_kib / (sum(wc) / len(wc))
    tsuite_of = {n: t for n, _, _, t in SUITE}
    alg = 2 * 4 * N * N
    rows = []
    for k in fetch:
        sk = suite_kernel(k)
        if not sk or k not in write:
            continue
        name, t = sk
        tsuite = tsuite_of[name]
        fb = sum(fetch[k]) / len(fetch[k]) * fs * 1024
        wb = sum(write[k]) / len(write[k]) * ws * 1024
        rec = dict(kernel=k.split("(")[0], steps_per_launch=t, launches_per_step=tsuite // t,
                   fetch_bytes=round(fb), write_bytes=round(wb),
                   hbm_bytes_per_launch=round(fb + wb), algorithmic_bytes=alg,
                   traffic_over_algorithmic=round((fb + wb) / alg, 4),
                   scale_fetch=round(fs, 4), scale_write=round(ws, 4))
        (ROOT / "profiles" / f"traffic_{name}_8192_f32.json").write_text(json.dumps(rec, indent=1) + "\n")
        rows.append((name, rec))
    md = ["# HBM traffic per launch, FP32 suite kernels on 8192x8192 (MI355X)", "",
          "Commands (two passes, counters in separate runs; T holds the tuner's choices, saved by "
          "`python3 bench.py --no-extras --no-cpu-baseline --tune-file <T>`):", "",
          "    rocprofv3 --pmc FETCH_SIZE --output-format csv -d <F> -o run -- python3 bench.py --steps 1 --warmup 0 --no-extras --no-cpu-baseline --tune-file <T>",
          "    rocprofv3 --pmc WRITE_SIZE --output-format csv -d <W> -o run -- python3 bench.py --steps 1 --warmup 0 --no-extras --no-cpu-baseline --tune-file <T>",
          "    python3 bench.py --ingest-pmc <F> <W>", "",
          f"Calibration: the 1 GiB stream copy reads FETCH_SIZE x{fs:.3f} and WRITE_SIZE x{ws:.3f} "
          "of its true bytes (KiB per dispatch); every figure below is scaled by these.", "",
          "| kernel | steps per launch | fetch MiB | write MiB | HBM / algorithmic (per launch) |",
          "|---|---|---|---|---|"]
    for name, rec in sorted(rows):
        md.append(f"| {name} | {rec['steps_per_launch']} | {rec['fetch_bytes'] / 2**20:.1f} | "
                  f"{rec['write_bytes'] / 2**20:.1f} | {rec['traffic_over_algorithmic']:.3f} |")
    (ROOT / "profiles" / "traffic_suite_8192_f32.md").write_text("\n".join(md) + "\n")
    print("\n".join(md))


def roof(flops: float, nbytes: float, ms: float, peak_tflops: float) -> dict:
    """The binding roof of one kernel call: HBM (algorithmic bytes at 8 TB/s) or
    compute (algorithmic flops at the published peak), whichever takes longer;
    frac = that time / the measured time."""
    t_hbm = nbytes / (HBM_PEAK_GBPS * 1e9)
    t_cmp = flops / (peak_tflops * 1e12)
    bound = "mfma" if t_cmp > t_hbm else "hbm"
    t_roof = max(t_hbm, t_cmp)
    achieved = (flops / (ms * 1e-3) / 1e12) if bound == "mfma" else (nbytes / (ms * 1e-3) / 1e9)
    return dict(bound=bound, flop_per_byte=round(flops / nbytes, 2),
                ridge=round(peak_tflops * 1e12 / (HBM_PEAK_GBPS * 1e9), 2),
                achieved=round(achieved, 1), unit="TFLOP/s" if bound == "mfma" else "GB/s",
                peak=peak_tflops if bound == "mfma" else HBM_PEAK_GBPS,
                frac=round(t_roof / (ms * 1e-3), 4))


def band_errors(got, seed: int, a: float, b: float, f, steps: int = 1, bands=BANDS) -> dict:
    """Error of a device result against the FP64 oracle on 16-row bands: the
    input rows each band needs are regenerated on the host (the same splitmix64
    stream as the device fill) and run through the oracle in FP64; max and RMS
    absolute error, and max |error| / max |gold| over the bands."""
    import numpy as np

    from oracle import oracle

    n_rows, n_cols = got.shape
    rt = (f.shape[0] // 2) * steps
    mx, sq, cnt, gmax = 0.0, 0.0, 0, 0.0
    for y0 in bands:
        y0 = min(y0, n_rows - 16)
        lo, hi = max(0, y0 - rt), min(n_rows, y0 + 16 + rt)
        x = oracle.uniform(seed, (hi - lo) * n_cols, a, b, np.float64, offset=lo * n_cols).reshape(hi - lo, n_cols)
        want = oracle.conv2d_iter(x, f, steps) if steps > 1 else oracle.conv2d(x, np.asarray(f, np.float64))
        a0 = y0 - lo
        if lo > 0 and a0 < rt:
            continue  # the slab's artificial top edge would reach the band
        w = want[a0:a0 + 16]
        g = got[y0:y0 + 16].double().cpu().numpy()
        d = np.abs(g - w)
        mx = max(mx, float(d.max()))
        sq += float((d * d).sum())
        cnt += d.size
        gmax = max(gmax, float(np.abs(w).max()))
    return dict(max_abs_err=mx, rms_err=(sq / max(cnt, 1)) ** 0.5, max_rel_err_normwise=mx / max(gmax, 1e-300),
                checked_points=cnt, against="FP64 oracle on row bands " + ",".join(map(str, bands)))


def main():
    a = parse()
    if a.ingest_pmc:
        ingest_pmc(*a.ingest_pmc)
        return
    if a.ingest_trace:
        ingest_trace(a.ingest_trace, a.steps, a.warmup)
        return
    if a.ingest_stats:
        ingest_stats(*a.ingest_stats)
        return
    if a.ingest_sq:
        ingest_sq(a.ingest_sq[0], a.ingest_sq[1], a.ingest_sq[2:])
        return
    if a.ingest_3d_traffic:
        ingest_3d_traffic(*a.ingest_3d_traffic)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the SSAM stencil path runs on a single MI355X here "
              "(the multi-GPU slab decomposition is a design, not run)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from ssam_amd import _lib as L
    from ssam_amd import api
    from ssam_amd import workloads as wl

    L.lib()  # fail loudly if libssam.so is missing
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    if a.side:
        copy_gbps = 5200.0  # nominal for the fractions; the full bench measures it
        fn = {"config0": lambda: side_config0(stream, copy_gbps),
              "config2": lambda: side_3d(stream, copy_gbps),
              "config3": lambda: side_conv(stream, copy_gbps, tune=not a.no_tune),
              "config4": lambda: side_batch(stream)}[a.side]
        print(json.dumps({a.side: fn()}))
        return
    suite = [s for s in SUITE if not a.only or s[0] in a.only.split(",")]

    def taps(shape, r):
        return wl.normalised_taps(shape, r, seed=1000 + 10 * r + (shape == "star"))

    # the host-side taps of every kernel, generated once (like the grids, inputs
    # prepared before the timed region)
    tapset = [taps(shape, r) for (_, shape, r, _) in suite]

    def run_suite(src, dst, events=None):
        for i, (name, shape, r, t) in enumerate(suite):
            if events is not None:
                events[i][0].record(stream)
            api.stencil2d(src, dst, tapset[i], shape=L.STAR if shape == "star" else L.BOX,
                          steps=t, work=work, stream=stream)
            if events is not None:
                events[i][1].record(stream)

    # ---------------------------------------------------------------- inputs in HBM
    src = torch.empty((N, N), dtype=torch.float32, device=dev)
    api.fill_uniform(src, 0, -1.0, 1.0)
    dst = torch.empty_like(src)
    work = torch.empty_like(src)  # fused steps split over launches ping-pong through it
    torch.cuda.synchronize()

    # autotune each kernel's tiling and temporal-blocking degree on these buffers
    # (before the timed region; the library caches the choice per request shape)
    tuning = {}
    tune_path = pathlib.Path(a.tune_file) if a.tune_file else None
    if tune_path is not None and tune_path.exists():
        api.tune_load(str(tune_path))
        tuning = {k: tuple(v) for k, v in json.loads(tune_path.with_suffix(".json").read_text()).items()}
    elif not a.no_tune:
        for name, shape, r, t in suite:
            tuning[name] = api.tune2d(src, dst, taps(shape, r),
                                      shape=L.STAR if shape == "star" else L.BOX, steps=t, work=work)
        if tune_path is not None:
            api.tune_save(str(tune_path))
            tune_path.with_suffix(".json").write_text(json.dumps(tuning))
    torch.cuda.synchronize()
    # a 16-byte copy marks the start of the warm-up + timed launches in a kernel
    # trace (bench.py --ingest-trace keeps the dispatches after it)
    mark = torch.zeros(8, dtype=torch.float32, device=dev)
    api.stream_copy(mark[:4], mark[4:])
    for _ in range(a.warmup):
        run_suite(src, dst)
    torch.cuda.synchronize()

    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in suite] for _ in range(a.steps)]
    t_start = torch.cuda.Event(enable_timing=True)
    t_stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_start.record(stream)
    for k in range(a.steps):
        run_suite(src, dst, ev[k])
    t_stop.record(stream)
    torch.cuda.synchronize()
    total_ms = t_start.elapsed_time(t_stop)
    ms_per_step = total_ms / a.steps

    pts = N * N
    per = []
    for i, (name, shape, r, t) in enumerate(suite):
        ms = float(np.mean([ev[k][i][0].elapsed_time(ev[k][i][1]) for k in range(a.steps)]))
        upd = pts * t
        tw, tc = tuning.get(name, (0, 0))
        # launches actually issued: the steps split into fused chunks of the tuned size
        launches = (t + tc - 1) // tc if t > 1 and tc else 1
        byts = 2 * 4 * pts * launches  # each point read once and written once per launch
        taps_n = (4 * r + 1) if shape == "star" else (2 * r + 1) ** 2
        flops = upd * (2 * taps_n - 1)
        per.append(dict(name=name, ms=round(ms, 4), gpts=round(upd / ms / 1e6, 2),
                        launches=launches, gbps=round(byts / ms / 1e6, 1),
                        hbm_frac=round(byts / ms / 1e6 / HBM_PEAK_GBPS, 3),
                        tflops=round(flops / ms / 1e9, 2),
                        roofline=roof(flops, byts, ms, FP32_PEAK_TFLOPS),
                        tuned_waves_per_simd=tw, tuned_steps_per_launch=tc))
    updates_per_step = sum(pts * t for (_, _, _, t) in suite)
    value = updates_per_step / (ms_per_step * 1e-3) / 1e9

    # the dominant kernel: the largest share of the step's time
    dom = max(per, key=lambda p: p["ms"])
    di = [p["name"] for p in per].index(dom["name"])
    d_shape, d_r, d_t = suite[di][1], suite[di][2], suite[di][3]
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_{dom['name']}_8192_f32.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    dr = dom["roofline"]
    roofline = dict(bound=dr["bound"], kernel=f"{dom['name']} f32 8192^2", achieved=dr["achieved"],
                    peak=dr["peak"], unit=dr["unit"], frac=dr["frac"], traffic=traffic,
                    algorithmic_bytes=2 * 4 * pts, algorithmic_flops=int(
                        pts * d_t * (2 * ((4 * d_r + 1) if d_shape == "star" else (2 * d_r + 1) ** 2) - 1)),
                    launches=dom["launches"], ms=dom["ms"],
                    note="frac = time at the binding roof (HBM 8 TB/s or FP32 157.3 TFLOP/s, whichever "
                         "is longer for the algorithmic bytes / flops) over the measured time")
    # FP32 suite errors against the FP64 oracle (bands of the 8192^2 output)
    for i, (name, shape, r, t) in enumerate(suite):
        api.stencil2d(src, dst, taps(shape, r), shape=L.STAR if shape == "star" else L.BOX,
                      steps=t, work=work, stream=stream)
        torch.cuda.synchronize()
        per[i]["error_vs_fp64_oracle"] = band_errors(dst, 0, -1.0, 1.0, taps(shape, r), t)

    # ---------------------------------------------------------------- stream bandwidth
    cp_a = torch.empty(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
    cp_b = torch.empty_like(cp_a)
    api.fill_uniform(cp_a, 1)
    for _ in range(2):
        api.stream_copy(cp_a, cp_b)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(5):
        api.stream_copy(cp_a, cp_b)
    e1.record(stream)
    torch.cuda.synchronize()
    copy_gbps = 2 * cp_a.numel() * 4 * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del cp_a, cp_b
    for p in per:
        p["stream_frac"] = round(p["gbps"] / copy_gbps, 3)

    extras = {}
    if not a.no_extras:
        extras = side_runs(suite, taps, stream, per, tune=not a.no_tune)
        extras["config0_star5_hold"] = side_config0(stream, copy_gbps)
        extras["config2_3d"] = side_3d(stream, copy_gbps)
        extras["config3_conv"] = side_conv(stream, copy_gbps, tune=not a.no_tune)
        extras["config4_batch"] = side_batch(stream)
        extras["microbench_cdna4"] = {k: round(v, 3) for k, v in api.microbench(stream).items()}

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(suite, taps)

    line = {
        "metric": "Gpoints/s + % HBM roofline (1 GPU); BF16-storage vs FP32 plain at stated err",
        "value": round(value, 2),
        "unit": "Gpoints/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: U[-1,1) grid from splitmix64 seed 0, taps U[0,1) normalised to 1",
        "config": {
            "workload": "2D SSAM suite on 8192x8192 FP32: box r1-5 (9/25/49/81/121 pt), "
                        "star r3-5 (13/17/21 pt), temporal 5-pt and 9-pt star T=8; "
                        "one sweep of each per step",
            "grid": [N, N],
            "kernels": [s[0] for s in suite],
            "point_updates_per_step": updates_per_step,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stream_copy_gbps": round(copy_gbps, 1),
        "per_kernel": per,
        **extras,
        "device": torch.cuda.get_device_name(0),
    }
    print(json.dumps(line))


def side_runs(suite, taps, stream, per_f32, tune=True):
    """FP64 and BF16/FP16-storage versions of the suite (same input values), each
    kernel autotuned (tiling, fused steps per launch, VALU or matrix-core form) and
    timed with events; errors against the FP64 oracle on row bands."""
    import numpy as np
    import torch

    from ssam_amd import _lib as L
    from ssam_amd import api

    out = {}
    n = N
    x64 = torch.empty((n, n), dtype=torch.float64, device="cuda")
    api.fill_uniform(x64, 0, -1.0, 1.0)
    for label, dt in (("fp64", torch.float64), ("bf16", torch.bfloat16), ("fp16", torch.float16)):
        src = x64.to(dt)
        dst = torch.empty_like(src)
        # runs over more than one launch (FP64 radius 2 fuses 4 steps; the tuner may
        # split BF16 / FP16 runs too) chain through work: the storage type for FP64,
        # two FP32 grids for 2-byte storage, whose running state stays FP32
        work = torch.empty_like(src) if dt == torch.float64 else \
            torch.empty(2 * n * n, dtype=torch.float32, device="cuda")
        kern = []
        for i, (name, shape, r, t) in enumerate(suite):
            f = taps(shape, r)
            sh = L.STAR if shape == "star" else L.BOX
            if tune:
                api.tune2d(src, dst, f, shape=sh, steps=t, work=work, stream=stream)
            api.stencil2d(src, dst, f, shape=sh, steps=t, work=work, stream=stream)  # warm
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(3):
                api.stencil2d(src, dst, f, shape=sh, steps=t, work=work, stream=stream)
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / 3
            entry = dict(name=name, ms=round(ms, 4), gpts=round(n * n * t / ms / 1e6, 2),
                         speedup_vs_f32=round(per_f32[i]["ms"] / ms, 3))
            entry.update(band_errors(dst, 0, -1.0, 1.0, f, t))
            kern.append(entry)
        upd = sum(n * n * t for (_, _, _, t) in suite)
        tot = sum(k["ms"] for k in kern)
        out[label] = dict(value_gpts=round(upd / tot / 1e6, 2), ms_per_step=round(tot, 4),
                          kernels=kern)
        del src, dst, work
    return out


def _time(fn, stream, reps=3, warm=1):
    import torch
    for _ in range(warm):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def side_config0(stream, copy_gbps):
    """Config 0: one sweep of the 5-point star (centre 0.5, N/S/E/W 0.125) on 2048^2
    FP32, values U[0,1) from seed 0, the 1-cell border held fixed (HOLD).  The
    32 MiB working set fits the 256 MB last-level cache, so it is timed two ways:
    cold -- a 1 GiB copy between repetitions evicts it, each sweep timed alone --
    and hot -- 20 back-to-back sweeps replayed from a graph.  The error is against
    the FP64 oracle on every point."""
    import numpy as np
    import torch

    from oracle import oracle
    from ssam_amd import _lib as L
    from ssam_amd import api

    n = 2048
    x = torch.empty((n, n), dtype=torch.float32, device="cuda")
    api.fill_uniform(x, 0, 0.0, 1.0)
    y = torch.empty_like(x)
    f = np.array([[0.0, 0.125, 0.0], [0.125, 0.5, 0.125], [0.0, 0.125, 0.0]])
    api.tune2d(x, y, f, shape=L.STAR, boundary=L.HOLD, stream=stream)
    run = lambda: api.stencil2d(x, y, f, shape=L.STAR, boundary=L.HOLD, stream=stream)  # noqa: E731
    # cold: evict the LLC before every timed sweep
    fa = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    fb = torch.empty_like(fa)
    cold = []
    for _ in range(12):
        api.stream_copy(fa, fb, stream=stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        run()
        e1.record(stream)
        torch.cuda.synchronize()
        cold.append(e0.elapsed_time(e1))
    del fa, fb
    ms_cold = float(np.median(cold[2:]))
    # hot: back-to-back sweeps from one graph
    run()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        for _ in range(20):
            api.stencil2d(x, y, f, shape=L.STAR, boundary=L.HOLD)
    g.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(5):
        g.replay()
    e1.record(stream)
    torch.cuda.synchronize()
    ms_hot = e0.elapsed_time(e1) / 100
    run()
    torch.cuda.synchronize()
    xs = oracle.uniform(0, n * n, 0.0, 1.0, np.float64).reshape(n, n)
    want = oracle.conv2d(xs, f)
    want[0, :], want[-1, :], want[:, 0], want[:, -1] = xs[0, :], xs[-1, :], xs[:, 0], xs[:, -1]
    d = np.abs(y.double().cpu().numpy() - want)
    out = {}
    for label, ms in (("cold", ms_cold), ("hot", ms_hot)):
        gbps = 8 * n * n / ms / 1e6
        out[label] = dict(ms=round(ms, 4), gpts=round(n * n / ms / 1e6, 2), gbps=round(gbps, 1),
                          hbm_frac=round(gbps / HBM_PEAK_GBPS, 3), stream_frac=round(gbps / copy_gbps, 3))
    out["cold"]["note"] = "LLC evicted by a 1 GiB copy before each sweep; median of 10"
    out["hot"]["note"] = "20 back-to-back sweeps per graph replay (the 32 MiB set stays in the LLC)"
    out.update(max_abs_err=float(d.max()), rms_err=float(np.sqrt((d * d).mean())),
               max_rel_err_normwise=float(d.max() / np.abs(want).max()), against="FP64 oracle, every point")
    return out


def side_3d(stream, copy_gbps):
    """512^3: 7-pt and 13-pt star sweeps (reference coefficients), the reference
    19-pt 'poisson', 10 sweeps of the 7-pt star, and 100 Jacobi sweeps of the
    19-point Laplacian with the FP64 residual of the last sweep -- FP32 and FP64,
    and BF16 / FP16 storage (FP32 accumulation, FP32 running state between
    launches in a work area of two FP32 grids)."""
    import torch

    from ssam_amd import _lib as L
    from ssam_amd import api
    from ssam_amd import workloads as wl

    n = 512
    out = {}
    for label, dt, es in (("fp32", torch.float32, 4), ("fp64", torch.float64, 8),
                          ("bf16", torch.bfloat16, 2), ("fp16", torch.float16, 2)):
        if es == 2:
            # the FP32 grid's values, rounded to the storage type
            src = torch.empty((n, n, n), dtype=torch.float32, device="cuda")
            api.fill_uniform(src, 2, 0.0, 1.0)
            src = src.to(dt)
            work = torch.empty(2 * n ** 3, dtype=torch.float32, device="cuda")
        else:
            src = torch.empty((n, n, n), dtype=dt, device="cuda")
            api.fill_uniform(src, 2, 0.0, 1.0)
            work = torch.empty_like(src)
        dst = torch.empty_like(src)
        kern = []
        for name, shape, r, coef in (("7pt", L.STAR, 1, wl.REF3D_7PT), ("13pt", L.STAR, 2, wl.REF3D_13PT),
                                     ("poisson19", L.FACE_EDGE, 1, wl.REF3D_POISSON)):
            ms = _time(lambda: api.stencil3d(src, dst, coef, shape=shape, radius=r, stream=stream), stream)
            gbps = 2 * es * n ** 3 / ms / 1e6
            ent = dict(name=name, ms=round(ms, 4), gpts=round(n ** 3 / ms / 1e6, 2),
                       gbps=round(gbps, 1), stream_frac=round(gbps / copy_gbps, 3),
                       hbm_frac=round(gbps / HBM_PEAK_GBPS, 3))
            ent.update(slab_errors3d(dst, 2, n, wl.taps3d("face_edge" if shape == L.FACE_EDGE else "star", r),
                                     coef, r, dt))
            kern.append(ent)
        # multi-sweep runs launch three sweeps at a time (3+3+2+2: one HBM read +
        # write per launch): per-sweep rates above the single-sweep bandwidth
        # bound; effective_gbps counts the bytes of single sweeps.  Three warm-up
        # runs: the first run after the single sweeps above measured 10-15% slower
        # than the ones after it.
        ms = _time(lambda: api.stencil3d(src, dst, wl.REF3D_7PT, shape=L.STAR, radius=1, steps=10, work=work,
                                         stream=stream), stream, reps=5, warm=3) / 10
        kern.append(dict(name="7pt_x10", ms_per_step=round(ms, 4), gpts=round(n ** 3 / ms / 1e6, 2),
                         effective_gbps=round(2 * es * n ** 3 / ms / 1e6, 1), launches=[3, 3, 2, 2]))
        steps = 100
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        res = api.jacobi19(src, dst, steps, work, residual=True, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        gbps = 2 * es * n ** 3 / ms / 1e6
        kern.append(dict(name="jacobi19_x100", ms_per_step=round(ms, 4),
                         gpts=round(n ** 3 / ms / 1e6, 2), effective_gbps=round(gbps, 1),
                         sweeps_per_launch=2, residual_l2_fp64=res))  # 19 points: two per launch
        if es == 2:
            for a, b in zip(kern, out["fp32"]):
                ka = "ms" if "ms" in a else "ms_per_step"
                a["speedup_vs_fp32"] = round(b[ka] / a[ka], 3)
        out[label] = kern
        del src, dst, work
        torch.cuda.empty_cache()
    return out


def slab_errors3d(got, seed: int, n: int, taps, coef, r: int, dt) -> dict:
    """Error of a one-sweep n^3 result against the oracle (interior-only gold, zero
    boundary) on three 8-plane z-slabs; the input planes are regenerated on the
    host from the device fill's stream."""
    import numpy as np
    import torch

    from oracle import oracle

    npdt = np.float32 if dt == torch.float32 else np.float64
    mx, sq, cnt, gmax = 0.0, 0.0, 0, 0.0
    for z0 in (0, n // 2 - 4, n - 8):
        lo, hi = max(0, z0 - r), min(n, z0 + 8 + r)
        x = oracle.uniform(seed, (hi - lo) * n * n, 0.0, 1.0, npdt, offset=lo * n * n).reshape(hi - lo, n, n)
        want = oracle.stencil3d_taps(x.astype(np.float64), taps, coef, r)
        a = z0 - lo
        w = want[a:a + 8]
        g = got[z0:z0 + 8].double().cpu().numpy()  # planes a..a+7 sit r planes inside the slab
        d = np.abs(g - w)
        mx = max(mx, float(d.max()))
        sq += float((d * d).sum())
        cnt += d.size
        gmax = max(gmax, float(np.abs(w).max()))
    return dict(max_abs_err=mx, rms_err=(sq / cnt) ** 0.5, max_rel_err_normwise=mx / gmax,
                against="FP64 oracle on z-slabs 0-7, n/2-4..n/2+3, n-8..n-1")


def side_conv(stream, copy_gbps, tune=True):
    """8192^2 image U[0,1), K x K taps U[-1,1)/K^2: FP32 plain vs BF16 / FP16 storage
    (FP32 accumulation); errors against the FP64 oracle on row bands."""
    import torch

    from ssam_amd import _lib as L
    from ssam_amd import api
    from ssam_amd import workloads as wl

    n = N
    x64 = torch.empty((n, n), dtype=torch.float64, device="cuda")
    api.fill_uniform(x64, 3, 0.0, 1.0)
    rows = []
    for k in (3, 5, 7, 9, 15, 19):
        f = wl.conv_taps(k, seed=50 + k)
        row = dict(K=k)
        for label, dt in (("fp64", torch.float64), ("fp32", torch.float32), ("bf16", torch.bfloat16),
                          ("fp16", torch.float16)):
            src = x64 if dt == torch.float64 else x64.to(dt)
            dst = torch.empty_like(src)
            if tune:
                api.tune2d(src, dst, f, shape=L.DENSE, stream=stream)
            ms = _time(lambda: api.stencil2d(src, dst, f, shape=L.DENSE, stream=stream), stream)
            flops = n * n * (2 * k * k - 1)
            ent = dict(ms=round(ms, 4), gpts=round(n * n / ms / 1e6, 2),
                       tflops=round(flops / ms / 1e9, 2))
            if dt in (torch.float32, torch.float64):
                ent["roofline"] = roof(flops, 2 * src.element_size() * n * n, ms,
                                       FP32_PEAK_TFLOPS if dt == torch.float32 else FP64_PEAK_TFLOPS)
            api.stencil2d(src, dst, f, shape=L.DENSE, stream=stream)
            torch.cuda.synchronize()
            ent.update(band_errors(dst, 3, 0.0, 1.0, f))
            row[label] = ent
            del dst
        row["bf16_speedup_vs_fp32"] = round(row["fp32"]["ms"] / row["bf16"]["ms"], 3)
        row["fp16_speedup_vs_fp32"] = round(row["fp32"]["ms"] / row["fp16"]["ms"], 3)
        rows.append(row)
    return rows


def side_batch(stream):
    """The uneven batch: 192 2D grids (5-pt star / 5x5 / 11x11
    box) plus 64 3D grids x 10 sweeps of the 7-point star, replayed as one
    hipGraph.  Efficiency = the time the same point counts would take at
    single-large-grid throughput (per 2D shape a 256 MiB grid: FP32 8192^2, BF16
    16384 x 8192, with BF16 8192^2 reported beside it; 512^3 x 10 sweeps of the
    7-point star through the same single-grid entry point, which launches two
    sweeps at a time like the batch; both measured here) / the batch time."""
    import torch

    from ssam_amd import _lib as L
    from ssam_amd import api
    from ssam_amd import workloads as wl

    g2, g3 = wl.config4_batch(seed=4)
    shapes = {("star", 1): L.STAR, ("box", 2): L.BOX, ("box", 5): L.BOX}
    out = {}
    # single-large-grid reference throughputs (Gpoints/s), on grids of 256 MiB per
    # buffer for both storage types: FP32 8192^2 and BF16 16384 x 8192.  A BF16
    # 8192^2 pair (256 MiB in + out) is partly served from the 256 MB last-level
    # cache when timed back to back, which the batch's 1.2 GiB of BF16 2D grids
    # are not; its rate is reported too ("bf16_8192sq").
    ref = {}
    for dt, lab, rows in ((torch.float32, "fp32", N), (torch.bfloat16, "bf16", 2 * N),
                          (torch.bfloat16, "bf16_8192sq", N)):
        big = torch.empty((rows, N), dtype=torch.float32, device="cuda")
        api.fill_uniform(big, 5, -1.0, 1.0)
        big = big.to(dt)
        bd = torch.empty_like(big)
        for (sh, r), code in shapes.items():
            f = wl.normalised_taps(sh, r, 60 + r)
            # the best single-grid rate: autotuned like the batch
            api.tune2d(big, bd, f, shape=code, stream=stream)
            ms = _time(lambda: api.stencil2d(big, bd, f, shape=code, stream=stream), stream)
            ref[(lab, sh, r)] = rows * N / ms / 1e6
        del big, bd
        torch.cuda.empty_cache()
    n3 = 512
    ref3 = {}
    for dt, lab in ((torch.float32, "fp32"), (torch.bfloat16, "bf16")):
        b3 = torch.empty((n3, n3, n3), dtype=torch.float32, device="cuda")
        api.fill_uniform(b3, 6, 0.0, 1.0)
        b3 = b3.to(dt)
        d3 = torch.empty_like(b3)
        # (BF16: the running state between launches stays FP32, two FP32 grids)
        w3 = torch.empty_like(b3) if dt == torch.float32 else torch.empty(2 * n3 ** 3, dtype=torch.float32,
                                                                          device="cuda")
        ms = _time(lambda: api.stencil3d(b3, d3, wl.REF3D_7PT, shape=L.STAR, radius=1, steps=10, work=w3,
                                         stream=stream), stream, reps=3)
        ref3[lab] = n3 ** 3 * 10 / ms / 1e6
        del b3, d3, w3
        torch.cuda.empty_cache()
    for dt, lab in ((torch.float32, "fp32"), (torch.bfloat16, "bf16")):
        items, keep = [], []
        exp_ms = 0.0
        pts2 = pts3 = 0
        for k, g in enumerate(g2):
            src = torch.empty((g["h"], g["w"]), dtype=torch.float32, device="cuda")
            api.fill_uniform(src, 1000 + k, -1.0, 1.0)
            src = src.to(dt)
            dst = torch.empty_like(src)
            keep += [src, dst]
            items.append(dict(ndim=2, src=src, dst=dst, shape=shapes[(g["shape"], g["r"])],
                              radius=g["r"], coeffs=wl.normalised_taps(g["shape"], g["r"], 60 + g["r"])))
            pts2 += g["w"] * g["h"]
            exp_ms += g["w"] * g["h"] / (ref[(lab, g["shape"], g["r"])] * 1e6)
        for k, g in enumerate(g3):
            src = torch.empty((g["nz"], g["ny"], g["nx"]), dtype=torch.float32, device="cuda")
            api.fill_uniform(src, 5000 + k, 0.0, 1.0)
            src = src.to(dt)
            dst = torch.empty_like(src)
            # BF16 items keep their state between passes in FP32 (two FP32 grids)
            work = torch.empty_like(src) if dt == torch.float32 else \
                torch.empty(2 * src.numel(), dtype=torch.float32, device="cuda")
            keep += [src, dst, work]
            items.append(dict(ndim=3, src=src, dst=dst, work=work, shape=L.STAR, radius=1,
                              steps=g["steps"], coeffs=wl.REF3D_7PT))
            n = g["nx"] * g["ny"] * g["nz"] * g["steps"]
            pts3 += n
            exp_ms += n / (ref3[lab] * 1e6)
        # the 2D and 3D halves alone, to attribute the time
        parts = {}
        for part, sel in (("2d_only", lambda it: it["ndim"] == 2), ("3d_only", lambda it: it["ndim"] == 3)):
            pp = api.BatchPlan([it for it in items if sel(it)])
            pp.tune(stream=stream)
            pp.run(stream=stream)
            torch.cuda.synchronize()
            parts[part] = round(_time(lambda: pp.run(stream=stream), stream, reps=3), 4)
            pp.close()
        exp2 = sum(g["w"] * g["h"] / (ref[(lab, g["shape"], g["r"])] * 1e6) for g in g2)
        extra = {}
        if lab == "bf16":
            exp2_sq = sum(g["w"] * g["h"] / (ref[("bf16_8192sq", g["shape"], g["r"])] * 1e6) for g in g2)
            extra = dict(single_grid_equivalent_2d_ms_8192sq=round(exp2_sq, 4),
                         single_grid_equivalent_ms_8192sq=round(exp_ms - exp2 + exp2_sq, 4))
        plan = api.BatchPlan(items)
        tiling = plan.tune(stream=stream)  # tiles chosen by the autotuner
        t2, t3, pu = plan.info()
        plan.run(stream=stream)  # captures the graph
        torch.cuda.synchronize()
        ms = _time(lambda: plan.run(stream=stream), stream, reps=5)
        ms_eager = _time(lambda: plan.run(eager=True, stream=stream), stream, reps=3)
        plan.run(stream=stream)
        torch.cuda.synchronize()
        errs = batch_errors(g2, g3, items)
        out[lab] = dict(ms=round(ms, 4), ms_eager=round(ms_eager, 4), point_updates=pu,
                        gpts=round(pu / ms / 1e6, 2), tiles_2d=t2, tiles_3d=t3,
                        points_2d=pts2, point_updates_3d=pts3,
                        single_grid_equivalent_ms=round(exp_ms, 4),
                        efficiency_vs_single_grid=round(exp_ms / ms, 3),
                        part_ms=parts, single_grid_equivalent_2d_ms=round(exp2, 4),
                        single_grid_equivalent_3d_ms=round(exp_ms - exp2, 4),
                        storage_2d=lab, storage_3d=lab, tiling=tiling, error_vs_fp64_oracle=errs)
        if extra:
            extra["efficiency_vs_single_grid_8192sq"] = round(extra["single_grid_equivalent_ms_8192sq"] / ms, 3)
            out[lab].update(extra)
        plan.close()
        del keep, items
        torch.cuda.empty_cache()
    out["single_grid_gpts"] = {f"{k[0]}_{k[1]}{k[2]}": round(v, 1) for k, v in ref.items()}
    for lab, v in ref3.items():
        out["single_grid_gpts"][f"{lab}_3d7pt_x10"] = round(v, 1)
    return out


def batch_errors(g2, g3, items) -> dict:
    """Max error of sampled batch items against the FP64 oracle: the six smallest
    2D grids of each shape (every point) and the four smallest 3D grids (every
    point, all 10 sweeps)."""
    import numpy as np

    from oracle import oracle
    from ssam_amd import workloads as wl

    out = {}
    for shape, r in (("star", 1), ("box", 2), ("box", 5)):
        ks = sorted((g["w"] * g["h"], k) for k, g in enumerate(g2) if (g["shape"], g["r"]) == (shape, r))[:6]
        mx, gm = 0.0, 0.0
        for _, k in ks:
            g = g2[k]
            x = oracle.uniform(1000 + k, g["w"] * g["h"], -1.0, 1.0, np.float64).reshape(g["h"], g["w"])
            want = oracle.conv2d(x, wl.normalised_taps(shape, r, 60 + r))
            got = items[k]["dst"].double().cpu().numpy()
            mx, gm = max(mx, float(np.abs(got - want).max())), max(gm, float(np.abs(want).max()))
        out[f"{shape}{2 * r + 1}"] = dict(items=len(ks), max_abs_err=mx, max_rel_err_normwise=mx / gm)
    ks = sorted((g["nx"] * g["ny"] * g["nz"], k) for k, g in enumerate(g3))[:4]
    mx, gm = 0.0, 0.0
    for _, k in ks:
        g = g3[k]
        x = oracle.uniform(5000 + k, g["nx"] * g["ny"] * g["nz"], 0.0, 1.0, np.float32).reshape(
            g["nz"], g["ny"], g["nx"]).astype(np.float64)
        for _ in range(g["steps"]):
            x = oracle.stencil3d_taps(x, wl.taps3d("star", 1), wl.REF3D_7PT, 1)
        got = items[len(g2) + k]["dst"].double().cpu().numpy()
        mx, gm = max(mx, float(np.abs(got - x).max())), max(gm, float(np.abs(x).max()))
    out["3d_7pt_x10"] = dict(items=len(ks), max_abs_err=mx, max_rel_err_normwise=mx / gm)
    out["against"] = "FP64 oracle, every point of the sampled items"
    return out


def cpu_baseline(suite, taps):
    """The oracle (restated reference gold, OpenMP over rows) on a bounded slab of
    the same workload: 1024 rows x 8192 columns per kernel."""
    import numpy as np

    from oracle import oracle
    from ssam_amd import workloads as wl

    rows = 1024
    x = wl.uniform(0, rows * N, -1.0, 1.0).astype(np.float32).reshape(rows, N)
    upd, sec = 0, 0.0
    for name, shape, r, t in suite:
        f = taps(shape, r).astype(np.float32)
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            if t > 1:
                oracle.conv2d_iter(x, f, t)
            else:
                oracle.conv2d(x, f)
            times.append(time.perf_counter() - t0)
        sec += float(np.median(times))
        upd += rows * N * t
    return dict(value=round(upd / sec / 1e9, 4), unit="Gpoints/s", cores=oracle.max_threads(),
                kind="port",
                sample=f"same 10-kernel FP32 suite on a {rows}x{N} slab (one sweep each; "
                       "temporal kernels iterated 8x), median of 3")


if __name__ == "__main__":
    main()
