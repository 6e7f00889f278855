#!/usr/bin/env python3
"""Benchmark: converged RVE solves/s of the EffPlast2D hot path on one MI355X.

Workload (one "step" = one converged solve): the 1024x1024 elastoplastic RVE
with all three moduli -- nGrid=32, Lx=Ly=20, one pore, porosity 0.01,
Y = 2.5e-3*K0, K0=1, G0=0.5, rho0=1, CFL=0.5, preload initLoadValue=-3e-3
along loadType [1,1,0], then the 0.5% hydrostatic increment and a pure xy
shear increment (shear loadType [0,0,1]), nTimeSteps=1, eIter=1e-8,
nIter=100000.  Each step creates the solver (pore masks, buffers) and runs
ComputeEffModuli to convergence through the C ABI; inputs are generated on
the device, nothing crosses PCIe but scalars and sampled boundary values.

Printed: ONE JSON line with value = solves/s of the FP64 fused scheme (the
configuration's precision), the dominant kernel's live HIP-event timing
against the HBM roofline, the other schemes (plain four-kernel FP64 and the
FP32/FP64 split) as extra fields, and the CPU oracle's rate on this host.

Other entry points:
  --probe      a fixed 200-iteration run of one scheme plus a copy kernel of
               known size (for rocprofv3 --pmc counter runs)
  --summarize  turn rocprofv3 csv output into the short text summaries kept
               under profiles/
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = ("converged RVE solves/s + cell-iter/s (% HBM peak), FP32-split vs FP64; p50/p99")
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per point per iteration: U and V read once and written once
BYTES_PER_POINT = {"fp64": 64, "plain": 64, "split": 48}

WORKLOAD = dict(Lx=20.0, nGrid=32, nPores=1, porosity=0.01, Y=2.5e-3, K0=1.0, G0=0.5,
                rho0=1.0, CFL=0.5)
LOAD = dict(initLoad=-3e-3, loadType=[1.0, 1.0, 0.0], shearLoadType=[0.0, 0.0, 1.0])


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_pa(nx):
    from effplast2d_amd import derive_params
    w = dict(WORKLOAD)
    w["nX"] = nx
    return derive_params(**w)


def solve_once(pa, nx, mode, check, timing_every, nIter=100_000, eIter=1e-8):
    from effplast2d_amd import EffPlast2D
    t0 = time.perf_counter()
    with EffPlast2D(pa, nx, nTasks=3, nIter=nIter, eIter=eIter, checkInterval=check, mode=mode,
                    shearLoadType=LOAD["shearLoadType"]) as g:
        g.set_timing(timing_every)
        res = g.ComputeEffModuli(LOAD["initLoad"], LOAD["loadType"])
        ms, n, name = g.timing()
        recs = g.records()
    dt = time.perf_counter() - t0
    return {"seconds": dt, "res": res, "kernel_ms": ms, "kernel_samples": n, "kernel": name,
            "iters": [int(r["iters"]) for r in recs], "totalIters": int(res["totalIters"])}


def run_mode(pa, nx, mode, steps, warmup, check, timing_every):
    for _ in range(warmup):
        solve_once(pa, nx, mode, check, timing_every)
    runs = []
    t0 = time.perf_counter()
    for k in range(steps):
        r = solve_once(pa, nx, mode, check, timing_every)
        runs.append(r)
        log(f"[bench] {mode} step {k + 1}/{steps}: {r['seconds'] * 1e3:.1f} ms, "
            f"iters {r['iters']}, kernel {r['kernel_ms'] * 1e3:.2f} us")
    elapsed = time.perf_counter() - t0
    lat = sorted(r["seconds"] for r in runs)
    tot_iters = sum(r["totalIters"] for r in runs)
    ksum = sum(r["kernel_ms"] * r["kernel_samples"] for r in runs)
    kn = sum(r["kernel_samples"] for r in runs)
    kms = ksum / kn if kn else float("nan")
    points = (nx + 1) * nx
    bytes_launch = BYTES_PER_POINT[mode] * points
    achieved = bytes_launch / (kms * 1e-3) / 1e9 if kn else float("nan")
    res = runs[-1]["res"]
    return {
        "mode": mode,
        "solves_per_s": steps / elapsed,
        "elapsed_s": elapsed,
        "ms_per_solve": 1e3 * elapsed / steps,
        "p50_ms": 1e3 * statistics.median(lat),
        "p99_ms": 1e3 * lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))],
        "iters_per_solve": tot_iters / steps,
        "iters_by_task": runs[-1]["iters"],
        "cell_iter_per_s": tot_iters * nx * nx / elapsed,
        "kernel": runs[-1]["kernel"],
        "kernel_avg_us": kms * 1e3,
        "kernel_samples": kn,
        "bytes_per_launch": bytes_launch,
        "achieved_GBs": achieved,
        "moduli": {"Kphi": res["num"]["Kphi"], "Kd": res["num"]["Kd"], "Gxy": res["Gxy"],
                   "Kphi_an": res["an"]["Kphi"], "Kd_an": res["an"]["Kd"], "G_an": res["an"]["G"]},
        "converged": bool(res["converged"]),
    }


def read_traffic(mode, nx):
    """Measured HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated)."""
    path = os.path.join(ROOT, "profiles", f"pmc_fused_{mode}_{nx}.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1].split()[0])
    return None


def cpu_baseline(pa, nx, iters_per_solve, seconds):
    """The CPU oracle (four passes, FP64, OpenMP) on this host, timed on a
    bounded sample of the same workload; solves/s extrapolated with the
    GPU-measured iterations per solve."""
    import numpy as np

    from oracle import ora
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    s = ora.State(nX=nx, nY=nx, pa=np.asarray(pa), nl=3, nThreads=threads)
    s.apply_affine(-3e-3, -3e-3, 0.0, 0.0)
    s.iterate(2)
    done, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        s.iterate(10)
        done += 10
    dt = time.perf_counter() - t0
    rate = done * nx * nx / dt
    return {"value": rate / (iters_per_solve * nx * nx), "unit": "solves/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle 4-pass FP64 stencil, {nx}x{nx}, {done} iterations in {dt:.1f} s "
                      f"({rate / 1e6:.1f} M cell-iter/s); solve time extrapolated to the "
                      f"{iters_per_solve:.0f} iterations per solve measured on the GPU",
            "cell_iter_per_s": rate}


def probe(mode, nx, iters, copy_mib):
    """Fixed workload for counter runs: `iters` iterations of one scheme on the
    preloaded RVE, then 5 copies of copy_mib MiB."""
    import ctypes as C

    from effplast2d_amd import EffPlast2D
    from effplast2d_amd._lib import check, lib
    pa, _ = make_pa(nx)
    with EffPlast2D(pa, nx, nTasks=3, mode=mode) as g:
        g.apply_affine(-3e-3, -3e-3, 0.0, 0.0)
        g.iterate(iters)
        g.sync()
    ms = C.c_double()
    check(lib().ep2d_diag_copy(0, copy_mib << 20, 5, C.byref(ms)))
    log(f"[probe] {mode} {nx}: {iters} iterations; copy {copy_mib} MiB: {ms.value:.3f} ms "
        f"({2 * (copy_mib << 20) / ms.value / 1e6:.0f} GB/s)")


def sweep(sizes, modes, tys, iters):
    """Average fused-kernel duration per (scheme, tile height, grid size) on the
    preloaded RVE, every launch timed with dispatch events.  This is the data
    the default tile-height table is chosen from."""
    from effplast2d_amd import EffPlast2D
    rows = []
    for nx in sizes:
        pa, _ = make_pa(nx)
        for mode in modes:
            for ty in tys:
                with EffPlast2D(pa, nx, nTasks=3, mode=mode, tileRows=ty) as g:
                    g.apply_affine(-3e-3, -3e-3, 0.0, 0.0)
                    g.iterate(20)
                    g.set_timing(1)
                    g.iterate(iters)
                    ms, n, name = g.timing()
                b = BYTES_PER_POINT[mode] * (nx + 1) * nx
                row = {"nx": nx, "mode": mode, "ty": ty, "kernel_us": ms * 1e3,
                       "GBs": b / (ms * 1e-3) / 1e9, "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                       "kernel": name}
                rows.append(row)
                log(f"[sweep] nx={nx} {mode} TY={ty}: {ms * 1e3:.2f} us, "
                    f"{row['GBs']:.0f} GB/s ({100 * row['frac']:.1f}%) {name}")
    print(json.dumps({"sweep": rows}), flush=True)


def summarize(outdir, dest):
    """Condense rocprofv3 csv output in outdir into text lines."""
    lines = []
    for path in sorted(glob.glob(os.path.join(outdir, "**", "*kernel_stats.csv"), recursive=True)):
        lines.append(f"# kernel stats: {os.path.relpath(path, outdir)}")
        with open(path) as f:
            rows = list(csv.DictReader(f))
        for r in rows[:12]:
            lines.append(f"{r.get('Name', '?')[:90]:90s} calls={r.get('Calls')} "
                         f"avg_ns={float(r.get('AverageNs', 0)):.0f} "
                         f"pct={float(r.get('Percentage', 0)):.2f}")
    for path in sorted(glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"),
                                 recursive=True)):
        lines.append(f"# counters: {os.path.relpath(path, outdir)}")
        agg = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                key = (r.get("Kernel_Name", "?")[:90], r.get("Counter_Name"))
                agg.setdefault(key, []).append(float(r.get("Counter_Value", 0)))
        for (k, c), v in sorted(agg.items()):
            lines.append(f"{k:90s} {c}: dispatches={len(v)} mean={statistics.mean(v):.1f} "
                         f"median={statistics.median(v):.1f}")
    text = "\n".join(lines) + "\n"
    if dest:
        with open(dest, "w") as f:
            f.write(text)
    print(text)


def _counter_means(path, counter):
    out = {}
    with open(path) as f:
        for r in csv.DictReader(f):
            if r.get("Counter_Name") != counter:
                continue
            out.setdefault(r.get("Kernel_Name", "?"), []).append(float(r.get("Counter_Value", 0)))
    return {k: statistics.mean(v) for k, v in out.items()}


def pmc_traffic(fetch_dir, write_dir, kernel_key, copy_mib, dest, command):
    """HBM bytes per launch of the kernel whose name contains kernel_key:
    FETCH_SIZE and WRITE_SIZE (KiB per dispatch) from two counter runs, each
    scaled by what the same counter reads for the copy kernel of known size
    (copy_mib MiB read and copy_mib MiB written per dispatch)."""
    f = _counter_means(glob.glob(os.path.join(fetch_dir, "**", "*counter_collection.csv"),
                                 recursive=True)[0], "FETCH_SIZE")
    w = _counter_means(glob.glob(os.path.join(write_dir, "**", "*counter_collection.csv"),
                                 recursive=True)[0], "WRITE_SIZE")
    pick = lambda d, key: next(v for k, v in d.items() if key in k)  # noqa: E731
    kib = copy_mib * 1024.0
    cf, cw = kib / pick(f, "k_copy"), kib / pick(w, "k_copy")
    kf, kw = pick(f, kernel_key), pick(w, kernel_key)
    fetch_b, write_b = kf * cf * 1024.0, kw * cw * 1024.0
    lines = [f"# rocprofv3 PMC summary: {kernel_key}",
             f"# command: {command}",
             f"copy kernel: {copy_mib} MiB read + {copy_mib} MiB written per dispatch",
             f"  FETCH_SIZE raw {pick(f, 'k_copy'):.1f} KiB -> calibration x{cf:.4f}",
             f"  WRITE_SIZE raw {pick(w, 'k_copy'):.1f} KiB -> calibration x{cw:.4f}",
             f"{kernel_key}: FETCH_SIZE raw {kf:.1f} KiB, WRITE_SIZE raw {kw:.1f} KiB per dispatch",
             f"fetch_bytes_per_launch = {fetch_b:.0f}",
             f"write_bytes_per_launch = {write_b:.0f}",
             f"traffic_bytes_per_launch = {fetch_b + write_b:.0f}"]
    text = "\n".join(lines) + "\n"
    if dest:
        with open(dest, "w") as fh:
            fh.write(text)
    print(text)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="fp64", choices=["fp64", "split", "plain"])
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--check-interval", type=int, default=500)
    ap.add_argument("--timing-every", type=int, default=16)
    ap.add_argument("--no-modes", action="store_true", help="skip the other schemes")
    ap.add_argument("--mode-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--probe", action="store_true")
    ap.add_argument("--probe-iters", type=int, default=200)
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--summarize", default=None)
    ap.add_argument("--summary-out", default=None)
    ap.add_argument("--sweep", action="store_true")
    ap.add_argument("--sweep-sizes", default="256,1024,4096")
    ap.add_argument("--sweep-iters", type=int, default=300)
    ap.add_argument("--pmc-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"), default=None)
    ap.add_argument("--pmc-kernel", default="k_fused<double")
    ap.add_argument("--pmc-command", default="")
    a = ap.parse_args()

    if a.pmc_traffic:
        pmc_traffic(a.pmc_traffic[0], a.pmc_traffic[1], a.pmc_kernel, a.copy_mib,
                    a.summary_out, a.pmc_command)
        return 0
    if a.summarize:
        summarize(a.summarize, a.summary_out)
        return 0
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "a multi-GPU run would shard independent RVEs across devices and is not "
              "part of this build", file=sys.stderr)
        return 2
    if a.probe:
        probe(a.mode, a.nx, a.probe_iters, a.copy_mib)
        return 0
    if a.sweep:
        sweep([int(x) for x in a.sweep_sizes.split(",")], ["fp64", "split"], [6, 8, 14, 30],
              a.sweep_iters)
        return 0

    from effplast2d_amd import device_count
    if device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    pa, nx = make_pa(a.nx)
    main_run = run_mode(pa, nx, a.mode, a.steps, a.warmup, a.check_interval, a.timing_every)
    modes = {}
    if not a.no_modes:
        for m in ("fp64", "split", "plain"):
            if m == a.mode:
                continue
            r = run_mode(pa, nx, m, a.mode_steps, 0, a.check_interval, a.timing_every)
            ref = main_run["moduli"]
            r["moduli_rel_diff_vs_" + a.mode] = {
                k: abs(r["moduli"][k] - ref[k]) / abs(ref[k]) if ref[k] else None
                for k in ("Kphi", "Kd", "Gxy")}
            modes[m] = r
    frac = main_run["achieved_GBs"] / HBM_PEAK_GBS
    traffic = read_traffic(a.mode, nx)
    out = {
        "metric": METRIC,
        "value": main_run["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main_run["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.mode != "split" else "f32/f64-split",
        "data": "synthetic",
        "config": {"workload": f"1024x1024 elastoplastic RVE, nTasks=3 (preload [1,1,0] "
                               f"-3e-3, hydrostatic +0.5%, xy shear +0.5%), one pore phi=0.01, "
                               f"Y=2.5e-3, eIter=1e-8, scheme={a.mode}",
                   "nx": nx, "nTasks": 3, "check_interval": a.check_interval,
                   "eIter": 1e-8, "nIter": 100000, "G0": 0.5, "shear_load_type": [0, 0, 1]},
        "roofline": {"bound": "hbm", "achieved": main_run["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": frac, "traffic": traffic,
                     "kernel": main_run["kernel"], "kernel_avg_us": main_run["kernel_avg_us"],
                     "bytes_per_launch": main_run["bytes_per_launch"],
                     "kernel_samples": main_run["kernel_samples"]},
        "cell_iter_per_s": main_run["cell_iter_per_s"],
        "p50_ms": main_run["p50_ms"],
        "p99_ms": main_run["p99_ms"],
        "iters_per_solve": main_run["iters_per_solve"],
        "iters_by_task": main_run["iters_by_task"],
        "moduli": main_run["moduli"],
        "converged": main_run["converged"],
        "modes": modes,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pa, nx, main_run["iters_per_solve"], a.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
