#!/usr/bin/env python3
"""Benchmark: converged RVE solves/s of the EffPlast2D hot path on one MI355X.

Headline workload (one "step" = one converged solve): the 1024x1024
elastoplastic RVE with all three moduli -- nGrid=32, Lx=Ly=20, one pore,
porosity 0.01, Y = 2.5e-3*K0, K0=1, G0=0.5, rho0=1, CFL=0.5, preload
initLoadValue=-3e-3 along loadType [1,1,0], then the 0.5% hydrostatic
increment and a pure xy shear increment (shear loadType [0,0,1]),
nTimeSteps=1, eIter=1e-8, nIter=100000, residual checked every 500
iterations.  The timed region submits K such requests to the batched engine
in the FP32/FP64 split scheme (graph-captured chunks, kernel shape tuned on
the running batch) and runs them to convergence; value = K / elapsed.  Inputs are
generated on the device; nothing crosses PCIe but scalars, tables and sampled
boundary values.

Beside it, in the same run:
  * single-RVE solves of the same request: fused FP64, fused split, and the
    plain four-kernel FP64 solver (the reference's structure);
  * the batched FP64 engine on the same K requests;
  * the batched sweep: 256 RVEs of seeded random size/porosity/pores/yield/
    preload (sizes 128..1024), split scheme, one engine (all retired solves/s
    and converged solves/s);
  * serving: Poisson arrivals of moduli queries of sizes 256..2048 under an
    HBM budget at several arrival rates, p50/p99 latency, solves/s and the
    highest sustainable rate, beside the sequential plain FP64 solver run on
    the first requests of the same stream;
  * the CPU oracle on this host's cores on a bounded sample.

Printed: ONE JSON line; value and roofline are the headline's, the roofline's
kernel time is the live HIP-event timing of the batched kernel's graph
replays on the engine's stream (the sampled single-dispatch figure beside it).

Other entry points:
  --probe      a fixed run of one scheme plus a copy kernel of known size
               (for rocprofv3 --pmc counter runs): the batched kernel of
               --probe-shape over --probe-batch RVEs, or one RVE with that
               shape ("auto": the handle's own choice)
  --probe-shapes  the batched probe once per listed kernel shape (for a
               rocprofv3 kernel trace by grid size)
  --sweep      fused-kernel duration per grid size, scheme and tile height
  --rve4096    4096^2, 16 pores, porosity 0.05: split against FP64 moduli
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

from effplast2d_amd.workloads import ITER_CAP_PER_N, mixed_requests, serving_stream  # noqa: E402

METRIC = ("converged RVE solves/s + cell-iter/s (% HBM peak), FP32-split vs FP64; p50/p99")
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12   # lane instructions per second, 10^12
# algorithmic bytes per point per iteration: U and V read once and written once
BYTES_PER_POINT = {"fp64": 64, "plain": 64, "split": 48}

WORKLOAD = dict(Lx=20.0, nGrid=32, nPores=1, porosity=0.01, Y=2.5e-3, K0=1.0, G0=0.5,
                rho0=1.0, CFL=0.5)
LOAD = dict(initLoad=-3e-3, loadType=[1.0, 1.0, 0.0], shearLoadType=[0.0, 0.0, 1.0])
# every run checks the residual every --check-interval iterations, or, with
# --adaptive-check, on the adaptive schedule (fewer iterations; the single-RVE
# legs the headline is compared with then use it too)
ADAPTIVE_CHECK = False


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
                    shearLoadType=LOAD["shearLoadType"], adaptiveCheck=ADAPTIVE_CHECK) as g:
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


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return float("nan")
    k = (len(xs) - 1) * q
    lo, hi = int(k), min(int(k) + 1, len(xs) - 1)
    return xs[lo] + (xs[hi] - xs[lo]) * (k - lo)


def shape_name(kind, rows):
    names = {1: 'tiles', 2: 'march', 3: 'persist', 4: 'march2x', 5: 'pair', 6: 'triple', 7: 'quad',
             8: 'wide3x', 9: 'wide4x', 11: 'triple2w', 12: 'quad2w'}
    return f"{names[kind]}{rows}"


def cand_us(st):
    """Last tuning's mean us per iteration, by candidate shape name."""
    return {shape_name(k, r): u for k, r, u in zip(st["candKind"], st["candRows"], st["candUs"])}


def batch_stats_summary(st, mode):
    """Kernel figures of a batched run.  The launch duration is the average
    over the timed graph replays (HIP events on the engine's stream around
    each replay of back-to-back launches of the chosen kernel) when graphs
    ran, else over the sampled single launches between dispatch events; a
    single bracketed launch also carries the dispatch latency of the launch
    after a replay, so it reads longer than the kernel itself.

    Bytes per launch: a launch of an n-iteration shape reads and writes U and
    V once for its n iterations, so the bytes it must move are 48 B (split) or
    64 B (FP64) per point once -- `bytes_per_launch`, the roofline's
    algorithmic bytes.  `effective_GBs` counts those bytes once per iteration
    instead: the HBM rate a one-iteration kernel would need to match the
    launch's iteration rate (above the peak for n > 1)."""
    nan = float("nan")
    ipl = {4: 2, 5: 2, 6: 3, 7: 4, 8: 3, 9: 4, 11: 3, 12: 4}.get(st["kernel"], 1)  # per launch
    it_disp = st["timedMs"] / st["timedLaunches"] if st["timedLaunches"] else nan
    it_graph = st["graphMs"] / st["graphLaunches"] if st["graphLaunches"] else None
    # timedBytes counts the algorithmic bytes of each iteration
    bytes_iter = st["timedBytes"] / st["timedLaunches"] if st["timedLaunches"] else nan
    kms = ipl * (it_graph if it_graph is not None else it_disp)   # ms per launch
    ct = "float" if mode == "split" else "double"
    kern = {1: f"k_fused_batch<{ct}, {ct}, TY>", 2: f"k_march<{ct}, {ct}>",
            3: f"k_fused_persist<{ct}, {ct}, TY>", 4: f"k_march2<{ct}, {ct}>",
            5: f"k_march2p<{ct}, {ct}>", 6: f"k_chain<{ct}, {ct}, 3, 1>",
            7: f"k_chain<{ct}, {ct}, 4, 1>", 8: "k_wide<3>", 9: "k_wide<4>",
            11: f"k_chain<{ct}, {ct}, 3, 2>", 12: f"k_chain<{ct}, {ct}, 4, 2>"}[st["kernel"]]
    return {"kernel": kern, "iterations_per_launch": ipl,
            "kernel_avg_us": kms * 1e3,
            "kernel_timing": "graph replays" if it_graph is not None else "dispatch events",
            "kernel_samples": (st["graphLaunches"] if it_graph is not None
                               else st["timedLaunches"]) // ipl,
            "kernel_avg_us_dispatch": ipl * it_disp * 1e3,
            "bytes_per_launch": bytes_iter,
            "achieved_GBs": bytes_iter / (kms * 1e-3) / 1e9,
            "effective_GBs": ipl * bytes_iter / (kms * 1e-3) / 1e9,
            "launches": st["launches"], "chunks": st["chunks"],
            "graph_avg_us_per_launch": ipl * 1e3 * it_graph if it_graph is not None else None,
            "kernel_shape": shape_name(st["kernel"], st["kernelRows"]),
            "tuning_us": cand_us(st),
            "cell_iterations": st["cellIterations"], "bytes_algorithmic": st["bytesAlgorithmic"],
            # time steps the split scheme finished in the FP64 arithmetic
            "fp64_finishes": st["fp64Finishes"],
            # share of the cell-iterations advanced in the FP64 arithmetic
            "fp64_cell_iter_frac": (st["fp64CellIterations"] / st["cellIterations"]
                                    if st["cellIterations"] else 0.0)}


def run_batch(reqs, mode, check, timing_every=1, **opts):
    """Submit every request at once and run the engine dry.  Returns results
    (by tag), engine statistics and elapsed seconds."""
    if ADAPTIVE_CHECK:
        opts.setdefault("adaptiveCheck", True)
    from effplast2d_amd import Batch
    with Batch(mode=mode, checkInterval=check, timingEvery=timing_every, **opts) as b:
        t0 = time.perf_counter()
        for r in reqs:
            b.submit(**r)
        b.run()
        dt = time.perf_counter() - t0
        res = {r["tag"]: r for r in b.poll()}
        st = b.stats()
    bad = [r for r in res.values() if r["status"] != 0]
    if bad:
        raise RuntimeError(f"batched requests failed: {bad[0]['message']}")
    return res, st, dt


def headline_request(pa, nx, tag):
    return dict(pa=pa, nX=nx, nTasks=3, nIter=100_000, eIter=1e-8, initLoad=LOAD["initLoad"],
                loadType=LOAD["loadType"], shearLoadType=LOAD["shearLoadType"], tag=tag)


def run_batched_headline(pa, nx, mode, steps, warmup, check, graphs=True):
    if warmup:
        run_batch([headline_request(pa, nx, k) for k in range(warmup)], mode, check,
                  useGraphs=graphs)
    res, st, dt = run_batch([headline_request(pa, nx, k) for k in range(steps)], mode, check,
                            useGraphs=graphs)
    lat = [r["tDone"] - r["tSubmit"] for r in res.values()]
    tot_iters = sum(r["result"]["totalIters"] for r in res.values())
    r0 = res[0]["result"]
    log(f"[bench] batched {mode} x{steps}: {dt:.2f} s, {steps / dt:.2f} solves/s, "
        f"{st['fp64Finishes']} FP64 finishes, tunings {st['tunings']}, chunks {st['chunks']}, "
        f"kernel {shape_name(st['kernel'], st['kernelRows'])} "
        f"{ {k: round(v, 1) for k, v in cand_us(st).items()} }")
    out = {"mode": mode, "batched": steps, "solves_per_s": steps / dt, "elapsed_s": dt,
           "ms_per_solve": 1e3 * dt / steps, "p50_ms": 1e3 * percentile(lat, 0.5),
           "p99_ms": 1e3 * percentile(lat, 0.99), "iters_per_solve": tot_iters / steps,
           "cell_iter_per_s": tot_iters * nx * nx / dt,
           "moduli": {"Kphi": r0["num"]["Kphi"], "Kd": r0["num"]["Kd"], "Gxy": r0["Gxy"]},
           "converged": all(bool(r["result"]["converged"]) for r in res.values())}
    out.update(batch_stats_summary(st, mode))
    return out


def run_elastic_128(steps=5):
    """128^2, one central pore (porosity 0.005), elastic (nTasks=1), plain
    FP64 (the reference's four kernels), preload -1e-5 along [1,1,0],
    eIter=1e-8: solves/s and the numerical moduli against the analytic
    elastic ones the reference logs (EffPlast2D.cu:643-654)."""
    from effplast2d_amd import EffPlast2D, derive_params
    pa, nx = derive_params(Lx=20.0, Ly=20.0, nX=128, nPores=1, porosity=0.005, Y=2.5e-3, K0=1.0,
                           G0=0.5, rho0=1.0, CFL=0.5)
    t0 = time.perf_counter()
    for _ in range(steps):
        with EffPlast2D(pa, nx, nTasks=1, nIter=200_000, eIter=1e-8, checkInterval=500,
                        mode="plain") as g:
            res = g.ComputeEffModuli(-1e-5, [1.0, 1.0, 0.0])
    dt = time.perf_counter() - t0
    out = {"workload": "128x128, one central pore, porosity 0.005, elastic (nTasks=1), plain "
                       "FP64, preload -1e-5 [1,1,0], eIter 1e-8, check 500",
           "solves_per_s": steps / dt, "iters": int(res["totalIters"]),
           "Kd": res["num"]["Kd"], "Kphi": res["num"]["Kphi"],
           "KdExactElast": res["KdExactElast"], "KphiExact": res["an"]["Kphi"],
           "Kd_rel_vs_exact": abs(res["num"]["Kd"] - res["KdExactElast"]) / res["KdExactElast"],
           "Kphi_rel_vs_exact": abs(res["num"]["Kphi"] - res["an"]["Kphi"]) / res["an"]["Kphi"]}
    log(f"[bench] elastic 128: {out['solves_per_s']:.1f} solves/s, Kd {out['Kd']:.5f} (exact "
        f"{out['KdExactElast']:.5f}), Kphi {out['Kphi']:.4f} (exact {out['KphiExact']:.4f})")
    return out


def run_mixed_batch(n, seed, check):
    reqs = mixed_requests(n, seed)
    res, st, dt = run_batch(reqs, "split", check)
    lat = [r["tDone"] - r["tSubmit"] for r in res.values()]
    cells = st["cellIterations"]
    by_size = {}
    for rq in reqs:
        r = res[rq["tag"]]["result"]
        d = by_size.setdefault(rq["nX"], {"n": 0, "converged": 0, "iters": []})
        d["n"] += 1
        d["converged"] += int(bool(r["converged"]))
        d["iters"].append(int(r["totalIters"]))
    for d in by_size.values():
        d["mean_iters"] = sum(d["iters"]) / len(d["iters"])
        del d["iters"]
    conv = sum(bool(r["result"]["converged"]) for r in res.values())
    out = {"workload": f"{n} RVEs, sizes U{{128,256,512,1024}}, porosity U[0.005,0.05], "
                       f"nPores U{{1,2,3}}, Y U[1e-3,5e-3], preload U[1e-3,5e-3] x random unit "
                       f"(Exx,Eyy,Exy), nTasks=3, seed {seed}, split, one engine",
           "solves": n, "elapsed_s": dt,
           # every request retired (converged or at its iteration cap) / the
           # converged ones only
           "solves_per_s": n / dt, "converged": conv, "capped": n - conv,
           "converged_solves_per_s": conv / dt, "cell_iter_per_s": cells / dt,
           "hbm_frac_whole_run": st["bytesAlgorithmic"] / dt / 1e9 / HBM_PEAK_GBS,
           "p50_ms": 1e3 * percentile(lat, 0.5), "p99_ms": 1e3 * percentile(lat, 0.99),
           "by_size": by_size, "iteration_cap": f"{ITER_CAP_PER_N} N per load step"}
    out.update(batch_stats_summary(st, "split"))
    log(f"[bench] mixed batch {n} RVEs: {dt:.2f} s, {n / dt:.2f} solves/s ({conv} converged, "
        f"{conv / dt:.2f} converged solves/s), {cells / dt / 1e9:.1f} G cell-iter/s, kernel "
        f"{out['achieved_GBs']:.0f} GB/s")
    return out, res


def serve_stream(stream, check, budget_gib):
    """Continuous batching under Poisson arrivals: requests are submitted at
    their arrival times between chunks; each chunk admits what the HBM budget
    allows, runs every admitted RVE to its next residual check, retires the
    finished ones and backfills.  Returns results by tag, engine statistics,
    the peak number of admitted RVEs and the stream's start on the batch
    clock."""
    from effplast2d_amd import Batch
    done = {}
    with Batch(mode="split", checkInterval=check, timingEvery=1,
               memoryBudget=int(budget_gib * (1 << 30)), adaptiveCheck=ADAPTIVE_CHECK) as b:
        t0 = b.clock()
        i, peak = 0, 0
        while i < len(stream) or len(done) < i:
            now = b.clock() - t0
            while i < len(stream) and stream[i][0] <= now:
                b.submit(**stream[i][1])
                i += 1
            if len(done) == i:  # idle until the next arrival
                time.sleep(max(0.0, stream[i][0] - (b.clock() - t0)))
                continue
            active, _ = b.step()
            peak = max(peak, active)
            for r in b.poll():
                done[r["tag"]] = r
        st = b.stats()
    bad = [r for r in done.values() if r["status"] != 0]
    if bad:
        raise RuntimeError(f"served request failed: {bad[0]['message']}")
    return done, st, peak, t0


def run_serving(rate, duration, seed, check, budget_gib):
    """One serving run at `rate` arrivals per second: p50/p99 latency (submit
    to retirement), solves/s over the span from the first arrival to the last
    retirement, converged counts, and the backlog left when arrivals end."""
    stream = serving_stream(rate, duration, seed)
    done, st, peak, t0 = serve_stream(stream, check, budget_gib)
    lat = [r["tDone"] - r["tSubmit"] for r in done.values()]
    span = max(r["tDone"] for r in done.values()) - t0 - stream[0][0]
    conv = sum(bool(r["result"]["converged"]) for r in done.values())
    backlog = sum(1 for r in done.values() if r["tDone"] - t0 > duration)
    out = {"rate_per_s": rate,
           "workload": f"Poisson arrivals {rate}/s for {duration} s (seed {seed}), sizes "
                       f"{{256,512,1024,2048}} w [.4,.3,.2,.1], nTasks U{{1,2,3}}, other "
                       f"parameters as the batched sweep, iteration cap {ITER_CAP_PER_N} N per "
                       f"load step, split, HBM budget {budget_gib} GiB",
           "requests": len(stream), "converged": conv, "solves_per_s": len(stream) / span,
           "converged_solves_per_s": conv / span, "span_s": span,
           "p50_ms": 1e3 * percentile(lat, 0.5), "p99_ms": 1e3 * percentile(lat, 0.99),
           "peak_active": peak, "unfinished_when_arrivals_end": backlog,
           "cell_iter_per_s": st["cellIterations"] / span}
    out.update(batch_stats_summary(st, "split"))
    log(f"[bench] serving {rate}/s: {len(stream)} requests ({conv} converged), "
        f"{out['solves_per_s']:.2f} solves/s, p50 {out['p50_ms']:.0f} ms p99 "
        f"{out['p99_ms']:.0f} ms, peak {peak} active, {backlog} unfinished at {duration} s")
    return out, stream, done


def run_sequential_plain(stream, results, count):
    """The plain FP64 solver (the reference's four kernels), one request at a
    time in arrival order, on the first `count` requests of a served stream:
    measured service times, FIFO queueing over the same arrival times."""
    from effplast2d_amd import EffPlast2D
    t_free, lat, svc_total, iters_equal = 0.0, [], 0.0, 0
    for (ta, rq) in stream[:count]:
        t0 = time.perf_counter()
        with EffPlast2D(rq["pa"], rq["nX"], nTasks=rq["nTasks"], nIter=rq["nIter"],
                        eIter=rq["eIter"], checkInterval=500, mode="plain") as g:
            res = g.ComputeEffModuli(rq["initLoad"], rq["loadType"])
        svc = time.perf_counter() - t0
        svc_total += svc
        start = max(ta, t_free)
        t_free = start + svc
        lat.append(t_free - ta)
        iters_equal += int(res["totalIters"] == results[rq["tag"]]["result"]["totalIters"])
    span = t_free - stream[0][0]
    out = {"requests": count, "solves_per_s": count / span, "service_s_total": svc_total,
           "p50_ms": 1e3 * percentile(lat, 0.5), "p99_ms": 1e3 * percentile(lat, 0.99),
           "same_iterations_as_served_split": iters_equal,
           "how": f"measured: the first {count} requests of the served stream, plain FP64 "
                  "single-RVE solver, one at a time, FIFO over their arrival times"}
    served = sorted(results[rq["tag"]]["tDone"] - results[rq["tag"]]["tSubmit"]
                    for _, rq in stream[:count])
    out["served_split_same_requests_p50_ms"] = 1e3 * percentile(served, 0.5)
    log(f"[bench] sequential plain FP64, first {count} requests: {out['solves_per_s']:.2f} "
        f"solves/s, p50 {out['p50_ms']:.0f} ms p99 {out['p99_ms']:.0f} ms")
    return out


def run_serving_sweep(rates, duration, seed, check, budget_gib, seq_count):
    """Serving at several arrival rates (same seed): the sustainable rate is
    the highest rate whose arrivals are all retired within 1.5 x the
    arrival window and whose p99 stays within 3 x the lowest rate's p99."""
    runs, first = [], None
    for rate in rates:
        out, stream, done = run_serving(rate, duration, seed, check, budget_gib)
        runs.append(out)
        if first is None:
            first = (stream, done)
    base = runs[0]["p99_ms"]
    ok = [r for r in runs if r["span_s"] <= 1.5 * duration and r["p99_ms"] <= 3 * base]
    out = {"runs": runs,
           "sustainable_rate_per_s": max((r["rate_per_s"] for r in ok), default=None),
           "sequential_plain_fp64": run_sequential_plain(first[0], first[1], seq_count)}
    return out


def profile_path(kind, shape, mode, nx, nrve):
    """Committed counter summary of the batched kernel in `shape` over `nrve`
    RVEs of nx^2 (kind "pmc": HBM traffic, "sq": instruction counters)."""
    return os.path.join(ROOT, "profiles", f"{kind}_batch_{shape}_{mode}_{nx}_x{nrve}.txt")


def read_profile_value(path, key):
    """`key = value` line of a committed summary; None when absent."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith(key + " ="):
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


def probe(mode, nx, iters, copy_mib, batch=0, shape="tiles8"):
    """Fixed workload for counter runs: `iters` iterations of one scheme on the
    preloaded RVE (or, with batch > 0, `iters` iterations of the batched
    kernel of `shape` (a name as shape_name gives it, e.g. wide3x64) over
    `batch` headline RVEs, launched one by one), then 5 copies of copy_mib
    MiB (the calibration kernel of the traffic counters)."""
    import ctypes as C

    from effplast2d_amd import EffPlast2D
    from effplast2d_amd._lib import check, lib
    pa, _ = make_pa(nx)
    kind = shape.rstrip("0123456789")
    rows = int(shape[len(kind):]) if shape != "auto" else 0
    if batch > 0:
        reqs = [dict(headline_request(pa, nx, k), nIter=iters, eIter=1e-300)
                for k in range(batch)]
        run_batch(reqs, mode, 10 ** 9, timing_every=0, useGraphs=False,
                  **{SHAPE_OPTION[kind]: rows})
    else:
        # one RVE: a shape as single_shape names it, or "auto" (the handle's
        # automatic choice)
        kw = {} if shape == "auto" else single_shape(shape)
        with EffPlast2D(pa, nx, nTasks=3, mode=mode, **kw) as g:
            g.apply_affine(-3e-3, -3e-3, 0.0, 0.0)
            g.iterate(iters)
            g.sync()
    ms = C.c_double()
    check(lib().ep2d_diag_copy(0, copy_mib << 20, 5, C.byref(ms)))
    log(f"[probe] {mode} {nx} batch {batch} {shape}: {iters} iterations; copy {copy_mib} MiB: "
        f"{ms.value:.3f} ms ({2 * (copy_mib << 20) / ms.value / 1e6:.0f} GB/s)")


SHAPE_OPTION = {"tiles": "tileRows", "march": "marchRows", "persist": "persistRows",
                "march2x": "march2Rows", "pair": "pairRows", "triple": "tripleRows",
                "quad": "quadRows", "wide3x": "wide3Rows", "wide4x": "wide4Rows",
                "triple2w": "triple2Rows", "quad2w": "quad2Rows"}


def probe_shapes(mode, nx, iters, batch, shapes):
    """Kernel-trace workload: for each shape name (as shape_name gives it),
    `iters` iterations of the batched engine over `batch` headline RVEs with
    that shape fixed, launched one by one (no graphs), so a trace averages
    each shape's launch duration by its grid size."""
    pa, _ = make_pa(nx)
    for name in shapes:
        kind = name.rstrip("0123456789")
        rows = int(name[len(kind):])
        reqs = [dict(headline_request(pa, nx, k), nIter=iters, eIter=1e-300)
                for k in range(batch)]
        run_batch(reqs, mode, 10 ** 9, timing_every=0, useGraphs=False,
                  **{SHAPE_OPTION[kind]: rows})
        log(f"[probe] {mode} {nx} batch {batch}: {name}, {iters} iterations")


# single-RVE kernel shapes of the sweep: name -> EffPlast2D options
def single_shape(name):
    kind = name.rstrip("0123456789")
    n = int(name[len(kind):])
    return {"tiles": {"tileRows": n}, "march": {"marchRows": n},
            "pair": {"marchRows": n, "itersPerLaunch": 2},
            "triple": {"marchRows": n, "itersPerLaunch": 3},
            "quad": {"marchRows": n, "itersPerLaunch": 4},
            "wide3x": {"marchRows": n, "itersPerLaunch": 3, "laneColumns": 2},
            "wide4x": {"marchRows": n, "itersPerLaunch": 4, "laneColumns": 2},
            "triple2w": {"marchRows": n, "itersPerLaunch": 3, "stageWaves": 2},
            "quad2w": {"marchRows": n, "itersPerLaunch": 4, "stageWaves": 2}}[kind]


def sweep(sizes, modes, shapes, iters):
    """Average kernel duration per iteration for each (scheme, kernel shape,
    grid size) on the preloaded single RVE, every launch timed with dispatch
    events (a multi-iteration launch counts as its iterations)."""
    from effplast2d_amd import EffPlast2D
    rows = []
    for nx in sizes:
        pa, _ = make_pa(nx)
        for mode in modes:
            for shape in shapes:
                kw = single_shape(shape)
                if mode != "split" and "laneColumns" in kw:
                    continue
                with EffPlast2D(pa, nx, nTasks=3, mode=mode, **kw) as g:
                    g.apply_affine(-3e-3, -3e-3, 0.0, 0.0)
                    g.iterate(24)
                    g.set_timing(1)
                    g.iterate(iters)
                    ms, n_, name = g.timing()
                b = BYTES_PER_POINT[mode] * (nx + 1) * nx
                row = {"nx": nx, "mode": mode, "shape": shape, "kernel_us": ms * 1e3,
                       "GBs_per_iteration": b / (ms * 1e-3) / 1e9, "kernel": name}
                rows.append(row)
                log(f"[sweep] nx={nx} {mode} {shape}: {ms * 1e3:.2f} us per iteration")
    print(json.dumps({"sweep": rows}), flush=True)


def summarize(outdir, dest, command=""):
    """Condense rocprofv3 csv output in outdir into text lines."""
    lines = [f"# rocprofv3 summary of {os.path.basename(outdir.rstrip('/'))}"]
    if command:
        lines.append(f"# command: {command}")
    for path in sorted(glob.glob(os.path.join(outdir, "**", "*kernel_stats.csv"), recursive=True)):
        lines.append(f"# kernel stats: {os.path.relpath(path, outdir)}")
        with open(path) as f:
            rows = list(csv.DictReader(f))
        for r in rows[:12]:
            lines.append(f"{r.get('Name', '?')[:90]:90s} calls={r.get('Calls')} "
                         f"avg_ns={float(r.get('AverageNs', 0)):.0f} "
                         f"pct={float(r.get('Percentage', 0)):.2f}")
    for path in sorted(glob.glob(os.path.join(outdir, "**", "*kernel_trace.csv"), recursive=True)):
        # per (kernel, grid size): launches over different batch sizes apart
        lines.append(f"# kernel trace by grid size: {os.path.relpath(path, outdir)}")
        agg = {}
        with open(path) as f:
            for r in csv.DictReader(f):
                key = (r.get("Kernel_Name", "?"), int(r.get("Grid_Size_X", 0)))
                agg.setdefault(key, []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        top = sorted(agg.items(), key=lambda kv: -sum(kv[1]))[:8]
        for (k, g), v in top:
            lines.append(f"{k[:90]:90s} grid_x={g} calls={len(v)} avg_ns={statistics.mean(v):.0f} "
                         f"median_ns={statistics.median(v):.0f}")
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


def sq_summary(sq_dir, kernel_key, dest, command, points=None):
    """VALU issue of the kernel whose name contains kernel_key, from one
    counter run of VALUBusy (rocprofv3's derived metric: SQ_ACTIVE_INST_VALU
    summed over the SIMDs / CUs / GRBM_GUI_ACTIVE, the percentage of cycles
    a SIMD issues VALU instructions), SQ_INSTS_VALU and SQ_WAVES, and, when
    the run collected them, SQ_WAIT_ANY and SQ_WAVE_CYCLES (the share of a
    wave's lifetime spent waiting on a counter or barrier).  points: point
    iterations one dispatch computes (the lane instructions per point
    iteration are then reported)."""
    path = glob.glob(os.path.join(sq_dir, "**", "*counter_collection.csv"), recursive=True)[0]
    vals = {}
    for c in ("VALUBusy", "SQ_INSTS_VALU", "SQ_WAVES", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"):
        m = _counter_means(path, c)
        vals[c] = next((v for k, v in m.items() if kernel_key in k), None)
    busy = vals["VALUBusy"]
    lines = [f"# rocprofv3 SQ summary: {kernel_key}",
             f"# command: {command}",
             f"VALUBusy (percent, mean over dispatches) = {busy}",
             f"SQ_INSTS_VALU per dispatch = {vals['SQ_INSTS_VALU']}",
             f"SQ_WAVES per dispatch = {vals['SQ_WAVES']}",
             f"valu_insts_per_wave = "
             f"{vals['SQ_INSTS_VALU'] / vals['SQ_WAVES'] if vals['SQ_WAVES'] else None}"]
    if vals["SQ_WAIT_ANY"] is not None and vals["SQ_WAVE_CYCLES"]:
        lines.append(f"SQ_WAIT_ANY / SQ_WAVE_CYCLES (share of wave lifetime waiting) = "
                     f"{vals['SQ_WAIT_ANY'] / vals['SQ_WAVE_CYCLES']:.4f}")
    if points and vals["SQ_INSTS_VALU"]:
        lines.append(f"point iterations per dispatch = {points}")
        lines.append(f"valu_lane_insts_per_point_iteration = "
                     f"{vals['SQ_INSTS_VALU'] * 64.0 / points:.1f}")
    lines.append(f"valu_busy_frac = {busy / 100.0 if busy is not None else None}")
    text = "\n".join(lines) + "\n"
    if dest:
        with open(dest, "w") as fh:
            fh.write(text)
    print(text)


def run_rve4096(check, nIter=100_000, initLoad=None):
    """4096^2, 16 pores (nPores=4), porosity 0.05, Y=2.5e-3, nTasks=3: the
    split scheme's moduli (including the periodic internal-cell ones) against
    an FP64 run of the same case.  Each load step runs to eIter=1e-8 or to
    nIter iterations (the reference's NITER, EffPlast2D.h:29-31), whichever
    comes first; both schemes get the same budget.  initLoad: the preload
    (default the headline's -3e-3; at that preload neither scheme reaches
    eIter at this size, at -1e-3 both do)."""
    initLoad = LOAD["initLoad"] if initLoad is None else initLoad
    from effplast2d_amd import EffPlast2D, derive_params
    pa, nx = derive_params(Lx=20.0, Ly=20.0, nX=4096, nPores=4, porosity=0.05, Y=2.5e-3, K0=1.0,
                           G0=0.5, rho0=1.0, CFL=0.5)
    out = {}
    for mode in ("split", "fp64"):
        t0 = time.perf_counter()
        # automatic single-RVE shape: four-stage chains on two waves per stage
        # (split), wave pairs (FP64)
        # verbose: the reference's console lines, one per residual check
        with EffPlast2D(pa, nx, nTasks=3, nIter=nIter, eIter=1e-8, checkInterval=check,
                        mode=mode, shearLoadType=LOAD["shearLoadType"], verbose=True) as g:
            g.set_timing(64)
            res = g.ComputeEffModuli(initLoad, LOAD["loadType"])
            ms, n, name = g.timing()
            recs = g.records()
        dt = time.perf_counter() - t0
        out[mode] = {"seconds": dt, "initLoad": initLoad, "kernel_avg_us": ms * 1e3, "kernel": name,
                     "iters": [int(r["iters"]) for r in recs],
                     "errors": [float(r["error"]) for r in recs],
                     "converged": bool(res["converged"]),
                     "num": res["num"], "num_per": res["num_per"], "Gxy": res["Gxy"],
                     "GxyPer": res["GxyPer"]}
        log(f"[bench] rve4096 {mode}: {dt:.1f} s, iters {out[mode]['iters']}, errors "
            f"{['%.2e' % e for e in out[mode]['errors']]}, converged {out[mode]['converged']}")
    err = {}
    for grp in ("num", "num_per"):
        for k in ("Kphi", "Kd", "G"):
            a, b = out["split"][grp][k], out["fp64"][grp][k]
            err[f"{grp}.{k}"] = abs(a - b) / abs(b) if b else None
    for k in ("Gxy", "GxyPer"):
        a, b = out["split"][k], out["fp64"][k]
        err[k] = abs(a - b) / abs(b) if b else None
    out["split_rel_err_vs_fp64"] = err
    print(json.dumps({"rve4096_16pores": out}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="requests in the timed batch")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="split", choices=["fp64", "split"],
                    help="scheme of the batched headline")
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--check-interval", type=int, default=500)
    ap.add_argument("--timing-every", type=int, default=16)
    ap.add_argument("--no-legs", action="store_true", help="headline and CPU baseline only")
    ap.add_argument("--adaptive-check", action="store_true",
                    help="batched runs place residual checks where the decay rate predicts "
                         "eIter (first check at --check-interval)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch the batched chunks on the stream instead of replaying graphs "
                         "(for kernel tracers that cannot follow graph replays)")
    ap.add_argument("--single-steps", type=int, default=2)
    ap.add_argument("--mixed-rves", type=int, default=256)
    ap.add_argument("--mixed-seed", type=int, default=2024)
    ap.add_argument("--serve-rates", default="1.5,3,6",
                    help="arrival rates (per second) of the serving runs")
    ap.add_argument("--serve-seconds", type=float, default=20.0)
    ap.add_argument("--serve-sequential", type=int, default=6,
                    help="requests of the first serving stream also run on the sequential "
                         "plain FP64 solver")
    ap.add_argument("--serve-seed", type=int, default=7)
    ap.add_argument("--serve-budget-gib", type=float, default=2.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--probe", action="store_true")
    ap.add_argument("--probe-shapes", default="",
                    help="kernel trace workload: the batched probe once per listed shape "
                         "(e.g. pair26,triple42,quad64)")
    ap.add_argument("--probe-mode", default="fp64", choices=["fp64", "split", "plain"])
    ap.add_argument("--probe-iters", type=int, default=200)
    ap.add_argument("--probe-batch", type=int, default=0,
                    help="probe the batched kernel with this many headline RVEs")
    ap.add_argument("--probe-shape", default="tiles8",
                    help="probe: kernel shape, e.g. tiles8, march32 (one RVE) or pair26, "
                         "triple42, wide3x64 (batched)")
    # the same as --probe-shape <kind><rows> (the forms the older summaries in
    # profiles/ were made with)
    for kind in ("tile", "march", "march2", "pair", "triple", "quad"):
        ap.add_argument(f"--probe-{kind}", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--summarize", default=None)
    ap.add_argument("--summary-out", default=None)
    ap.add_argument("--sweep", action="store_true")
    ap.add_argument("--sweep-sizes", default="256,1024,4096")
    ap.add_argument("--sweep-iters", type=int, default=300)
    ap.add_argument("--sweep-modes", default="fp64,split")
    ap.add_argument("--sweep-shapes", default="tiles8,tiles14,march16,march32,pair26,triple42,"
                                               "quad16,wide3x32,wide4x16")
    ap.add_argument("--rve4096", action="store_true")
    ap.add_argument("--rve4096-load", type=float, default=None,
                    help="preload of the 4096^2 run (default -3e-3)")
    ap.add_argument("--rve4096-iters", type=int, default=100_000,
                    help="iteration cap per load step of the 4096^2 run (the reference's "
                         "NITER is 100000)")
    ap.add_argument("--pmc-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"), default=None)
    ap.add_argument("--pmc-kernel", default="k_fused<double")
    ap.add_argument("--sq-summary", default=None, metavar="SQ_DIR",
                    help="write the VALU issue summary of --pmc-kernel from a counter run")
    ap.add_argument("--sq-points", type=float, default=0,
                    help="point iterations per dispatch of --pmc-kernel (for --sq-summary)")
    ap.add_argument("--pmc-command", default="")
    a = ap.parse_args()

    if a.pmc_traffic:
        pmc_traffic(a.pmc_traffic[0], a.pmc_traffic[1], a.pmc_kernel, a.copy_mib,
                    a.summary_out, a.pmc_command)
        return 0
    if a.sq_summary:
        sq_summary(a.sq_summary, a.pmc_kernel, a.summary_out, a.pmc_command,
                   a.sq_points or None)
        return 0
    if a.summarize:
        summarize(a.summarize, a.summary_out, a.pmc_command)
        return 0
    if a.gpus != 1:
        print(f"bench.py: the solver runs on one GPU only (asked for --gpus {a.gpus}); "
              "a multi-GPU run would shard independent RVEs across devices and is not "
              "implemented", file=sys.stderr)
        return 2
    if a.probe_shapes:
        probe_shapes(a.probe_mode, a.nx, a.probe_iters, a.probe_batch or 8,
                     a.probe_shapes.split(","))
        return 0
    if a.probe:
        shape = a.probe_shape
        for kind, name in (("tile", "tiles"), ("march", "march"), ("march2", "march2x"),
                           ("pair", "pair"), ("triple", "triple"), ("quad", "quad")):
            rows = getattr(a, f"probe_{kind}")
            if rows:
                shape = f"{name}{rows}"
        probe(a.probe_mode, a.nx, a.probe_iters, a.copy_mib, a.probe_batch, shape)
        return 0
    if a.sweep:
        sweep([int(x) for x in a.sweep_sizes.split(",")], a.sweep_modes.split(","),
              a.sweep_shapes.split(","), a.sweep_iters)
        return 0
    global ADAPTIVE_CHECK
    ADAPTIVE_CHECK = a.adaptive_check
    if a.rve4096:
        run_rve4096(a.check_interval, a.rve4096_iters, a.rve4096_load)
        return 0

    from effplast2d_amd import device_count
    if device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    pa, nx = make_pa(a.nx)
    head = run_batched_headline(pa, nx, a.mode, a.steps, a.warmup, a.check_interval,
                                graphs=not a.no_graphs)
    legs = {}
    if not a.no_legs:
        for m in ("fp64", "split", "plain"):
            legs[f"single_{m}"] = run_mode(pa, nx, m, a.single_steps, 1 if m != "plain" else 0,
                                           a.check_interval, a.timing_every)
        other = "fp64" if a.mode == "split" else "split"
        legs[f"batched_{other}"] = run_batched_headline(pa, nx, other, a.steps, 0,
                                                        a.check_interval)
        ref = legs["single_fp64"]["moduli"]
        legs["moduli_rel_diff_vs_single_fp64"] = {
            name: {k: abs(r["moduli"][k] - ref[k]) / abs(ref[k]) for k in ("Kphi", "Kd", "Gxy")}
            for name, r in [("batched_" + a.mode, head)] + list(legs.items())
            if isinstance(r, dict) and "moduli" in r}
        legs["speedup_vs_single_plain_fp64"] = head["solves_per_s"] / \
            legs["single_plain"]["solves_per_s"]
        legs["elastic_128_plain"] = run_elastic_128()
        legs["mixed_batch"], _ = run_mixed_batch(a.mixed_rves, a.mixed_seed, a.check_interval)
        legs["serving"] = run_serving_sweep([float(x) for x in a.serve_rates.split(",")],
                                            a.serve_seconds, a.serve_seed, a.check_interval,
                                            a.serve_budget_gib, a.serve_sequential)
    frac = head["achieved_GBs"] / HBM_PEAK_GBS
    # HBM traffic and instruction counters of this kernel shape at this batch
    # size, from the committed rocprofv3 summaries (None when none matches)
    pmc = profile_path("pmc", head["kernel_shape"], a.mode, nx, a.steps)
    sq = profile_path("sq", head["kernel_shape"], a.mode, nx, a.steps)
    traffic = read_profile_value(pmc, "traffic_bytes_per_launch")
    valu_busy = read_profile_value(sq, "valu_busy_frac")
    # VALU issue: lane instructions per point-iteration (the counter profile)
    # times the launch's point-iterations over its duration, against the
    # MI355X peak of one lane instruction per lane per cycle (256 CUs x 4
    # SIMDs x 16 lanes x 2.4 GHz = 39.3 T/s; a fused multiply-add counts one)
    valu_per_pt = read_profile_value(sq, "valu_lane_insts_per_point_iteration")
    pt_per_launch = head["bytes_per_launch"] / BYTES_PER_POINT[a.mode] * \
        head["iterations_per_launch"]
    valu_t = (valu_per_pt * pt_per_launch / (head["kernel_avg_us"] * 1e-6) / 1e12
              if valu_per_pt else None)
    out = {
        "metric": METRIC,
        "value": head["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_solve"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64-split" if a.mode == "split" else "f64",
        "data": "synthetic",
        "config": {"workload": f"1024x1024 elastoplastic RVE, nTasks=3 (preload [1,1,0] -3e-3, "
                               f"hydrostatic +0.5%, xy shear +0.5%), one pore phi=0.01, "
                               f"Y=2.5e-3, G0=0.5, eIter=1e-8; {a.steps} requests batched, "
                               f"scheme={a.mode}",
                   "nx": nx, "nTasks": 3, "check_interval": a.check_interval, "eIter": 1e-8,
                   "check_schedule": "adaptive (first check at check_interval, then at "
                   "the iteration the residual's decay rate predicts)" if a.adaptive_check
                   else "fixed",
                   "nIter": 100000, "batched_requests": a.steps,
                   "graphs": not a.no_graphs,
                   "shear_load_type": [0, 0, 1]},
        "roofline": {"bound": "hbm", "achieved": head["achieved_GBs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": frac, "traffic": traffic,
                     "kernel": head["kernel"], "kernel_avg_us": head["kernel_avg_us"],
                     "kernel_timing": head["kernel_timing"],
                     "kernel_avg_us_dispatch": head["kernel_avg_us_dispatch"],
                     "bytes_per_launch": head["bytes_per_launch"],
                     "kernel_samples": head["kernel_samples"],
                     "kernel_shape": head["kernel_shape"],
                     "iterations_per_launch": head["iterations_per_launch"],
                     # achieved: the bytes the launch must move (U and V read
                     # and written once for all its iterations) over its
                     # duration; effective: the same bytes counted once per
                     # iteration (the rate a one-iteration kernel would need)
                     "effective_GBs": head["effective_GBs"],
                     "effective_frac": head["effective_GBs"] / HBM_PEAK_GBS,
                     # measured HBM bytes per launch over its duration
                     "traffic_GBs": (traffic / (head["kernel_avg_us"] * 1e-6) / 1e9
                                     if traffic else None),
                     "traffic_profile": os.path.relpath(pmc, ROOT) if traffic else None,
                     # fraction of SIMD cycles issuing VALU instructions
                     "valu_busy_frac": valu_busy,
                     "valu_profile": os.path.relpath(sq, ROOT) if valu_busy else None,
                     "valu_lane_insts_per_point_iteration": valu_per_pt,
                     "valu_achieved_Tops": valu_t,
                     "valu_peak_Tops": VALU_PEAK_TOPS,
                     "valu_frac": valu_t / VALU_PEAK_TOPS if valu_t else None},
        "cell_iter_per_s": head["cell_iter_per_s"],
        "p50_ms": head["p50_ms"],
        "p99_ms": head["p99_ms"],
        "iters_per_solve": head["iters_per_solve"],
        "moduli": head["moduli"],
        "converged": head["converged"],
        "legs": legs,
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pa, nx, head["iters_per_solve"], a.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
