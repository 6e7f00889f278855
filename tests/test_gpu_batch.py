"""GPU: the batched multi-RVE engine against the single-RVE path and the
CPU oracle.

An RVE's arithmetic in the batched launch is the single-RVE kernel's, so its
result must not depend on the other RVEs it shares launches with: moduli,
per-step records and iteration counts are compared for exact equality with
the single-RVE solve of the same request and check interval ("alone vs in the
mix"), with and without graph-captured chunks, and under admission caps (a
count, and a device-memory budget in bytes with requests arriving while
others run) that force retirement and backfill.  Seeded sampled requests
(mixed sizes, porosities, pore counts, yield stresses, nTasks 1 to 3 and
preloads along random directions in (Exx, Eyy, Exy), so the shear term of
the affine increment is exercised) are checked against the CPU oracle: the
FP64 batch within 1e-8 relative on the moduli with equal iteration counts,
the split batch within 1e-4 relative of the FP64 oracle on Kd, Kphi and G
(residual checks every 200 iterations) and iteration counts within 10% plus
one check interval per load step (checks every 20 iterations).
"""
import sys

import numpy as np
import pytest

from effplast2d_amd import Batch, EffPlast2D, request_bytes
from effplast2d_amd.workloads import mixed_requests
from tests.pairs import make_pa, oracle_state

pytestmark = pytest.mark.gpu

LT = [1.0, 1.0, 0.0]

# (nX, nPores, porosity, Y, nTasks, initLoad)
REQS = [
    (64, 1, 0.01, 2.5e-3, 3, -3e-3),
    (96, 2, 0.03, 3e-3, 3, -2e-3),
    (128, 1, 0.005, 2.0e-3, 1, -1e-4),
    (70, 3, 0.04, 4e-3, 2, -2.5e-3),
]
K = 200


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _single(req, mode):
    nX, nP, phi, Y, nl, load = req
    pa = make_pa(nX, nPores=nP, porosity=phi, Y=Y)
    with EffPlast2D(pa, nX, nTasks=nl, nIter=200_000, eIter=1e-8, checkInterval=K,
                    mode=mode) as g:
        res = g.ComputeEffModuli(load, LT)
        return res


def _batch(reqs, mode, **kw):
    out = {}
    with Batch(mode=mode, checkInterval=K, **kw) as b:
        for t, (nX, nP, phi, Y, nl, load) in enumerate(reqs):
            pa = make_pa(nX, nPores=nP, porosity=phi, Y=Y)
            b.submit(pa, nX, nTasks=nl, nIter=200_000, eIter=1e-8, initLoad=load, loadType=LT,
                     tag=t)
        b.run()
        for r in b.poll():
            out[r["tag"]] = r
        st = b.stats()
    return out, st


def _same(a, b):
    for grp in ["num", "an", "num_per"]:
        for k in ["Kphi", "Kd", "G"]:
            x, y = a[grp][k], b[grp][k]
            assert (x == y) or (np.isnan(x) and np.isnan(y)), (grp, k, x, y)
    assert a["totalIters"] == b["totalIters"]
    assert a["nRecords"] == b["nRecords"]


@pytest.mark.parametrize("mode", ["fp64", "split"])
@pytest.mark.parametrize("graphs", [True, False])
def test_batch_equals_single(mode, graphs):
    _note(f"batch mode={mode} graphs={graphs}")
    got, st = _batch(REQS, mode, useGraphs=graphs)
    assert st["finished"] == len(REQS) and st["failed"] == 0
    for t, rq in enumerate(REQS):
        assert got[t]["status"] == 0, got[t]["message"]
        _same(got[t]["result"], _single(rq, mode))


def test_alone_vs_in_the_mix():
    _note("alone vs mix")
    alone, _ = _batch(REQS[:1], "split")
    mix, _ = _batch(REQS[::-1], "split")
    _same(alone[0]["result"], mix[len(REQS) - 1]["result"])


def test_budget_forces_backfill():
    _note("budget backfill")
    got, st = _batch(REQS, "fp64", maxActive=2)
    assert st["finished"] == len(REQS)
    ref, _ = _batch(REQS, "fp64")
    for t in range(len(REQS)):
        _same(got[t]["result"], ref[t]["result"])
    # admission is FIFO: the third request is admitted only after a retirement
    admits = sorted(got.values(), key=lambda r: r["tAdmit"])
    assert admits[2]["tAdmit"] >= min(admits[0]["tDone"], admits[1]["tDone"])


def test_failed_request_reports_and_others_finish():
    _note("bad request")
    with Batch(mode="fp64", checkInterval=K) as b:
        pa = make_pa(64)
        bad = pa.copy()
        bad[3] = np.nan  # K0: the first residual check sees non-finite velocities
        b.submit(pa, 64, nTasks=1, nIter=50_000, eIter=1e-8, initLoad=-1e-4, tag=0)
        b.submit(bad, 64, nTasks=1, nIter=50_000, eIter=1e-8, initLoad=-1e-4, tag=1)
        b.run()
        res = {r["tag"]: r for r in b.poll()}
    assert res[0]["status"] == 0
    assert res[1]["status"] != 0 and res[1]["message"]


def test_tuned_kernel_equals_fixed():
    """The kernel shape is tuned on real iterations of the running batch; the
    shape (tile kernel or row-marching kernel, any size) does not change
    results."""
    _note("tuned vs fixed kernel shapes")
    tuned, st = _batch(REQS, "split")
    assert st["tunings"] >= 1 and st["kernel"] in (1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12)
    assert all(u > 0 for u in st["candUs"])
    for kw in ({"tileRows": 6}, {"tileRows": 30}, {"marchRows": 16}, {"marchRows": 128},
               {"persistRows": 6}, {"persistRows": 14}, {"march2Rows": 8}, {"march2Rows": 48},
               {"pairRows": 8}, {"pairRows": 40}, {"tripleRows": 8}, {"tripleRows": 42},
               {"quadRows": 8}, {"quadRows": 64}, {"wide3Rows": 8}, {"wide3Rows": 64},
               {"wide4Rows": 8}, {"wide4Rows": 64}):
        fixed, st2 = _batch(REQS, "split", **kw)
        assert st2["tunings"] == 0
        assert st2["kernelRows"] == list(kw.values())[0]
        for t in range(len(REQS)):
            _same(tuned[t]["result"], fixed[t]["result"])


@pytest.mark.parametrize("mode", ["fp64", "split"])
def test_march_batch_equals_single(mode):
    _note(f"march batch {mode}")
    got, st = _batch(REQS, mode, marchRows=32)
    assert st["kernel"] == 2
    for t, rq in enumerate(REQS):
        _same(got[t]["result"], _single(rq, mode))


def test_request_check_interval():
    """A request's own check interval overrides the engine's and gives the
    single-RVE result at that interval, whatever the others use."""
    _note("per-request check interval")
    nX, nP, phi, Y, nl, load = REQS[1]
    pa = make_pa(nX, nPores=nP, porosity=phi, Y=Y)
    with EffPlast2D(pa, nX, nTasks=nl, nIter=200_000, eIter=1e-8, checkInterval=300,
                    mode="split") as g:
        ref = g.ComputeEffModuli(load, LT)
    with Batch(mode="split", checkInterval=K) as b:
        for t, (n2, p2, f2, y2, l2, ld2) in enumerate(REQS):
            b.submit(make_pa(n2, nPores=p2, porosity=f2, Y=y2), n2, nTasks=l2, nIter=200_000,
                     eIter=1e-8, initLoad=ld2, loadType=LT, tag=t,
                     checkInterval=300 if t == 1 else 0)
        b.run()
        got = {r["tag"]: r for r in b.poll()}
    _same(got[1]["result"], ref)
    _same(got[0]["result"], _single(REQS[0], "split"))


def test_timing_with_graphs_counts_launches():
    _note("graph timing")
    _, st = _batch(REQS[:2], "fp64", timingEvery=1)
    assert st["timedLaunches"] >= st["chunks"] - st["tunings"] - 1
    assert st["timedMs"] > 0 and st["timedBytes"] > 0


def test_persistent_kernel_tiny_batch():
    """Fewer tiles than XCDs: most persistent blocks find no work and leave."""
    _note("persistent tiny")
    reqs = [(8, 1, 0.01, 2.5e-3, 3, -3e-3), (12, 1, 0.02, 2.5e-3, 1, -1e-4)]
    a, _ = _batch(reqs, "fp64", persistRows=14)
    b, _ = _batch(reqs, "fp64", tileRows=14)
    for t in range(len(reqs)):
        _same(a[t]["result"], b[t]["result"])


@pytest.mark.parametrize("mode", ["fp64", "split"])
@pytest.mark.parametrize("graphs", [True, False])
@pytest.mark.parametrize("shape", ["march2Rows", "pairRows", "tripleRows", "quadRows",
                                   "wide3Rows", "wide4Rows", "triple2Rows", "quad2Rows"])
def test_march2_batch_equals_single(mode, graphs, shape):
    """Multi-iteration launches in the engine (two iterations per launch on one
    wave or on a pair of waves per segment, three or four on a chain of
    stages with one or two waves per stage, or (split scheme) one wave with
    two columns per lane; then one to four single iterations per chunk; odd
    check intervals included) give the single-RVE results."""
    if shape.startswith("wide") and mode != "split":
        pytest.skip("the wide chains run the split scheme only")
    _note(f"{shape} batch {mode} graphs={graphs}")
    got, st = _batch(REQS, mode, useGraphs=graphs, timingEvery=1, **{shape: 32})
    assert st["kernel"] == {"march2Rows": 4, "pairRows": 5, "tripleRows": 6, "quadRows": 7,
                            "wide3Rows": 8, "wide4Rows": 9, "triple2Rows": 11,
                            "quad2Rows": 12}[shape]
    for t, rq in enumerate(REQS):
        _same(got[t]["result"], _single(rq, mode))
    nX, nP, phi, Y, nl, load = REQS[0]
    pa = make_pa(nX, nPores=nP, porosity=phi, Y=Y)
    with EffPlast2D(pa, nX, nTasks=nl, nIter=200_000, eIter=1e-8, checkInterval=199,
                    mode=mode) as g:
        ref = g.ComputeEffModuli(load, LT)
    with Batch(mode=mode, checkInterval=199, useGraphs=graphs, **{shape: 16}) as b:
        b.submit(pa, nX, nTasks=nl, nIter=200_000, eIter=1e-8, initLoad=load, loadType=LT, tag=0)
        b.run()
        r = b.poll()[0]
    _same(r["result"], ref)


@pytest.mark.parametrize("mode", ["fp64", "split"])
def test_adaptive_check_schedule(mode):
    """Residual checks placed by the decay rate: every load step ends below
    eIter, the solve takes at most one check interval more than the fixed
    schedule (in practice fewer), the moduli agree with the fixed schedule's
    to 1e-4, and the batch gives the single-RVE results with the same
    schedule."""
    _note(f"adaptive check {mode}")
    nX, nP, phi, Y, nl, load = REQS[0]
    pa = make_pa(nX, nPores=nP, porosity=phi, Y=Y)
    runs = {}
    for adaptive in (False, True):
        with EffPlast2D(pa, nX, nTasks=nl, nIter=200_000, eIter=1e-8, checkInterval=K,
                        mode=mode, adaptiveCheck=adaptive) as g:
            runs[adaptive] = (g.ComputeEffModuli(load, LT), g.records())
    (rf, recf), (ra, reca) = runs[False], runs[True]
    assert all(r["error"] < 1e-8 for r in reca)
    assert sum(a["iters"] for a in reca) <= sum(f["iters"] for f in recf) + K
    # both stop below eIter, at different iterations: the moduli agree to the
    # solver's own convergence (Kphi, a small difference quotient, to ~1e-5)
    for k in ("Kphi", "Kd"):
        assert ra["num"][k] == pytest.approx(rf["num"][k], rel=1e-4), k
    got, _ = _batch(REQS, mode, adaptiveCheck=True)
    _same(got[0]["result"], ra)


@pytest.mark.parametrize("mode", ["fp64", "split"])
def test_chains_full_size_equal_tiles(mode):
    """1024^2 (interior and edge segments, the headline size): 3 x 300 plastic
    iterations through the three- and four-stage chains, one or two waves per
    stage, give the tile kernel's records and moduli bit for bit."""
    _note(f"chains 1024 {mode}")
    nX = 1024
    pa = make_pa(nX, nPores=1, porosity=0.01, Y=2.5e-3)

    def run(**shape):
        with Batch(mode=mode, checkInterval=100, **shape) as b:
            b.submit(pa, nX, nTasks=3, nIter=300, eIter=1e-300, initLoad=-3e-3, loadType=LT,
                     tag=0)
            b.run()
            return b.poll()[0], b.stats()

    ref, _ = run(tileRows=14)
    kinds = {"tripleRows": 6, "quadRows": 7, "triple2Rows": 11, "quad2Rows": 12}
    for shape in ({"tripleRows": 86}, {"tripleRows": 42}, {"quadRows": 64}, {"quadRows": 256},
                  {"triple2Rows": 64}, {"quad2Rows": 64}, {"quad2Rows": 200}):
        got, st = run(**shape)
        assert st["kernel"] == kinds[next(iter(shape))]
        _same(got["result"], ref["result"])
        assert got["result"]["totalIters"] == 900


# ---- sampled requests against the oracle --------------------------------------------
def _sampled():
    """Eight seeded requests at 128^2 and 192^2, nTasks 1 to 3, preload along a
    random unit direction in (Exx, Eyy, Exy)."""
    return mixed_requests(8, 31, sizes=(128, 192), tasks=(1, 2, 3))


# Residual check intervals of the sampled requests.  A load step stops at
# the first check below eIter, and near eIter the pseudo-transient residual
# rises and falls between checks (in FP64 too), with the FP32 stress
# rounding on top in the split scheme.  At a coarse interval a split step
# can therefore pass several checks later than FP64 (tag 0 at 200
# iterations: 2600 against 1400 in its first step; 1440 against 1400 at 20),
# and at a fine one the two stop a few iterations apart, where Kphi, a small
# difference quotient, differs by the convergence error at eIter (tag 7 at
# 20: 1.7e-4; the FP64 solve itself moves that much between two stopping
# points below eIter).  Iteration counts are compared at 20, the moduli at
# 200, where both schemes stop at the same checks.
KS_ITERS, KS_MODULI = 20, 200


def _oracle_sampled(check):
    out = {}
    for rq in _sampled():
        s = oracle_state(rq["nX"], rq["pa"], rq["nTasks"], None, nIter=rq["nIter"],
                         eIter=rq["eIter"], outputStep=check)
        res, recs = s.compute_eff_moduli(rq["initLoad"], rq["loadType"], 1)
        out[rq["tag"]] = (res, [r for st in recs for r in st])
    return out


@pytest.fixture(scope="module")
def sampled_oracle():
    return {k: _oracle_sampled(k) for k in (KS_ITERS, KS_MODULI)}


def _run_sampled(mode, check, **kw):
    out = {}
    with Batch(mode=mode, checkInterval=check, **kw) as b:
        for rq in _sampled():
            b.submit(**rq)
        b.run()
        for r in b.poll():
            out[r["tag"]] = r
    return out


def test_sampled_requests_have_shear_preloads():
    reqs = _sampled()
    assert sum(abs(rq["loadType"][2]) > 0.1 for rq in reqs) >= 4
    assert {rq["nTasks"] for rq in reqs} == {1, 2, 3}
    assert len({rq["nX"] for rq in reqs}) == 2


@pytest.mark.parametrize("check", [KS_ITERS, KS_MODULI])
def test_fp64_batch_matches_oracle(sampled_oracle, check):
    got = _run_sampled("fp64", check)
    for tag, (res_o, recs_o) in sampled_oracle[check].items():
        r = got[tag]
        assert r["status"] == 0, r["message"]
        res_g = r["result"]
        _note(f"fp64 batch check={check} tag={tag} iters={res_g['totalIters']}")
        assert res_g["totalIters"] == sum(int(x["iters"]) for x in recs_o)
        for grp in ("num", "num_per", "an"):
            for k in ("Kphi", "Kd", "G"):
                a, b = res_g[grp][k], res_o[grp][k]
                if np.isfinite(b) and b != 0:
                    assert a == pytest.approx(b, rel=1e-8), (tag, grp, k)


def test_split_batch_iterations_close_to_fp64_oracle(sampled_oracle):
    """Iteration counts within 10% of the FP64 oracle's, plus one check
    interval per load step (checks every 20 iterations)."""
    got = _run_sampled("split", KS_ITERS)
    for tag, (res_o, recs_o) in sampled_oracle[KS_ITERS].items():
        res_g = got[tag]["result"]
        it_o = sum(int(x["iters"]) for x in recs_o)
        _note(f"split batch tag={tag} iters {res_g['totalIters']} (FP64 oracle {it_o})")
        assert abs(res_g["totalIters"] - it_o) <= 0.1 * it_o + KS_ITERS * len(recs_o), tag


def test_split_batch_moduli_close_to_fp64_oracle(sampled_oracle):
    """Converged requests: Kd, Kphi and G within 1e-4 of the FP64 oracle
    (checks every 200 iterations)."""
    got = _run_sampled("split", KS_MODULI)
    checked = 0
    for tag, (res_o, recs_o) in sampled_oracle[KS_MODULI].items():
        res_g = got[tag]["result"]
        if not (res_g["converged"] and all(x["error"] < 1e-8 for x in recs_o)):
            continue
        checked += 1
        for k in ("Kd", "Kphi", "G"):
            a, b = res_g["num"][k], res_o["num"][k]
            if np.isfinite(b) and b != 0:
                e = abs(a - b) / abs(b)
                _note(f"split batch tag={tag} {k}: rel {e:.2e}")
                assert e <= 1e-4, (tag, k)
    assert checked >= 4


def test_serving_byte_budget_admission_and_backfill():
    """Continuous batching under a device-memory budget in bytes that holds two
    RVEs: two requests start, two more arrive after two chunks while the first
    two run; the late ones are admitted only as memory frees up, the engine
    never holds more than the budget, and every result equals the request's
    solve alone."""
    reqs = mixed_requests(4, 77, sizes=(96, 128), tasks=(1, 2, 3))
    need = [request_bytes("split", rq["pa"], rq["nX"]) for rq in reqs]
    budget = sorted(need)[-1] + sorted(need)[-2] + min(need) // 2
    _note(f"serving budget {budget} bytes, needs {need}")
    done, peak = {}, 0
    with Batch(mode="split", checkInterval=K, memoryBudget=budget) as b:
        for rq in reqs[:2]:
            b.submit(**rq)
        steps = 0
        while len(done) < len(reqs):
            if steps == 2:
                for rq in reqs[2:]:
                    b.submit(**rq)
            active, queued = b.step()
            st = b.stats()
            assert st["deviceBytes"] <= budget
            assert active <= 2
            peak = max(peak, active)
            for r in b.poll():
                done[r["tag"]] = r
            steps += 1
            assert steps < 10_000
    assert peak == 2
    first_done = min(done[0]["tDone"], done[1]["tDone"])
    for t in (2, 3):
        assert done[t]["tSubmit"] > done[0]["tAdmit"]       # arrived mid-flight
        assert done[t]["tAdmit"] >= first_done               # waited for memory
    for rq in reqs:
        alone = _run_alone(rq)
        _same(done[rq["tag"]]["result"], alone)


def _run_alone(rq, mode="split"):
    with EffPlast2D(rq["pa"], rq["nX"], nTasks=rq["nTasks"], nIter=rq["nIter"],
                    eIter=rq["eIter"], checkInterval=K, mode=mode) as g:
        return g.ComputeEffModuli(rq["initLoad"], rq["loadType"])


@pytest.mark.parametrize("mode", ["split", "fp64"])
def test_cut_work_lists_equal_single(mode):
    """A batch whose chain work list spans more than one round of the GPU:
    24 RVEs of 161^2 on 8-row segments of the two-wave four-stage chains
    (42 segments each, 1008 in all; each XCD's share of 126 is three RVEs).
    The list is reordered and the end of each share cut into half- and
    quarter-height segments; the last band of each RVE has one row, so its
    segments are listed several times instead of cut.  Every result equals
    the request's solve alone (tile kernel) bit for bit."""
    _note(f"cut work lists {mode}")
    reqs = []
    for t in range(24):
        pa = make_pa(161, nPores=1 + t % 3, porosity=0.01 + 0.002 * (t % 5), Y=2.5e-3 + 1e-4 * t)
        reqs.append(dict(pa=pa, nX=161, nTasks=3, nIter=600, eIter=1e-300,
                         initLoad=-2e-3 - 1e-4 * (t % 4), loadType=LT, tag=t))
    with Batch(mode=mode, checkInterval=K, quad2Rows=8) as b:
        for rq in reqs:
            b.submit(**rq)
        b.run()
        got = {r["tag"]: r for r in b.poll()}
        st = b.stats()
    assert st["kernel"] == 12 and st["kernelRows"] == 8
    for rq in reqs:
        assert got[rq["tag"]]["status"] == 0
        _same(got[rq["tag"]]["result"], _run_alone(rq, mode))


def test_admission_estimate_is_what_an_rve_holds():
    """The bytes counted at admission are the bytes an admitted RVE holds."""
    reqs = mixed_requests(3, 5, sizes=(64, 96, 160))
    with Batch(mode="fp64", checkInterval=K) as b:
        for rq in reqs:
            b.submit(**rq)
        b.step()
        st = b.stats()
        assert st["active"] == 3
        assert st["deviceBytes"] == sum(request_bytes("fp64", rq["pa"], rq["nX"]) for rq in reqs)


def test_split_stall_finishes_in_fp64_arithmetic():
    """A sampled request whose residual normaliser max(|e0|, max(e1, e2)) is
    37 times below the load's largest strain (tag 4 of the sampled set): the
    FP32 stress noise holds the split residual near 3e-7, above eIter =
    1e-8.  With the FP64 finish (the default) every time step reaches eIter
    and the moduli stay within 1e-4 of the FP64 solve; with the FP32 stress
    kept, the steps run to the iteration cap.  In the batched engine, beside
    a request that does not stall, the FP64 finish runs on the chain kernels
    and gives the single handle's bits."""
    rq = _sampled()[4]
    assert abs(rq["initLoad"] * rq["loadType"][0]) < abs(rq["initLoad"] * rq["loadType"][1]) / 30
    _note("split stall")
    out = {}
    for mode, keep in (("fp64", False), ("split", False), ("split", True)):
        with EffPlast2D(rq["pa"], rq["nX"], nTasks=rq["nTasks"], nIter=4 * rq["nIter"],
                        eIter=rq["eIter"], checkInterval=K, mode=mode,
                        keepFp32Stress=keep) as g:
            res = g.ComputeEffModuli(rq["initLoad"], rq["loadType"])
            out[(mode, keep)] = (res, g.records())
    res64, recs64 = out[("fp64", False)]
    res_s, recs_s = out[("split", False)]
    res_k, recs_k = out[("split", True)]
    assert all(r["error"] < rq["eIter"] for r in recs64)
    assert all(r["error"] < rq["eIter"] for r in recs_s)
    assert not all(r["error"] < rq["eIter"] for r in recs_k)
    for k in ("Kd", "Kphi", "G"):
        e = abs(res_s["num"][k] - res64["num"][k]) / abs(res64["num"][k])
        _note(f"  {k}: rel {e:.2e}")
        assert e <= 1e-4, k
    # in the engine, beside a request that does not stall, the stalled steps
    # finish on the chain kernels in the FP64 arithmetic (the single handle
    # runs the tile kernel there): the same bits
    other = _sampled()[0]
    with Batch(mode="split", checkInterval=K) as b:
        for q in (rq, other):
            b.submit(q["pa"], q["nX"], nTasks=q["nTasks"], nIter=4 * q["nIter"],
                     eIter=q["eIter"], initLoad=q["initLoad"], loadType=q["loadType"],
                     tag=q["tag"])
        b.run()
        got = {r["tag"]: r["result"] for r in b.poll()}
        st = b.stats()
    assert st["fp64Finishes"] >= 1 and st["fp64CellIterations"] > 0
    _same(got[rq["tag"]], res_s)
    # a single handle on a multi-iteration shape finishes on that shape
    for shape in (dict(marchRows=16, itersPerLaunch=4, stageWaves=2),
                  dict(marchRows=16, itersPerLaunch=3),
                  dict(marchRows=14, itersPerLaunch=2)):
        with EffPlast2D(rq["pa"], rq["nX"], nTasks=rq["nTasks"], nIter=4 * rq["nIter"],
                        eIter=rq["eIter"], checkInterval=K, mode="split", **shape) as g:
            _same(g.ComputeEffModuli(rq["initLoad"], rq["loadType"]), res_s)

