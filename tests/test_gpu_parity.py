"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle on
identical seeded inputs.

Tolerances (written per test):
  * plain and fp64 schemes evaluate the reference expressions in the same
    order with IEEE division/sqrt and no contraction: fields after a fixed
    number of iterations agree to <= 1e-12 relative (max-norm) -- in practice
    bit for bit -- and the masks, which are integer, exactly.
  * split scheme (FP32 stress/velocity, FP64 displacement): fields after a
    fixed number of iterations within 1e-4 relative (max-norm); the converged
    moduli Kd, Kphi and G within 1e-4 relative of the FP64 solve, field
    relative L2 <= 1e-3, iteration counts within 10%.
"""
import sys

import numpy as np
import pytest
# torch before the library: torch's HIP libraries name the runtime without
# its version, so once the library has loaded /opt/rocm's runtime, a later
# torch import loads a second runtime and finds no GPU; imported first, the
# library binds to torch's runtime and the process has one
import torch

from effplast2d_amd import EffPlast2D
from oracle import ora
from tests.pairs import (PLASTIC_FIELDS, STATE_FIELDS, STRESS_FIELDS, l2_rel, make_pa,
                         oracle_state, random_state, rel_err)

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("nX,nPores,porosity", [(32, 1, 0.01), (64, 2, 0.04), (100, 3, 0.05),
                                                (128, 4, 0.05), (128, 1, 0.005)])
def test_mask_bit_exact(nX, nPores, porosity):
    _note(f"mask nX={nX} nPores={nPores}")
    pa = make_pa(nX, nPores=nPores, porosity=porosity)
    s = oracle_state(nX, pa, nl=3)
    with EffPlast2D(pa, nX, mode="fp64") as g:
        m = g.mask()
        K = g.get("K")
        G = g.get("G")
    assert np.array_equal((m & 1).astype(bool), s["K"] < pa[ora.PA_K0])
    assert np.array_equal(K, s["K"])
    assert np.array_equal(G, s["G"])
    assert np.array_equal(((m >> 1) & 1).astype(np.uint8), s["cellPore"])
    vert = ((m >> 2) & 1).reshape(nX, nX)[: nX - 1, : nX - 1].ravel()
    assert np.array_equal(vert.astype(np.uint8), s["vertPore"])


@pytest.mark.parametrize("mode", ["plain", "fp64"])
@pytest.mark.parametrize("nX,nPores,nl,iters,seed", [
    (32, 1, 3, 1, 0), (64, 2, 3, 7, 1), (72, 1, 1, 5, 2), (130, 3, 3, 50, 3),
    (128, 4, 3, 20, 4), (96, 1, 2, 3, 5)])
def test_iterations_match_oracle(mode, nX, nPores, nl, iters, seed):
    _note(f"iterate mode={mode} nX={nX} nPores={nPores} nl={nl} iters={iters} seed={seed}")
    pa = make_pa(nX, nPores=nPores, porosity=0.02)
    init = random_state(nX, seed, pa=pa, boundary_v=(seed % 2 == 1))
    s = oracle_state(nX, pa, nl, init)
    s.iterate(iters)
    with EffPlast2D(pa, nX, nTasks=nl, mode=mode) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        fields = STATE_FIELDS + STRESS_FIELDS + (PLASTIC_FIELDS if nl > 1 else [])
        worst = 0.0
        for k in fields:
            got = g.get(k)
            e = rel_err(got, s[k])
            worst = max(worst, e)
            assert e <= 1e-12, f"{k}: rel err {e:.3e} ({np.count_nonzero(got != s[k])} differ)"
        _note(f"  worst rel err {worst:.3e}")


@pytest.mark.parametrize("mode", ["fp64", "split"])
@pytest.mark.parametrize("ty", [6, 8, 14, 30])
def test_tile_heights_match_oracle(mode, ty):
    """Every fused tile shape, on a grid with interior and edge tiles."""
    nX, nPores, nl, iters = 200, 2, 3, 15
    _note(f"tile mode={mode} ty={ty}")
    pa = make_pa(nX, nPores=nPores, porosity=0.03)
    init = random_state(nX, 100 + ty, pa=pa, boundary_v=True)
    if mode == "split":
        init["Vx"] = init["Vx"].astype(np.float32).astype(np.float64)
        init["Vy"] = init["Vy"].astype(np.float32).astype(np.float64)
    s = oracle_state(nX, pa, nl, init)
    s.iterate(iters)
    with EffPlast2D(pa, nX, nTasks=nl, mode=mode, tileRows=ty) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        for k in STATE_FIELDS + STRESS_FIELDS:
            got = g.get(k)
            if mode == "fp64":
                e = rel_err(got, s[k])
                _note(f"  {k}: rel {e:.2e}, {np.count_nonzero(got != s[k])} differ")
                assert e <= 1e-12, k
            else:
                e = l2_rel(got, s[k])
                assert e <= (1e-6 if k in ("Ux", "Uy") else 1e-4), f"{k}: {e:.3e}"


@pytest.mark.parametrize("rows", [1, 2, 3, 16, 64])
@pytest.mark.parametrize("nX,nPores", [(200, 2), (61, 1), (125, 3), (8, 1)])
def test_march_matches_oracle(rows, nX, nPores):
    """Row-marching kernel, FP64: every segment length (1 row: every segment
    is all halo) on grids with one and several strips, edge and interior
    segments, bit for bit against the oracle."""
    nl, iters = 3, 12
    _note(f"march fp64 rows={rows} nX={nX}")
    pa = make_pa(nX, nPores=nPores, porosity=0.03)
    init = random_state(nX, 300 + rows + nX, pa=pa, boundary_v=True)
    s = oracle_state(nX, pa, nl, init)
    s.iterate(iters)
    with EffPlast2D(pa, nX, nTasks=nl, mode="fp64", marchRows=rows) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        for k in STATE_FIELDS + STRESS_FIELDS + PLASTIC_FIELDS:
            got = g.get(k)
            e = rel_err(got, s[k])
            assert e <= 1e-12, f"{k}: rel err {e:.3e} ({np.count_nonzero(got != s[k])} differ)"


@pytest.mark.parametrize("rows,iters", [(1, 7), (3, 12), (16, 1), (16, 2), (64, 9)])
@pytest.mark.parametrize("nX,nPores", [(200, 2), (61, 1), (125, 3), (8, 1)])
def test_march2_matches_oracle(rows, iters, nX, nPores):
    """Two iterations per launch (temporal blocking, each segment on a pair of
    waves), FP64: odd and even iteration counts, every segment length, bit for
    bit against the oracle; the stress fields come from U^{n-1}, which the
    last (single) launch read."""
    nl = 3
    _note(f"march2 fp64 rows={rows} iters={iters} nX={nX}")
    pa = make_pa(nX, nPores=nPores, porosity=0.03)
    init = random_state(nX, 500 + rows + nX, pa=pa, boundary_v=True)
    s = oracle_state(nX, pa, nl, init)
    s.iterate(iters)
    with EffPlast2D(pa, nX, nTasks=nl, mode="fp64", marchRows=rows, itersPerLaunch=2) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        for k in STATE_FIELDS + STRESS_FIELDS + PLASTIC_FIELDS:
            got = g.get(k)
            e = rel_err(got, s[k])
            assert e <= 1e-12, f"{k}: rel err {e:.3e} ({np.count_nonzero(got != s[k])} differ)"


@pytest.mark.parametrize("waves", [1, 2])
@pytest.mark.parametrize("ipl", [3, 4])
@pytest.mark.parametrize("rows,iters", [(1, 7), (5, 13), (16, 3), (64, 10), (24, 8)])
@pytest.mark.parametrize("nX,nPores", [(200, 2), (61, 1), (8, 1), (300, 3)])
def test_chain_matches_oracle(ipl, rows, iters, nX, nPores, waves):
    """Three or four iterations per launch on a chain of stages, one or two
    waves per stage, FP64: every remainder of the iteration count (the chain
    launches, then one to four closing iterations, three of which run as a
    wave-pair launch and four as a three-stage launch before the single
    one), short and long segments,
    strips that all touch a boundary (8^2, 61^2, 200^2) and interior ones
    (300^2), bit for bit against the oracle."""
    nl = 3
    _note(f"chain{ipl}x{waves} fp64 rows={rows} iters={iters} nX={nX}")
    pa = make_pa(nX, nPores=nPores, porosity=0.03)
    init = random_state(nX, 900 + rows + nX + ipl, pa=pa, boundary_v=True)
    s = oracle_state(nX, pa, nl, init)
    s.iterate(iters)
    with EffPlast2D(pa, nX, nTasks=nl, mode="fp64", marchRows=rows, itersPerLaunch=ipl,
                    stageWaves=waves) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        for k in STATE_FIELDS + STRESS_FIELDS + PLASTIC_FIELDS:
            got = g.get(k)
            e = rel_err(got, s[k])
            assert e <= 1e-12, f"{k}: rel err {e:.3e} ({np.count_nonzero(got != s[k])} differ)"


@pytest.mark.parametrize("nl", [1, 3])
def test_march2_split_equals_tiles(nl):
    nX, iters = 190, 21
    _note(f"march2 split nl={nl}")
    pa = make_pa(nX, nPores=2, porosity=0.04)
    init = random_state(nX, 700 + nl, pa=pa, boundary_v=True)
    out = []
    for kw in ({"marchRows": 24, "itersPerLaunch": 2}, {"tileRows": 8}):
        with EffPlast2D(pa, nX, nTasks=nl, mode="split", **kw) as g:
            for k in STATE_FIELDS:
                g.set(k, init[k])
            g.iterate(iters)
            g.iterate(iters + 1)
            out.append({k: g.get(k) for k in STATE_FIELDS + STRESS_FIELDS})
    for k in STATE_FIELDS + STRESS_FIELDS:
        assert np.array_equal(out[0][k], out[1][k]), \
            f"{k}: {np.count_nonzero(out[0][k] != out[1][k])} differ"


@pytest.mark.parametrize("rows", [2, 32])
@pytest.mark.parametrize("nl", [1, 3])
def test_march_split_equals_tiles(rows, nl):
    """The split scheme through the row-marching kernel and through the tile
    kernel: the same expressions, so identical fields."""
    nX, iters = 190, 20
    _note(f"march split rows={rows} nl={nl}")
    pa = make_pa(nX, nPores=2, porosity=0.04)
    init = random_state(nX, 400 + rows, pa=pa, boundary_v=True)
    out = []
    for kw in ({"marchRows": rows}, {"tileRows": 8}):
        with EffPlast2D(pa, nX, nTasks=nl, mode="split", **kw) as g:
            for k in STATE_FIELDS:
                g.set(k, init[k])
            g.iterate(iters)
            out.append({k: g.get(k) for k in STATE_FIELDS})
    for k in STATE_FIELDS:
        assert np.array_equal(out[0][k], out[1][k]), \
            f"{k}: {np.count_nonzero(out[0][k] != out[1][k])} differ"


@pytest.mark.parametrize("nX,nPores,nl,iters,seed", [(64, 1, 3, 10, 11), (130, 3, 3, 40, 12),
                                                     (96, 2, 1, 25, 13)])
def test_split_iterations_close_to_oracle(nX, nPores, nl, iters, seed):
    _note(f"split iterate nX={nX} nPores={nPores} nl={nl} iters={iters}")
    pa = make_pa(nX, nPores=nPores, porosity=0.02)
    init = random_state(nX, seed, pa=pa)
    # the split scheme stores V in FP32: give the oracle the same rounded start
    init["Vx"] = init["Vx"].astype(np.float32).astype(np.float64)
    init["Vy"] = init["Vy"].astype(np.float32).astype(np.float64)
    s = oracle_state(nX, pa, nl, init)
    s.iterate(iters)
    with EffPlast2D(pa, nX, nTasks=nl, mode="split") as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        for k in ["Ux", "Uy"]:
            assert rel_err(g.get(k), s[k]) <= 1e-6, k
        for k in ["Vx", "Vy"] + STRESS_FIELDS:
            e = l2_rel(g.get(k), s[k])
            assert e <= 1e-4, f"{k}: L2 rel err {e:.3e}"


@pytest.mark.parametrize("mode", ["plain", "fp64"])
def test_affine_residual_averages_boundary(mode):
    nX, nl = 96, 3
    _note(f"pieces mode={mode}")
    pa = make_pa(nX, nPores=3, porosity=0.05)
    init = random_state(nX, 21, pa=pa)
    s = oracle_state(nX, pa, nl, init)
    with EffPlast2D(pa, nX, nTasks=nl, mode=mode) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        # affine increment (EffPlast2D.cu:370-381): bit for bit
        g.apply_affine(-1e-3, -2e-3, 5e-4, 5e-4)
        s.apply_affine(-1e-3, -2e-3, 5e-4, 5e-4)
        assert np.array_equal(g.get("Ux"), s["Ux"])
        assert np.array_equal(g.get("Uy"), s["Uy"])
        g.iterate(9)
        s.iterate(9)
        cur = [-3e-3, -2e-3, 1e-3]
        err, bad = g.residual(cur)
        assert bad == 0
        assert err == s.residual(cur)     # max is exact, the formula is host FP64
        a_g, a_o = g.averages(), s.averages()
        for k in ["sigma", "sigmaPer", "epsilonPer"]:
            np.testing.assert_allclose(a_g[k], a_o[k], rtol=1e-11, atol=1e-16, err_msg=k)
        np.testing.assert_allclose(a_g["sigmaXYclean"], a_o["sigmaXYclean"], rtol=1e-11)
        # boundary means and hole area: host sums in the reference order over
        # gathered values, so they are exact whenever the fields are
        b_g, b_o = g.boundary(), s.boundary()
        for k in b_o:
            assert b_g[k] == pytest.approx(b_o[k], rel=1e-12, abs=1e-300), k
            _note(f"  {k} exact={b_g[k] == b_o[k]}")
        h_g, h_o = g.hole_area(), s.hole_area()
        for k in ["holeAreaPi", "internalHoleAreaPi", "Phi0", "Phi", "PhiPer", "dPhi", "dPhiPer"]:
            assert h_g[k] == pytest.approx(h_o[k], rel=1e-12, abs=1e-300), k


def test_nonfinite_velocity_reported():
    nX = 64
    pa = make_pa(nX)
    init = random_state(nX, 3, pa=pa)
    init["Vx"][5 * (nX + 1) + 7] = np.nan
    with EffPlast2D(pa, nX, nTasks=3, mode="fp64") as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        err, bad = g.residual([-1e-3, -1e-3, 0.0])
        assert bad == 1


MODULI_CASES = [
    # nX, nPores, porosity, Y, nl, load, loadType, checkInterval
    (64, 1, 0.01, 2.5e-3, 3, -3e-3, [1.0, 1.0, 0.0], 100),
    (64, 1, 0.005, 2.5e-3, 1, -1e-5, [1.0, 1.0, 0.0], 100),
    (96, 3, 0.05, 2.5e-3, 3, -3e-3, [1.0, 0.6, 0.2], 200),
    (64, 2, 0.03, 3e-3, 2, -2e-3, [0.8, 1.0, 0.0], 100),
    # the 128^2 elastic RVE with one pore of porosity 0.005 under -1e-5
    # isotropic preload (nGrid = 4), checked every 500 iterations
    (128, 1, 0.005, 2.5e-3, 1, -1e-5, [1.0, 1.0, 0.0], 500),
]


def _moduli_pair(nX, nPores, porosity, Y, nl, load, lt, K, mode, shear=None):
    pa = make_pa(nX, nPores=nPores, porosity=porosity, Y=Y)
    s = oracle_state(nX, pa, nl, None, nIter=200_000, eIter=1e-8, outputStep=K,
                     shearLoadType=shear)
    res_o, recs_o = s.compute_eff_moduli(load, lt, 1)
    with EffPlast2D(pa, nX, nTasks=nl, nIter=200_000, eIter=1e-8, checkInterval=K, mode=mode,
                    shearLoadType=shear) as g:
        res_g = g.ComputeEffModuli(load, lt)
        recs_g = g.records()
        fields = {k: g.get(k) for k in ["P", "tauXX", "Ux", "Uy"]}
    return res_o, recs_o, res_g, recs_g, fields, s


@pytest.mark.parametrize("mode", ["plain", "fp64"])
@pytest.mark.parametrize("case", MODULI_CASES)
def test_moduli_match_oracle(mode, case):
    _note(f"moduli mode={mode} case={case}")
    res_o, recs_o, res_g, recs_g, fields, s = _moduli_pair(*case, mode)
    flat_o = [r for st in recs_o for r in st]
    assert len(recs_g) == len(flat_o)
    for rg, ro in zip(recs_g, flat_o):
        assert rg["iters"] == ro["iters"]
        for k in ["deltaP", "deltaPper", "tauInfty", "dPhi", "holeAreaPi"]:
            assert rg[k] == pytest.approx(ro[k], rel=1e-10, abs=1e-300), k
        np.testing.assert_allclose(rg["sigma"], ro["sigma"], rtol=1e-10, atol=1e-18)
    for grp in ["num", "an"]:
        for k in ["Kphi", "Kd", "G"]:
            a, b = res_g[grp][k], res_o[grp][k]
            if np.isfinite(b) and b != 0:
                assert a == pytest.approx(b, rel=1e-8), (grp, k)
    for k in ["P", "tauXX", "Ux", "Uy"]:
        assert rel_err(fields[k], s[k]) <= 1e-12, k


@pytest.mark.parametrize("case", MODULI_CASES)
def test_split_moduli_close_to_fp64_oracle(case):
    _note(f"split moduli case={case}")
    res_o, recs_o, res_g, recs_g, fields, s = _moduli_pair(*case, "split")
    flat_o = [r for st in recs_o for r in st]
    for rg, ro in zip(recs_g, flat_o):
        assert abs(rg["iters"] - ro["iters"]) <= 0.1 * ro["iters"] + case[-1]
    # relative error of Kd, Kphi and G against the FP64 oracle <= 1e-4
    for k in ["Kd", "G", "Kphi"]:
        b = res_o["num"][k]
        if np.isfinite(b) and b != 0:
            assert res_g["num"][k] == pytest.approx(b, rel=1e-4), k
            _note(f"  {k}: rel {abs(res_g['num'][k] - b) / abs(b):.2e}")
    for k in ["P", "tauXX", "Ux", "Uy"]:
        assert l2_rel(fields[k], s[k]) <= 1e-3, k


@pytest.mark.parametrize("kw", [{"marchRows": 26, "itersPerLaunch": 2}, {"tileRows": 8},
                                {"marchRows": 56, "itersPerLaunch": 3},
                                {"marchRows": 16, "itersPerLaunch": 4}],
                         ids=["pairs26", "tiles8", "chain3x56", "chain4x16"])
def test_full_size_fp64_matches_oracle(kw):
    """The headline's RVE at full size (1024^2, one pore, porosity 0.01,
    Y = 2.5e-3, G0 = 0.5, three tasks): from a noisy affine compression, 101
    plastic iterations (an odd count: two-iteration launches then one single),
    FP64, every field against the oracle (OpenMP) at <= 1e-12 relative, bit
    for bit in practice."""
    import os
    nX, iters = 1024, 101
    _note(f"full size fp64 {kw} iters={iters}")
    pa = make_pa(nX, nPores=1, porosity=0.01)
    init = random_state(nX, 4242, pa=pa, boundary_v=True)
    s = oracle_state(nX, pa, 3, init, nThreads=min(16, os.cpu_count() or 1))
    s.iterate(iters)
    with EffPlast2D(pa, nX, nTasks=3, mode="fp64", **kw) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        for k in STATE_FIELDS + STRESS_FIELDS + PLASTIC_FIELDS:
            got = g.get(k)
            e = rel_err(got, s[k])
            assert e <= 1e-12, f"{k}: rel err {e:.3e} ({np.count_nonzero(got != s[k])} differ)"
    # the run must have exercised the return mapping
    assert np.count_nonzero(s["J2"] > pa[ora.PA_COH] * (1 - 1e-12)) > 0


def test_full_size_split_pairs_equal_tiles():
    """Split scheme at 1024^2: the two-iteration kernel on wave pairs and the
    three- and four-iteration chains give the tile kernel's fields bit for bit
    (201 iterations)."""
    nX, iters = 1024, 201
    _note(f"full size split pairs and chains vs tiles iters={iters}")
    pa = make_pa(nX, nPores=1, porosity=0.01)
    init = random_state(nX, 4243, pa=pa, boundary_v=True)
    out = []
    for kw in ({"tileRows": 8}, {"marchRows": 26, "itersPerLaunch": 2},
               {"marchRows": 56, "itersPerLaunch": 3}, {"marchRows": 16, "itersPerLaunch": 4}):
        with EffPlast2D(pa, nX, nTasks=3, mode="split", **kw) as g:
            for k in STATE_FIELDS:
                g.set(k, init[k])
            g.iterate(iters)
            out.append({k: g.get(k) for k in STATE_FIELDS + STRESS_FIELDS})
    for o in out[1:]:
        for k in STATE_FIELDS + STRESS_FIELDS:
            assert np.array_equal(o[k], out[0][k]), \
                f"{k}: {np.count_nonzero(o[k] != out[0][k])} differ"


@pytest.mark.parametrize("f32", [True, False])
def test_residual_vector_and_element_reads_agree(f32):
    """The residual kernel reads velocity fields that start 16-byte aligned
    as 16-byte vectors and others element by element: on the same values
    (fields of (nX+1) nX elements, not a whole number of vectors, two NaNs
    planted) both give the same residual bits and non-finite count, and the
    largest value, put in the last element past the vectors, is seen."""
    import ctypes as C
    from effplast2d_amd._lib import check, lib
    dt = torch.float32 if f32 else torch.float64
    stream = torch.cuda.current_stream().cuda_stream
    e = (C.c_double * 3)(-2e-3, -1e-3, 5e-4)

    def residual(vx, vy, nX):
        err = torch.zeros(1, dtype=torch.float64, device="cuda")
        bad = torch.zeros(1, dtype=torch.int64, device="cuda")
        check(lib().ep2d_residual_noalloc(
            C.c_void_p(vx.data_ptr()), C.c_void_p(vy.data_ptr()), nX, nX, 1 if f32 else 0,
            C.c_double(0.02), C.c_double(0.02), C.c_double(0.007), e,
            C.c_void_p(err.data_ptr()), C.c_void_p(bad.data_ptr()), C.c_void_p(stream)))
        torch.cuda.current_stream().synchronize()
        return float(err.item()), int(bad.item())

    for nX in (97, 1023):
        n = (nX + 1) * nX
        _note(f"residual reads f32={f32} nX={nX}")
        gen = torch.Generator().manual_seed(nX)
        vals = torch.randn(2, n, generator=gen, dtype=torch.float64)
        vals[0, n - 1] = 40.0             # the largest |Vx|, past the last whole vector
        vals[1, 0] = -30.0                # the largest |Vy|, in the first vector
        vals[0, 7] = float("nan")
        vals[1, n // 2] = float("nan")
        buf = torch.zeros(2, (n + 11) // 4 * 4, dtype=dt, device="cuda")
        assert buf.data_ptr() % 16 == 0 and buf[1].data_ptr() % 16 == 0
        out = {}
        for off in (0, 1, 3):             # 16-byte aligned, then element-aligned only
            buf.zero_()
            buf[:, off:off + n] = vals.to(dt).cuda()
            out[off] = residual(buf[0, off:off + n], buf[1, off:off + n], nX)
        assert out[0] == out[1] == out[3], out
        assert out[0][1] == 2
        vals[0, n - 1] = 0.0
        buf.zero_()
        buf[:, :n] = vals.to(dt).cuda()
        smaller = residual(buf[0, :n], buf[1, :n], nX)
        assert smaller[0] < out[0][0], (smaller, out[0])
