"""GPU: the wide chains (two columns per lane, split scheme) and the chains
with two waves per stage against the chains of one wave per stage and the
tile kernel.

Each wide stage computes every element with the split scheme's functions in
the same order as the one-column chains, so the fields must agree bit for
bit: after any number of iterations (chain launches, then one to four single
iterations), for segments of every height, on grids where every strip touches
a boundary (8^2, 61^2), where interior strips exist (200^2, 300^2), with pores
inside segments (the mask path) and away from them (the mask-free path), and
at 1024^2 and 4096^2 with 16 pores.  Tolerance: exact equality.
"""
import sys

import numpy as np
import pytest

from effplast2d_amd import Batch, EffPlast2D
from tests.pairs import PLASTIC_FIELDS, STATE_FIELDS, STRESS_FIELDS, make_pa, random_state

pytestmark = pytest.mark.gpu

LT = [1.0, 1.0, 0.0]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fields(pa, nX, nl, init, iters, **kw):
    with EffPlast2D(pa, nX, nTasks=nl, mode="split", **kw) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        out = {k: g.get(k) for k in STATE_FIELDS + STRESS_FIELDS}
        if nl > 1:
            out.update({k: g.get(k) for k in PLASTIC_FIELDS})
    return out


def _same_fields(a, b):
    for k in b:
        assert np.array_equal(a[k], b[k]), f"{k}: {np.count_nonzero(a[k] != b[k])} differ"


@pytest.mark.parametrize("ipl", [3, 4])
@pytest.mark.parametrize("rows,iters", [(1, 7), (5, 13), (16, 3), (64, 10)])
@pytest.mark.parametrize("nX,nPores", [(200, 2), (61, 1), (8, 1), (300, 3)])
def test_wide_equals_chain(ipl, rows, iters, nX, nPores):
    nl = 3
    _note(f"wide{ipl} rows={rows} iters={iters} nX={nX} nPores={nPores}")
    pa = make_pa(nX, nPores=nPores, porosity=0.03)
    init = random_state(nX, 1300 + rows + nX + ipl, pa=pa, boundary_v=True)
    one = _fields(pa, nX, nl, init, iters, marchRows=rows, itersPerLaunch=ipl, laneColumns=1)
    two = _fields(pa, nX, nl, init, iters, marchRows=rows, itersPerLaunch=ipl, laneColumns=2)
    _same_fields(two, one)


@pytest.mark.parametrize("nl", [1, 3])
def test_wide_equals_tiles(nl):
    """Elastic and plastic, against the tile kernel (a different kernel
    family with the same split arithmetic)."""
    nX, iters = 190, 22
    _note(f"wide vs tiles nl={nl}")
    pa = make_pa(nX, nPores=2, porosity=0.04)
    init = random_state(nX, 1700 + nl, pa=pa, boundary_v=True)
    tiles = _fields(pa, nX, nl, init, iters, tileRows=8)
    wide = _fields(pa, nX, nl, init, iters, marchRows=24, itersPerLaunch=3, laneColumns=2)
    _same_fields(wide, tiles)


@pytest.mark.parametrize("other", [{"laneColumns": 2}, {"stageWaves": 2}],
                         ids=["wide", "two_waves"])
def test_wide_4096_sixteen_pores_equals_chain(other):
    """4096^2 with 4 x 4 pores, porosity 0.05: interior segments with and
    without pore cells, 30 plastic iterations; the wide chains and the chains
    with two waves per stage against the chains of one wave per stage."""
    nX, iters = 4096, 30
    _note(f"4096 16 pores {other}")
    pa = make_pa(nX, nPores=4, porosity=0.05, Y=2.5e-3)
    init = random_state(nX, 4096, pa=pa, boundary_v=True)
    one = _fields(pa, nX, 3, init, iters, marchRows=56, itersPerLaunch=3, laneColumns=1)
    two = _fields(pa, nX, 3, init, iters, marchRows=56, itersPerLaunch=3, **other)
    _same_fields(two, one)


def test_wide_rejects_fp64():
    pa = make_pa(64)
    with pytest.raises(Exception):
        EffPlast2D(pa, 64, mode="fp64", marchRows=16, itersPerLaunch=3, laneColumns=2)
    with pytest.raises(Exception):
        Batch(mode="fp64", wide3Rows=32)


def test_wide_batch_full_size_equals_tiles():
    """1024^2 (interior, edge, pore and pore-free segments), 3 x 300 plastic
    iterations through the batched wide chains: the tile kernel's records
    and moduli bit for bit."""
    nX = 1024
    _note("wide batch 1024")
    pa = make_pa(nX, nPores=1, porosity=0.01, Y=2.5e-3)

    def run(**shape):
        with Batch(mode="split", checkInterval=100, **shape) as b:
            for tag in (0, 1):
                b.submit(pa, nX, nTasks=3, nIter=300, eIter=1e-300, initLoad=-3e-3,
                         loadType=LT, tag=tag)
            b.run()
            return {r["tag"]: r for r in b.poll()}, b.stats()

    ref, _ = run(tileRows=14)
    for shape in ({"wide3Rows": 24}, {"wide3Rows": 64}, {"wide4Rows": 64}, {"wide4Rows": 128}):
        got, st = run(**shape)
        assert st["kernel"] == (8 if "wide3Rows" in shape else 9)
        for tag in (0, 1):
            a, b = got[tag]["result"], ref[tag]["result"]
            for grp in ("num", "an", "num_per"):
                for k in ("Kphi", "Kd", "G"):
                    x, y = a[grp][k], b[grp][k]
                    assert (x == y) or (np.isnan(x) and np.isnan(y)), (shape, grp, k, x, y)
            assert a["totalIters"] == b["totalIters"] == 900
