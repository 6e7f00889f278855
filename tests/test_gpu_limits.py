"""GPU: the largest grid the multi-iteration kernels take.

The chain kernels address every field through buffer resources with 32-bit
byte offsets, and lanes outside a field read past a bound at offset 2^31, so
every field must stay under 2 GiB: nX = nY <= 16382 (Ux holds 16383 x 16382
doubles, 2 147 090 448 bytes).  At that size the split scheme's four-stage
chains on two waves per stage must give the tile kernel's fields bit for bit
(the tile kernel's 32-bit offsets hold up to 4 GiB per field)."""
import sys

import numpy as np
import pytest

from effplast2d_amd import EffPlast2D
from tests.pairs import make_pa

pytestmark = pytest.mark.gpu

NMAX = 16382


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _fields(pa, **shape):
    # affine compression -3e-3 with one pore (the return mapping yields at
    # its rim); 16 iterations = three four-iteration launches, a three-
    # iteration closing launch and the final single iteration
    with EffPlast2D(pa, NMAX, nTasks=3, mode="split", **shape) as g:
        g.apply_affine(-3e-3, -3e-3, 0.0, 0.0)
        g.iterate(16)
        return {k: g.get(k) for k in ("Ux", "Uy", "Vy")}


def test_largest_multi_iteration_grid_equals_tiles():
    _note(f"nX={NMAX} split chains on two waves per stage vs tiles")
    pa = make_pa(NMAX, nPores=1, porosity=0.01, Y=2.5e-3)
    chain = _fields(pa, marchRows=64, itersPerLaunch=4, stageWaves=2)
    tiles = _fields(pa, tileRows=8)
    for k in chain:
        assert chain[k].shape == tiles[k].shape
        nd = int(np.count_nonzero(chain[k] != tiles[k]))
        assert nd == 0, f"{k}: {nd} values differ"
    # the state moved: the velocity is not zero anywhere near the loaded edges
    assert np.count_nonzero(tiles["Vy"]) > 1000
