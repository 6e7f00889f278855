"""GPU: SaveAnStatic2D (EffPlast2D.cu:1060-1292) through the C ABI against the
oracle on the same converged state.  The library's host code and the oracle
evaluate the same expressions in the same order (C++ std::complex and C99
complex over the same libm): fields within 1e-12 relative, statistics too;
the files follow the reference's names and raw FP64 layout."""
import os
import sys

import numpy as np
import pytest

from effplast2d_amd import EffPlast2D
from oracle import ora
from tests.pairs import make_pa, rel_err

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _pair(nX, nPores, lt, shear=None):
    pa = make_pa(nX, nPores=nPores, porosity=0.01, Y=2.5e-3)
    s = ora.State(nX=nX, nY=nX, pa=pa, nl=3, nIter=200_000, eIter=1e-8, outputStep=200,
                  shearLoadType=shear)
    res_o, recs_o = s.compute_eff_moduli(-3e-3, lt, 1)
    g = EffPlast2D(pa, nX, nTasks=3, nIter=200_000, eIter=1e-8, checkInterval=200, mode="fp64",
                   shearLoadType=shear)
    g.ComputeEffModuli(-3e-3, lt)
    return s, recs_o, g


@pytest.mark.parametrize("lt", [[1.0, 1.0, 0.0], [1.0, 0.6, 0.0]])
def test_an_static_matches_oracle(lt):
    _note(f"an_static lt={lt}")
    s, recs, g = _pair(64, 1, lt)
    with g:
        dP, tau = recs[0][0]["deltaP"], recs[0][0]["tauInfty"]
        got = g.an_static(dP, tau, lt)
    exp = s.an_static(dP, tau, lt)
    assert got["analytic"] == exp["analytic"] == 1
    for k in ora.AN_ARRAYS:
        assert rel_err(got[k], exp[k]) <= 1e-12, k
    for k in ora.AN_STATS:
        assert got[k] == pytest.approx(exp[k], rel=1e-12), k


def test_save_an_static_files(tmp_path):
    _note("save_an_static")
    lt = [1.0, 1.0, 0.0]
    s, recs, g = _pair(64, 1, lt)
    with g:
        dP, tau = recs[0][0]["deltaP"], recs[0][0]["tauInfty"]
        st = g.save_an_static(dP, tau, lt, tmp_path)
        arr = g.an_static(dP, tau, lt)
    assert st["analytic"] == 1 and st["errorJ1Avg"] == arr["errorJ1Avg"]
    for k in ora.AN_ARRAYS:
        p = tmp_path / f"{k}_64_.dat"
        data = np.fromfile(p, dtype="<f8")
        assert np.array_equal(data, arr[k]), k


def test_an_static_several_pores_numerical_only(tmp_path):
    _note("an_static nPores=2")
    lt = [1.0, 1.0, 0.0]
    s, recs, g = _pair(64, 2, lt)
    with g:
        dP, tau = recs[0][0]["deltaP"], recs[0][0]["tauInfty"]
        got = g.an_static(dP, tau, lt)
        g.save_an_static(dP, tau, lt, tmp_path)
    exp = s.an_static(dP, tau, lt)
    assert got["analytic"] == 0
    for k in ("UnuAbs", "J1nu", "J2nu"):
        assert rel_err(got[k], exp[k]) <= 1e-12, k
    assert sorted(os.listdir(tmp_path)) == ["J1nu_64_.dat", "J2nu_64_.dat", "UnuAbs_64_.dat"]
