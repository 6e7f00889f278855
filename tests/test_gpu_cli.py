"""GPU: the reference's command line end to end (main.cu argv order,
data/pa.dat in, data/*.dat and EffPlast2D.log out), through the 64x64 build
of the CLI (ep2d_cli64: NGRID=2, NL=3, EITER=1e-8, residual check every 200
iterations), against the oracle run with the same settings."""
import os
import subprocess
import sys

import numpy as np
import pytest

from effplast2d_amd import write_pa
from effplast2d_amd.datfile import read_data_2D
from oracle import ora
from tests.pairs import make_pa, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "effplast2d_amd", "ep2d_cli64")


def test_cli_end_to_end(tmp_path, monkeypatch):
    print("cli end to end", flush=True)
    sys.stdout.flush()
    nX = 64
    pa = make_pa(nX, nPores=1, porosity=0.01, Y=2.5e-3)
    os.makedirs(tmp_path / "data")
    write_pa(tmp_path / "data" / "pa.dat", pa)
    r = subprocess.run([CLI, "-3e-3", "1", "1", "0", "1"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    log = open(tmp_path / "EffPlast2D.log").read()
    for line in ["LOAD STEP 1 FROM 3: PRELOADING", "SMALL HYDROSTATIC INCREMENT",
                 "SMALL DEVIATORIC INCREMENT", "Grid resolution is 64x64",
                 "EFFECTIVE MODULI ESTIMATED", "Uabs max error", "Calculation time is"]:
        assert line in log, line
    s = ora.State(nX=nX, nY=nX, pa=pa, nl=3, nIter=100_000, eIter=1e-8, outputStep=200)
    res, recs = s.compute_eff_moduli(-3e-3, [1.0, 1.0, 0.0], 1)
    monkeypatch.chdir(tmp_path)
    P = read_data_2D("data/Pc", nX, nX, nX).ravel(order="F")
    T = read_data_2D("data/tauXXc", nX, nX, nX).ravel(order="F")
    assert rel_err(P, s["P"]) <= 1e-12
    assert rel_err(T, s["tauXX"]) <= 1e-12
    an = s.an_static(recs[0][0]["deltaP"], recs[0][0]["tauInfty"], [1.0, 1.0, 0.0])
    for k in ora.AN_ARRAYS:
        got = np.fromfile(tmp_path / "data" / f"{k}_{nX}_.dat", dtype="<f8")
        assert rel_err(got, an[k]) <= 1e-12, k
