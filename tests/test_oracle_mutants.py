"""Planted mistakes: the oracle pins must reject a wrong oracle.

Each case compiles oracle/ep2d_oracle.c with one planted mistake
(-DORA_MUTANT=k: 1 wrong sign, 2 index shifted by one vertex, 3 dropped
term, 4 transposed operand) into a temporary library and shows that the
independent checks used by test_oracle_pins.py -- the MATLAB whole-array
solver, the isotropic-load symmetries and the uniform Hooke state --
report it, while the unmodified oracle passes the same checks.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import ora
from oracle.matlab_path import MatlabSolver, to_flat
from tests.pairs import make_pa, oracle_state, random_state, rel_err

SRC = os.path.join(os.path.dirname(ora.__file__), "ep2d_oracle.c")


def build(tmp_path, mutant):
    out = os.path.join(tmp_path, f"libora_m{mutant}.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-D_GNU_SOURCE", "-fPIC", "-fopenmp",
                    "-ffp-contract=off", f"-DORA_MUTANT={mutant}", "-shared", "-o", out, SRC,
                    "-lm"], check=True)
    return out


@pytest.fixture
def use_lib(monkeypatch):
    def _use(path):
        monkeypatch.setattr(ora, "LIB_PATH", path)
        monkeypatch.setattr(ora, "_lib", None)
    yield _use
    ora._lib = None


def matlab_gap():
    nX = 32
    pa = make_pa(nX, nPores=1, porosity=0.02, Y=1e-3)
    m = MatlabSolver(pa, nX)
    init = random_state(nX, 1, pa=pa, noise=1.0)
    m.set_flat(init["Ux"], init["Uy"], init["Vx"], init["Vy"])
    s = oracle_state(nX, pa, 3, init)
    m.iterate(20)
    s.iterate(20)
    return max(rel_err(s[k], to_flat(A)) for k, A in
               (("Ux", m.Ux), ("Uy", m.Uy), ("P", m.P), ("tauXX", m.tauxx), ("tauXY", m.tauxy)))


def symmetry_gap():
    nX = 40
    pa = make_pa(nX, nPores=1, porosity=0.03, Y=1.5e-3)
    s = oracle_state(nX, pa, 3)
    s.apply_affine(-3e-3, -3e-3, 0.0, 0.0)
    s.iterate(100)
    Ux = s["Ux"].reshape(nX, nX + 1)
    Uy = s["Uy"].reshape(nX + 1, nX)
    P = s["P"].reshape(nX, nX)
    sc = np.max(np.abs(Ux))
    return max(np.max(np.abs(Ux + Ux[:, ::-1])) / sc, np.max(np.abs(Ux - Uy.T)) / sc,
               rel_err(P, P.T))


def hooke_gap():
    nX = 24
    pa = make_pa(nX, nPores=1, porosity=0.0, Y=1.0)
    s = oracle_state(nX, pa, 3)
    e = (-2e-3, 1e-3, 5e-4)
    s.apply_affine(e[0], e[1], e[2], e[2])
    s.iterate(1)
    K0, G0 = pa[ora.PA_K0], pa[ora.PA_G0]
    div = e[0] + e[1]
    return max(rel_err(s["P"], np.full(nX * nX, -K0 * div)),
               rel_err(s["tauXX"], np.full(nX * nX, 2 * G0 * (e[0] - div / 3))),
               rel_err(s["tauXY"], np.full((nX - 1) ** 2, 2 * G0 * e[2])))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("mutant", [0, 1, 2, 3, 4])
def test_pins_reject_planted_mistakes(tmp_path, use_lib, mutant):
    use_lib(build(str(tmp_path), mutant))
    gaps = {"matlab": matlab_gap(), "symmetry": symmetry_gap(), "hooke": hooke_gap()}
    if mutant == 0:
        assert gaps["matlab"] < 1e-9 and gaps["symmetry"] < 1e-10 and gaps["hooke"] < 1e-11, gaps
    else:
        # the MATLAB restatement catches every planted mistake by a wide margin,
        assert gaps["matlab"] > 1e-3, gaps
        # and each is also caught by at least one closed form or invariant
        assert gaps["symmetry"] > 1e-6 or gaps["hooke"] > 1e-6, gaps


# Planted mistakes in the post-processing (-DORA_MUTANT=5..15) and the
# closed-form pin of tests/test_oracle_postproc_pins.py that must reject each.
POSTPROC_MUTANTS = {
    5: ("row 1 for row 0 in getDeltaP_honest", ["delta_p_honest"]),
    6: ("column N(nP-1)/nP for N(nP-2)/nP in getDeltaP_periodic", ["delta_p_periodic"]),
    7: ("integer y bound in getDeltaP_periodic", ["delta_p_periodic"]),
    8: ("yy - xx on one line of getTauInfty_honest", ["delta_p_honest"]),
    9: ("sigma[0] over the summed cells, not N^2", ["getters_uniform_state"]),
    10: ("nu0 fold dropped from sigma[2]", ["getters_uniform_state"]),
    11: ("measured strain in Kd", ["getters_records", "getters_uniform_state"]),
    12: ("sign of the yy difference in G", ["getters_records", "getters_uniform_state"]),
    13: ("-1 dropped from the plastic exponent", ["plastic_analytic_reduction"]),
    14: ("right pore-edge window [-1, +3]", ["hole_area_windows"]),
    15: ("magnitude for the signed element in FindMaxAbs", ["hole_area_windows"]),
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("mutant", sorted(POSTPROC_MUTANTS))
def test_postprocessing_pins_reject_planted_mistakes(tmp_path, use_lib, mutant):
    from tests.test_oracle_postproc_pins import PINS
    what, pins = POSTPROC_MUTANTS[mutant]
    use_lib(build(str(tmp_path), mutant))
    for name in pins:
        with pytest.raises(AssertionError):
            PINS[name]()
