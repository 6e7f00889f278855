"""Pins of the oracle's post-processing that do not go through the oracle's
own formulas: boundary means, periodic boundary means, averages, the moduli
getters, the plastic analytic moduli and the pore-edge hole area, each
checked against a closed form derived here from the reference's index sets
and formulas (EffPlast2D.cu:453-503, 514-620, 643-696, 825-882, 925-942,
1412-1539).

Each check is a plain function that raises AssertionError, so
test_oracle_mutants.py can show that the planted mistake it was written
against (a shifted row, a mirrored column, an integer bound, a sign, a
dropped fold, a wrong strain, a wrong window, the magnitude for the signed
element) makes it fail.
"""
import math

import numpy as np
import pytest

from oracle import ora
from tests.pairs import make_pa, oracle_state


def _grid_fields(s, n):
    """Cell index grids (i along x, j along y) in the reference layout."""
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    return i.ravel().astype(float), j.ravel().astype(float)


def _set_cells(s, P, XX, YY):
    s["P"][:] = P
    s["tauXX"][:] = XX
    s["tauYY"][:] = YY


# ---- getDeltaP_honest / getTauInfty_honest (EffPlast2D.cu:825-853, 925-942) ----------
def pin_delta_p_honest():
    """Linear fields: each boundary line (rows 0 and N-1 over i in [1, N-2],
    columns 0 and N-1 over j in [1, N-2]) has mean index (N-1)/2 along the line
    and the two lines of a pair average to (N-1)/2 across, so -1/8 of the four
    line means of (sxx + syy) is -(sxx + syy)/2 at the centre cell, and
    -1/8 of those of (txx - tyy) is -(txx - tyy)/2 there.  A load only on row
    0 gives 1/4 of its value (one of eight line terms, corners excluded)."""
    n = 40
    s = oracle_state(n, make_pa(n), 3)
    i, j = _grid_fields(s, n)
    P = 0.3 + 0.011 * i - 0.007 * j
    XX = -0.2 + 0.005 * i + 0.013 * j
    YY = 0.1 - 0.009 * i + 0.004 * j
    _set_cells(s, P, XX, YY)
    c = (n - 1) / 2
    Pc, XXc, YYc = (0.3 + 0.011 * c - 0.007 * c, -0.2 + 0.005 * c + 0.013 * c,
                    0.1 - 0.009 * c + 0.004 * c)
    b = s.boundary()
    assert b["deltaP"] == pytest.approx(-0.5 * ((XXc - Pc) + (YYc - Pc)), rel=1e-13), b
    assert b["tauInfty"] == pytest.approx(-0.5 * (XXc - YYc), rel=1e-13), b
    # P = 1 on row 0 only (txx = tyy = 0): one of the eight line terms is -2
    P = np.where(j == 0, 1.0, 0.0)
    _set_cells(s, P, 0 * P, 0 * P)
    assert s.boundary()["deltaP"] == pytest.approx(0.25, rel=1e-15)
    # txx = 1 on row 0 only: tauInfty = -1/8 * 1 (bottom line mean)
    _set_cells(s, 0 * P, P, 0 * P)
    assert s.boundary()["tauInfty"] == pytest.approx(-0.125, rel=1e-15)
    # the four corner cells belong to no line
    corner = np.where(((i == 0) | (i == n - 1)) & ((j == 0) | (j == n - 1)), 1.0, 0.0)
    _set_cells(s, corner, corner, 0 * corner)
    b = s.boundary()
    assert b["deltaP"] == 0.0 and b["tauInfty"] == 0.0


# ---- getDeltaP_periodic (EffPlast2D.cu:854-882) -------------------------------------
def pin_delta_p_periodic():
    """Integer index sets of the periodic variant, counted by hand:
      N = 100, nP = 3, P = 1: rows 33 and 66, i in [33, 66): 33 terms; the y
        loop runs j in [33, 200/3.0 = 66.67): 34 terms; both divide by
        N(nP-2)/nP = 33, so dPper = -1/8 (-4 - 4*34/33) = 1/2 + 17/33.
      N = 100, nP = 4, P = i: rows 25 and 75 over i in [25, 75) give
        -4 * (25 + ... + 74) / 50 = -198; the y loop (j in [25, 75), 50
        terms) reads columns N/nP = 25 and N(nP-2)/nP = 50: -2 (25 + 50) =
        -150; dPper = -1/8 (-198 - 150) = 43.5.
      nP <= 2: 0."""
    n = 100
    s = oracle_state(n, make_pa(n, nPores=3, porosity=0.02), 3)
    one = np.ones(n * n)
    _set_cells(s, one, 0 * one, 0 * one)
    assert s.boundary()["deltaPper"] == pytest.approx(0.5 + 17.0 / 33.0, rel=1e-15)
    s = oracle_state(n, make_pa(n, nPores=4, porosity=0.02), 3)
    i, j = _grid_fields(s, n)
    _set_cells(s, i, 0 * i, 0 * i)
    assert s.boundary()["deltaPper"] == pytest.approx(43.5, rel=1e-15)
    s = oracle_state(n, make_pa(n, nPores=2, porosity=0.02), 3)
    _set_cells(s, one, one, one)
    assert s.boundary()["deltaPper"] == 0.0


# ---- getters on a uniform no-pore state (EffPlast2D.cu:453-503, 1412-1539) -----------
def _uniform_moduli(shear):
    n, nP = 48, 3
    pa = make_pa(n, nPores=nP, porosity=0.0, Y=1.0)
    s = oracle_state(n, pa, 3, nIter=200, eIter=1e-8, outputStep=10, shearLoadType=shear)
    res, recs = s.compute_eff_moduli(-2e-3, [1.0, 1.0, 0.0], 1)
    return n, nP, pa, res, recs


def pin_getters_uniform_state():
    """No pores, no yield (Y = 1): every load step is a uniform Hooke state
    (V stays at rounding level, the first check converges).  With e the
    isotropic strain increment of the hydrostatic task, P = -2 K0 e,
    txx = tyy = 2 G0 e / 3, txy = 0.  The reference sums (N-2)^2 cells
    (i, j in [1, N-2]) but divides sigma[0], sigma[1] by N^2 and sigma[2] by
    (N-1)^2, and folds nu0 (txx + tyy - 2P) into sigma[2]; Kd divides
    -0.33333333 (dsigma0 + dsigma1 + dsigma2) by the applied 2e:
      Kd = 0.33333333 (r1 (2 G0/3 + 2 K0) + 2 r2 nu0 (G0/3 + K0)),
      r1 = (N-2)^2 / N^2, r2 = (N-2)^2 / (N-1)^2.
    The periodic window (N/nP, N - N/nP) sums (N - 2 off - 1)^2 cells over
    (N - 2 off)^2 in stress and measured strain alike, which cancels:
      KdPer = 0.33333333 (2 G0/3 + 2 K0)  (sigmaPer[2] is xy, zero here).
    The shear task [1, -1, 0] (strain e, -e): dtxx = 2 G0 e, dP = 0, so
      G = 0.5 (r1 4 G0 e) / (2 e) = r1 G0, GPer = G0;
    the shear task [0, 0, 1]: txy = 2 G0 e at every vertex, Gxy = G0."""
    n, nP, pa, res, recs = _uniform_moduli((1.0, -1.0, 0.0))
    K0, G0 = pa[ora.PA_K0], pa[ora.PA_G0]
    nu0 = (1.5 * K0 - G0) / (3.0 * K0 + G0)
    r1, r2 = (n - 2) ** 2 / n ** 2, (n - 2) ** 2 / (n - 1) ** 2
    for st in recs:
        assert st[0]["iters"] == 10 and st[0]["error"] < 1e-8
    kd = 0.33333333 * (r1 * (2 * G0 / 3 + 2 * K0) + 2 * r2 * nu0 * (G0 / 3 + K0))
    assert res["num"]["Kd"] == pytest.approx(kd, rel=1e-9), res["num"]
    assert res["num_per"]["Kd"] == pytest.approx(0.33333333 * (2 * G0 / 3 + 2 * K0), rel=1e-9)
    assert res["num"]["G"] == pytest.approx(r1 * G0, rel=1e-9), res["num"]
    assert res["num_per"]["G"] == pytest.approx(G0, rel=1e-9)
    n, nP, pa, res, recs = _uniform_moduli((0.0, 0.0, 1.0))
    assert res["Gxy"] == pytest.approx(G0, rel=1e-9)
    assert res["GxyPer"] == pytest.approx(G0, rel=1e-9)


def pin_getters_records():
    """The getters on synthetic records: Kd = -(-0.33333333 sum dsigma) /
    (de0 + de1) with the applied strain, KdPer with the measured one,
    G = 0.5 (dsxx - dsyy) / (dexx - deyy) between tasks 2 and 3, and the
    elastic Kphi = deltaP / dPhi of a single task."""
    r0 = dict(sigma=[1.0, 2.0, 0.5, 0.0], sigmaPer=[3.0, 1.0, 0.0, 0.0],
              epsilon=[-1.0, -1.0, 0.0], epsilonPer=[-2.0, -2.0, 0.0], deltaP=2.0, dPhi=0.5,
              deltaPper=4.0, dPhiPer=2.0)
    r1 = dict(sigma=[1.5, 2.5, 1.0, 0.0], sigmaPer=[3.5, 1.5, 0.25, 0.0],
              epsilon=[-1.25, -1.25, 0.0], epsilonPer=[-3.0, -3.0, 0.0], deltaP=3.0, dPhi=0.75,
              deltaPper=5.0, dPhiPer=2.5)
    r2 = dict(sigma=[2.5, 1.5, 1.0, 0.0], sigmaPer=[4.5, 0.5, 0.25, 0.0],
              epsilon=[-1.0, -1.5, 0.0], epsilonPer=[-2.0, -4.0, 0.0])
    res = ora.moduli_getters(3, r0, r1, r2)
    assert res["num"]["Kd"] == pytest.approx(0.33333333 * 1.5 / -0.5, rel=1e-15)
    assert res["num_per"]["Kd"] == pytest.approx(0.33333333 * 1.25 / -2.0, rel=1e-15)
    assert res["num"]["G"] == pytest.approx(0.5 * (1.0 + 1.0) / (0.25 + 0.25), rel=1e-15)
    assert res["num_per"]["G"] == pytest.approx(0.5 * (1.0 + 1.0) / (1.0 + 1.0), rel=1e-15)
    assert res["num"]["Kphi"] == pytest.approx(1.0 / 0.25, rel=1e-15)
    assert res["num_per"]["Kphi"] == pytest.approx(1.0 / 0.5, rel=1e-15)
    res = ora.moduli_getters(1, r0)
    assert res["num"]["Kphi"] == 4.0 and res["num_per"]["Kphi"] == 2.0
    assert res["num"]["Kd"] == pytest.approx(-0.33333333 * 3.5 / -2.0 * -1, rel=1e-15)


# ---- plastic analytic moduli (EffPlast2D.cu:643-696) ---------------------------------
def pin_plastic_analytic_reduction():
    """Kphi = G0 / Phi / exp(|dP|/Y - 1) / (1 + tau^2/Y^2): at |dP| = Y and
    tau = 0 it is the elastic G0 / Phi (Lame hole, EffPlast2D.cu:649), each
    further Y of pressure divides it by e, tau = Y halves it; G = G0 / (1 +
    Phi exp(|dP|/Y - 1)) is G0 / (1 + Phi) at |dP| = Y (the logged elastic
    value, :683); Kd = (1 - phi) / (1/K0 + 1/Kphi) with phi = pi~ rad^2 /
    (Lx Ly), pi~ = 3.1415926."""
    n = 64
    pa = make_pa(n, nPores=1, porosity=0.01, Y=2.5e-3)
    s = oracle_state(n, pa, 3)
    Y = pa[ora.PA_COH] / math.sqrt(2.0)
    G0, K0, rad = pa[ora.PA_G0], pa[ora.PA_K0], pa[ora.PA_RAD]
    Phi = 0.012
    phi = 3.1415926 * rad ** 2 / (pa[ora.PA_DX] * (n - 1)) ** 2
    for dP, tau, kphi in [(Y, 0.0, G0 / Phi), (-Y, 0.0, G0 / Phi), (2 * Y, 0.0, G0 / Phi / math.e),
                          (Y, Y, 0.5 * G0 / Phi), (3 * Y, 0.5 * Y, G0 / Phi / math.e ** 2 / 1.25)]:
        res = s.analytic_moduli(0, dict(deltaP=dP, deltaPper=dP, tauInfty=tau, Phi=Phi, Phi0=Phi))
        assert res["an"]["Kphi"] == pytest.approx(kphi, rel=1e-14), (dP, tau)
        assert res["an_per"]["Kphi"] == pytest.approx(kphi, rel=1e-14)
        assert res["an"]["Kd"] == pytest.approx((1 - phi) / (1 / K0 + 1 / kphi), rel=1e-14)
        g = G0 / (1 + Phi * math.exp(abs(dP) / Y - 1))
        assert res["an"]["G"] == pytest.approx(g, rel=1e-14)
        if abs(dP) == Y:
            assert res["an"]["G"] == pytest.approx(res["GexactElast"], rel=1e-14)


# ---- pore-edge hole area (EffPlast2D.cu:514-620) -------------------------------------
def _hole_case():
    n = 64
    pa = make_pa(n, nPores=1, porosity=0.02)
    s = oracle_state(n, pa, 3)
    dX, rad = pa[ora.PA_DX], pa[ora.PA_RAD]
    # one central pore: cxdX = cydY = 0; samples on Ux row floor((N-1)/2), the
    # left window starts one node left of floor(N/2 - rad/dX), the right one two
    # nodes left of floor(N/2 + rad/dX); both five nodes long
    row = (n - 1) // 2
    left0 = int(math.floor(n / 2 - rad / dX)) - 1
    right0 = int(math.floor(n / 2 + rad / dX)) - 2
    return n, pa, s, row, left0, right0, rad


def pin_hole_area_windows():
    """Spikes in an otherwise zero Ux: a value c at the outermost node of the
    right window and -c at the outermost node of the left window give
    dRx = -1/2 (-c - c) = c and HoleAreaPi = (rad + c) rad; a node just past
    either window is not sampled; the signed element is taken (-c on the
    right gives dRx = -c/2).  Uniform dilation U = a (x, y): the outermost
    samples sit one to three nodes outside the pore edge, so a rad < dRx,
    dRy < a (rad + 3 dX)."""
    n, pa, s, row, left0, right0, rad = _hole_case()
    c = 0.05
    m = n + 1

    def area(spikes):
        s["Ux"][:] = 0.0
        s["Uy"][:] = 0.0
        for i, v in spikes:
            s["Ux"][row * m + i] = v
        return s.hole_area()["holeAreaPi"]

    assert area([]) == pytest.approx(rad * rad, rel=1e-15)
    assert area([(right0 + 4, c), (left0, -c)]) == pytest.approx((rad + c) * rad, rel=1e-14)
    assert area([(right0 + 5, c)]) == pytest.approx(rad * rad, rel=1e-15)
    assert area([(left0 - 1, -c)]) == pytest.approx(rad * rad, rel=1e-15)
    assert area([(right0 + 4, -c)]) == pytest.approx((rad - 0.5 * c) * rad, rel=1e-14)
    # uniform dilation
    dX = pa[ora.PA_DX]
    a = 1e-3
    xs = (np.arange(n + 1) - 0.5 * n) * dX
    ys = (np.arange(n) - 0.5 * (n - 1)) * dX
    s["Ux"][:] = (a * xs[None, :] + 0 * ys[:, None]).ravel()
    ys2 = (np.arange(n + 1) - 0.5 * n) * dX
    xs2 = (np.arange(n) - 0.5 * (n - 1)) * dX
    s["Uy"][:] = (a * ys2[:, None] + 0 * xs2[None, :]).ravel()
    h = s.hole_area()["holeAreaPi"]
    assert (rad + a * rad) ** 2 < h < (rad + a * (rad + 3 * dX)) ** 2


PINS = {
    "delta_p_honest": pin_delta_p_honest,
    "delta_p_periodic": pin_delta_p_periodic,
    "getters_uniform_state": pin_getters_uniform_state,
    "getters_records": pin_getters_records,
    "plastic_analytic_reduction": pin_plastic_analytic_reduction,
    "hole_area_windows": pin_hole_area_windows,
}


@pytest.mark.parametrize("name", sorted(PINS))
def test_postprocessing_pin(name):
    PINS[name]()
