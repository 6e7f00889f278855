"""Pins of the CPU oracle (oracle/ep2d_oracle.c) -- CPU only.

The reference ships no tests, fixtures or golden vectors, so each pin comes
from something the oracle does not compute itself:
  * golden pa[] vectors from an independent Python restatement of
    get_sigma_2D.m:3-32 (tests/golden/make_pa_golden.py),
  * the reference's second implementation, the MATLAB CPU solver
    (get_sigma_2D.m:45-166), restated as whole-array numpy code
    (oracle/matlab_path.py) and compared field by field,
  * closed forms: the uniform Hooke state without pores, the affine increment,
    the residual of a known velocity field, the averages' normalisations,
  * invariants: exact linearity of the elastic scheme, mirror and transpose
    symmetry of an isotropic load on centred pores, the yield bound after the
    return mapping, zero load stays zero,
  * the analytic elastic moduli of a dilute pore (EffPlast2D.cu:643-654),
  * the analytic static fields (SaveAnStatic2D, EffPlast2D.cu:988-1292): the
    complex-potential branch reduces to the hydrostatic closed form without
    shear, the elastic and plastic stresses agree on the elliptic plastic-zone
    boundary, the displacement is radial without shear and keeps the load's
    mirror symmetries with it, and the numerical-vs-analytic J1 error falls
    with the grid.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

from oracle import ora
from oracle.matlab_path import MatlabSolver, to_flat
from tests.pairs import make_pa, oracle_state, random_state, rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- parameters ---------------------------------------------------------------
def test_pa_matches_golden_bit_exact():
    gold = json.load(open(os.path.join(GOLDEN, "pa_golden.json")))
    for name, g in gold.items():
        i = g["inputs"]
        pa = ora.derive_params(i["Lx"], i["Ly"], g["nX"], i["nPores"], i["porosity"], i["Y"],
                               i["K0"], i["G0"], i["rho0"], i["CFL"])
        want = [float.fromhex(h) for h in g["pa_hex"]]
        assert list(pa) == want, name
        raw = open(os.path.join(GOLDEN, f"pa_{name}.dat"), "rb").read()
        assert list(struct.unpack("<11d", raw)) == want


def test_pa_closed_form_relations():
    pa = ora.derive_params(20.0, 20.0, 1024, 1, 0.01, 2.5e-3, 1.0, 0.5, 1.0, 0.5)
    dX, dt, K0, G0, rho = pa[ora.PA_DX], pa[ora.PA_DT], pa[ora.PA_K0], pa[ora.PA_G0], pa[ora.PA_RHO]
    assert dX == pytest.approx(20.0 / 1023, rel=1e-15)
    assert dt * math.sqrt((K0 + 4 * G0 / 3) / rho) == pytest.approx(0.5 * dX, rel=1e-15)
    assert pa[ora.PA_DAMPX] * dt * 1024 == pytest.approx(4.0, rel=1e-15)
    rad = pa[ora.PA_RAD]
    assert math.pi * rad * rad / 400.0 == pytest.approx(0.01, rel=1e-14)
    assert pa[ora.PA_COH] == pytest.approx(2.5e-3 * math.sqrt(2), rel=1e-15)


# ---- closed forms -------------------------------------------------------------
def _uniform_state(nX, e, nl=3, Y=1.0):
    """No pores (porosity 0 -> rad 0) and an affine strain (exx, eyy, exy)."""
    pa = make_pa(nX, nPores=1, porosity=0.0, Y=Y)
    s = oracle_state(nX, pa, nl)
    s.apply_affine(e[0], e[1], e[2], e[2])
    return s, pa


def test_uniform_hooke_state_closed_form():
    nX = 40
    e = (-2e-3, 1e-3, 5e-4)
    s, pa = _uniform_state(nX, e)
    K0, G0 = pa[ora.PA_K0], pa[ora.PA_G0]
    assert not s["cellPore"].any() and not s["vertPore"].any()
    s.iterate(1)
    div = e[0] + e[1]
    # EffPlast2D.cu:54-75 for a uniform strain
    np.testing.assert_allclose(s["P"], -K0 * div, rtol=1e-11)
    np.testing.assert_allclose(s["tauXX"], 2 * G0 * (e[0] - div / 3), rtol=1e-11)
    np.testing.assert_allclose(s["tauYY"], 2 * G0 * (e[1] - div / 3), rtol=1e-11)
    np.testing.assert_allclose(s["tauXY"], G0 * 2 * e[2], rtol=1e-11)
    # a uniform stress has no divergence: the velocity stays at rounding level
    assert np.max(np.abs(s["Vx"])) < 1e-14 and np.max(np.abs(s["Vy"])) < 1e-14


def test_averages_normalisations_closed_form():
    """EffPlast2D.cu:453-503: sums over the (N-2)^2 interior divided by N^2
    (xx, yy) and (N-1)^2 (xy + nu0 * trace term)."""
    nX = 48
    e = (-2e-3, -1e-3, 3e-4)
    s, pa = _uniform_state(nX, e)
    s.iterate(1)
    K0, G0 = pa[ora.PA_K0], pa[ora.PA_G0]
    nu0 = (1.5 * K0 - G0) / (3.0 * K0 + G0)
    div = e[0] + e[1]
    P, txx, tyy, txy = -K0 * div, 2 * G0 * (e[0] - div / 3), 2 * G0 * (e[1] - div / 3), 2 * G0 * e[2]
    a = s.averages()
    n2 = (nX - 2) ** 2
    assert a["sigma"][0] == pytest.approx((txx - P) * n2 / nX ** 2, rel=1e-10)
    assert a["sigma"][1] == pytest.approx((tyy - P) * n2 / nX ** 2, rel=1e-10)
    assert a["sigma"][2] == pytest.approx((txy + nu0 * (txx + tyy - 2 * P)) * n2 / (nX - 1) ** 2,
                                          rel=1e-10)
    assert a["sigma"][3] == 0.0
    assert a["sigmaXYclean"] == pytest.approx(txy, rel=1e-10)
    # one pore per side: the periodic window is empty (perOffset = nX)
    assert a["sigmaPer"] == [0.0, 0.0, 0.0, 0.0]


def test_affine_increment_closed_form():
    nX = 24
    pa = make_pa(nX)
    s = oracle_state(nX, pa, 3)
    dX = pa[ora.PA_DX]
    s.apply_affine(1e-3, 2e-3, 3e-4, 3e-4)
    Ux = s["Ux"].reshape(nX, nX + 1)
    Uy = s["Uy"].reshape(nX + 1, nX)
    xs = -0.5 * dX * nX + dX * np.arange(nX + 1)
    ys = -0.5 * dX * (nX - 1) + dX * np.arange(nX)
    np.testing.assert_allclose(Ux, xs[None, :] * 1e-3 + ys[:, None] * 3e-4, rtol=1e-13, atol=1e-18)
    xs2 = -0.5 * dX * (nX - 1) + dX * np.arange(nX)
    ys2 = -0.5 * dX * nX + dX * np.arange(nX + 1)
    np.testing.assert_allclose(Uy, ys2[:, None] * 2e-3 + xs2[None, :] * 3e-4, rtol=1e-13, atol=1e-18)


def test_residual_closed_form():
    nX = 32
    pa = make_pa(nX)
    s = oracle_state(nX, pa, 3)
    s["Vx"][:] = 0.0
    s["Vy"][:] = 0.0
    s["Vx"][7] = -3e-6
    s["Vy"][11] = 2e-6
    s["Vy"][12] = np.nan      # FindMaxAbs skips NaN (EffPlast2D.cu:805-813)
    dX, dT = pa[ora.PA_DX], pa[ora.PA_DT]
    cur = [-4e-3, -2e-3, 1e-3]
    want = (3e-6 / (dX * (nX - 1)) + 2e-6 / (dX * (nX - 1))) * dT / max(abs(cur[0]), max(cur[1], cur[2]))
    assert s.residual(cur) == pytest.approx(want, rel=1e-15)


def test_find_max_abs_overloads():
    L = ora.lib()
    v = np.array([1.0, -5.0, 4.0, np.nan, 5.0])
    p = v.ctypes.data_as(ora._dp)
    assert L.ora_find_max_abs(p, 5) == 5.0
    assert L.ora_find_max_abs_signed(p, 5) == -5.0    # first of equal magnitude, signed


# ---- invariants ------------------------------------------------------------------
def test_zero_load_stays_zero():
    nX = 32
    s = oracle_state(nX, make_pa(nX, nPores=2, porosity=0.04), 3)
    s.iterate(20)
    for k in ("Ux", "Uy", "Vx", "Vy", "P", "tauXX", "tauYY", "tauXY"):
        assert not np.any(s[k]), k


def test_elastic_scheme_is_exactly_linear():
    """Without plasticity every operation is linear: doubling the state doubles
    every field bit for bit (scaling by 2 commutes with rounding)."""
    nX = 36
    pa = make_pa(nX, nPores=2, porosity=0.05)
    init = random_state(nX, 5, pa=pa, boundary_v=True)
    a = oracle_state(nX, pa, 1, init)
    b = oracle_state(nX, pa, 1, {k: 2.0 * v for k, v in init.items()})
    a.iterate(25)
    b.iterate(25)
    for k in ("Ux", "Uy", "Vx", "Vy", "P", "tauXX", "tauYY", "tauXY"):
        assert np.array_equal(b[k], 2.0 * a[k]), k


@pytest.mark.parametrize("nl", [1, 3])
def test_isotropic_load_symmetries(nl):
    """Isotropic preload on a centred pore: mirror x -> -x and transpose x <-> y."""
    nX = 40
    pa = make_pa(nX, nPores=1, porosity=0.03, Y=1.5e-3)
    s = oracle_state(nX, pa, nl)
    s.apply_affine(-3e-3, -3e-3, 0.0, 0.0)
    s.iterate(150)
    Ux = s["Ux"].reshape(nX, nX + 1)
    Uy = s["Uy"].reshape(nX + 1, nX)
    P = s["P"].reshape(nX, nX)
    txx = s["tauXX"].reshape(nX, nX)
    tyy = s["tauYY"].reshape(nX, nX)
    txy = s["tauXY"].reshape(nX - 1, nX - 1)
    sc = np.max(np.abs(Ux))
    tol = 1e-10
    # mirror in x: Ux odd, Uy even; cell fields even; vertex shear odd
    assert np.max(np.abs(Ux + Ux[:, ::-1])) <= tol * sc
    assert np.max(np.abs(Uy - Uy[:, ::-1])) <= tol * sc
    assert rel_err(P, P[:, ::-1]) <= tol
    assert np.max(np.abs(txy + txy[:, ::-1])) <= tol * np.max(np.abs(P))
    # transpose: Ux(i, j) = Uy(j, i), tauXX = tauYY^T, P = P^T, tauXY = tauXY^T
    assert np.max(np.abs(Ux - Uy.T)) <= tol * sc
    assert rel_err(P, P.T) <= tol
    assert rel_err(txx, tyy.T) <= tol
    assert np.max(np.abs(txy - txy.T)) <= tol * np.max(np.abs(P))
    if nl > 1:
        assert np.any(s["J2"] > 0)


def test_return_mapping_bound_and_pore_zeroing():
    nX = 48
    pa = make_pa(nX, nPores=2, porosity=0.05, Y=1e-3)
    init = random_state(nX, 9, pa=pa, noise=2.0)
    s = oracle_state(nX, pa, 3, init)
    s.iterate(3)
    coh = pa[ora.PA_COH]
    # after ComputePlasticity every cell is on or inside the yield circle
    assert np.max(s["J2"]) <= coh * (1 + 1e-14)
    # and many cells actually yielded (the test exercises the return mapping)
    assert np.count_nonzero(np.isclose(s["J2"], coh, rtol=1e-12)) > nX
    cp = s["cellPore"].astype(bool)
    assert cp.any()
    assert not np.any(s["P"][cp]) and not np.any(s["tauXX"][cp]) and not np.any(s["tauYY"][cp])
    assert not np.any(s["tauXY"][s["vertPore"].astype(bool)])


def test_pore_masks_match_material_mask():
    """The ComputeStress cell test and the SetMaterials test are the same
    predicate (EffPlast2D.cu:79-82 vs :713-720)."""
    for nX, nP, por in [(64, 1, 0.01), (96, 3, 0.05), (128, 4, 0.05)]:
        pa = make_pa(nX, nPores=nP, porosity=por)
        s = oracle_state(nX, pa, 3)
        assert np.array_equal(s["cellPore"].astype(bool), s["K"] < pa[ora.PA_K0])
        assert int(s["cellPore"].sum()) == pytest.approx(
            por * nX * nX, rel=0.15)


# ---- the reference's second implementation: the MATLAB CPU solver -----------------
@pytest.mark.parametrize("nX,nPores,por,Y,seed,iters", [(32, 1, 0.02, 1e-3, 1, 40),
                                                        (48, 2, 0.04, 2e-3, 2, 30),
                                                        (40, 3, 0.05, 5e-4, 3, 25)])
def test_matches_matlab_whole_array_solver(nX, nPores, por, Y, seed, iters):
    pa = make_pa(nX, nPores=nPores, porosity=por, Y=Y)
    m = MatlabSolver(pa, nX)
    assert m.margin > 1e-9     # no cell within rounding distance of a pore edge
    init = random_state(nX, seed, pa=pa, noise=1.0)
    m.set_flat(init["Ux"], init["Uy"], init["Vx"], init["Vy"])
    s = oracle_state(nX, pa, 3, init)     # nl = 3: plasticity on, as in MATLAB
    assert np.array_equal(to_flat(m.cell).astype(np.uint8), s["cellPore"])
    assert np.array_equal(to_flat(m.vert).astype(np.uint8), s["vertPore"])
    m.iterate(iters)
    s.iterate(iters)
    # MATLAB keeps the trial J2 (get_sigma_2D.m:126); CUDA recomputes it after
    # the return mapping (EffPlast2D.cu:169), which puts yielded cells at coh
    J2m = np.minimum(m.J2, pa[ora.PA_COH])
    for k, A in (("Ux", m.Ux), ("Uy", m.Uy), ("Vx", m.Vx), ("Vy", m.Vy), ("P", m.P),
                 ("tauXX", m.tauxx), ("tauYY", m.tauyy), ("tauXY", m.tauxy), ("J2", J2m)):
        e = rel_err(s[k], to_flat(A))
        assert e <= 1e-9, f"{k}: {e:.3e}"
    # the yield test is active in this case
    assert np.count_nonzero(np.isclose(s["J2"], pa[ora.PA_COH], rtol=1e-12)) > 0


def test_matlab_residual_matches_for_step_zero():
    nX = 32
    pa = make_pa(nX, nPores=1, porosity=0.02, Y=1.0)
    m = MatlabSolver(pa, nX)
    s = oracle_state(nX, pa, 3)
    s.apply_affine(-1e-3, -1e-3, 0.0, 0.0)
    m.set_flat(s["Ux"], s["Uy"], s["Vx"], s["Vy"])
    m.iterate(60)
    s.iterate(60)
    L = pa[ora.PA_DX] * (nX - 1)
    # CUDA: max(|e0|, max(e1, e2)) of the macro strain; MATLAB: max|load*loadType|
    e_cuda = s.residual([-1e-3, -1e-3, 0.0])
    e_mat = m.residual(-1e-3, [1.0, 1.0, 0.0], L, L)
    assert e_cuda == pytest.approx(e_mat, rel=1e-9)


# ---- analytic elastic moduli of a dilute pore ----------------------------------------
def test_elastic_dilute_pore_moduli_against_analytic():
    """nTasks = 1 (pure elastic), one pore: K_d approaches the logged exact
    value (1-phi)/(1/K0 + 1/Kphi), Kphi = G0/phi (EffPlast2D.cu:643-654), up to
    the reference's (N-2)^2/N^2 normalisation of the mean stress."""
    nX = 128
    pa = make_pa(nX, nPores=1, porosity=0.005, Y=1.0)
    s = oracle_state(nX, pa, 1, nIter=20_000, eIter=1e-10, outputStep=500)
    res, recs = s.compute_eff_moduli(-1e-5, [1.0, 1.0, 0.0], 1)
    r = recs[0][0]
    assert r["error"] < 1e-10
    phi = r["Phi0"]
    assert phi == pytest.approx(0.005, rel=1e-6)
    assert res["an"]["Kphi"] == pytest.approx(pa[ora.PA_G0] / phi, rel=1e-12)
    kd = res["num"]["Kd"] / ((nX - 2) / nX) ** 2
    assert kd == pytest.approx(res["KdExactElast"], rel=0.02)
    # the pore opens under compression: dPhi > 0, Kphi of the right sign and scale
    assert r["dPhi"] > 0
    assert 0.5 * res["an"]["Kphi"] < res["num"]["Kphi"] < 1.2 * res["an"]["Kphi"]


def test_driver_records_and_load_sequence():
    """ComputeEffModuli task sequence (EffPlast2D.cu:179-229): cumulative macro
    strain, 0.5% increments, checks every outputStep."""
    nX = 48
    pa = make_pa(nX, nPores=1, porosity=0.01)
    s = oracle_state(nX, pa, 3, nIter=50_000, eIter=1e-8, outputStep=200,
                     shearLoadType=(0.0, 0.0, 1.0))
    res, recs = s.compute_eff_moduli(-3e-3, [1.0, 0.6, 0.0], 1)
    r0, r1, r2 = recs[0][0], recs[1][0], recs[2][0]
    assert r0["epsilon"] == pytest.approx([-3e-3, -1.8e-3, 0.0], rel=1e-15)
    inc = -3e-3 * 0.005
    assert r1["epsilon"] == pytest.approx([-3e-3 + 0.8 * inc, -1.8e-3 + 0.8 * inc, 0.0], rel=1e-12)
    assert r2["epsilon"][2] == pytest.approx(inc, rel=1e-12)
    for r in (r0, r1, r2):
        assert r["iters"] % 200 == 0 and r["error"] < 1e-8
    # xy shear modulus: of the order of G0, lowered by the pore
    assert 0.3 < res["Gxy"] < pa[ora.PA_G0]


# ---- analytic static fields (SaveAnStatic2D, EffPlast2D.cu:988-1292) ----------------
def _an_case():
    Y = 2.5e-3
    rad, G0, K0, dP = 1.0, 0.5, 1.0, 2.4 * 2.5e-3
    c0 = rad * np.exp(abs(dP) / 2 / Y - 0.5)
    return Y, rad, G0, K0, dP, c0


def test_analytic_kappa_zero_reduces_to_closed_form():
    """With no far-field shear (kappa = 0) the complex-potential stresses of the
    elastic zone must equal the hydrostatic closed form (the reference's two
    branches of SaveAnStatic2D describe the same field)."""
    nX = 64
    pa = make_pa(nX, nPores=1, porosity=0.01, Y=2.5e-3)
    s = ora.State(nX=nX, nY=nX, pa=pa, nl=3, nIter=200000, eIter=1e-8, outputStep=200)
    res, recs = s.compute_eff_moduli(-3e-3, [1.0, 1.0, 0.0], 1)
    dP = recs[0][0]["deltaP"]
    assert dP > 0
    hyd = s.an_static(dP, 0.0, [1.0, 1.0, 0.0])
    cpx = s.an_static(dP, 0.0, [1.0, 0.999, 0.0])
    m = hyd["J1an"] != 0
    assert m.sum() > 1000
    for k in ("J1an", "J2an"):
        rel = np.abs(cpx[k][m] - hyd[k][m]) / np.abs(hyd[k][m])
        assert rel.max() <= 1e-12, (k, rel.max())


@pytest.mark.parametrize("kappa", [0.0, 0.1, 0.25])
def test_analytic_stress_continuous_on_plastic_boundary(kappa):
    """J1, J2 of the elastic-zone solution and of the plastic-zone solution
    agree on the elliptic boundary x^2/Rx^2 + y^2/Ry^2 = 1 (Rx, Ry = c0 (1 -+
    kappa)): the elastoplastic solution is continuous there."""
    Y, rad, G0, K0, dP, c0 = _an_case()
    Rx, Ry = c0 * (1 - kappa), c0 * (1 + kappa)
    for t in np.linspace(0.05, 2 * np.pi, 13):
        x, y = Rx * np.cos(t), Ry * np.sin(t)
        je = ora.an_point("ora_an_j_elast", x * (1 + 1e-9), y * (1 + 1e-9), 1.0, kappa, c0, Y, rad)
        jp = ora.an_point("ora_an_j_plast", np.hypot(x, y), 1.0, Y, rad)
        assert je[0] == pytest.approx(jp[0], rel=1e-7)
        assert je[1] == pytest.approx(jp[1], rel=1e-7)


def test_analytic_displacement_symmetries():
    """kappa = 0, no shear: the displacement is radial and depends on r only;
    with shear it keeps the mirror symmetries of the load."""
    Y, rad, G0, K0, dP, c0 = _an_case()

    def U(x, y, tau, kappa):
        re, im = ora.an_point("ora_an_u_elast", x, y, tau, 1.0, kappa, c0, Y, rad, G0, K0)
        return complex(re, im)

    for x, y in [(5, 1), (-3, 4), (0.5, -6)]:
        z, u = complex(x, y), U(x, y, 0.0, 0.0)
        assert abs((u * z.conjugate()).imag) <= 1e-12 * abs(u) * abs(z)
        assert abs(u) == pytest.approx(abs(U(abs(z), 0.0, 0.0, 0.0)), rel=1e-12)
    for kappa in (0.1, 0.25):
        tau = kappa * Y
        for x, y in [(5, 1), (-3, 4), (2.5, -6)]:
            u = U(x, y, tau, kappa)
            assert abs(U(x, -y, tau, kappa) - u.conjugate()) <= 1e-13 * abs(u)
            assert abs(U(-x, -y, tau, kappa) + u) <= 1e-13 * abs(u)
            assert abs(U(-x, y, tau, kappa) + u.conjugate()) <= 1e-13 * abs(u)


def test_analytic_j1_error_shrinks_with_resolution():
    """Numerical mean stress against the analytic one around a converged
    elastoplastic pore: the average relative J1 error falls with the grid."""
    errs = []
    for nX in (64, 128):
        pa = make_pa(nX, nPores=1, porosity=0.01, Y=2.5e-3)
        s = ora.State(nX=nX, nY=nX, pa=pa, nl=3, nIter=200000, eIter=1e-8, outputStep=200)
        res, recs = s.compute_eff_moduli(-3e-3, [1.0, 1.0, 0.0], 1)
        o = s.an_static(recs[0][0]["deltaP"], recs[0][0]["tauInfty"], [1.0, 1.0, 0.0])
        assert o["analytic"] == 1
        errs.append(o["errorJ1Avg"])
    assert errs[1] < 0.5 * errs[0] and errs[1] < 0.02, errs
