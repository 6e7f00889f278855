"""GPU properties that fix the result without an oracle run, at sizes up to
the headline's 1024^2, for every scheme:

  * zero state stays zero (no load, no motion);
  * exact linearity of the elastic scheme: doubling U and V doubles every
    field bit for bit (every operation commutes with a power-of-two scale);
  * mirror symmetry of a centred pore under an isotropic load (x -> -x, and
    y -> -y), and the x <-> y transpose (Ux(i, j) <-> Uy(j, i));
  * the yield bound after the return mapping: J2 <= coh (1 + 4 eps);
  * a long run (1000 iterations) bit for bit against the oracle,

for the tile, marching, wave-pair and wave-chain kernels.
"""
import os
import sys

import numpy as np
import pytest

from effplast2d_amd import EffPlast2D
from oracle import ora
from tests.pairs import (PLASTIC_FIELDS, STATE_FIELDS, STRESS_FIELDS, make_pa, oracle_state,
                         random_state, rel_err)

pytestmark = pytest.mark.gpu

SHAPES = {"tiles8": {"tileRows": 8}, "march16": {"marchRows": 16},
          "pairs26": {"marchRows": 26, "itersPerLaunch": 2},
          "chain3": {"marchRows": 42, "itersPerLaunch": 3}}
# the plain scheme has one kernel set; the fused schemes run every shape
MODE_SHAPES = [("plain", "tiles8")] + [(m, sh) for m in ("fp64", "split") for sh in SHAPES]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(pa, nX, nl, mode, shape, init, iters):
    opts = {} if mode == "plain" else SHAPES[shape]
    with EffPlast2D(pa, nX, nTasks=nl, mode=mode, **opts) as g:
        for k in STATE_FIELDS:
            g.set(k, init[k])
        g.iterate(iters)
        out = {k: g.get(k) for k in STATE_FIELDS + STRESS_FIELDS}
        if nl > 1:
            out.update({k: g.get(k) for k in PLASTIC_FIELDS})
    return out


@pytest.mark.parametrize("mode,shape", MODE_SHAPES)
def test_zero_state_stays_zero(mode, shape):
    nX = 256
    _note(f"zero state mode={mode} shape={shape}")
    pa = make_pa(nX, nPores=2, porosity=0.03)
    z = {"Ux": np.zeros((nX + 1) * nX), "Uy": np.zeros(nX * (nX + 1)),
         "Vx": np.zeros((nX + 1) * nX), "Vy": np.zeros(nX * (nX + 1))}
    out = _run(pa, nX, 3, mode, shape, z, 37)
    for k, v in out.items():
        assert not np.any(v), f"{k}: {np.count_nonzero(v)} nonzero"


@pytest.mark.parametrize("mode", ["fp64", "split"])
@pytest.mark.parametrize("shape", ["tiles8", "pairs26", "chain3"])
def test_elastic_linearity_exact(mode, shape):
    """Elastic (one task), 1024^2: doubling the state doubles every field."""
    nX, iters = 1024, 41
    _note(f"linearity mode={mode} shape={shape}")
    pa = make_pa(nX, nPores=1, porosity=0.01)
    a = random_state(nX, 77, pa=pa, boundary_v=True)
    b = {k: 2.0 * v for k, v in a.items()}
    ra = _run(pa, nX, 1, mode, shape, a, iters)
    rb = _run(pa, nX, 1, mode, shape, b, iters)
    for k in ra:
        assert np.array_equal(2.0 * ra[k], rb[k]), \
            f"{k}: {np.count_nonzero(2.0 * ra[k] != rb[k])} differ"


def _centred_affine_state(pa, nX, e):
    """U of the isotropic strain e (no noise), V = 0, in the reference layout;
    coordinates as in EffPlast2D.cu:370-381 (origin at the grid centre)."""
    dX = pa[ora.PA_DX]
    xs = (np.arange(nX + 1) - 0.5 * nX) * dX
    ys = (np.arange(nX) - 0.5 * (nX - 1)) * dX
    Ux = np.broadcast_to(e * xs[None, :], (nX, nX + 1)).copy()
    xs2 = (np.arange(nX) - 0.5 * (nX - 1)) * dX
    ys2 = (np.arange(nX + 1) - 0.5 * nX) * dX
    Uy = np.broadcast_to(e * ys2[:, None], (nX + 1, nX)).copy()
    return {"Ux": Ux.ravel(), "Uy": Uy.ravel(), "Vx": np.zeros((nX + 1) * nX),
            "Vy": np.zeros(nX * (nX + 1))}


@pytest.mark.parametrize("mode", ["fp64", "split"])
@pytest.mark.parametrize("shape", ["tiles8", "pairs26", "chain3"])
@pytest.mark.parametrize("nX", [128, 1024])
def test_centred_pore_symmetries(mode, shape, nX):
    """One centred pore, isotropic compression (plastic, three tasks): P is
    mirror-symmetric in x and in y and symmetric under x <-> y, and
    Ux(i, j) = -Ux(nX-i, j) = Uy(j, i).  Rounding differs between mirrored
    cells (sums taken in another order), so the tolerance is relative:
    1e-9 for FP64, 1e-4 for the FP32 stress of the split scheme."""
    iters = 150
    _note(f"symmetry mode={mode} shape={shape} nX={nX}")
    pa = make_pa(nX, nPores=1, porosity=0.01)
    st = _centred_affine_state(pa, nX, -3e-3)
    out = _run(pa, nX, 3, mode, shape, st, iters)
    tol = 1e-9 if mode == "fp64" else 1e-4
    P = out["P"].reshape(nX, nX)
    sc = np.max(np.abs(P))
    assert sc > 0
    assert np.max(np.abs(P - P[:, ::-1])) <= tol * sc
    assert np.max(np.abs(P - P[::-1, :])) <= tol * sc
    assert np.max(np.abs(P - P.T)) <= tol * sc
    Ux = out["Ux"].reshape(nX, nX + 1)
    Uy = out["Uy"].reshape(nX + 1, nX)
    su = np.max(np.abs(Ux))
    utol = 1e-9 if mode == "fp64" else 1e-6
    assert np.max(np.abs(Ux + Ux[:, ::-1])) <= utol * su
    assert np.max(np.abs(Ux - Uy.T)) <= utol * su


@pytest.mark.parametrize("mode,shape", [ms for ms in MODE_SHAPES if ms[1] != "march16"])
def test_yield_bound_after_return_mapping(mode, shape):
    """J2 never exceeds the yield radius coh after the return mapping
    (EffPlast2D.cu:147-177, which recomputes J2 from the scaled stress), up
    to rounding.  (J2XY keeps the trial value there.)"""
    nX, iters = 512, 300
    _note(f"yield bound mode={mode} shape={shape}")
    pa = make_pa(nX, nPores=2, porosity=0.05)
    init = random_state(nX, 91, amp=4e-3, noise=0.5, pa=pa, boundary_v=True)
    out = _run(pa, nX, 3, mode, shape, init, iters)
    coh = pa[ora.PA_COH]
    eps = np.finfo(np.float64).eps
    assert np.max(out["J2"]) <= coh * (1 + 4 * eps)
    assert np.count_nonzero(out["J2"] >= coh * (1 - 1e-9)) > 0, "nothing yielded"


@pytest.mark.parametrize("shape", ["tiles8", "march16", "pairs26", "chain3"])
def test_long_run_matches_oracle(shape):
    """1000 plastic iterations at 128^2 (three pores), FP64, against the
    oracle at <= 1e-12 relative (bit for bit in practice)."""
    nX, iters = 128, 1000
    _note(f"long run shape={shape}")
    pa = make_pa(nX, nPores=3, porosity=0.04)
    init = random_state(nX, 1000, pa=pa, boundary_v=True)
    s = oracle_state(nX, pa, 3, init, nThreads=min(16, os.cpu_count() or 1))
    s.iterate(iters)
    out = _run(pa, nX, 3, "fp64", shape, init, iters)
    for k, v in out.items():
        e = rel_err(v, s[k])
        assert e <= 1e-12, f"{k}: rel err {e:.3e} ({np.count_nonzero(v != s[k])} differ)"
