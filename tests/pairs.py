"""Helpers that drive the HIP library and the CPU oracle on identical inputs.

The oracle (oracle/ora.py) is the checker; the library under test is
effplast2d_amd (libep2d.so)."""
from __future__ import annotations

import numpy as np

from oracle import ora


def make_pa(nX, nPores=1, porosity=0.01, Y=2.5e-3, Lx=20.0, K0=1.0, G0=0.5, rho0=1.0, CFL=0.5):
    return ora.derive_params(Lx, Lx, nX, nPores, porosity, Y, K0, G0, rho0, CFL)


def random_state(nX, seed, amp=3e-3, noise=0.3, vamp=1e-4, pa=None, boundary_v=False):
    """Affine compression plus noise: U and V arrays in the reference layout."""
    rng = np.random.default_rng(seed)
    dX = pa[ora.PA_DX] if pa is not None else 20.0 / (nX - 1)
    xs = (np.arange(nX + 1) - 0.5 * nX) * dX
    ys = (np.arange(nX) - 0.5 * (nX - 1)) * dX
    Ux = (-amp * xs[None, :] + 0 * ys[:, None])
    Ux = Ux + noise * amp * dX * rng.standard_normal(Ux.shape)
    xs2 = (np.arange(nX) - 0.5 * (nX - 1)) * dX
    ys2 = (np.arange(nX + 1) - 0.5 * nX) * dX
    Uy = (-amp * ys2[:, None] + 0 * xs2[None, :])
    Uy = Uy + noise * amp * dX * rng.standard_normal(Uy.shape)
    Vx = vamp * rng.standard_normal((nX, nX + 1))
    Vy = vamp * rng.standard_normal((nX + 1, nX))
    if not boundary_v:
        Vx[:, 0] = 0; Vx[:, -1] = 0; Vx[0, :] = 0; Vx[-1, :] = 0
        Vy[:, 0] = 0; Vy[:, -1] = 0; Vy[0, :] = 0; Vy[-1, :] = 0
    return {"Ux": Ux.ravel(), "Uy": Uy.ravel(), "Vx": Vx.ravel(), "Vy": Vy.ravel()}


def oracle_state(nX, pa, nl, init=None, **kw):
    s = ora.State(nX=nX, nY=nX, pa=np.asarray(pa, dtype=np.float64), nl=nl, **kw)
    if init:
        for k, v in init.items():
            s[k][:] = v
    return s


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / den)


def l2_rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


STRESS_FIELDS = ["P", "tauXX", "tauYY", "tauXY"]
PLASTIC_FIELDS = ["tauXYav", "J2", "J2XY"]
STATE_FIELDS = ["Ux", "Uy", "Vx", "Vy"]
