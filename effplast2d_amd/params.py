"""Parameter derivation and the data/pa.dat format.

``derive_params`` is get_sigma_2D.m:3-32 evaluated by the C ABI
(ep2d_derive_params); ``read_pa``/``write_pa`` are the 11-double
little-endian files of get_sigma_2D.m:41-43 and ReadParams
(EffPlast2D.cu:700-707).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib


def derive_params(Lx=20.0, Ly=20.0, nGrid=4, nX=0, nPores=1, porosity=0.005, Y=2.5e-3,
                  K0=1.0, G0=0.5, rho0=1.0, CFL=0.5):
    """Return (pa[11], nX)."""
    ph = L.Physics(Lx=Lx, Ly=Ly, nGrid=int(nGrid), nX=int(nX), nPores=int(nPores),
                   porosity=porosity, Y=Y, K0=K0, G0=G0, rho0=rho0, CFL=CFL)
    pa = np.zeros(L.NPARS)
    n = C.c_int()
    check(lib().ep2d_derive_params(C.byref(ph), pa.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(n)))
    return pa, n.value


def read_pa(path):
    pa = np.zeros(L.NPARS)
    check(lib().ep2d_read_pa(str(path).encode(), pa.ctypes.data_as(C.POINTER(C.c_double))))
    return pa


def write_pa(path, pa):
    a = np.ascontiguousarray(pa, dtype=np.float64)
    if a.shape != (L.NPARS,):
        raise ValueError("pa must hold 11 values")
    check(lib().ep2d_write_pa(str(path).encode(), a.ctypes.data_as(C.POINTER(C.c_double))))
