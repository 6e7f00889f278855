"""The reference's binary field files: data/<var>_<Nx>_.dat.

SaveMatrix (EffPlast2D.cu:795-799) writes a field as raw little-endian
float64, x fastest (A[j*m + i]), no header; read_data_2D.m:1-12 reads it back
with ``reshape(field, nX, nY)`` (column-major), i.e. MATLAB's field(i+1, j+1)
is A[j*m + i].
"""
from __future__ import annotations

import os

import numpy as np


def dat_name(var_name, size):
    """read_data_2D.m:2 -- strcat(var_name, '_', int2str(size), '_.dat')."""
    return f"{var_name}_{int(size)}_.dat"


def write_dat(path, flat):
    np.ascontiguousarray(flat, dtype="<f8").tofile(path)


def read_data_2D(var_name, size, nX, nY):
    """read_data_2D.m: returns the (nX, nY) array with [i, j] = A[j*nX + i];
    zeros (with a message) when the file is missing, as in the reference."""
    fn = dat_name(var_name, size)
    if not os.path.isfile(fn):
        print(f"Error! File {fn} does not exist!")
        return np.zeros((nX, nY))
    a = np.fromfile(fn, dtype="<f8")
    return a.reshape(nY, nX).T
