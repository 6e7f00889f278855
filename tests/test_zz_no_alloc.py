"""GPU: ep2d_residual_noalloc, the residual on caller-owned velocity fields
with no device allocation, against ep2d_residual on the same fields.

Two calls in a row through the C ABI -- a 160^2 split RVE, then a smaller
96^2 FP64 RVE -- each compared bit for bit with the handle's own residual
(the maximum is exact and the formula is evaluated in the same order), and
the non-finite count with it.  The outputs live in device memory the test
allocates with torch (the entry point writes them on the caller's stream).
"""
import ctypes as C
import sys

import numpy as np
import pytest
# torch before the library (one HIP runtime in the process)
import torch

from effplast2d_amd import EffPlast2D
from effplast2d_amd._lib import check, lib
from tests.pairs import make_pa, random_state

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc_residual(g, nX, mode, strain, stream):
    pa = g.pa
    err = torch.zeros(1, dtype=torch.float64, device="cuda")
    bad = torch.full((1,), 7, dtype=torch.int64, device="cuda")
    e = (C.c_double * 3)(*strain)
    rc = lib().ep2d_residual_noalloc(
        C.c_void_p(g.device_ptr("Vx")), C.c_void_p(g.device_ptr("Vy")), nX, nX,
        1 if mode == "split" else 0, C.c_double(pa[0]), C.c_double(pa[1]), C.c_double(pa[2]),
        e, C.c_void_p(err.data_ptr()), C.c_void_p(bad.data_ptr()), C.c_void_p(stream))
    check(rc)
    torch.cuda.current_stream().synchronize()
    return float(err.item()), int(bad.item())


def test_noalloc_residual_equals_handle_residual():
    stream = torch.cuda.current_stream().cuda_stream
    for nX, mode, seed, strain in ((160, "split", 5, [-3e-3, -2e-3, 1e-3]),
                                   (96, "fp64", 6, [-1e-3, -1e-3, 0.0])):
        _note(f"noalloc residual nX={nX} mode={mode}")
        pa = make_pa(nX, nPores=2, porosity=0.03)
        init = random_state(nX, seed, pa=pa, boundary_v=True)
        with EffPlast2D(pa, nX, nTasks=3, mode=mode) as g:
            for k in ("Ux", "Uy", "Vx", "Vy"):
                g.set(k, init[k])
            g.iterate(25)
            g.sync()
            ref, ref_bad = g.residual(strain)
            got, got_bad = _noalloc_residual(g, nX, mode, strain, stream)
        assert got == ref, (got, ref)
        assert got_bad == ref_bad == 0


def test_noalloc_rejects_sizes_beyond_the_maximum():
    e = (C.c_double * 3)(-1e-3, -1e-3, 0.0)
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = C.c_void_p(dummy.data_ptr())
    for n in (2, 32769):
        rc = lib().ep2d_residual_noalloc(p, p, n, 64, 0, C.c_double(0.1), C.c_double(0.1),
                                         C.c_double(0.05), e, p, None, None)
        assert rc != 0
        assert b"EP2D_NOALLOC_MAX_SIDE" in lib().ep2d_last_error()
    assert np.isfinite(dummy.cpu().numpy()).all()
