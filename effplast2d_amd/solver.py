"""Python mirror of the reference's EffPlast2D class over libep2d.so.

Reference interface (EffPlast2D.h:53-65):
    EffPlast2D(nTimeSteps=1)                      -> EffPlast2D(...)
    ComputeEffModuli(initLoadValue, loadType, loadValue=0.0)
    ComputeEffDamping(initLoadValue, loadType, loadValue=0.0)
The reference takes NGRID/NL/NITER/EITER/DEVICE_IDX as compile-time macros and
pa[] from data/pa.dat; here they are constructor arguments.  Results are
returned as dicts (the reference only prints them).  Every call goes through
the C ABI; device errors raise ``Ep2dError`` instead of exiting.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def default_options(**kw) -> L.Options:
    o = L.Options()
    lib().ep2d_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "mode" and isinstance(v, str):
            v = L.MODES[v]
        if k == "shearLoadType":
            if v is None:
                o.useShearLoadType = 0
                continue
            o.useShearLoadType = 1
            for i in range(3):
                o.shearLoadType[i] = float(v[i])
            continue
        if k == "logPath" and v is not None:
            v = str(v).encode()
        setattr(o, k, v)
    return o


class EffPlast2D:
    """One RVE on the GPU.  ``pa`` is the 11-double parameter vector of
    data/pa.dat (get_sigma_2D.m:32); ``nX`` the grid size (32*NGRID)."""

    def __init__(self, pa, nX, nTimeSteps=1, nTasks=3, nIter=100_000, eIter=1e-10,
                 checkInterval=10_000, mode="fp64", device=0, shearLoadType=None,
                 verbose=False, logPath=None, tileRows=0, marchRows=0, itersPerLaunch=0,
                 adaptiveCheck=False, stream=None, laneColumns=0, keepFp32Stress=False,
                 stageWaves=0):
        self.pa = np.ascontiguousarray(pa, dtype=np.float64)
        if self.pa.shape != (L.NPARS,):
            raise ValueError("pa must hold 11 values")
        self.nX = self.nY = int(nX)
        self.nTasks = int(nTasks)
        self.mode = mode
        self._opts = default_options(nTasks=nTasks, nIter=nIter, eIter=eIter,
                                     checkInterval=checkInterval, mode=mode, device=device,
                                     nTimeSteps=nTimeSteps, shearLoadType=shearLoadType,
                                     verbose=1 if verbose else 0, logPath=logPath,
                                     tileRows=tileRows, marchRows=marchRows,
                                     itersPerLaunch=itersPerLaunch,
                                     adaptiveCheck=1 if adaptiveCheck else 0,
                                     laneColumns=laneColumns,
                                     keepFp32Stress=1 if keepFp32Stress else 0,
                                     stageWaves=int(stageWaves))
        h = C.c_void_p()
        # stream: a hipStream_t (as an integer) the handle orders its work on
        check(lib().ep2d_create_on_stream(_dptr(self.pa), self.nX, self.nY, C.byref(self._opts),
                                          C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h
        L.track(self)

    @classmethod
    def from_physics(cls, Lx=20.0, Ly=None, nGrid=4, nX=0, nPores=1, porosity=0.005, Y=2.5e-3,
                     K0=1.0, G0=0.5, rho0=1.0, CFL=0.5, **kw):
        from .params import derive_params
        pa, n = derive_params(Lx=Lx, Ly=Lx if Ly is None else Ly, nGrid=nGrid, nX=nX,
                              nPores=nPores, porosity=porosity, Y=Y, K0=K0, G0=G0, rho0=rho0,
                              CFL=CFL)
        return cls(pa, n, **kw)

    # -- lifetime ------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().ep2d_destroy(self._h)
            self._h = None

    def __del__(self):
        if L.finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- reference API -------------------------------------------------------
    def ComputeEffModuli(self, initLoadValue, loadType, loadValue=0.0):
        """EffPlast2D::ComputeEffModuli (EffPlast2D.cu:179-229)."""
        lt = np.ascontiguousarray(loadType, dtype=np.float64)
        res = L.Result()
        check(lib().ep2d_compute_eff_moduli(self._h, float(initLoadValue), _dptr(lt),
                                            float(loadValue), C.byref(res)))
        return res.as_dict()

    def ComputeEffDamping(self, initLoadValue, loadType, loadValue=0.0):
        """EffPlast2D::ComputeEffDamping (EffPlast2D.cu:230-310)."""
        lt = np.ascontiguousarray(loadType, dtype=np.float64)
        res = L.Result()
        check(lib().ep2d_compute_eff_damping(self._h, float(initLoadValue), _dptr(lt),
                                             float(loadValue), C.byref(res)))
        return res.as_dict()

    def records(self):
        n = C.c_int()
        check(lib().ep2d_get_records(self._h, None, 0, C.byref(n)))
        arr = (L.Record * max(n.value, 1))()
        check(lib().ep2d_get_records(self._h, arr, n.value, C.byref(n)))
        return [arr[k].as_dict() for k in range(n.value)]

    # -- fields --------------------------------------------------------------
    def field_shape(self, name):
        m, n = C.c_int(), C.c_int()
        check(lib().ep2d_field_shape(self._h, L.FIELD_ID[name], C.byref(m), C.byref(n)))
        return m.value, n.value

    def get(self, name):
        """Field in the reference layout, flat A[j*m + i]."""
        m, n = self.field_shape(name)
        out = np.empty(m * n, dtype=np.float64)
        check(lib().ep2d_get_field(self._h, L.FIELD_ID[name], _dptr(out), out.size))
        return out

    def device_ptr(self, name):
        """Device address of a field of the current state (ep2d_field_device)."""
        p = C.c_void_p()
        check(lib().ep2d_field_device(self._h, L.FIELD_ID[name], C.byref(p)))
        return p.value

    def set(self, name, values):
        m, n = self.field_shape(name)
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if a.size != m * n:
            raise ValueError(f"{name} needs {m * n} values")
        check(lib().ep2d_set_field(self._h, L.FIELD_ID[name], _dptr(a), a.size))

    def save(self, name, path):
        """SaveMatrix (EffPlast2D.cu:795-799)."""
        check(lib().ep2d_save_field(self._h, L.FIELD_ID[name], str(path).encode()))

    def an_static(self, deltaP, tauInfty, loadType):
        """SaveAnStatic2D's fields (EffPlast2D.cu:1060-1292) for the current
        state, as arrays, with the error statistics; no files."""
        nC, nV = self.nX * self.nY, (self.nX - 1) * (self.nY - 1)
        arrs = {n: np.zeros(nC if n in ("UnuAbs", "UanAbs", "errorUabs") else nV)
                for n in L.AN_ARRAYS}
        o = L.AnStatic(**{n: _dptr(a) for n, a in arrs.items()})
        lt = np.ascontiguousarray(loadType, dtype=np.float64)
        check(lib().ep2d_an_static_fields(self._h, float(deltaP), float(tauInfty), _dptr(lt),
                                          C.byref(o)))
        out = dict(arrs)
        out.update({n: getattr(o, n) for n in L.AN_STATS})
        out["analytic"] = o.analytic
        return out

    def save_an_static(self, deltaP, tauInfty, loadType, directory):
        """The same written as the reference's data/<name>_<nX>_.dat files."""
        o = L.AnStatic()
        lt = np.ascontiguousarray(loadType, dtype=np.float64)
        check(lib().ep2d_save_an_static(self._h, float(deltaP), float(tauInfty), _dptr(lt),
                                        str(directory).encode(), C.byref(o)))
        return {n: getattr(o, n) for n in L.AN_STATS} | {"analytic": o.analytic}

    def mask(self):
        out = np.empty(self.nX * self.nY, dtype=np.uint8)
        check(lib().ep2d_get_mask(self._h, out.ctypes.data_as(C.POINTER(C.c_ubyte)), out.size))
        return out

    # -- step-level ----------------------------------------------------------
    def iterate(self, n):
        check(lib().ep2d_iterate(self._h, int(n)))

    def apply_affine(self, dUxdx, dUydy, dUxdy, dUydx):
        check(lib().ep2d_apply_affine(self._h, dUxdx, dUydy, dUxdy, dUydx))

    def residual(self, curEffStrain):
        e = np.ascontiguousarray(curEffStrain, dtype=np.float64)
        err = C.c_double()
        bad = C.c_long()
        check(lib().ep2d_residual(self._h, _dptr(e), C.byref(err), C.byref(bad)))
        return err.value, bad.value

    def averages(self):
        r = L.Record()
        check(lib().ep2d_averages(self._h, C.byref(r)))
        return r.as_dict()

    def boundary(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        check(lib().ep2d_boundary(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"deltaP": a.value, "deltaPper": b.value, "tauInfty": c.value}

    def hole_area(self):
        r = L.Record()
        check(lib().ep2d_hole_area(self._h, C.byref(r)))
        return r.as_dict()

    def set_timing(self, sample_every):
        check(lib().ep2d_set_timing(self._h, int(sample_every)))

    def timing(self):
        ms, n, name = C.c_double(), C.c_long(), C.c_char_p()
        check(lib().ep2d_get_timing(self._h, C.byref(ms), C.byref(n), C.byref(name)))
        return ms.value, n.value, name.value.decode()

    def sync(self):
        check(lib().ep2d_sync(self._h))

    def device_bytes(self):
        return int(lib().ep2d_device_bytes(self._h))


def device_count():
    return int(lib().ep2d_device_count())
