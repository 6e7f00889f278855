"""ctypes binding of the CPU oracle (oracle/libep2d_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
bench's cpu_baseline leg, never by the product package.  The oracle restates
the reference algorithm (EffPlast2D.cu, get_sigma_2D.m); see ep2d_oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libep2d_oracle.so")

NPARS = 11
PA_DX, PA_DY, PA_DT, PA_K0, PA_G0, PA_RHO, PA_DAMPX, PA_DAMPY, PA_COH, PA_RAD, PA_NPORES = range(11)

_dp = C.POINTER(C.c_double)
_ucp = C.POINTER(C.c_ubyte)


class Fields(C.Structure):
    _fields_ = [(n, _dp) for n in ("K", "G", "P0", "P", "tauXX", "tauYY", "tauXYav", "J2",
                                   "tauXY", "J2XY", "Ux", "Vx", "Uy", "Vy")] + \
               [("cellPore", _ucp), ("vertPore", _ucp)]


class Config(C.Structure):
    _fields_ = [("nX", C.c_long), ("nY", C.c_long), ("pa", C.c_double * NPARS),
                ("nl", C.c_int), ("nIter", C.c_long), ("eIter", C.c_double),
                ("outputStep", C.c_long), ("nThreads", C.c_int),
                ("useShearLoadType", C.c_int), ("shearLoadType", C.c_double * 3)]


REC_FIELDS = [("epsilon", 3), ("epsilonPer", 3), ("sigma", 4), ("sigmaPer", 4),
              ("deltaP", 1), ("deltaPper", 1), ("tauInfty", 1), ("dPhi", 1), ("dPhiPer", 1),
              ("holeAreaPi", 1), ("internalHoleAreaPi", 1), ("Phi0", 1), ("Phi", 1),
              ("PhiPer", 1), ("error", 1), ("iters", 1), ("sigmaXYclean", 1)]


class StepRecord(C.Structure):
    _fields_ = [(n, C.c_double * k) if k > 1 else (n, C.c_double) for n, k in REC_FIELDS]

    def as_dict(self):
        out = {}
        for n, k in REC_FIELDS:
            v = getattr(self, n)
            out[n] = list(v) if k > 1 else float(v)
        return out


AN_ARRAYS = ["UnuAbs", "J1nu", "J2nu", "UanAbs", "errorUabs", "J1an", "J2an", "errorJ1",
             "errorJ2"]
AN_STATS = ["errorUabsMax", "errorUabsAvg", "errorJ1Max", "errorJ1Avg", "errorJ2Max",
            "errorJ2Avg"]


class AnOut(C.Structure):
    _fields_ = [(n, _dp) for n in AN_ARRAYS] + [(n, C.c_double) for n in AN_STATS]


def an_point(fn, *args):
    """Point functions of the analytic static solution (two outputs)."""
    a, b = C.c_double(), C.c_double()
    getattr(lib(), fn)(*[float(x) for x in args], C.byref(a), C.byref(b))
    return a.value, b.value


class Moduli(C.Structure):
    _fields_ = [("Kphi", C.c_double), ("Kd", C.c_double), ("G", C.c_double)]

    def as_dict(self):
        return {"Kphi": self.Kphi, "Kd": self.Kd, "G": self.G}


class Result(C.Structure):
    _fields_ = [("an", Moduli), ("an_per", Moduli), ("num", Moduli), ("num_per", Moduli)] + \
               [(n, C.c_double) for n in (
                   "KphiExactElast", "KdExactElast", "GexactElast",
                   "KphiExactPlast", "KdExactPlast", "GexactPlast",
                   "KphiExactPlastPer", "KdExactPlastPer", "GexactPlastPer", "Gxy", "GxyPer")]

    def as_dict(self):
        d = {k: getattr(self, k).as_dict() for k in ("an", "an_per", "num", "num_per")}
        for n, _ in self._fields_[4:]:
            d[n] = getattr(self, n)
        return d


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.ora_derive_params.argtypes = [C.c_double, C.c_double, C.c_long, C.c_long, C.c_double,
                                        C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, _dp]
        L.ora_derive_params.restype = None
        for fn in ("ora_set_materials", "ora_set_pore_masks", "ora_set_init_pressure",
                   "ora_compute_stress", "ora_compute_j2", "ora_compute_plasticity",
                   "ora_compute_disp"):
            getattr(L, fn).argtypes = [C.POINTER(Config), C.POINTER(Fields)]
            getattr(L, fn).restype = None
        L.ora_iterate.argtypes = [C.POINTER(Config), C.POINTER(Fields), C.c_long]
        L.ora_iterate.restype = None
        L.ora_find_max_abs.argtypes = [_dp, C.c_long]
        L.ora_find_max_abs.restype = C.c_double
        L.ora_find_max_abs_signed.argtypes = [_dp, C.c_long]
        L.ora_find_max_abs_signed.restype = C.c_double
        L.ora_residual.argtypes = [C.POINTER(Config), C.POINTER(Fields), _dp]
        L.ora_residual.restype = C.c_double
        L.ora_apply_affine.argtypes = [C.POINTER(Config), C.POINTER(Fields)] + [C.c_double] * 4
        L.ora_apply_affine.restype = None
        L.ora_averages.argtypes = [C.POINTER(Config), C.POINTER(Fields), C.POINTER(StepRecord)]
        L.ora_averages.restype = None
        for fn in ("ora_delta_p_honest", "ora_delta_p_periodic", "ora_tau_infty_honest"):
            getattr(L, fn).argtypes = [C.POINTER(Config), C.POINTER(Fields)]
            getattr(L, fn).restype = C.c_double
        L.ora_pore_in_u_node.argtypes = [C.POINTER(Config), C.c_int, C.c_long, C.c_long]
        L.ora_pore_in_u_node.restype = C.c_int
        L.ora_hole_area.argtypes = [C.POINTER(Config), C.POINTER(Fields), C.POINTER(StepRecord)]
        L.ora_hole_area.restype = C.c_int
        L.ora_compute_eff_moduli.argtypes = [C.POINTER(Config), C.POINTER(Fields), C.c_double,
                                             _dp, C.c_long, C.POINTER(StepRecord),
                                             C.POINTER(Result)]
        L.ora_compute_eff_moduli.restype = C.c_int
        L.ora_compute_eff_damping.argtypes = [C.POINTER(Config), C.POINTER(Fields), C.c_double,
                                              _dp, C.c_double, C.c_long,
                                              C.POINTER(StepRecord), C.POINTER(Result)]
        L.ora_compute_eff_damping.restype = C.c_int
        L.ora_analytic_moduli.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(StepRecord),
                                          C.POINTER(Result)]
        L.ora_analytic_moduli.restype = None
        L.ora_moduli_getters.argtypes = [C.c_int, C.POINTER(StepRecord), C.POINTER(StepRecord),
                                         C.POINTER(StepRecord), C.POINTER(Result)]
        L.ora_moduli_getters.restype = None
        L.ora_an_ur_hydro.argtypes = [C.c_double] * 5
        L.ora_an_ur_hydro.restype = C.c_double
        L.ora_an_u_elast.argtypes = [C.c_double] * 10 + [_dp, _dp]
        L.ora_an_u_elast.restype = None
        L.ora_an_j_elast.argtypes = [C.c_double] * 7 + [_dp, _dp]
        L.ora_an_j_elast.restype = None
        L.ora_an_j_plast.argtypes = [C.c_double] * 4 + [_dp, _dp]
        L.ora_an_j_plast.restype = None
        L.ora_an_static.argtypes = [C.POINTER(Config), C.POINTER(Fields), C.c_double, C.c_double,
                                    _dp, C.POINTER(AnOut)]
        L.ora_an_static.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def make_record(**kw):
    """A StepRecord with the given fields (the rest zero)."""
    r = StepRecord()
    for n, k in REC_FIELDS:
        if n in kw:
            v = kw[n]
            if k > 1:
                arr = getattr(r, n)
                for i in range(k):
                    arr[i] = float(v[i]) if i < len(v) else 0.0
            else:
                setattr(r, n, float(v))
    return r


def moduli_getters(nl, r0, r1=None, r2=None):
    """The numerical moduli of EffPlast2D.cu:1412-1539 from step records
    (dicts of record fields)."""
    recs = [make_record(**(r or {})) for r in (r0, r1, r2)]
    res = Result()
    lib().ora_moduli_getters(int(nl), C.byref(recs[0]), C.byref(recs[1]), C.byref(recs[2]),
                             C.byref(res))
    return res.as_dict()


def derive_params(Lx, Ly, nX, nPores, porosity, Y, K0=1.0, G0=0.5, rho0=1.0, CFL=0.5, nY=None):
    """get_sigma_2D.m:3-32 restated in C; returns the 11-double pa vector."""
    pa = np.zeros(NPARS)
    lib().ora_derive_params(Lx, Ly, nX, nX if nY is None else nY, float(nPores), porosity, Y,
                            K0, G0, rho0, CFL, _ptr(pa))
    return pa


@dataclass
class State:
    """Oracle configuration plus numpy-owned field arrays."""
    nX: int
    nY: int
    pa: np.ndarray
    nl: int = 3
    nIter: int = 100_000
    eIter: float = 1e-10
    outputStep: int = 10_000
    nThreads: int = 0
    shearLoadType: tuple | None = None
    arrays: dict = field(default_factory=dict)

    def __post_init__(self):
        nX, nY = self.nX, self.nY
        shapes = {"K": nX * nY, "G": nX * nY, "P0": nX * nY, "P": nX * nY, "tauXX": nX * nY,
                  "tauYY": nX * nY, "tauXYav": nX * nY, "J2": nX * nY,
                  "tauXY": (nX - 1) * (nY - 1), "J2XY": (nX - 1) * (nY - 1),
                  "Ux": (nX + 1) * nY, "Vx": (nX + 1) * nY, "Uy": nX * (nY + 1), "Vy": nX * (nY + 1)}
        for k, n in shapes.items():
            self.arrays[k] = np.zeros(n)
        self.arrays["cellPore"] = np.zeros(nX * nY, dtype=np.uint8)
        self.arrays["vertPore"] = np.zeros((nX - 1) * (nY - 1), dtype=np.uint8)
        self.fields = Fields(**{k: (v.ctypes.data_as(_ucp) if v.dtype == np.uint8 else _ptr(v))
                                for k, v in self.arrays.items()})
        self.config = Config()
        self.sync_config()
        L = lib()
        L.ora_set_materials(C.byref(self.config), C.byref(self.fields))
        L.ora_set_pore_masks(C.byref(self.config), C.byref(self.fields))
        L.ora_set_init_pressure(C.byref(self.config), C.byref(self.fields))

    def sync_config(self):
        c = self.config
        c.nX, c.nY = self.nX, self.nY
        for k in range(NPARS):
            c.pa[k] = float(self.pa[k])
        c.nl, c.nIter, c.eIter = self.nl, self.nIter, self.eIter
        c.outputStep, c.nThreads = self.outputStep, self.nThreads
        c.useShearLoadType = 1 if self.shearLoadType is not None else 0
        if self.shearLoadType is not None:
            for k in range(3):
                c.shearLoadType[k] = float(self.shearLoadType[k])

    def __getitem__(self, k):
        return self.arrays[k]

    # -- kernels -------------------------------------------------------------
    def stress(self):
        lib().ora_compute_stress(C.byref(self.config), C.byref(self.fields))

    def j2(self):
        lib().ora_compute_j2(C.byref(self.config), C.byref(self.fields))

    def plasticity(self):
        lib().ora_compute_plasticity(C.byref(self.config), C.byref(self.fields))

    def disp(self):
        lib().ora_compute_disp(C.byref(self.config), C.byref(self.fields))

    def iterate(self, n):
        self.sync_config()
        lib().ora_iterate(C.byref(self.config), C.byref(self.fields), int(n))

    def residual(self, cur):
        cur = np.asarray(cur, dtype=np.float64)
        return lib().ora_residual(C.byref(self.config), C.byref(self.fields), _ptr(cur))

    def apply_affine(self, dUxdx, dUydy, dUxdy, dUydx):
        lib().ora_apply_affine(C.byref(self.config), C.byref(self.fields),
                               dUxdx, dUydy, dUxdy, dUydx)

    def averages(self):
        r = StepRecord()
        lib().ora_averages(C.byref(self.config), C.byref(self.fields), C.byref(r))
        return r.as_dict()

    def boundary(self):
        L = lib()
        return {"deltaP": L.ora_delta_p_honest(C.byref(self.config), C.byref(self.fields)),
                "deltaPper": L.ora_delta_p_periodic(C.byref(self.config), C.byref(self.fields)),
                "tauInfty": L.ora_tau_infty_honest(C.byref(self.config), C.byref(self.fields))}

    def hole_area(self):
        r = StepRecord()
        rc = lib().ora_hole_area(C.byref(self.config), C.byref(self.fields), C.byref(r))
        if rc:
            raise ValueError("pore-edge sample outside the grid")
        return r.as_dict()

    def an_static(self, deltaP, tauInfty, loadType):
        """SaveAnStatic2D's fields and error statistics (no files)."""
        self.sync_config()
        nC, nV = self.nX * self.nY, (self.nX - 1) * (self.nY - 1)
        arrs = {n: np.zeros(nC if n in ("UnuAbs", "UanAbs", "errorUabs") else nV)
                for n in AN_ARRAYS}
        o = AnOut(**{n: _ptr(a) for n, a in arrs.items()})
        lt = np.asarray(loadType, dtype=np.float64)
        an = lib().ora_an_static(C.byref(self.config), C.byref(self.fields), float(deltaP),
                                 float(tauInfty), _ptr(lt), C.byref(o))
        out = dict(arrs)
        out.update({n: getattr(o, n) for n in AN_STATS})
        out["analytic"] = int(an)
        return out

    def analytic_moduli(self, step, rec):
        """Per-step analytic moduli (EffPlast2D.cu:643-696) of one record."""
        self.sync_config()
        res = Result()
        r = make_record(**rec)
        lib().ora_analytic_moduli(C.byref(self.config), int(step), C.byref(r), C.byref(res))
        return res.as_dict()

    # -- drivers -------------------------------------------------------------
    def compute_eff_moduli(self, initLoad, loadType, nTimeSteps=1):
        self.sync_config()
        T = max(int(nTimeSteps), 1)
        recs = (StepRecord * (max(self.nl, 1) * T))()
        res = Result()
        lt = np.asarray(loadType, dtype=np.float64)
        rc = lib().ora_compute_eff_moduli(C.byref(self.config), C.byref(self.fields),
                                          float(initLoad), _ptr(lt), T, recs, C.byref(res))
        if rc:
            raise ValueError(f"ora_compute_eff_moduli failed rc={rc}")
        return res.as_dict(), [[recs[s * T + t].as_dict() for t in range(T if s == 0 else 1)]
                               for s in range(self.nl)]

    def compute_eff_damping(self, initLoad, loadType, loadValue, nTimeSteps=1):
        self.sync_config()
        T = max(int(nTimeSteps), 1)
        recs = (StepRecord * (self.nl * T))()
        res = Result()
        lt = np.asarray(loadType, dtype=np.float64)
        rc = lib().ora_compute_eff_damping(C.byref(self.config), C.byref(self.fields),
                                           float(initLoad), _ptr(lt), float(loadValue), T, recs,
                                           C.byref(res))
        if rc:
            raise ValueError(f"ora_compute_eff_damping failed rc={rc}")
        return res.as_dict(), [[recs[s * T + t].as_dict() for t in range(T if s == 0 else 1)]
                               for s in range(self.nl)]


def make_state(nX, Lx=20.0, nPores=1, porosity=0.005, Y=2.5e-3, K0=1.0, G0=0.5, rho0=1.0,
               CFL=0.5, **kw):
    pa = derive_params(Lx, Lx, nX, nPores, porosity, Y, K0, G0, rho0, CFL)
    return State(nX=nX, nY=nX, pa=pa, **kw)
