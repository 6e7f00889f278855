"""ctypes binding of libep2d.so (include/ep2d.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C effplast2d_amd/csrc``).  Loading fails loudly when it is missing:
there is no fallback implementation.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libep2d.so")

NPARS = 11
(PA_DX, PA_DY, PA_DT, PA_K0, PA_G0, PA_RHO, PA_DAMPX, PA_DAMPY, PA_COH, PA_RAD,
 PA_NPORES) = range(NPARS)

MODE_PLAIN, MODE_FP64, MODE_SPLIT = 0, 1, 2
MODES = {"plain": MODE_PLAIN, "fp64": MODE_FP64, "split": MODE_SPLIT}

FIELDS = ["P", "tauXX", "tauYY", "tauXY", "tauXYav", "J2", "J2XY", "Ux", "Uy", "Vx", "Vy",
          "K", "G"]
FIELD_ID = {n: k for k, n in enumerate(FIELDS)}

STATUS = {0: "EP2D_OK", -1: "EP2D_ERR_ARG", -2: "EP2D_ERR_HIP", -3: "EP2D_ERR_IO",
          -4: "EP2D_ERR_GEOMETRY", -5: "EP2D_ERR_NONFINITE", -6: "EP2D_ERR_MEMORY",
          -7: "EP2D_ERR_STATE"}

_dp = C.POINTER(C.c_double)


class Physics(C.Structure):
    _fields_ = [("Lx", C.c_double), ("Ly", C.c_double), ("nGrid", C.c_int), ("nX", C.c_int),
                ("nPores", C.c_int), ("porosity", C.c_double), ("Y", C.c_double),
                ("K0", C.c_double), ("G0", C.c_double), ("rho0", C.c_double),
                ("CFL", C.c_double)]


class Options(C.Structure):
    _fields_ = [("nTasks", C.c_int), ("nIter", C.c_long), ("eIter", C.c_double),
                ("checkInterval", C.c_long), ("mode", C.c_int), ("device", C.c_int),
                ("nTimeSteps", C.c_int), ("useShearLoadType", C.c_int),
                ("shearLoadType", C.c_double * 3), ("verbose", C.c_int),
                ("logPath", C.c_char_p), ("tileRows", C.c_int), ("marchRows", C.c_int),
                ("itersPerLaunch", C.c_int), ("adaptiveCheck", C.c_int),
                ("laneColumns", C.c_int), ("keepFp32Stress", C.c_int), ("stageWaves", C.c_int)]


class Moduli(C.Structure):
    _fields_ = [("Kphi", C.c_double), ("Kd", C.c_double), ("G", C.c_double)]

    def as_dict(self):
        return {"Kphi": self.Kphi, "Kd": self.Kd, "G": self.G}


REC_FIELDS = [("epsilon", 3), ("epsilonPer", 3), ("sigma", 4), ("sigmaPer", 4),
              ("deltaP", 1), ("deltaPper", 1), ("tauInfty", 1), ("dPhi", 1), ("dPhiPer", 1),
              ("holeAreaPi", 1), ("internalHoleAreaPi", 1), ("Phi0", 1), ("Phi", 1),
              ("PhiPer", 1), ("error", 1), ("iters", 1), ("sigmaXYclean", 1)]


class Record(C.Structure):
    _fields_ = [(n, C.c_double * k) if k > 1 else (n, C.c_double) for n, k in REC_FIELDS]

    def as_dict(self):
        out = {}
        for n, k in REC_FIELDS:
            v = getattr(self, n)
            out[n] = list(v) if k > 1 else float(v)
        return out


_RES_SCALARS = ["KphiExactElast", "KdExactElast", "GexactElast", "KphiExactPlast",
                "KdExactPlast", "GexactPlast", "KphiExactPlastPer", "KdExactPlastPer",
                "GexactPlastPer", "Gxy", "GxyPer"]


class Result(C.Structure):
    _fields_ = [("an", Moduli), ("an_per", Moduli), ("num", Moduli), ("num_per", Moduli)] + \
               [(n, C.c_double) for n in _RES_SCALARS] + \
               [("nRecords", C.c_int), ("totalIters", C.c_long), ("seconds", C.c_double),
                ("converged", C.c_int)]

    def as_dict(self):
        d = {k: getattr(self, k).as_dict() for k in ("an", "an_per", "num", "num_per")}
        for n in _RES_SCALARS + ["nRecords", "totalIters", "seconds", "converged"]:
            d[n] = getattr(self, n)
        return d


class BatchOptions(C.Structure):
    _fields_ = [("mode", C.c_int), ("checkInterval", C.c_long), ("tileRows", C.c_int),
                ("marchRows", C.c_int), ("persistRows", C.c_int), ("march2Rows", C.c_int),
                ("pairRows", C.c_int), ("tripleRows", C.c_int), ("quadRows", C.c_int),
                ("device", C.c_int), ("memoryBudget", C.c_size_t), ("maxActive", C.c_int),
                ("useGraphs", C.c_int), ("timingEvery", C.c_long), ("adaptiveCheck", C.c_int),
                ("wide3Rows", C.c_int), ("wide4Rows", C.c_int), ("keepFp32Stress", C.c_int),
                ("triple2Rows", C.c_int), ("quad2Rows", C.c_int)]


class Request(C.Structure):
    _fields_ = [("pa", C.c_double * NPARS), ("nX", C.c_int), ("nTasks", C.c_int),
                ("nIter", C.c_long), ("eIter", C.c_double), ("nTimeSteps", C.c_int),
                ("initLoadValue", C.c_double), ("loadType", C.c_double * 3),
                ("loadValue", C.c_double), ("useShearLoadType", C.c_int),
                ("shearLoadType", C.c_double * 3), ("checkInterval", C.c_long),
                ("tag", C.c_longlong)]


class BatchResult(C.Structure):
    _fields_ = [("id", C.c_longlong), ("tag", C.c_longlong), ("status", C.c_int),
                ("result", Result), ("tSubmit", C.c_double), ("tAdmit", C.c_double),
                ("tDone", C.c_double), ("message", C.c_char * 160)]

    def as_dict(self):
        return {"id": self.id, "tag": self.tag, "status": self.status,
                "result": self.result.as_dict(), "tSubmit": self.tSubmit,
                "tAdmit": self.tAdmit, "tDone": self.tDone,
                "message": self.message.decode(errors="replace")}


MAX_CAND = 32  # EP2D_MAX_CAND


class BatchStats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in ("launches", "chunks", "tilesLaunched", "submitted",
                                            "finished", "failed", "cellIterations")] + \
               [("bytesAlgorithmic", C.c_double), ("timedMs", C.c_double),
                ("timedBytes", C.c_double), ("timedLaunches", C.c_longlong),
                ("graphMs", C.c_double), ("graphLaunches", C.c_longlong),
                ("active", C.c_int), ("queued", C.c_int), ("deviceBytes", C.c_size_t),
                ("kernel", C.c_int), ("kernelRows", C.c_int), ("tunings", C.c_int),
                ("nCand", C.c_int), ("candKind", C.c_int * MAX_CAND),
                ("candRows", C.c_int * MAX_CAND), ("candUs", C.c_double * MAX_CAND),
                ("fp64Finishes", C.c_longlong), ("fp64CellIterations", C.c_longlong)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        n = self.nCand
        for k in ("candKind", "candRows", "candUs"):
            d[k] = list(getattr(self, k))[:n]
        return d


AN_ARRAYS = ["UnuAbs", "J1nu", "J2nu", "UanAbs", "errorUabs", "J1an", "J2an", "errorJ1",
             "errorJ2"]
AN_STATS = ["errorUabsMax", "errorUabsAvg", "errorJ1Max", "errorJ1Avg", "errorJ2Max",
            "errorJ2Avg"]


class AnStatic(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in AN_ARRAYS] + \
               [(n, C.c_double) for n in AN_STATS] + [("analytic", C.c_int)]


class Ep2dError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


# Every symbol include/ep2d.h declares, with its ctypes signature.
SIGNATURES = {
    "ep2d_last_error": ([], C.c_char_p),
    "ep2d_struct_size": ([C.c_char_p], C.c_size_t),
    "ep2d_version": ([], C.c_char_p),
    "ep2d_device_count": ([], C.c_int),
    "ep2d_derive_params": ([C.POINTER(Physics), _dp, C.POINTER(C.c_int)], C.c_int),
    "ep2d_read_pa": ([C.c_char_p, _dp], C.c_int),
    "ep2d_write_pa": ([C.c_char_p, _dp], C.c_int),
    "ep2d_default_options": ([C.POINTER(Options)], None),
    "ep2d_create": ([_dp, C.c_int, C.c_int, C.POINTER(Options), C.POINTER(C.c_void_p)], C.c_int),
    "ep2d_create_on_stream": ([_dp, C.c_int, C.c_int, C.POINTER(Options), C.c_void_p,
                               C.POINTER(C.c_void_p)], C.c_int),
    "ep2d_field_device": ([C.c_void_p, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "ep2d_create_from_physics": ([C.POINTER(Physics), C.POINTER(Options),
                                  C.POINTER(C.c_void_p)], C.c_int),
    "ep2d_destroy": ([C.c_void_p], None),
    "ep2d_compute_eff_moduli": ([C.c_void_p, C.c_double, _dp, C.c_double, C.POINTER(Result)],
                                C.c_int),
    "ep2d_compute_eff_damping": ([C.c_void_p, C.c_double, _dp, C.c_double, C.POINTER(Result)],
                                 C.c_int),
    "ep2d_get_records": ([C.c_void_p, C.POINTER(Record), C.c_int, C.POINTER(C.c_int)], C.c_int),
    "ep2d_field_shape": ([C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "ep2d_get_field": ([C.c_void_p, C.c_int, _dp, C.c_size_t], C.c_int),
    "ep2d_set_field": ([C.c_void_p, C.c_int, _dp, C.c_size_t], C.c_int),
    "ep2d_save_field": ([C.c_void_p, C.c_int, C.c_char_p], C.c_int),
    "ep2d_get_mask": ([C.c_void_p, C.POINTER(C.c_ubyte), C.c_size_t], C.c_int),
    "ep2d_iterate": ([C.c_void_p, C.c_long], C.c_int),
    "ep2d_apply_affine": ([C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double], C.c_int),
    "ep2d_residual": ([C.c_void_p, _dp, _dp, C.POINTER(C.c_long)], C.c_int),
    "ep2d_residual_noalloc": ([C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                               C.c_double, C.c_double, _dp, C.c_void_p, C.c_void_p,
                               C.c_void_p], C.c_int),
    "ep2d_averages": ([C.c_void_p, C.POINTER(Record)], C.c_int),
    "ep2d_boundary": ([C.c_void_p, _dp, _dp, _dp], C.c_int),
    "ep2d_hole_area": ([C.c_void_p, C.POINTER(Record)], C.c_int),
    "ep2d_set_timing": ([C.c_void_p, C.c_long], C.c_int),
    "ep2d_get_timing": ([C.c_void_p, _dp, C.POINTER(C.c_long), C.POINTER(C.c_char_p)], C.c_int),
    "ep2d_diag_copy": ([C.c_int, C.c_size_t, C.c_int, _dp], C.c_int),
    "ep2d_diag_chain_order": ([C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                               C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "ep2d_an_static_fields": ([C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double),
                               C.POINTER(AnStatic)], C.c_int),
    "ep2d_save_an_static": ([C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_double),
                             C.c_char_p, C.POINTER(AnStatic)], C.c_int),
    "ep2d_sync": ([C.c_void_p], C.c_int),
    "ep2d_device_bytes": ([C.c_void_p], C.c_size_t),
    "ep2d_batch_default_options": ([C.POINTER(BatchOptions)], None),
    "ep2d_batch_request_bytes": ([C.POINTER(BatchOptions), C.POINTER(Request),
                                  C.POINTER(C.c_size_t)], C.c_int),
    "ep2d_batch_create": ([C.POINTER(BatchOptions), C.POINTER(C.c_void_p)], C.c_int),
    "ep2d_batch_destroy": ([C.c_void_p], None),
    "ep2d_batch_submit": ([C.c_void_p, C.POINTER(Request), C.POINTER(C.c_longlong)], C.c_int),
    "ep2d_batch_step": ([C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "ep2d_batch_run": ([C.c_void_p], C.c_int),
    "ep2d_batch_poll": ([C.c_void_p, C.POINTER(BatchResult), C.c_int, C.POINTER(C.c_int)],
                        C.c_int),
    "ep2d_batch_clock": ([C.c_void_p], C.c_double),
    "ep2d_batch_stats_get": ([C.c_void_p, C.POINTER(BatchStats)], C.c_int),
}

_lib = None


def lib():
    """Load libep2d.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libep2d.so not built: {LIB_PATH} is missing "
                               "(run python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise Ep2dError(rc, lib().ep2d_last_error().decode())
    return rc


# Handles (EffPlast2D, Batch) still open when the interpreter exits are
# closed by an exit handler, while the HIP runtime is still up; a close that
# would only happen during the interpreter's own finalisation (garbage
# collection of module globals) is skipped instead, so no HIP call runs while
# the runtime or a profiler attached to it is shutting down.
_live = weakref.WeakSet()


def track(obj):
    _live.add(obj)


def finalizing():
    return sys.is_finalizing()


@atexit.register
def _close_all():
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass
