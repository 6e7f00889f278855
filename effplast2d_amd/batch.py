"""Many RVEs on one GPU through the batched engine of libep2d.so.

Each request is one ComputeEffModuli (nTasks <= 3) or ComputeEffDamping
(nTasks > 3) of its own RVE.  All admitted RVEs advance together, one launch
per iteration over their tiles; finished RVEs retire and queued requests are
admitted while the device-memory budget allows (continuous batching).  A
request's result is the same as the single-RVE path's for the same inputs and
check interval.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib


class Batch:
    def __init__(self, mode="split", checkInterval=500, tileRows=0, marchRows=0, persistRows=0,
                 march2Rows=0, pairRows=0, tripleRows=0, quadRows=0, device=0, memoryBudget=0,
                 maxActive=0, useGraphs=True,
                 timingEvery=0, adaptiveCheck=False, wide3Rows=0, wide4Rows=0,
                 keepFp32Stress=False, triple2Rows=0, quad2Rows=0):
        o = L.BatchOptions()
        lib().ep2d_batch_default_options(C.byref(o))
        o.mode = L.MODES[mode] if isinstance(mode, str) else int(mode)
        o.checkInterval = int(checkInterval)
        o.tileRows = int(tileRows)
        o.marchRows = int(marchRows)
        o.persistRows = int(persistRows)
        o.march2Rows = int(march2Rows)
        o.pairRows = int(pairRows)
        o.tripleRows = int(tripleRows)
        o.quadRows = int(quadRows)
        o.wide3Rows = int(wide3Rows)
        o.wide4Rows = int(wide4Rows)
        o.keepFp32Stress = 1 if keepFp32Stress else 0
        o.triple2Rows = int(triple2Rows)
        o.quad2Rows = int(quad2Rows)
        o.device = int(device)
        o.memoryBudget = int(memoryBudget)
        o.maxActive = int(maxActive)
        o.useGraphs = 1 if useGraphs else 0
        o.timingEvery = int(timingEvery)
        o.adaptiveCheck = 1 if adaptiveCheck else 0
        self._opts = o
        h = C.c_void_p()
        check(lib().ep2d_batch_create(C.byref(o), C.byref(h)))
        self._h = h
        L.track(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().ep2d_batch_destroy(self._h)
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

    def submit(self, pa, nX, nTasks=3, nIter=100_000, eIter=1e-8, nTimeSteps=1, initLoad=-3e-3,
               loadType=(1.0, 1.0, 0.0), loadValue=0.0, shearLoadType=None, checkInterval=0,
               tag=0):
        rq = make_request(pa, nX, nTasks, nIter, eIter, nTimeSteps, initLoad, loadType,
                          loadValue, shearLoadType, checkInterval, tag)
        rid = C.c_longlong()
        check(lib().ep2d_batch_submit(self._h, C.byref(rq), C.byref(rid)))
        return rid.value

    def request_bytes(self, pa, nX, **kw):
        """Device bytes a request counts against memoryBudget when admitted."""
        return request_bytes(self._opts.mode, pa, nX, **kw)

    def step(self):
        a, q = C.c_int(), C.c_int()
        check(lib().ep2d_batch_step(self._h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def run(self):
        check(lib().ep2d_batch_run(self._h))

    def poll(self, capacity=4096):
        buf = (L.BatchResult * capacity)()
        n = C.c_int()
        check(lib().ep2d_batch_poll(self._h, buf, capacity, C.byref(n)))
        return [buf[k].as_dict() for k in range(n.value)]

    def clock(self):
        return float(lib().ep2d_batch_clock(self._h))

    def stats(self):
        st = L.BatchStats()
        check(lib().ep2d_batch_stats_get(self._h, C.byref(st)))
        return st.as_dict()


def make_request(pa, nX, nTasks=3, nIter=100_000, eIter=1e-8, nTimeSteps=1, initLoad=-3e-3,
                 loadType=(1.0, 1.0, 0.0), loadValue=0.0, shearLoadType=None, checkInterval=0,
                 tag=0):
    rq = L.Request()
    pa = np.ascontiguousarray(pa, dtype=np.float64)
    for k in range(L.NPARS):
        rq.pa[k] = float(pa[k])
    rq.nX, rq.nTasks, rq.nIter, rq.eIter = int(nX), int(nTasks), int(nIter), float(eIter)
    rq.nTimeSteps, rq.initLoadValue, rq.loadValue = int(nTimeSteps), float(initLoad), float(loadValue)
    for k in range(3):
        rq.loadType[k] = float(loadType[k])
    if shearLoadType is not None:
        rq.useShearLoadType = 1
        for k in range(3):
            rq.shearLoadType[k] = float(shearLoadType[k])
    rq.checkInterval = int(checkInterval)
    rq.tag = int(tag)
    return rq


def request_bytes(mode, pa, nX, **kw):
    """Device bytes the engine counts for one request of this grid and pore
    count in the given scheme (host only, ep2d_batch_request_bytes)."""
    o = L.BatchOptions()
    lib().ep2d_batch_default_options(C.byref(o))
    o.mode = L.MODES[mode] if isinstance(mode, str) else int(mode)
    rq = make_request(pa, nX, **kw)
    n = C.c_size_t()
    check(lib().ep2d_batch_request_bytes(C.byref(o), C.byref(rq), C.byref(n)))
    return int(n.value)
