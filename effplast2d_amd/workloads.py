"""Seeded synthetic request streams for the batched engine.

Each request is one ComputeEffModuli (EffPlast2D.cu:179-229) of its own RVE,
as keyword arguments of ``Batch.submit``: a square grid, porosity, pore
count, yield stress and a preload of given magnitude along a unit direction
in (Exx, Eyy, Exy) (loadType, EffPlast2D.cu:349-352).  Parameters are derived
with get_sigma_2D.m:3-32 (``derive_params``) at Lx = Ly = 20, K0 = 1,
G0 = 0.5, rho0 = 1, CFL = 0.5.
"""
from __future__ import annotations

import numpy as np

from .params import derive_params

# Iteration cap per load step of sampled RVEs: 40 N, about five times what a
# converging RVE of size N needs (6.3 N for the preload at 1024^2).  Some
# sampled loads exceed the limit load (the plastic zone never stops growing)
# and would otherwise run to the reference's 100000-iteration default.
ITER_CAP_PER_N = 40


def _sample_rve(rng, N, nTasks, tag):
    phi = float(rng.uniform(0.005, 0.05))
    npores = int(rng.integers(1, 4))
    Y = float(rng.uniform(1e-3, 5e-3))
    mag = float(rng.uniform(1e-3, 5e-3))
    u = rng.standard_normal(3)
    u /= np.linalg.norm(u)
    pa, nx = derive_params(Lx=20.0, Ly=20.0, nX=N, nPores=npores, porosity=phi, Y=Y, K0=1.0,
                           G0=0.5, rho0=1.0, CFL=0.5)
    if not pa[9] < 20.0 / (2 * npores):
        raise ValueError("pore radius exceeds half the pore spacing")
    return dict(pa=pa, nX=nx, nTasks=nTasks, nIter=ITER_CAP_PER_N * nx, eIter=1e-8,
                initLoad=-mag, loadType=[float(x) for x in u], tag=tag)


def mixed_requests(n, seed, sizes=(128, 256, 512, 1024), tasks=(3,)):
    """Batched sweep: sizes uniform in `sizes`, porosity U[0.005, 0.05],
    nPores uniform in {1, 2, 3}, Y/K0 U[1e-3, 5e-3], preload magnitude
    U[1e-3, 5e-3] along a random unit direction in (Exx, Eyy, Exy), nTasks
    uniform in `tasks`."""
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n):
        N = int(rng.choice(list(sizes)))
        nl = int(rng.choice(list(tasks))) if len(tasks) > 1 else int(tasks[0])
        reqs.append(_sample_rve(rng, N, nl, k))
    return reqs


def serving_stream(rate, duration, seed, sizes=(256, 512, 1024, 2048),
                   weights=(0.4, 0.3, 0.2, 0.1), tasks=(1, 2, 3)):
    """Poisson arrivals at `rate` per second over `duration` seconds: list of
    (arrival time, request); sizes drawn with `weights`, other parameters as
    mixed_requests, nTasks uniform in `tasks`."""
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration:
            break
        N = int(rng.choice(list(sizes), p=list(weights)))
        rq = _sample_rve(rng, N, 0, len(reqs))
        rq["nTasks"] = int(rng.integers(min(tasks), max(tasks) + 1))
        reqs.append((t, rq))
    return reqs
