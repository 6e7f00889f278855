"""Golden pa.dat vectors for the RVEs the benchmark and tests use.

Independent restatement of get_sigma_2D.m:3-32 in Python floats (IEEE
binary64), evaluated in MATLAB's left-to-right order:
    coh  = Y * sqrt(2.0)                                  (:6)
    rad  = sqrt(porosity * Lx * Ly / (pi * N * N))        (:9)
    dX   = Lx / (Nx - 1), dY = Ly / (Ny - 1)              (:17-18)
    dt   = CFL * min(dX, dY) / sqrt((K0 + 4*G0/3) / rho0) (:27)
    dampX = 4.0 / dt / Nx, dampY = 4.0 / dt / Ny          (:28-29)
    pa = [dX, dY, dt, K0, G0, rho0, dampX, dampY, coh, rad, N]  (:32)
written as 11 little-endian float64 (fwrite(fil, pa(:), 'double'), :41-43).

Run:  python tests/golden/make_pa_golden.py   (rewrites tests/golden/pa_*.dat
and tests/golden/pa_golden.json).
"""
import json
import math
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    # name: Lx, Ly, nGrid, nPores, porosity, Y, K0, G0, rho0, CFL
    "rve128_elastic": (20.0, 20.0, 4, 1, 0.005, 2.5e-3, 1.0, 0.5, 1.0, 0.5),
    "rve1024_plastic": (20.0, 20.0, 32, 1, 0.01, 2.5e-3, 1.0, 0.5, 1.0, 0.5),
    "rve4096_16pores": (20.0, 20.0, 128, 4, 0.05, 2.5e-3, 1.0, 0.5, 1.0, 0.5),
    "driver_defaults_2048": (20.0, 20.0, 64, 1, 0.005, 0.00001, 1.0, 0.01, 1.0, 0.5),
}


def pa_of(Lx, Ly, nGrid, N, porosity, Y, K0, G0, rho0, CFL):
    coh = Y * math.sqrt(2.0)
    rad = math.sqrt(porosity * Lx * Ly / (math.pi * N * N))
    Nx = 32 * nGrid
    Ny = 32 * nGrid
    dX = Lx / (Nx - 1)
    dY = Ly / (Ny - 1)
    dt = CFL * min(dX, dY) / math.sqrt((K0 + 4 * G0 / 3) / rho0)
    dampX = 4.0 / dt / Nx
    dampY = 4.0 / dt / Ny
    return [dX, dY, dt, K0, G0, rho0, dampX, dampY, coh, rad, float(N)], Nx


def main():
    out = {}
    for name, args in CASES.items():
        pa, nx = pa_of(*args)
        raw = struct.pack("<11d", *pa)
        with open(os.path.join(HERE, f"pa_{name}.dat"), "wb") as f:
            f.write(raw)
        out[name] = {"inputs": dict(zip(["Lx", "Ly", "nGrid", "nPores", "porosity", "Y", "K0",
                                         "G0", "rho0", "CFL"], args)),
                     "nX": nx, "pa_hex": [x.hex() for x in pa]}
    with open(os.path.join(HERE, "pa_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
