"""Whole-array restatement of the reference's CPU path (get_sigma_2D.m:45-166).

TEST INFRASTRUCTURE ONLY.  A second, independent formulation of the same
scheme as oracle/ep2d_oracle.c: the C oracle restates the CUDA kernels cell by
cell (EffPlast2D.cu:3-177); this module restates the MATLAB solver's
vectorised form (diff / av4 over whole arrays, boolean-mask assignment).  The
tests compare the two on identical inputs; an index, sign or transposition
mistake in either shows up as an O(1) difference.

Arrays are indexed [i, j] with i along x, as MATLAB's (Nx, Ny) arrays; the flat
reference layout A[j*m + i] is ``a.ravel(order="F")``.

Differences from the CUDA path that the tests account for (they vanish for
the inputs used): MATLAB always applies plasticity (get_sigma_2D.m:126-138),
adds no x-term to Uy in the affine increment (:86), normalises the residual by
max|loadValue*loadType| and divides by the input Lx, Ly (:152).  Node
coordinates are taken as x_i = -Lx/2 + i*dX (the exact rounding of MATLAB's
colon operator is not reproduced; the tests use geometries whose pore tests
have no cell within rounding distance of the pore edge).
"""
from __future__ import annotations

import numpy as np


def av4(A):
    """av4.m:1-3."""
    return 0.25 * (A[:-1, :-1] + A[:-1, 1:] + A[1:, :-1] + A[1:, 1:])


def from_flat(a, m, n):
    return np.asarray(a, dtype=np.float64).reshape(n, m).T.copy()


def to_flat(A):
    return np.asarray(A).ravel(order="F").copy()


def pore_masks(Nx, Ny, dX, dY, rad, N):
    """set_mats_2D.m:8-13 and get_sigma_2D.m:103-110: boolean masks of the
    cells (x, y) and vertices (xC, yC) inside any of the N x N pores."""
    Lx, Ly = dX * (Nx - 1), dY * (Ny - 1)
    x = -Lx / 2 + dX * np.arange(Nx)
    y = -Ly / 2 + dY * np.arange(Ny)
    X, Yg = np.meshgrid(x, y, indexing="ij")
    XC, YC = av4(X), av4(Yg)
    cell = np.zeros((Nx, Ny), dtype=bool)
    vert = np.zeros((Nx - 1, Ny - 1), dtype=bool)
    dist_c = np.full((Nx, Ny), np.inf)
    for i in range(int(N)):
        for j in range(int(N)):
            cx = -(0.5 * Lx * (1 - 1 / N)) + (Lx / N) * i
            cy = -(0.5 * Ly * (1 - 1 / N)) + (Ly / N) * j
            d = np.sqrt((X + cx) ** 2 + (Yg + cy) ** 2)
            dist_c = np.minimum(dist_c, np.abs(d - rad))
            cell |= d < rad
            vert |= np.sqrt((XC + cx) ** 2 + (YC + cy) ** 2) < rad
    return cell, vert, float(dist_c.min())


class MatlabSolver:
    """State and iteration of get_sigma_2D.m:45-166."""

    def __init__(self, pa, Nx, Ny=None):
        Ny = Nx if Ny is None else Ny
        (self.dX, self.dY, self.dt, self.K0, self.G0, self.rho0, self.dampX, self.dampY,
         self.coh, self.rad, self.N) = [float(v) for v in pa]
        self.Nx, self.Ny = Nx, Ny
        self.cell, self.vert, self.margin = pore_masks(Nx, Ny, self.dX, self.dY, self.rad, self.N)
        self.K = np.where(self.cell, 0.01 * self.K0, self.K0)          # set_mats_2D.m:6-11
        self.G = np.where(self.cell, 0.01 * self.G0, self.G0)
        self.Pinit = np.zeros((Nx, Ny))                                # :52,55 (P0 = 0)
        self.Ux = np.zeros((Nx + 1, Ny))
        self.Uy = np.zeros((Nx, Ny + 1))
        self.Vx = np.zeros((Nx + 1, Ny))
        self.Vy = np.zeros((Nx, Ny + 1))
        self.tauxyAv = np.zeros((Nx, Ny))

    def set_flat(self, Ux, Uy, Vx, Vy):
        Nx, Ny = self.Nx, self.Ny
        self.Ux = from_flat(Ux, Nx + 1, Ny)
        self.Uy = from_flat(Uy, Nx, Ny + 1)
        self.Vx = from_flat(Vx, Nx + 1, Ny)
        self.Vy = from_flat(Vy, Nx, Ny + 1)

    def iterate(self, n):
        dX, dY, dt, coh, rho0 = self.dX, self.dY, self.dt, self.coh, self.rho0
        for _ in range(n):
            Ux, Uy = self.Ux, self.Uy
            divU = np.diff(Ux, axis=0) / dX + np.diff(Uy, axis=1) / dY          # :94
            P = self.Pinit - self.K * divU                                       # :97
            tauxx = 2.0 * self.G * (np.diff(Ux, axis=0) / dX - divU / 3.0)       # :99
            tauyy = 2.0 * self.G * (np.diff(Uy, axis=1) / dY - divU / 3.0)       # :100
            tauxy = av4(self.G) * (np.diff(Ux[1:-1, :], axis=1) / dY +
                                   np.diff(Uy[:, 1:-1], axis=0) / dX)            # :101
            P[self.cell] = 0.0                                                   # :103-110
            tauxx[self.cell] = 0.0
            tauyy[self.cell] = 0.0
            tauxy[self.vert] = 0.0
            tA = self.tauxyAv                                                    # :114-123
            tA[1:-1, 1:-1] = av4(tauxy)
            tA[0, 1:-1] = tA[1, 1:-1]
            tA[-1, 1:-1] = tA[-2, 1:-1]
            tA[1:-1, 0] = tA[1:-1, 1]
            tA[1:-1, -1] = tA[1:-1, -2]
            tA[0, 0] = 0.5 * (tA[0, 1] + tA[1, 0])
            tA[-1, 0] = 0.5 * (tA[-1, 1] + tA[-2, 0])
            tA[0, -1] = 0.5 * (tA[1, -1] + tA[0, -2])
            tA[-1, -1] = 0.5 * (tA[-1, -2] + tA[-2, -1])
            J2 = np.sqrt(tauxx * tauxx + tauyy * tauyy + 2.0 * tA * tA)          # :126
            J2xy = np.sqrt(av4(tauxx) ** 2 + av4(tauyy) ** 2 + 2.0 * tauxy * tauxy)  # :127
            ip = J2 > coh                                                        # :128-133
            tauxx[ip] = tauxx[ip] * coh / J2[ip]
            tauyy[ip] = tauyy[ip] * coh / J2[ip]
            ipx = J2xy > coh                                                     # :134-138
            tauxy[ipx] = tauxy[ipx] * coh / J2xy[ipx]
            dVxdt = (np.diff(-P[:, 1:-1] + tauxx[:, 1:-1], axis=0) / dX / rho0 +
                     np.diff(tauxy, axis=1) / dY / rho0)                         # :141
            self.Vx[1:-1, 1:-1] = self.Vx[1:-1, 1:-1] * (1 - dt * self.dampX) + dVxdt * dt
            dVydt = (np.diff(-P[1:-1, :] + tauyy[1:-1, :], axis=1) / dY / rho0 +
                     np.diff(tauxy, axis=0) / dX / rho0)                         # :143
            self.Vy[1:-1, 1:-1] = self.Vy[1:-1, 1:-1] * (1 - dt * self.dampY) + dVydt * dt
            self.Ux = Ux + self.Vx * dt                                          # :147-148
            self.Uy = Uy + self.Vy * dt
            self.P, self.tauxx, self.tauyy, self.tauxy, self.J2 = P, tauxx, tauyy, tauxy, J2

    def residual(self, loadValue, loadType, Lx, Ly):
        """get_sigma_2D.m:152."""
        den = np.max(np.abs(loadValue * np.asarray(loadType, dtype=np.float64)))
        return (np.max(np.abs(self.Vx)) / Lx + np.max(np.abs(self.Vy)) / Ly) * self.dt / den
