/*
 * ep2d_oracle.c -- CPU FP64 restatement of the EffPlast2D hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the timed CPU baseline).
 * The product library never links or calls this file.
 *
 * Each function restates the reference function cited beside it, in the
 * reference's own operation order.  Built with -ffp-contract=off.
 */
#include "ep2d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORA_MUTANT
#define ORA_MUTANT 0   /* >0 only in the pin tests' planted-mistake builds */
#endif

#define IDX_C(i, j) ((j) * nX + (i))             /* cell           */
#define IDX_V(i, j) ((j) * (nX - 1) + (i))       /* vertex         */
#define IDX_UX(i, j) ((j) * (nX + 1) + (i))      /* Ux / Vx        */
#define IDX_UY(i, j) ((j) * nX + (i))            /* Uy / Vy        */

static int nthreads_of(const ora_config* c) {
#ifdef _OPENMP
  return c->nThreads > 0 ? c->nThreads : omp_get_max_threads();
#else
  (void)c;
  return 1;
#endif
}

/* std::max semantics: returns a unless a < b. */
static double std_max(double a, double b) { return (a < b) ? b : a; }

/* get_sigma_2D.m:3-32 (MATLAB evaluates left to right). */
void ora_derive_params(double Lx, double Ly, long nX, long nY, double nPores,
                       double porosity, double Y, double K0, double G0,
                       double rho0, double CFL, double pa[ORA_NPARS]) {
  const double coh = Y * sqrt(2.0);                                   /* :6  */
  const double rad = sqrt(porosity * Lx * Ly / (M_PI * nPores * nPores)); /* :9 */
  const double dX = Lx / (double)(nX - 1);                            /* :17 */
  const double dY = Ly / (double)(nY - 1);                            /* :18 */
  const double mn = dX < dY ? dX : dY;                                /* min(dX, dY) */
  const double dt = CFL * mn / sqrt((K0 + 4.0 * G0 / 3.0) / rho0);    /* :27 */
  const double dampX = 4.0 / dt / (double)nX;                         /* :28 */
  const double dampY = 4.0 / dt / (double)nY;                         /* :29 */
  pa[ORA_PA_DX] = dX;
  pa[ORA_PA_DY] = dY;
  pa[ORA_PA_DT] = dt;
  pa[ORA_PA_K0] = K0;
  pa[ORA_PA_G0] = G0;
  pa[ORA_PA_RHO] = rho0;
  pa[ORA_PA_DAMPX] = dampX;
  pa[ORA_PA_DAMPY] = dampY;
  pa[ORA_PA_COH] = coh;
  pa[ORA_PA_RAD] = rad;
  pa[ORA_PA_NPORES] = nPores;
}

/* SetMaterials, EffPlast2D.cu:708-728 (the empty_spaces set only feeds the
 * deprecated getdPhi and is not restated). */
void ora_set_materials(const ora_config* c, ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double K0 = c->pa[ORA_PA_K0], G0 = c->pa[ORA_PA_G0];
  const double rad = c->pa[ORA_PA_RAD], nPores = c->pa[ORA_PA_NPORES];
  for (int i = 0; i < nX; i++) {
    for (int j = 0; j < nY; j++) {
      f->K[IDX_C(i, j)] = K0;
      f->G[IDX_C(i, j)] = G0;
      double x = -0.5 * dX * (nX - 1) + dX * i;
      double y = -0.5 * dY * (nY - 1) + dY * j;
      double Lx = dX * (nX - 1);
      double Ly = dY * (nY - 1);
      for (int k = 0; k < nPores; k++) {
        for (int l = 0; l < nPores; l++) {
          const double ex = x - 0.5 * Lx * (1.0 - 1.0 / nPores) + (Lx / nPores) * k;
          const double ey = y - 0.5 * Ly * (1.0 - 1.0 / nPores) + (Ly / nPores) * l;
          if (sqrt(ex * ex + ey * ey) < rad) {
            f->K[IDX_C(i, j)] = 0.01 * K0;
            f->G[IDX_C(i, j)] = 0.01 * G0;
          }
        }
      }
    }
  }
}

/* The pore test inside ComputeStress (EffPlast2D.cu:77-97) is a pure
 * function of (i, j) and the geometry; it is evaluated once here with the
 * kernel's expression and reused by every iteration. */
void ora_set_pore_masks(const ora_config* c, ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double rad = c->pa[ORA_PA_RAD];
  const double N = c->pa[ORA_PA_NPORES];
  for (int j = 0; j < nY; j++) {
    for (int i = 0; i < nX; i++) {
      unsigned char cell = 0, vert = 0;
      for (int k = 0; k < N; k++) {
        for (int l = 0; l < N; l++) {
          const double ex = -0.5 * dX * (nX - 1) + dX * i - 0.5 * dX * (nX - 1) * (1.0 - 1.0 / N) + (dX * (nX - 1) / N) * k;
          const double ey = -0.5 * dY * (nY - 1) + dY * j - 0.5 * dY * (nY - 1) * (1.0 - 1.0 / N) + (dY * (nY - 1) / N) * l;
          if (sqrt(ex * ex + ey * ey) < rad) cell = 1;
          if (i < nX - 1 && j < nY - 1) {
            const double vx = -0.5 * dX * (nX - 2) + dX * i - 0.5 * dX * (nX - 1) * (1.0 - 1.0 / N) + (dX * (nX - 1) / N) * k;
            const double vy = -0.5 * dY * (nY - 2) + dY * j - 0.5 * dY * (nY - 1) * (1.0 - 1.0 / N) + (dY * (nY - 1) / N) * l;
            if (sqrt(vx * vx + vy * vy) < rad) vert = 1;
          }
        }
      }
      f->cellPore[IDX_C(i, j)] = cell;
      if (i < nX - 1 && j < nY - 1) f->vertPore[IDX_V(i, j)] = vert;
    }
  }
}

/* SetInitPressure, EffPlast2D.cu:770-783: P0 = 0 everywhere (the in-pore
 * branch also assigns 0.0). */
void ora_set_init_pressure(const ora_config* c, ora_fields* f) {
  const long n = c->nX * c->nY;
  for (long k = 0; k < n; k++) f->P0[k] = 0.0;
}

/* ComputeStress, EffPlast2D.cu:37-98. */
void ora_compute_stress(const ora_config* c, ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double* Ux = f->Ux; const double* Uy = f->Uy;
  const double* K = f->K; const double* G = f->G; const double* P0 = f->P0;
  double* P = f->P; double* tauXX = f->tauXX; double* tauYY = f->tauYY; double* tauXY = f->tauXY;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(c))
  for (long j = 0; j < nY; j++) {
    for (long i = 0; i < nX; i++) {
#if ORA_MUTANT == 3   /* planted mistake for the pin tests: dropped eyy term */
      P[j * nX + i] = P0[j * nX + i] - K[j * nX + i] * (
        (Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / dX);
#else
      P[j * nX + i] = P0[j * nX + i] - K[j * nX + i] * (
        (Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / dX + (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / dY);
#endif
#if ORA_MUTANT == 4   /* planted mistake: transposed operand (eyy for exx) */
      tauXX[j * nX + i] = 2.0 * G[j * nX + i] * (
        (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / dY -
        ((Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / dX + (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / dY) / 3.0);
#else
      tauXX[j * nX + i] = 2.0 * G[j * nX + i] * (
        (Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / dX -
        ((Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / dX + (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / dY) / 3.0);
#endif
      tauYY[j * nX + i] = 2.0 * G[j * nX + i] * (
        (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / dY -
        ((Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / dX + (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / dY) / 3.0);
      if (i < nX - 1 && j < nY - 1) {
#if ORA_MUTANT == 1   /* planted mistake: wrong sign of dUy/dx */
        tauXY[j * (nX - 1) + i] = 0.25 * (G[j * nX + i] + G[j * nX + i + 1] + G[(j + 1) * nX + i] + G[(j + 1) * nX + i + 1]) * (
          (Ux[(j + 1) * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i + 1]) / dY - (Uy[(j + 1) * nX + i + 1] - Uy[(j + 1) * nX + i]) / dX);
#else
        tauXY[j * (nX - 1) + i] = 0.25 * (G[j * nX + i] + G[j * nX + i + 1] + G[(j + 1) * nX + i] + G[(j + 1) * nX + i + 1]) * (
          (Ux[(j + 1) * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i + 1]) / dY + (Uy[(j + 1) * nX + i + 1] - Uy[(j + 1) * nX + i]) / dX);
#endif
      }
      if (f->cellPore[j * nX + i]) {
        P[j * nX + i] = 0.0;
        tauXX[j * nX + i] = 0.0;
        tauYY[j * nX + i] = 0.0;
      }
      if (i < nX - 1 && j < nY - 1 && f->vertPore[j * (nX - 1) + i]) {
        tauXY[j * (nX - 1) + i] = 0.0;
      }
    }
  }
}

/* ComputeJ2, EffPlast2D.cu:100-145. */
void ora_compute_j2(const ora_config* c, ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double* tauXX = f->tauXX; const double* tauYY = f->tauYY; const double* tauXY = f->tauXY;
  double* tauXYav = f->tauXYav; double* J2 = f->J2; double* J2XY = f->J2XY;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(c))
  for (long j = 0; j < nY; j++) {
    for (long i = 0; i < nX; i++) {
      if (i > 0 && i < nX - 1 && j > 0 && j < nY - 1) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[(j - 1) * (nX - 1) + i - 1] + tauXY[(j - 1) * (nX - 1) + i] + tauXY[j * (nX - 1) + i - 1] + tauXY[j * (nX - 1) + i]);
      } else if (i == 0 && j > 0 && j < nY - 1) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[(j - 1) * (nX - 1) + i] + tauXY[(j - 1) * (nX - 1) + i + 1] + tauXY[j * (nX - 1) + i] + tauXY[j * (nX - 1) + i + 1]);
      } else if (i == nX - 1 && j > 0 && j < nY - 1) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[(j - 1) * (nX - 1) + i - 2] + tauXY[(j - 1) * (nX - 1) + i - 1] + tauXY[j * (nX - 1) + i - 2] + tauXY[j * (nX - 1) + i - 1]);
      } else if (i > 0 && i < nX - 1 && j == 0) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[j * (nX - 1) + i - 1] + tauXY[j * (nX - 1) + i] + tauXY[(j + 1) * (nX - 1) + i - 1] + tauXY[(j + 1) * (nX - 1) + i]);
      } else if (i > 0 && i < nX - 1 && j == nY - 1) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[(j - 2) * (nX - 1) + i - 1] + tauXY[(j - 2) * (nX - 1) + i] + tauXY[(j - 1) * (nX - 1) + i - 1] + tauXY[(j - 1) * (nX - 1) + i]);
      } else if (i == 0 && j == 0) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[j * (nX - 1) + i] + tauXY[j * (nX - 1) + i + 1] + tauXY[(j + 1) * (nX - 1) + i] + tauXY[(j + 1) * (nX - 1) + i + 1]);
      } else if (i == 0 && j == nY - 1) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[(j - 2) * (nX - 1) + i] + tauXY[(j - 2) * (nX - 1) + i + 1] + tauXY[(j - 1) * (nX - 1) + i] + tauXY[(j - 1) * (nX - 1) + i + 1]);
      } else if (i == nX - 1 && j == 0) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[j * (nX - 1) + i - 2] + tauXY[j * (nX - 1) + i - 1] + tauXY[(j + 1) * (nX - 1) + i - 2] + tauXY[(j + 1) * (nX - 1) + i - 1]);
      } else if (i == nX - 1 && j == nY - 1) {
        tauXYav[j * nX + i] = 0.25 * (tauXY[(j - 2) * (nX - 1) + i - 2] + tauXY[(j - 2) * (nX - 1) + i - 1] + tauXY[(j - 1) * (nX - 1) + i - 2] + tauXY[(j - 1) * (nX - 1) + i - 1]);
      }

      J2[j * nX + i] = sqrt(tauXX[j * nX + i] * tauXX[j * nX + i] + tauYY[j * nX + i] * tauYY[j * nX + i] + 2.0 * tauXYav[j * nX + i] * tauXYav[j * nX + i]);
      if (i < nX - 1 && j < nY - 1) {
        const double sxx = tauXX[j * nX + i] + tauXX[j * nX + i + 1] + tauXX[(j + 1) * nX + i] + tauXX[(j + 1) * nX + i + 1];
        const double syy = tauYY[j * nX + i] + tauYY[j * nX + i + 1] + tauYY[(j + 1) * nX + i] + tauYY[(j + 1) * nX + i + 1];
        J2XY[j * (nX - 1) + i] = sqrt(
          0.0625 * sxx * sxx +
          0.0625 * syy * syy +
          2.0 * tauXY[j * (nX - 1) + i] * tauXY[j * (nX - 1) + i]);
      }
    }
  }
}

/* ComputePlasticity, EffPlast2D.cu:147-177 (coh = pa[8]). */
void ora_compute_plasticity(const ora_config* c, ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double coh = c->pa[ORA_PA_COH];
  double* tauXX = f->tauXX; double* tauYY = f->tauYY; double* tauXY = f->tauXY;
  double* tauXYav = f->tauXYav; double* J2 = f->J2; const double* J2XY = f->J2XY;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(c))
  for (long j = 0; j < nY; j++) {
    for (long i = 0; i < nX; i++) {
      if (J2[j * nX + i] > coh) {
        tauXX[j * nX + i] *= coh / J2[j * nX + i];
        tauYY[j * nX + i] *= coh / J2[j * nX + i];
        tauXYav[j * nX + i] *= coh / J2[j * nX + i];
        J2[j * nX + i] = sqrt(tauXX[j * nX + i] * tauXX[j * nX + i] + tauYY[j * nX + i] * tauYY[j * nX + i] + 2.0 * tauXYav[j * nX + i] * tauXYav[j * nX + i]);
      }
      if (i < nX - 1 && j < nY - 1) {
        if (J2XY[j * (nX - 1) + i] > coh) {
          tauXY[j * (nX - 1) + i] *= coh / J2XY[j * (nX - 1) + i];
        }
      }
    }
  }
}

/* ComputeDisp, EffPlast2D.cu:3-35. */
void ora_compute_disp(const ora_config* c, ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double dT = c->pa[ORA_PA_DT];
  const double rho = c->pa[ORA_PA_RHO];
  const double dampX = c->pa[ORA_PA_DAMPX], dampY = c->pa[ORA_PA_DAMPY];
  double* Ux = f->Ux; double* Uy = f->Uy; double* Vx = f->Vx; double* Vy = f->Vy;
  const double* P = f->P; const double* tauXX = f->tauXX; const double* tauYY = f->tauYY;
  const double* tauXY = f->tauXY;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(c))
  for (long j = 0; j < nY; j++) {
    for (long i = 0; i < nX; i++) {
      if (i > 0 && i < nX && j > 0 && j < nY - 1) {
        Vx[j * (nX + 1) + i] = Vx[j * (nX + 1) + i] * (1.0 - dT * dampX) + (dT / rho) * ((
          -P[j * nX + i] + P[j * nX + i - 1] + tauXX[j * nX + i] - tauXX[j * nX + i - 1]
          ) / dX + (
#if ORA_MUTANT == 2   /* planted mistake: shear term one vertex to the right (kept in bounds) */
          tauXY[j * (nX - 1) + (i < nX - 1 ? i : i - 1)] - tauXY[(j - 1) * (nX - 1) + (i < nX - 1 ? i : i - 1)]
#else
          tauXY[j * (nX - 1) + i - 1] - tauXY[(j - 1) * (nX - 1) + i - 1]
#endif
          ) / dY);
      }
      if (i > 0 && i < nX - 1 && j > 0 && j < nY) {
        Vy[j * nX + i] = Vy[j * nX + i] * (1.0 - dT * dampY) + (dT / rho) * ((
          -P[j * nX + i] + P[(j - 1) * nX + i] + tauYY[j * nX + i] - tauYY[(j - 1) * nX + i]
          ) / dY + (
          tauXY[(j - 1) * (nX - 1) + i] - tauXY[(j - 1) * (nX - 1) + i - 1]
          ) / dX);
      }
      Ux[j * (nX + 1) + i] = Ux[j * (nX + 1) + i] + Vx[j * (nX + 1) + i] * dT;
      Uy[j * nX + i] = Uy[j * nX + i] + Vy[j * nX + i] * dT;
    }
  }
}

/* EffPlast2D.cu:390-399. */
void ora_iterate(const ora_config* c, ora_fields* f, long n) {
  for (long it = 0; it < n; it++) {
    ora_compute_stress(c, f);
    if (c->nl > 1) {
      ora_compute_j2(c, f);
      ora_compute_plasticity(c, f);
    }
    ora_compute_disp(c, f);
  }
}

/* FindMaxAbs(const double*, int), EffPlast2D.cu:805-813. */
double ora_find_max_abs(const double* arr, long size) {
  double max_el = 0.0;
  for (long i = 0; i < size; i++) {
    if (fabs(arr[i]) > fabs(max_el)) max_el = fabs(arr[i]);
  }
  return max_el;
}

/* FindMaxAbs(const std::vector<double>&), EffPlast2D.cu:814-822: returns
 * the signed element of largest magnitude. */
double ora_find_max_abs_signed(const double* v, long size) {
  double max_el = 0.0;
  for (long i = 0; i < size; i++) {
#if ORA_MUTANT == 15   /* planted mistake: the magnitude instead of the signed element */
    if (fabs(v[i]) > fabs(max_el)) max_el = fabs(v[i]);
#else
    if (fabs(v[i]) > fabs(max_el)) max_el = v[i];
#endif
  }
  return max_el;
}

/* EffPlast2D.cu:411-412. */
double ora_residual(const ora_config* c, const ora_fields* f, const double e[3]) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY], dT = c->pa[ORA_PA_DT];
  return (ora_find_max_abs(f->Vx, (nX + 1) * nY) / (dX * (nX - 1)) +
          ora_find_max_abs(f->Vy, nX * (nY + 1)) / (dY * (nY - 1))) * dT /
         (std_max(fabs(e[0]), std_max(e[1], e[2])));
}

/* EffPlast2D.cu:370-381. */
void ora_apply_affine(const ora_config* c, ora_fields* f, double dUxdx,
                      double dUydy, double dUxdy, double dUydx) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  for (int i = 0; i < nX + 1; i++) {
    for (int j = 0; j < nY; j++) {
      f->Ux[j * (nX + 1) + i] += (-0.5 * dX * nX + dX * i) * dUxdx + (-0.5 * dY * (nY - 1) + dY * j) * dUxdy;
    }
  }
  for (int i = 0; i < nX; i++) {
    for (int j = 0; j < nY + 1; j++) {
      f->Uy[j * nX + i] += (-0.5 * dY * nY + dY * j) * dUydy + (-0.5 * dX * (nX - 1) + dX * i) * dUydx;
    }
  }
}

/* Volume averages, EffPlast2D.cu:453-503 (nu0 from :1569).  Also the clean
 * xy mean used by the xy shear-modulus extension. */
void ora_averages(const ora_config* c, const ora_fields* f, ora_step_record* r) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double K0 = c->pa[ORA_PA_K0], G0 = c->pa[ORA_PA_G0];
  const double nPores = c->pa[ORA_PA_NPORES];
  const double nu0 = (1.5 * K0 - G0) / (3.0 * K0 + G0);
  const double* P = f->P; const double* tauXX = f->tauXX; const double* tauYY = f->tauYY;
  const double* tauXY = f->tauXY; const double* Ux = f->Ux; const double* Uy = f->Uy;
  double* sigma = r->sigma; double* sigmaPer = r->sigmaPer; double* epsilonPer = r->epsilonPer;
  for (int k = 0; k < 4; k++) { sigma[k] = 0.0; sigmaPer[k] = 0.0; }
  for (int k = 0; k < 3; k++) epsilonPer[k] = 0.0;

  const int perOffsetX = nX / nPores;
  const int perOffsetY = nY / nPores;

  for (int i = 1; i < nX - 1; i++) {
    for (int j = 1; j < nY - 1; j++) {
      sigma[0] += tauXX[j * nX + i] - P[j * nX + i];
      sigma[1] += tauYY[j * nX + i] - P[j * nX + i];
      if (i < nX - 1 && j < nY - 1) {
        sigma[2] += tauXY[j * (nX - 1) + i];
      }
#if ORA_MUTANT != 10   /* planted mistake 10: the nu0 (zz) fold dropped */
      sigma[2] += nu0 * (tauXX[j * nX + i] + tauYY[j * nX + i] - 2.0 * P[j * nX + i]);
#endif

      if (i > perOffsetX && i < nX - perOffsetX && j > perOffsetY && j < nY - perOffsetY) {
        sigmaPer[0] += tauXX[j * nX + i] - P[j * nX + i];
        sigmaPer[1] += tauYY[j * nX + i] - P[j * nX + i];
        sigmaPer[2] += 0.25 * (
          tauXY[(j - 1) * (nX - 1) + i - 1] +
          tauXY[(j - 1) * (nX - 1) + i] +
          tauXY[j * (nX - 1) + i - 1] +
          tauXY[j * (nX - 1) + i]);
        sigmaPer[3] += nu0 * (tauXX[j * nX + i] + tauYY[j * nX + i] - 2.0 * P[j * nX + i]);

        epsilonPer[0] += (Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / dX;
        epsilonPer[1] += (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / dY;
        epsilonPer[2] += 0.125 * (
          (Ux[j * (nX + 1) + i] - Ux[(j - 1) * (nX + 1) + i]) / dY + (Uy[j * nX + i] - Uy[j * nX + i - 1]) / dX +
          (Ux[j * (nX + 1) + i + 1] - Ux[(j - 1) * (nX + 1) + i + 1]) / dY + (Uy[j * nX + i + 1] - Uy[j * nX + i]) / dX +
          (Ux[(j + 1) * (nX + 1) + i] - Ux[j * (nX + 1) + i]) / dY + (Uy[(j + 1) * nX + i] - Uy[(j + 1) * nX + i - 1]) / dX +
          (Ux[(j + 1) * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i + 1]) / dY + (Uy[(j + 1) * nX + i + 1] - Uy[(j + 1) * nX + i]) / dX);
      }
    }
  }
#if ORA_MUTANT == 9   /* planted mistake: divided by the summed cell count */
  sigma[0] /= (nX - 2) * (nY - 2);
#else
  sigma[0] /= nX * nY;
#endif
  sigma[1] /= nX * nY;
  sigma[2] /= (nX - 1) * (nY - 1);
  sigma[3] /= nX * nY;

  const long perArea = (nX - 2 * perOffsetX) * (nY - 2 * perOffsetY);
  for (int k = 0; k < 4; k++) sigmaPer[k] /= perArea;
  for (int k = 0; k < 3; k++) epsilonPer[k] /= perArea;

  /* extension: plain mean of the vertex shear stress (no nu0 term) */
  double sxy = 0.0;
  for (long k = 0; k < (nX - 1) * (nY - 1); k++) sxy += tauXY[k];
  r->sigmaXYclean = sxy / (double)((nX - 1) * (nY - 1));
}

/* getDeltaP_honest, EffPlast2D.cu:825-853 (right column indexed nY-1). */
double ora_delta_p_honest(const ora_config* c, const ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double* P = f->P; const double* tauXX = f->tauXX; const double* tauYY = f->tauYY;
  double deltaPx = 0.0, deltaPy = 0.0;
#if ORA_MUTANT == 5   /* planted mistake: bottom row 1 instead of row 0 */
  const long jb = 1;
#else
  const long jb = 0;
#endif
  for (int i = 1; i < nX - 1; i++) {
    deltaPx += tauXX[jb * nX + i] - P[jb * nX + i];
    deltaPx += tauYY[jb * nX + i] - P[jb * nX + i];
    deltaPx += tauXX[(nY - 1) * nX + i] - P[(nY - 1) * nX + i];
    deltaPx += tauYY[(nY - 1) * nX + i] - P[(nY - 1) * nX + i];
  }
  deltaPx /= (nX - 2);
  for (int j = 1; j < nY - 1; j++) {
    deltaPy += tauXX[j * nX + 0] - P[j * nX + 0];
    deltaPy += tauYY[j * nX + 0] - P[j * nX + 0];
    deltaPy += tauXX[j * nX + nY - 1] - P[j * nX + nY - 1];
    deltaPy += tauYY[j * nX + nY - 1] - P[j * nX + nY - 1];
  }
  deltaPy /= (nY - 2);
  return -0.125 * (deltaPx + deltaPy);
}

/* getDeltaP_periodic, EffPlast2D.cu:854-882 (integer divisions as in the
 * reference; the y-loop bound divides by the double nPores). */
double ora_delta_p_periodic(const ora_config* c, const ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double nPores = c->pa[ORA_PA_NPORES];
  const double* P = f->P; const double* tauXX = f->tauXX; const double* tauYY = f->tauYY;
  if (nPores <= 2) return 0.0;
  double deltaPx = 0.0, deltaPy = 0.0;
  const int nP = (int)nPores;
  for (int i = nX / nP; i < nX * (nP - 1) / nP; i++) {
    deltaPx += tauXX[nY / nP * nX + i] - P[nY / nP * nX + i];
    deltaPx += tauYY[nY / nP * nX + i] - P[nY / nP * nX + i];
    deltaPx += tauXX[nY * (nP - 1) / nP * nX + i] - P[nY * (nP - 1) / nP * nX + i];
    deltaPx += tauYY[nY * (nP - 1) / nP * nX + i] - P[nY * (nP - 1) / nP * nX + i];
  }
  deltaPx /= nX * (nP - 2) / nP;
#if ORA_MUTANT == 6   /* planted mistake: the mirrored column N(nP-1)/nP */
  const long cr = nX * (nP - 1) / nP;
#else
  const long cr = nX * (nP - 2) / nP;                /* :875-876 */
#endif
#if ORA_MUTANT == 7   /* planted mistake: integer bound of the y loop */
  for (int j = nY / nP; j < nY * (nP - 1) / nP; j++) {
#else
  for (int j = nY / nP; j < nY * (nP - 1) / nPores; j++) {
#endif
    deltaPy += tauXX[j * nX + nX / nP] - P[j * nX + nX / nP];
    deltaPy += tauYY[j * nX + nX / nP] - P[j * nX + nX / nP];
    deltaPy += tauXX[j * nX + cr] - P[j * nX + cr];
    deltaPy += tauYY[j * nX + cr] - P[j * nX + cr];
  }
  deltaPy /= nY * (nP - 2) / nP;
  return -0.125 * (deltaPx + deltaPy);
}

/* getTauInfty_honest, EffPlast2D.cu:925-942. */
double ora_tau_infty_honest(const ora_config* c, const ora_fields* f) {
  const long nX = c->nX, nY = c->nY;
  const double* tauXX = f->tauXX; const double* tauYY = f->tauYY;
  double tx = 0.0, ty = 0.0;
  for (int i = 1; i < nX - 1; i++) {
#if ORA_MUTANT == 8   /* planted mistake: yy - xx on the bottom row */
    tx += tauYY[0 * nX + i] - tauXX[0 * nX + i];
#else
    tx += tauXX[0 * nX + i] - tauYY[0 * nX + i];
#endif
    tx += tauXX[(nY - 1) * nX + i] - tauYY[(nY - 1) * nX + i];
  }
  tx /= (nX - 2);
  for (int j = 1; j < nY - 1; j++) {
    ty += tauXX[j * nX + 0] - tauYY[j * nX + 0];
    ty += tauXX[j * nX + nY - 1] - tauYY[j * nX + nY - 1];
  }
  ty /= (nY - 2);
  return -0.125 * (tx + ty);
}

/* Pore test of the host-side U zeroing, EffPlast2D.cu:514-547.  isUy = 0:
 * Ux node (i in [0,nX], j in [0,nY)); isUy = 1: Uy node. */
int ora_pore_in_u_node(const ora_config* c, int isUy, long i, long j) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double rad = c->pa[ORA_PA_RAD], nPores = c->pa[ORA_PA_NPORES];
  const double Lx = dX * (nX - 1);
  const double Ly = dY * (nY - 1);
  double x, y;
  if (!isUy) {
    x = -0.5 * dX * nX + dX * (int)i;
    y = -0.5 * dY * (nY - 1) + dY * (int)j;
  } else {
    x = -0.5 * dX * (nX - 1) + dX * (int)i;
    y = -0.5 * dY * nY + dY * (int)j;
  }
  int in = 0;
  for (int k = 0; k < nPores; k++) {
    for (int l = 0; l < nPores; l++) {
      if (sqrt((x - 0.5 * Lx * (1.0 - 1.0 / nPores) + (Lx / nPores) * k) * (x - 0.5 * Lx * (1.0 - 1.0 / nPores) + (Lx / nPores) * k) +
               (y - 0.5 * Ly * (1.0 - 1.0 / nPores) + (Ly / nPores) * l) * (y - 0.5 * Ly * (1.0 - 1.0 / nPores) + (Ly / nPores) * l)) < rad) {
        in = 1;
      }
    }
  }
  return in;
}

/* Sample of the host U copy after the in-pore zeroing (:517-547). */
static int sample_u(const ora_config* c, const ora_fields* f, int isUy,
                    size_t flat, double* out) {
  const long nX = c->nX, nY = c->nY;
  const size_t m = isUy ? (size_t)nX : (size_t)(nX + 1);
  const size_t n = isUy ? (size_t)(nX * (nY + 1)) : (size_t)((nX + 1) * nY);
  if (flat >= n) return -1;
  const long i = (long)(flat % m), j = (long)(flat / m);
  const double v = isUy ? f->Uy[flat] : f->Ux[flat];
  *out = ora_pore_in_u_node(c, isUy, i, j) ? 0.0 : v;
  return 0;
}

/* Pore-edge hole area and porosity change, EffPlast2D.cu:549-620. */
int ora_hole_area(const ora_config* c, const ora_fields* f, ora_step_record* r) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double rad = c->pa[ORA_PA_RAD], nPores = c->pa[ORA_PA_NPORES];
  double HoleAreaPi = 0.0, InternalHoleAreaPi = 0.0;
  for (int k = 0; k < nPores; k++) {
    for (int l = 0; l < nPores; l++) {
      const double cxdX = 0.5 * (nX - 1) * (1.0 - 1.0 / nPores) - ((double)(nX - 1) / nPores) * k;
      const double cydY = 0.5 * (nY - 1) * (1.0 - 1.0 / nPores) - ((double)(nY - 1) / nPores) * l;
      double left[5], right[5], bottom[5], top[5];
      const size_t cyIdx = (size_t)(cydY + 0.5 * (nY - 1));
      size_t rxIdx = (size_t)(cxdX - rad / dX + 0.5 * nX);
      for (int i = 0; i < 5; i++)
        if (sample_u(c, f, 0, cyIdx * (nX + 1) + rxIdx - 1 + i, &left[i])) return -1;
      rxIdx = (size_t)(cxdX + rad / dX + 0.5 * nX);
#if ORA_MUTANT == 14   /* planted mistake: right window [-1, +3] like the left one */
      for (int i = 0; i < 5; i++)
        if (sample_u(c, f, 0, cyIdx * (nX + 1) + rxIdx - 1 + i, &right[i])) return -1;
#else
      for (int i = 0; i < 5; i++)
        if (sample_u(c, f, 0, cyIdx * (nX + 1) + rxIdx - 2 + i, &right[i])) return -1;
#endif
      const size_t cxIdx = (size_t)(cxdX + 0.5 * (nX - 1));
      size_t ryIdx = (size_t)(cydY - rad / dY + 0.5 * nY);
      for (int j = 0; j < 5; j++)
        if (sample_u(c, f, 1, (ryIdx - 1 + j) * nX + cxIdx, &bottom[j])) return -1;
      ryIdx = (size_t)(cydY + rad / dY + 0.5 * nY);
      for (int j = 0; j < 5; j++)
        if (sample_u(c, f, 1, (ryIdx - 2 + j) * nX + cxIdx, &top[j])) return -1;
      const double dRx = -0.5 * (ora_find_max_abs_signed(left, 5) - ora_find_max_abs_signed(right, 5));
      const double dRy = -0.5 * (ora_find_max_abs_signed(bottom, 5) - ora_find_max_abs_signed(top, 5));
      HoleAreaPi += (rad + dRx) * (rad + dRy);
      if (k > 0 && l > 0 && k < nPores - 1 && l < nPores - 1) {
        InternalHoleAreaPi += (rad + dRx) * (rad + dRy);
      }
    }
  }
  r->holeAreaPi = HoleAreaPi;
  r->internalHoleAreaPi = InternalHoleAreaPi;
  r->Phi0 = 3.1415926 * rad * rad * nPores * nPores / (dX * (nX - 1) * dY * (nY - 1));
  r->Phi = 3.1415926 * HoleAreaPi / (dX * (nX - 1) * dY * (nY - 1));
  r->PhiPer = nPores > 2 ?
    3.1415926 * InternalHoleAreaPi / (dX * (nX - 1) * dY * (nY - 1) * (nPores - 2) * (nPores - 2) / nPores / nPores) :
    0.0;
  r->dPhi = 3.1415926 * fabs(HoleAreaPi - rad * rad * nPores * nPores) / (dX * (nX - 1) * dY * (nY - 1));
  r->dPhiPer = fabs(r->PhiPer - r->Phi0);
  return 0;
}

/* Per-step analytic moduli, EffPlast2D.cu:643-696: elastic (NL <= 1) from
 * Phi0; elastoplastic from the load step 0 record (Yarushina's formulas with
 * the boundary means deltaP, deltaPper and tauInfty). */
void ora_analytic_moduli(const ora_config* c, int step, const ora_step_record* r,
                         ora_result* res) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double K0 = c->pa[ORA_PA_K0], G0 = c->pa[ORA_PA_G0];
  const double rad = c->pa[ORA_PA_RAD];
  const double Y = c->pa[ORA_PA_COH] / sqrt(2.0);   /* :1572 */
  const int NL = c->nl;
  const double Phi0 = r->Phi0, Phi = r->Phi;
  const double phi = 3.1415926 * rad * rad / (dX * (nX - 1) * dY * (nY - 1));  /* :643 */
  if (NL <= 1) {                                    /* :645-655 */
    const double KphiExactElast = G0 / Phi0;
    res->KphiExactElast = KphiExactElast;
    res->an.Kphi = G0 / Phi0;
    res->KdExactElast = (1.0 - phi) / (1.0 / K0 + 1.0 / KphiExactElast);
    res->an.Kd = G0 / Phi0;
  } else if (step == 0) {                           /* :656-680 */
#if ORA_MUTANT == 13   /* planted mistake: the -1 dropped from the exponent */
    const double KphiExactPlast = G0 / Phi / exp(fabs(r->deltaP) / Y) /
      (1.0 + r->tauInfty * r->tauInfty / Y / Y);
#else
    const double KphiExactPlast = G0 / Phi / exp(fabs(r->deltaP) / Y - 1.0) /
      (1.0 + r->tauInfty * r->tauInfty / Y / Y);
#endif
    res->KphiExactPlast = KphiExactPlast;
    res->an.Kphi = KphiExactPlast;
    res->KdExactPlast = (1.0 - phi) / (1.0 / K0 + 1.0 / KphiExactPlast);
    res->an.Kd = res->KdExactPlast;
    const double KphiExactPlastPer = G0 / Phi0 / exp(fabs(r->deltaPper) / Y - 1.0) /
      (1.0 + r->tauInfty * r->tauInfty / Y / Y);
    res->KphiExactPlastPer = KphiExactPlastPer;
    res->an_per.Kphi = KphiExactPlastPer;
    res->KdExactPlastPer = (1.0 - phi) / (1.0 / K0 + 1.0 / KphiExactPlastPer);
    res->an_per.Kd = res->KdExactPlastPer;
  }
  if (NL == 3 && step == 0) {                       /* :682-696 */
    res->GexactElast = G0 / (1.0 + Phi0);
    res->GexactPlast = G0 / (1.0 + Phi * exp(fabs(r->deltaP) / Y - 1.0));
    res->an.G = res->GexactPlast;
    res->GexactPlastPer = G0 / (1.0 + Phi0 * exp(fabs(r->deltaPper) / Y - 1.0));
    res->an_per.G = res->GexactPlastPer;
  }
}

/* ComputeEffParams, EffPlast2D.cu:311-698. */
int ora_compute_eff_params(const ora_config* c, ora_fields* f, int step,
                           double loadStepValue, const double loadType[3],
                           long nTimeSteps, double curEffStrain[3],
                           ora_step_record* recs, ora_result* res) {
  const long nX = c->nX, nY = c->nY;

  if (step == 0) {                                    /* :329-333 */
    curEffStrain[0] = curEffStrain[1] = curEffStrain[2] = 0.0;
    memset(f->Ux, 0, sizeof(double) * (nX + 1) * nY);
    memset(f->Uy, 0, sizeof(double) * nX * (nY + 1));
  }
  for (long it = 0; it < nTimeSteps; it++) {
    ora_step_record* r = &recs[it];
    memset(r, 0, sizeof(*r));
    const double dUxdx = loadStepValue * loadType[0] / (double)nTimeSteps;   /* :349-352 */
    const double dUydy = loadStepValue * loadType[1] / (double)nTimeSteps;
    const double dUxdy = loadStepValue * loadType[2] / (double)nTimeSteps;
    const double dUydx = dUxdy;
    curEffStrain[0] += dUxdx;                         /* :355-358 */
    curEffStrain[1] += dUydy;
    curEffStrain[2] += dUxdy;
    for (int k = 0; k < 3; k++) r->epsilon[k] = curEffStrain[k];
    ora_apply_affine(c, f, dUxdx, dUydy, dUxdy, dUydx);

    double error = 0.0;                               /* :386 */
    long iters = 0;
    for (long iter = 0; iter < c->nIter; iter++) {    /* :389-428 */
      ora_iterate(c, f, 1);
      iters = iter + 1;
      if ((iter + 1) % c->outputStep == 0) {
        error = ora_residual(c, f, curEffStrain);
        if (error < c->eIter) break;
      }
    }
    r->error = error;
    r->iters = (double)iters;

    ora_averages(c, f, r);                            /* :453-503 */
    r->deltaP = ora_delta_p_honest(c, f);             /* :506-512 */
    r->deltaPper = ora_delta_p_periodic(c, f);
    r->tauInfty = ora_tau_infty_honest(c, f);
    if (ora_hole_area(c, f, r)) return -2;            /* :514-620 */

    ora_analytic_moduli(c, step, r, res);            /* :643-696 */
  }
  return 0;
}

/* The moduli getters, EffPlast2D.cu:1412-1539: r0 is the last record of
 * load step 0, r1 the record of the hydrostatic increment (NL > 1), r2 the
 * record of the deviatoric increment (NL == 3). */
void ora_moduli_getters(int NL, const ora_step_record* r0, const ora_step_record* r1,
                        const ora_step_record* r2, ora_result* res) {
  if (NL == 1) {                                         /* :1412-1428 */
    res->num.Kphi = r0->deltaP / r0->dPhi;
    res->num_per.Kphi = r0->deltaPper / r0->dPhiPer;
    {
      const double pInc = -0.33333333 * (r0->sigma[0] + r0->sigma[1] + r0->sigma[2]);
      const double epsInc = r0->epsilon[0] + r0->epsilon[1];
      res->num.Kd = -pInc / epsInc;
    }
    {
      const double pIncPer = -0.33333333 * (r0->sigmaPer[0] + r0->sigmaPer[1] + r0->sigmaPer[2]);
      const double epsIncPer = r0->epsilonPer[0] + r0->epsilonPer[1];
      res->num_per.Kd = -pIncPer / epsIncPer;
    }
    return;
  }
  res->num.Kphi = (r1->deltaP - r0->deltaP) / (r1->dPhi - r0->dPhi);   /* :1483-1514 */
  res->num_per.Kphi = (r1->deltaPper - r0->deltaPper) / (r1->dPhiPer - r0->dPhiPer);
  {
    const double pInc = -0.33333333 * (r1->sigma[0] + r1->sigma[1] + r1->sigma[2] - r0->sigma[0] - r0->sigma[1] - r0->sigma[2]);
#if ORA_MUTANT == 11   /* planted mistake: measured instead of applied strain */
    const double epsInc = r1->epsilonPer[0] + r1->epsilonPer[1] - r0->epsilonPer[0] - r0->epsilonPer[1];
#else
    const double epsInc = r1->epsilon[0] + r1->epsilon[1] - r0->epsilon[0] - r0->epsilon[1];
#endif
    res->num.Kd = -pInc / epsInc;
  }
  {
    const double pIncPer = -0.33333333 * (r1->sigmaPer[0] + r1->sigmaPer[1] + r1->sigmaPer[2] - r0->sigmaPer[0] - r0->sigmaPer[1] - r0->sigmaPer[2]);
    const double epsIncPer = r1->epsilonPer[0] + r1->epsilonPer[1] - r0->epsilonPer[0] - r0->epsilonPer[1];
    res->num_per.Kd = -pIncPer / epsIncPer;
  }
  if (NL != 3) return;
#if ORA_MUTANT == 12   /* planted mistake: wrong sign of the yy stress difference */
  res->num.G = 0.5 * (r2->sigma[0] - r1->sigma[0] + r2->sigma[1] - r1->sigma[1]) /
               (r2->epsilon[0] - r1->epsilon[0] - r2->epsilon[1] + r1->epsilon[1]);
#else
  res->num.G = 0.5 * (r2->sigma[0] - r1->sigma[0] - r2->sigma[1] + r1->sigma[1]) /   /* :1516-1539 */
               (r2->epsilon[0] - r1->epsilon[0] - r2->epsilon[1] + r1->epsilon[1]);
#endif
  res->num_per.G = 0.5 * (r2->sigmaPer[0] - r1->sigmaPer[0] - r2->sigmaPer[1] + r1->sigmaPer[1]) /
                   (r2->epsilonPer[0] - r1->epsilonPer[0] - r2->epsilonPer[1] + r1->epsilonPer[1]);
  /* extension: xy shear modulus, G = d(sigma_xy) / (2 d(eps_xy)) */
  res->Gxy = 0.5 * (r2->sigmaXYclean - r1->sigmaXYclean) / (r2->epsilon[2] - r1->epsilon[2]);
  res->GxyPer = 0.5 * (r2->sigmaPer[2] - r1->sigmaPer[2]) / (r2->epsilonPer[2] - r1->epsilonPer[2]);
}

/* ComputeEffModuli (EffPlast2D.cu:179-229) and the getters (:1412-1539). */
int ora_compute_eff_moduli(const ora_config* c, ora_fields* f,
                           double initLoadValue, const double loadType[3],
                           long nTimeSteps, ora_step_record* recs,
                           ora_result* res) {
  const int NL = c->nl;
  if (NL > 3 || NL < 1 || nTimeSteps < 1) return -1;   /* :182-184 */
  const long T = nTimeSteps;
  const double incPercent = 0.005;                      /* EffPlast2D.h:105 */
  const double sph[3] = {0.5 * (loadType[0] + loadType[1]), 0.5 * (loadType[0] + loadType[1]), 0.0};
  double dev[3] = {loadType[0] - sph[0], loadType[1] - sph[1], loadType[2]};
  if (c->useShearLoadType) {
    dev[0] = c->shearLoadType[0]; dev[1] = c->shearLoadType[1]; dev[2] = c->shearLoadType[2];
  }
  double cur[3] = {0.0, 0.0, 0.0};
  memset(res, 0, sizeof(*res));
  int rc = ora_compute_eff_params(c, f, 0, initLoadValue, loadType, T, cur, &recs[0], res);
  if (rc) return rc;
  const ora_step_record* r0 = &recs[T - 1];
  if (NL > 1) {
    rc = ora_compute_eff_params(c, f, 1, initLoadValue * incPercent, sph, 1, cur, &recs[T], res);
    if (rc) return rc;
  }
  if (NL == 3) {
    rc = ora_compute_eff_params(c, f, 2, initLoadValue * incPercent, dev, 1, cur, &recs[2 * T], res);
    if (rc) return rc;
  }
  ora_moduli_getters(NL, r0, NL > 1 ? &recs[T] : NULL, NL == 3 ? &recs[2 * T] : NULL, res);
  return 0;
}

/* ComputeEffDamping, EffPlast2D.cu:230-310: the preload, then nl-1 steps of
 * a sinusoidal load increment.  recs is [nl][max(nTimeSteps,1)]. */
int ora_compute_eff_damping(const ora_config* c, ora_fields* f,
                            double initLoadValue, const double loadType[3],
                            double loadValue, long nTimeSteps,
                            ora_step_record* recs, ora_result* res) {
  const int NL = c->nl;
  if (NL < 4 || nTimeSteps < 1) return -1;            /* :233-235 */
  const long T = nTimeSteps;
  double cur[3] = {0.0, 0.0, 0.0};
  memset(res, 0, sizeof(*res));
  int rc = ora_compute_eff_params(c, f, 0, initLoadValue, loadType, T, cur, &recs[0], res);
  if (rc) return rc;
  for (long i = 1; i < NL; i++) {                      /* :240-243 */
    const double v = loadValue * (sin(i * 2.0 * 3.1415926 / (NL - 1)) - sin((i - 1) * 2.0 * 3.1415926 / (NL - 1)));
    rc = ora_compute_eff_params(c, f, (int)i, v, loadType, 1, cur, &recs[i * T], res);
    if (rc) return rc;
  }
  return 0;
}

/* ---- analytic static fields, EffPlast2D.cu:988-1292 ---------------------- */
#include <complex.h>
#include <float.h>
#include <stdlib.h>

/* getAnalytic, :999-1015: conformal map of the elliptic plastic-zone boundary
 * and the complex potentials Phi, Psi of the elastic zone. */
static void an_maps(double x, double y, double xi, double kappa, double c0, double Y, double rad,
                    double complex* zeta, double complex* w, double complex* dw,
                    double complex* wv, double complex* Phi, double complex* Psi) {
  const double complex z = x + I * y;
  double sgn = x > 0.0 ? 1.0 : -1.0;
  if (fabs(x) < DBL_EPSILON) sgn = 1.0;
  *zeta = (z + sgn * csqrt(z * z + 4.0 * c0 * c0 * kappa)) / 2.0 / c0;
  *w = c0 * (*zeta - kappa / *zeta);
  *dw = c0 * (1.0 + kappa / (*zeta * *zeta));
  *wv = c0 * (1.0 / *zeta - kappa * *zeta);
  *Phi = -Y * xi / 2.0 - Y * xi * clog(*w / *zeta / rad);
  *Psi = -Y * xi / *zeta * *wv / *dw;
}

double ora_an_ur_hydro(double r, double deltaP, double Y, double rad, double G0) {
  return -0.5 * Y * rad * rad * exp((deltaP - Y) / Y) / (G0 * r);
}

void ora_an_u_elast(double x, double y, double tauInfty, double xi, double kappa, double c0,
                    double Y, double rad, double G0, double K0, double* re, double* im) {
  double complex zeta, w, dw, wv, Phi, Psi;
  an_maps(x, y, xi, kappa, c0, Y, rad, &zeta, &w, &dw, &wv, &Phi, &Psi);
  const double complex phi = -Y * xi * w * (clog(w / zeta / rad) + 0.5) - 2.0 * c0 * tauInfty / zeta;
  const double complex psi = c0 * Y * xi * (1.0 / zeta + kappa * zeta);
  const double complex dphi = Phi * dw;
  const double complex U =
      ((1.0 + 6.0 * G0 / (G0 + 3.0 * K0)) * phi - w / conj(dw) * conj(dphi) - conj(psi)) / 2.0 / G0;
  *re = creal(U);
  *im = cimag(U);
}

static double an_j1(double s11, double s22) { return 0.5 * (s11 + s22); }
static double an_j2(double s11, double s22, double s12) {
  return 0.5 * ((s11 - s22) * (s11 - s22) + 4.0 * s12 * s12);
}

void ora_an_j_elast(double x, double y, double xi, double kappa, double c0, double Y,
                    double rad, double* J1, double* J2) {
  double complex zeta, w, dw, wv, Phi, Psi;
  an_maps(x, y, xi, kappa, c0, Y, rad, &zeta, &w, &dw, &wv, &Phi, &Psi);
  const double complex z = x + I * y;
  const double complex dPhi = -2.0 * xi * Y * kappa / zeta / (zeta * zeta - kappa);
  const double complex F = 2.0 * (conj(w) / dw * dPhi + Psi) / cexp(-2.0 * carg(z) * I);
  const double Srr = 2.0 * creal(Phi) - creal(F) / 2.0;
  const double Sff = 2.0 * creal(Phi) + creal(F) / 2.0;
  const double Srf = cimag(F) / 2.0;
  *J1 = an_j1(Srr, Sff);
  *J2 = an_j2(Srr, Sff, Srf);
}

void ora_an_j_plast(double r, double xi, double Y, double rad, double* J1, double* J2) {
  const double Srr = -2.0 * xi * Y * log(r / rad);
  const double Sff = -2.0 * xi * Y * (1.0 + log(r / rad));
  *J1 = an_j1(Srr, Sff);
  *J2 = an_j2(Srr, Sff, 0.0);
}

/* SaveAnStatic2D, :1060-1292 (fields and error statistics, no files). */
int ora_an_static(const ora_config* c, const ora_fields* f, double deltaP, double tauInfty,
                  const double lt[3], ora_an_out* o) {
  const long nX = c->nX, nY = c->nY;
  const double dX = c->pa[ORA_PA_DX], dY = c->pa[ORA_PA_DY];
  const double K0 = c->pa[ORA_PA_K0], G0 = c->pa[ORA_PA_G0];
  const double rad = c->pa[ORA_PA_RAD];
  const int nPores = (int)c->pa[ORA_PA_NPORES];
  const double Y = c->pa[ORA_PA_COH] / sqrt(2.0);
  const int hydro = fabs(lt[0] - lt[1]) < DBL_EPSILON && fabs(lt[2]) < DBL_EPSILON;
  const double Rmin = rad;
  const double Rmax = 0.5 * dX * (nX - 1) - dX * 20.0;
  const double eps = 1.0e-18;
  const double xi = deltaP > 0.0 ? 1.0 : -1.0;
  const double kappa = tauInfty / Y * xi;
  const double c0 = rad * exp(fabs(deltaP) / 2.0 / Y - 0.5);
  const double rx = rad + ora_an_ur_hydro(rad, deltaP, Y, rad, G0), ry = rx;
  const double Rx = c0 * (1.0 - kappa), Ry = c0 * (1.0 + kappa);
  const int an = nPores == 1;

  /* host copies of U with the pore nodes zeroed (:514-547) */
  const long nUx = (nX + 1) * nY, nUy = nX * (nY + 1);
  double* Ux = (double*)malloc(sizeof(double) * nUx);
  double* Uy = (double*)malloc(sizeof(double) * nUy);
  for (long j = 0; j < nY; j++)
    for (long i = 0; i < nX + 1; i++)
      Ux[j * (nX + 1) + i] = ora_pore_in_u_node(c, 0, i, j) ? 0.0 : f->Ux[j * (nX + 1) + i];
  for (long j = 0; j < nY + 1; j++)
    for (long i = 0; i < nX; i++)
      Uy[j * nX + i] = ora_pore_in_u_node(c, 1, i, j) ? 0.0 : f->Uy[j * nX + i];

  double eUmax = 0.0, eUavg = 0.0, eJ1max = 0.0, eJ1avg = 0.0, eJ2max = 0.0, eJ2avg = 0.0;
  size_t nU = 0, nJ = 0;
#define CUT(e) do { if (fabs(e) > 0.5) (e) = -0.5; } while (0)
  for (long i = 0; i < nX; i++) {
    for (long j = 0; j < nY; j++) {
      const double ux = 0.5 * (Ux[(nX + 1) * j + i] + Ux[(nX + 1) * j + (i + 1)]);
      const double uy = 0.5 * (Uy[nX * j + i] + Uy[nX * (j + 1) + i]);
      const double unu = sqrt(ux * ux + uy * uy);
      if (o->UnuAbs) o->UnuAbs[j * nX + i] = unu;
      double j1nu = 0.0, j2nu = 0.0;
      const long v = j * (nX - 1) + i;
      if (i < nX - 1 && j < nY - 1) {
        const double* P = f->P;
        const double* txx = f->tauXX;
        const double* tyy = f->tauYY;
        const double Sxx = 0.25 * (-P[j * nX + i] + txx[j * nX + i] + -P[j * nX + (i + 1)] +
                                   txx[j * nX + (i + 1)] + -P[(j + 1) * nX + i] +
                                   txx[(j + 1) * nX + i] + -P[(j + 1) * nX + (i + 1)] +
                                   txx[(j + 1) * nX + (i + 1)]);
        const double Syy = 0.25 * (-P[j * nX + i] + tyy[j * nX + i] + -P[j * nX + (i + 1)] +
                                   tyy[j * nX + (i + 1)] + -P[(j + 1) * nX + i] +
                                   tyy[(j + 1) * nX + i] + -P[(j + 1) * nX + (i + 1)] +
                                   tyy[(j + 1) * nX + (i + 1)]);
        const double Sxy = f->tauXY[v];
        j1nu = an_j1(Sxx, Syy);
        j2nu = an_j2(Sxx, Syy, Sxy);
        if (o->J1nu) o->J1nu[v] = j1nu;
        if (o->J2nu) o->J2nu[v] = j2nu;
      }
      if (!an) continue;
      {
        const double x = -0.5 * dX * (nX - 1) + dX * i;
        const double y = -0.5 * dY * (nY - 1) + dY * j;
        const double r = sqrt(x * x + y * y);
        double uan;
        if (x * x / (Rx * Rx) + y * y / (Ry * Ry) > 1.0) {
          if (hydro) {
            uan = fabs(ora_an_ur_hydro(r, deltaP, Y, rad, G0));
          } else {
            double re, im;
            ora_an_u_elast(x, y, tauInfty, xi, kappa, c0, Y, rad, G0, K0, &re, &im);
            uan = cabs(re + I * im);
          }
        } else if (x * x / (rx * rx) + y * y / (ry * ry) > 1.0) {
          uan = fabs(ora_an_ur_hydro(r, deltaP, Y, rad, G0));
        } else {
          uan = 0.0;
        }
        double eu = 0.0;
        if (x * x + y * y > Rmin * Rmin && x * x + y * y < Rmax * Rmax && fabs(unu) > eps) {
          eu = fabs((uan - unu) / uan);
          eUmax = fmax(eUmax, eu);
          eUavg += eu;
          nU++;
          CUT(eu);
        }
        if (o->UanAbs) o->UanAbs[j * nX + i] = uan;
        if (o->errorUabs) o->errorUabs[j * nX + i] = eu;
      }
      if (i < nX - 1 && j < nY - 1) {
        const double x = -0.5 * dX * (nX - 1) + dX * i + 0.5 * dX;
        const double y = -0.5 * dY * (nY - 1) + dY * j + 0.5 * dY;
        const double r = sqrt(x * x + y * y);
        double j1an, j2an;
        if (x * x / (Rx * Rx) + y * y / (Ry * Ry) > 1.0) {
          if (hydro) {
            const double relR = rad / r;
            const double Srr = -deltaP + relR * relR * Y * exp(deltaP / Y - 1);
            const double Sff = -deltaP - relR * relR * Y * exp(deltaP / Y - 1);
            j1an = an_j1(Srr, Sff);
            j2an = an_j2(Srr, Sff, 0.0);
          } else {
            ora_an_j_elast(x, y, xi, kappa, c0, Y, rad, &j1an, &j2an);
          }
        } else if (x * x / (rx * rx) + y * y / (ry * ry) > 1.0) {
          ora_an_j_plast(r, xi, Y, rad, &j1an, &j2an);
        } else {
          j1an = 0.0;
          j2an = 0.0;
        }
        double e1 = 0.0, e2 = 0.0;
        if (x * x + y * y > Rmin * Rmin && x * x + y * y < Rmax * Rmax) {
          if (fabs(j1nu) > eps) {
            e1 = fabs((j1an - j1nu) / j1an);
            eJ1max = fmax(eJ1max, e1);
            eJ1avg += e1;
            CUT(e1);
          }
          if (fabs(j2nu) > eps) {
            e2 = fabs((j2an - j2nu) / j2an);
            eJ2max = fmax(eJ2max, e2);
            eJ2avg += e2;
            CUT(e2);
          }
          nJ++;
        }
        if (o->J1an) o->J1an[v] = j1an;
        if (o->J2an) o->J2an[v] = j2an;
        if (o->errorJ1) o->errorJ1[v] = e1;
        if (o->errorJ2) o->errorJ2[v] = e2;
      }
    }
  }
#undef CUT
  free(Ux);
  free(Uy);
  o->errorUabsMax = eUmax;
  o->errorUabsAvg = an ? eUavg / (double)nU : 0.0;
  o->errorJ1Max = eJ1max;
  o->errorJ1Avg = an ? eJ1avg / (double)nJ : 0.0;
  o->errorJ2Max = eJ2max;
  o->errorJ2Avg = an ? eJ2avg / (double)nJ : 0.0;
  return an;
}
