// ep2d_kernels.hip -- gfx950 kernels of the EffPlast2D pseudo-transient
// elastoplastic RVE solver.
//
// The file is compiled with -ffp-contract=off: every FP64 expression of the
// plain and fused-FP64 paths is evaluated in the reference's operation order
// with IEEE division and square root, so those paths reproduce the
// reference arithmetic (EffPlast2D.cu:3-177) operation for operation.
//
// Layout: A[j*m + i], x fastest (EffPlast2D.cu:3-177).  Cells nX*nY,
// vertices (nX-1)*(nY-1), Ux/Vx (nX+1)*nY, Uy/Vy nX*(nY+1).
#include "ep2d_internal.h"

#include <cmath>

namespace ep2d {

namespace {

constexpr int TX = FUSED_TX;
constexpr int TY = FUSED_TY;
constexpr int NT = FUSED_NT;

__device__ __forceinline__ double kmat(unsigned char m, double v0, double vp) {
  return (m & M_CELL_MAT) ? vp : v0;
}

// ---------------------------------------------------------------------------
// Pore masks.  SetMaterials (EffPlast2D.cu:708-728) and the two pore tests of
// ComputeStress (EffPlast2D.cu:77-97) evaluated once, with the reference's
// expressions, instead of in every iteration.
__global__ void k_build_mask(unsigned char* __restrict__ mask, int nX, int nY, double dX,
                             double dY, double rad, double nP) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nX || j >= nY) return;
  const double N = nP;
  const double x = -0.5 * dX * (nX - 1) + dX * i;
  const double y = -0.5 * dY * (nY - 1) + dY * j;
  const double Lx = dX * (nX - 1);
  const double Ly = dY * (nY - 1);
  unsigned char m = 0;
  for (int k = 0; k < N; k++) {
    for (int l = 0; l < N; l++) {
      const double ex = x - 0.5 * Lx * (1.0 - 1.0 / N) + (Lx / N) * k;
      const double ey = y - 0.5 * Ly * (1.0 - 1.0 / N) + (Ly / N) * l;
      if (sqrt(ex * ex + ey * ey) < rad) m |= M_CELL_MAT;
      const double cx = -0.5 * dX * (nX - 1) + dX * i - 0.5 * dX * (nX - 1) * (1.0 - 1.0 / N) + (dX * (nX - 1) / N) * k;
      const double cy = -0.5 * dY * (nY - 1) + dY * j - 0.5 * dY * (nY - 1) * (1.0 - 1.0 / N) + (dY * (nY - 1) / N) * l;
      if (sqrt(cx * cx + cy * cy) < rad) m |= M_CELL_ZERO;
      if (i < nX - 1 && j < nY - 1) {
        const double vx = -0.5 * dX * (nX - 2) + dX * i - 0.5 * dX * (nX - 1) * (1.0 - 1.0 / N) + (dX * (nX - 1) / N) * k;
        const double vy = -0.5 * dY * (nY - 2) + dY * j - 0.5 * dY * (nY - 1) * (1.0 - 1.0 / N) + (dY * (nY - 1) / N) * l;
        if (sqrt(vx * vx + vy * vy) < rad) m |= M_VERT_ZERO;
      }
    }
  }
  mask[(size_t)j * nX + i] = m;
}

__global__ void k_fill(double* __restrict__ p, size_t n, double v) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = v;
}

__global__ void k_fill_f32(float* __restrict__ p, size_t n, float v) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    p[k] = v;
}

__global__ void k_f32_to_f64(const float* __restrict__ s, double* __restrict__ d, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    d[k] = (double)s[k];
}

__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    d[k] = (float)s[k];
}

// Affine displacement increment of the macro strain, EffPlast2D.cu:370-381.
__global__ void k_affine(double* __restrict__ Ux, double* __restrict__ Uy, int nX, int nY,
                         double dX, double dY, double dUxdx, double dUydy, double dUxdy,
                         double dUydx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i <= nX && j < nY)
    Ux[(size_t)j * (nX + 1) + i] += (-0.5 * dX * nX + dX * i) * dUxdx + (-0.5 * dY * (nY - 1) + dY * j) * dUxdy;
  if (i < nX && j <= nY)
    Uy[(size_t)j * nX + i] += (-0.5 * dY * nY + dY * j) * dUydy + (-0.5 * dX * (nX - 1) + dX * i) * dUydx;
}

// ---------------------------------------------------------------------------
// Plain scheme: the reference's four passes (EffPlast2D.cu:3-177), one thread
// per cell, stress fields in HBM.  K and G come from the mask (their only two
// values), the pore tests from the precomputed mask bits.

// ComputeStress, EffPlast2D.cu:37-98.
__global__ void k_plain_stress(RveDev r, const double* __restrict__ Ux,
                               const double* __restrict__ Uy, StressDev f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nX = r.nX, nY = r.nY;
  if (i >= nX || j >= nY) return;
  const unsigned char m = r.mask[j * nX + i];
  const double K = kmat(m, r.K0, r.Kp);
  const double G = kmat(m, r.G0, r.Gp);
  const double exx = (Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / r.dX;
  const double eyy = (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / r.dY;
  double p = r.P0 - K * (exx + eyy);
  double txx = 2.0 * G * (exx - (exx + eyy) / 3.0);
  double tyy = 2.0 * G * (eyy - (exx + eyy) / 3.0);
  if (m & M_CELL_ZERO) {
    p = 0.0;
    txx = 0.0;
    tyy = 0.0;
  }
  f.P[j * nX + i] = p;
  f.tauXX[j * nX + i] = txx;
  f.tauYY[j * nX + i] = tyy;
  if (i < nX - 1 && j < nY - 1) {
    const double g = 0.25 * (kmat(m, r.G0, r.Gp) + kmat(r.mask[j * nX + i + 1], r.G0, r.Gp) +
                             kmat(r.mask[(j + 1) * nX + i], r.G0, r.Gp) +
                             kmat(r.mask[(j + 1) * nX + i + 1], r.G0, r.Gp));
    double txy = g * ((Ux[(j + 1) * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i + 1]) / r.dY +
                      (Uy[(j + 1) * nX + i + 1] - Uy[(j + 1) * nX + i]) / r.dX);
    if (m & M_VERT_ZERO) txy = 0.0;
    f.tauXY[j * (nX - 1) + i] = txy;
  }
}

// ComputeJ2, EffPlast2D.cu:100-145.  The nine edge/corner branches of the
// reference all average the 2x2 vertex block starting at
// (clamp(i-1, 0, nX-3), clamp(j-1, 0, nY-3)), in the same summation order.
__global__ void k_plain_j2(RveDev r, StressDev f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nX = r.nX, nY = r.nY;
  if (i >= nX || j >= nY) return;
  const int vi = min(max(i - 1, 0), nX - 3);
  const int vj = min(max(j - 1, 0), nY - 3);
  const double* t = f.tauXY;
  const double av = 0.25 * (t[vj * (nX - 1) + vi] + t[vj * (nX - 1) + vi + 1] +
                            t[(vj + 1) * (nX - 1) + vi] + t[(vj + 1) * (nX - 1) + vi + 1]);
  f.tauXYav[j * nX + i] = av;
  const double txx = f.tauXX[j * nX + i], tyy = f.tauYY[j * nX + i];
  f.J2[j * nX + i] = sqrt(txx * txx + tyy * tyy + 2.0 * av * av);
  if (i < nX - 1 && j < nY - 1) {
    const double sxx = f.tauXX[j * nX + i] + f.tauXX[j * nX + i + 1] + f.tauXX[(j + 1) * nX + i] + f.tauXX[(j + 1) * nX + i + 1];
    const double syy = f.tauYY[j * nX + i] + f.tauYY[j * nX + i + 1] + f.tauYY[(j + 1) * nX + i] + f.tauYY[(j + 1) * nX + i + 1];
    const double txy = t[j * (nX - 1) + i];
    f.J2XY[j * (nX - 1) + i] = sqrt(0.0625 * sxx * sxx + 0.0625 * syy * syy + 2.0 * txy * txy);
  }
}

// ComputePlasticity, EffPlast2D.cu:147-177 (coh = pa[8]).
__global__ void k_plain_plasticity(RveDev r, StressDev f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nX = r.nX, nY = r.nY;
  if (i >= nX || j >= nY) return;
  const double coh = r.coh;
  const double J2 = f.J2[j * nX + i];
  if (J2 > coh) {
    const double s = coh / J2;
    const double txx = f.tauXX[j * nX + i] * s;
    const double tyy = f.tauYY[j * nX + i] * s;
    const double av = f.tauXYav[j * nX + i] * s;
    f.tauXX[j * nX + i] = txx;
    f.tauYY[j * nX + i] = tyy;
    f.tauXYav[j * nX + i] = av;
    f.J2[j * nX + i] = sqrt(txx * txx + tyy * tyy + 2.0 * av * av);
  }
  if (i < nX - 1 && j < nY - 1) {
    const double J2XY = f.J2XY[j * (nX - 1) + i];
    if (J2XY > coh) f.tauXY[j * (nX - 1) + i] *= coh / J2XY;
  }
}

// ComputeDisp, EffPlast2D.cu:3-35.
__global__ void k_plain_disp(RveDev r, double* __restrict__ Ux, double* __restrict__ Uy,
                             StressDev f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nX = r.nX, nY = r.nY;
  if (i >= nX || j >= nY) return;
  double* Vx = static_cast<double*>(r.Vx);
  double* Vy = static_cast<double*>(r.Vy);
  const double* P = f.P;
  const double* tauXX = f.tauXX;
  const double* tauYY = f.tauYY;
  const double* tauXY = f.tauXY;
  double vx = Vx[j * (nX + 1) + i];
  double vy = Vy[j * nX + i];
  if (i > 0 && i < nX && j > 0 && j < nY - 1) {
    vx = vx * r.vdampX + r.dTrho * ((-P[j * nX + i] + P[j * nX + i - 1] + tauXX[j * nX + i] - tauXX[j * nX + i - 1]) / r.dX +
                                    (tauXY[j * (nX - 1) + i - 1] - tauXY[(j - 1) * (nX - 1) + i - 1]) / r.dY);
    Vx[j * (nX + 1) + i] = vx;
  }
  if (i > 0 && i < nX - 1 && j > 0 && j < nY) {
    vy = vy * r.vdampY + r.dTrho * ((-P[j * nX + i] + P[(j - 1) * nX + i] + tauYY[j * nX + i] - tauYY[(j - 1) * nX + i]) / r.dY +
                                    (tauXY[(j - 1) * (nX - 1) + i] - tauXY[(j - 1) * (nX - 1) + i - 1]) / r.dX);
    Vy[j * nX + i] = vy;
  }
  Ux[j * (nX + 1) + i] = Ux[j * (nX + 1) + i] + vx * r.dT;
  Uy[j * nX + i] = Uy[j * nX + i] + vy * r.dT;
}

// ---------------------------------------------------------------------------
// Fused iteration: one workgroup computes one TX x TY tile of points
// (i0 <= i < i0+TX, j0 <= j < j0+TY) for Vx, Vy, Ux, Uy.
//
//   phase 0  U^n box [i0-2, i0+TX+2) x [j0-2, j0+TY+2) and mask -> LDS
//   phase 1  trial stress: cells [i0-1, i0+TX+1)^2, vertices [i0-2, i0+TX)^2
//   phase 2  return mapping: cells and vertices [i0-1, i0+TX)^2
//   phase 3  damped velocity and displacement for the tile -> HBM
//
// U is read from U[cur] and written to U[cur^1] (a neighbour tile may still
// read U^n); V is updated in place (each point reads only its own V).  The
// tile whose x range holds nX-1 also copies column nX of Ux, and the tile
// whose y range holds nY-1 copies row nY of Uy: those nodes are never
// updated (ComputeDisp runs nX x nY threads), so U^{n+1} = U^n there.
//
// CT is the arithmetic type of the stress and velocity update, VT the
// storage type of V.  With CT = double the expressions are those of the
// plain kernels, in the same order.
template <typename CT, typename VT, bool PLASTIC>
struct FusedTile {
  static constexpr int BX = TX + 4, BY = TY + 4;  // U box
  static constexpr int CX = TX + 2, CY = TY + 2;  // trial cells / vertices
  static constexpr int QX = TX + 1, QY = TY + 1;  // corrected cells / vertices
  static constexpr size_t SMEM =
      sizeof(double) * 2 * BX * BY + sizeof(CT) * (4 * CX * CY + 3 * QX * QY) + BX * BY;

  static __device__ __forceinline__ CT diffx(double a, double b, const RveDev& r) {
    if constexpr (sizeof(CT) == 8) return (a - b) / r.dX;
    else return (CT)(a - b) * (CT)r.rdX;
  }
  static __device__ __forceinline__ CT diffy(double a, double b, const RveDev& r) {
    if constexpr (sizeof(CT) == 8) return (a - b) / r.dY;
    else return (CT)(a - b) * (CT)r.rdY;
  }
  static __device__ __forceinline__ CT divx(CT a, const RveDev& r) {
    if constexpr (sizeof(CT) == 8) return a / r.dX;
    else return a * (CT)r.rdX;
  }
  static __device__ __forceinline__ CT divy(CT a, const RveDev& r) {
    if constexpr (sizeof(CT) == 8) return a / r.dY;
    else return a * (CT)r.rdY;
  }
  static __device__ __forceinline__ CT third(CT a) {
    if constexpr (sizeof(CT) == 8) return a / 3.0;
    else return a * (CT)(1.0 / 3.0);
  }

  static __device__ __forceinline__ void run(const RveDev& r, int cur, int i0, int j0,
                                             char* smem) {
    const int nX = r.nX, nY = r.nY;
    const int tid = threadIdx.x;
    double* sUx = reinterpret_cast<double*>(smem);
    double* sUy = sUx + BX * BY;
    CT* sP = reinterpret_cast<CT*>(sUy + BX * BY);
    CT* sTxx = sP + CX * CY;
    CT* sTyy = sTxx + CX * CY;
    CT* sTxy = sTyy + CX * CY;
    CT* cTxx = sTxy + CX * CY;
    CT* cTyy = cTxx + QX * QY;
    CT* cTxy = cTyy + QX * QY;
    unsigned char* sM = reinterpret_cast<unsigned char*>(cTxy + QX * QY);

    const double* __restrict__ Uxi = r.Ux[cur];
    const double* __restrict__ Uyi = r.Uy[cur];

    // phase 0
    for (int e = tid; e < BX * BY; e += NT) {
      const int a = e % BX, b = e / BX;
      const int gi = i0 - 2 + a, gj = j0 - 2 + b;
      double ux = 0.0, uy = 0.0;
      unsigned char m = 0;
      if (gi >= 0 && gj >= 0) {
        if (gi <= nX && gj < nY) ux = Uxi[(size_t)gj * (nX + 1) + gi];
        if (gi < nX && gj <= nY) uy = Uyi[(size_t)gj * nX + gi];
        if (gi < nX && gj < nY) m = r.mask[(size_t)gj * nX + gi];
      }
      sUx[e] = ux;
      sUy[e] = uy;
      sM[e] = m;
    }
    __syncthreads();

    // phase 1: trial stress (ComputeStress, EffPlast2D.cu:37-98)
    for (int e = tid; e < CX * CY; e += NT) {
      const int a = e % CX, b = e / CX;
      {  // cell (i0-1+a, j0-1+b) <-> box (a+1, b+1)
        const int ci = i0 - 1 + a, cj = j0 - 1 + b;
        CT p = 0, txx = 0, tyy = 0;
        if (ci >= 0 && ci < nX && cj >= 0 && cj < nY) {
          const unsigned char m = sM[(b + 1) * BX + a + 1];
          const CT K = (CT)kmat(m, r.K0, r.Kp);
          const CT G = (CT)kmat(m, r.G0, r.Gp);
          const CT exx = diffx(sUx[(b + 1) * BX + a + 2], sUx[(b + 1) * BX + a + 1], r);
          const CT eyy = diffy(sUy[(b + 2) * BX + a + 1], sUy[(b + 1) * BX + a + 1], r);
          p = (CT)r.P0 - K * (exx + eyy);
          txx = (CT)2.0 * G * (exx - third(exx + eyy));
          tyy = (CT)2.0 * G * (eyy - third(exx + eyy));
          if (m & M_CELL_ZERO) {
            p = 0;
            txx = 0;
            tyy = 0;
          }
        }
        sP[e] = p;
        sTxx[e] = txx;
        sTyy[e] = tyy;
      }
      {  // vertex (i0-2+a, j0-2+b) <-> box (a, b)
        const int vi = i0 - 2 + a, vj = j0 - 2 + b;
        CT txy = 0;
        if (vi >= 0 && vi < nX - 1 && vj >= 0 && vj < nY - 1) {
          const unsigned char m00 = sM[b * BX + a];
          const CT g = (CT)(0.25 * (kmat(m00, r.G0, r.Gp) + kmat(sM[b * BX + a + 1], r.G0, r.Gp) +
                                    kmat(sM[(b + 1) * BX + a], r.G0, r.Gp) +
                                    kmat(sM[(b + 1) * BX + a + 1], r.G0, r.Gp)));
          txy = g * (diffy(sUx[(b + 1) * BX + a + 1], sUx[b * BX + a + 1], r) +
                     diffx(sUy[(b + 1) * BX + a + 1], sUy[(b + 1) * BX + a], r));
          if (m00 & M_VERT_ZERO) txy = 0;
        }
        sTxy[e] = txy;
      }
    }
    __syncthreads();

    // phase 2: return mapping (ComputeJ2 + ComputePlasticity, EffPlast2D.cu:100-177)
    if constexpr (PLASTIC) {
      const CT coh = (CT)r.coh;
      for (int e = tid; e < QX * QY; e += NT) {
        const int a = e % QX, b = e / QX;
        {  // cell (i0-1+a, j0-1+b) <-> trial cell (a, b)
          const int ci = i0 - 1 + a, cj = j0 - 1 + b;
          CT txx = sTxx[b * CX + a], tyy = sTyy[b * CX + a];
          if (ci >= 0 && ci < nX && cj >= 0 && cj < nY) {
            const int va = min(max(ci - 1, 0), nX - 3) - (i0 - 2);
            const int vb = min(max(cj - 1, 0), nY - 3) - (j0 - 2);
            const CT av = (CT)0.25 * (sTxy[vb * CX + va] + sTxy[vb * CX + va + 1] +
                                      sTxy[(vb + 1) * CX + va] + sTxy[(vb + 1) * CX + va + 1]);
            const CT J2 = sqrt(txx * txx + tyy * tyy + (CT)2.0 * av * av);
            if (J2 > coh) {
              const CT s = coh / J2;
              txx = txx * s;
              tyy = tyy * s;
            }
          }
          cTxx[e] = txx;
          cTyy[e] = tyy;
        }
        {  // vertex (i0-1+a, j0-1+b) <-> trial vertex (a+1, b+1), trial cells (a..a+1, b..b+1)
          const int vi = i0 - 1 + a, vj = j0 - 1 + b;
          CT txy = sTxy[(b + 1) * CX + a + 1];
          if (vi >= 0 && vi < nX - 1 && vj >= 0 && vj < nY - 1) {
            const CT sxx = sTxx[b * CX + a] + sTxx[b * CX + a + 1] + sTxx[(b + 1) * CX + a] + sTxx[(b + 1) * CX + a + 1];
            const CT syy = sTyy[b * CX + a] + sTyy[b * CX + a + 1] + sTyy[(b + 1) * CX + a] + sTyy[(b + 1) * CX + a + 1];
            const CT J2XY = sqrt((CT)0.0625 * sxx * sxx + (CT)0.0625 * syy * syy + (CT)2.0 * txy * txy);
            if (J2XY > coh) txy *= coh / J2XY;
          }
          cTxy[e] = txy;
        }
      }
      __syncthreads();
    }

    // phase 3: damped velocity and displacement (ComputeDisp, EffPlast2D.cu:3-35)
    VT* __restrict__ Vx = static_cast<VT*>(r.Vx);
    VT* __restrict__ Vy = static_cast<VT*>(r.Vy);
    double* __restrict__ Uxo = r.Ux[cur ^ 1];
    double* __restrict__ Uyo = r.Uy[cur ^ 1];
    const CT vdampX = (CT)r.vdampX, vdampY = (CT)r.vdampY, dTrho = (CT)r.dTrho;
    for (int e = tid; e < TX * TY; e += NT) {
      const int a = e % TX, b = e / TX;
      const int i = i0 + a, j = j0 + b;
      if (i >= nX || j >= nY) continue;
      // corrected cell (i,j) <-> (a+1, b+1) in both trial (P) and corrected arrays;
      // corrected vertex (i-1, j-1) <-> (a, b) corrected / (a+1, b+1) trial.
      auto Pc = [&](int x, int y) -> CT { return sP[y * CX + x]; };
      auto Txx = [&](int x, int y) -> CT {
        if constexpr (PLASTIC) return cTxx[y * QX + x];
        else return sTxx[y * CX + x];
      };
      auto Tyy = [&](int x, int y) -> CT {
        if constexpr (PLASTIC) return cTyy[y * QX + x];
        else return sTyy[y * CX + x];
      };
      auto Txy = [&](int x, int y) -> CT {
        if constexpr (PLASTIC) return cTxy[y * QX + x];
        else return sTxy[(y + 1) * CX + x + 1];
      };
      const size_t ix = (size_t)j * (nX + 1) + i;
      const size_t iy = (size_t)j * nX + i;
      CT vx = (CT)Vx[ix];
      CT vy = (CT)Vy[iy];
      if (i > 0 && j > 0 && j < nY - 1) {
        vx = vx * vdampX + dTrho * (divx(-Pc(a + 1, b + 1) + Pc(a, b + 1) + Txx(a + 1, b + 1) - Txx(a, b + 1), r) +
                                    divy(Txy(a, b + 1) - Txy(a, b), r));
        Vx[ix] = (VT)vx;
      }
      if (i > 0 && i < nX - 1 && j > 0) {
        vy = vy * vdampY + dTrho * (divy(-Pc(a + 1, b + 1) + Pc(a + 1, b) + Tyy(a + 1, b + 1) - Tyy(a + 1, b), r) +
                                    divx(Txy(a + 1, b) - Txy(a, b), r));
        Vy[iy] = (VT)vy;
      }
      Uxo[ix] = sUx[(b + 2) * BX + a + 2] + (double)vx * r.dT;
      Uyo[iy] = sUy[(b + 2) * BX + a + 2] + (double)vy * r.dT;
      if (i == nX - 1) Uxo[ix + 1] = sUx[(b + 2) * BX + a + 3];
      if (j == nY - 1) Uyo[iy + nX] = sUy[(b + 3) * BX + a + 2];
    }
  }
};

// Block -> tile map.  Workgroups are dispatched round-robin over the 8 XCDs;
// when the tile count divides by 8, XCD k receives the contiguous band of
// tiles [k*nb/8, (k+1)*nb/8), so vertically adjacent tiles (which share halo
// rows) run on the same XCD and share its L2.
__device__ __forceinline__ int tile_of_block(int b, int nb) {
  if ((nb & 7) == 0) return (b & 7) * (nb >> 3) + (b >> 3);
  return b;
}

template <typename CT, typename VT, bool PLASTIC>
__global__ __launch_bounds__(NT) void k_fused(RveDev r, int cur, int ntx, int nb) {
  using Tile = FusedTile<CT, VT, PLASTIC>;
  __shared__ __attribute__((aligned(16))) char smem[Tile::SMEM];
  const int t = tile_of_block(blockIdx.x, nb);
  Tile::run(r, cur, (t % ntx) * TX, (t / ntx) * TY, smem);
}

// ---------------------------------------------------------------------------
// Residual (EffPlast2D.cu:407-412, FindMaxAbs :805-813): max |Vx|, max |Vy|
// and a count of non-finite values.  Wave64 shuffle reduction, then one
// 64-bit atomicMax per block on the bit pattern (monotone for values >= 0).
// NaN is skipped as in FindMaxAbs, and counted.
template <typename VT>
__global__ __launch_bounds__(256) void k_maxabs(const VT* __restrict__ Vx, size_t nx,
                                                const VT* __restrict__ Vy, size_t ny,
                                                unsigned long long* out) {
  double mx = 0.0, my = 0.0;
  unsigned long long bad = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < nx; k += stride) {
    const double v = fabs((double)Vx[k]);
    if (v > mx) mx = v;
    if (!(v <= 1.7976931348623157e308)) bad++;
  }
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < ny; k += stride) {
    const double v = fabs((double)Vy[k]);
    if (v > my) my = v;
    if (!(v <= 1.7976931348623157e308)) bad++;
  }
  for (int off = 32; off > 0; off >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, off, 64));
    my = fmax(my, __shfl_xor(my, off, 64));
    bad += __shfl_xor(bad, off, 64);
  }
  __shared__ double smx[4], smy[4];
  __shared__ unsigned long long sbad[4];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0) {
    smx[w] = mx;
    smy[w] = my;
    sbad[w] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
      mx = fmax(mx, smx[k]);
      my = fmax(my, smy[k]);
      bad += sbad[k];
    }
    atomicMax(&out[0], (unsigned long long)__double_as_longlong(mx));
    atomicMax(&out[1], (unsigned long long)__double_as_longlong(my));
    if (bad) atomicAdd(&out[2], bad);
  }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------------------
// Volume averages (EffPlast2D.cu:453-503) in FP64.  Each block reduces an
// AVG_TX x AVG_TY tile of cells in a fixed order into NAVG partials; a
// single block then reduces the partials in a fixed order.  The result is
// deterministic and independent of other work on the device.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  return ((sh[0] + sh[1]) + (sh[2] + sh[3]));
}

__global__ __launch_bounds__(256) void k_avg_partial(RveDev r, StressDev f,
                                                     const double* __restrict__ Ux,
                                                     const double* __restrict__ Uy, double nu0,
                                                     int pox, int poy, int ntx,
                                                     double* __restrict__ partial) {
  const int nX = r.nX, nY = r.nY;
  const int bx = blockIdx.x % ntx, by = blockIdx.x / ntx;
  const int i = bx * AVG_TX + (threadIdx.x % AVG_TX);
  const int j = by * AVG_TY + (threadIdx.x / AVG_TX);
  double acc[NAVG];
#pragma unroll
  for (int k = 0; k < NAVG; k++) acc[k] = 0.0;
  if (i < nX && j < nY) {
    const double* P = f.P;
    const double* tauXX = f.tauXX;
    const double* tauYY = f.tauYY;
    const double* tauXY = f.tauXY;
    if (i >= 1 && i < nX - 1 && j >= 1 && j < nY - 1) {
      const double txx = tauXX[j * nX + i], tyy = tauYY[j * nX + i], p = P[j * nX + i];
      acc[A_S0] = txx - p;
      acc[A_S1] = tyy - p;
      acc[A_S2] = tauXY[j * (nX - 1) + i] + nu0 * (txx + tyy - 2.0 * p);
      if (i > pox && i < nX - pox && j > poy && j < nY - poy) {
        const double dX = r.dX, dY = r.dY;
        acc[A_SP0] = txx - p;
        acc[A_SP1] = tyy - p;
        acc[A_SP2] = 0.25 * (tauXY[(j - 1) * (nX - 1) + i - 1] + tauXY[(j - 1) * (nX - 1) + i] +
                             tauXY[j * (nX - 1) + i - 1] + tauXY[j * (nX - 1) + i]);
        acc[A_SP3] = nu0 * (txx + tyy - 2.0 * p);
        acc[A_E0] = (Ux[j * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i]) / dX;
        acc[A_E1] = (Uy[(j + 1) * nX + i] - Uy[j * nX + i]) / dY;
        acc[A_E2] = 0.125 * (
          (Ux[j * (nX + 1) + i] - Ux[(j - 1) * (nX + 1) + i]) / dY + (Uy[j * nX + i] - Uy[j * nX + i - 1]) / dX +
          (Ux[j * (nX + 1) + i + 1] - Ux[(j - 1) * (nX + 1) + i + 1]) / dY + (Uy[j * nX + i + 1] - Uy[j * nX + i]) / dX +
          (Ux[(j + 1) * (nX + 1) + i] - Ux[j * (nX + 1) + i]) / dY + (Uy[(j + 1) * nX + i] - Uy[(j + 1) * nX + i - 1]) / dX +
          (Ux[(j + 1) * (nX + 1) + i + 1] - Ux[j * (nX + 1) + i + 1]) / dY + (Uy[(j + 1) * nX + i + 1] - Uy[(j + 1) * nX + i]) / dX);
      }
    }
    if (i < nX - 1 && j < nY - 1) acc[A_SXY] = tauXY[j * (nX - 1) + i];
  }
  __shared__ double sh[4];
#pragma unroll
  for (int k = 0; k < NAVG; k++) {
    const double s = block_sum256(acc[k], sh);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * NAVG + k] = s;
  }
}

__global__ __launch_bounds__(256) void k_avg_final(const double* __restrict__ partial, int nb,
                                                   double* __restrict__ result) {
  __shared__ double sh[4];
  for (int k = 0; k < NAVG; k++) {
    double v = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) v += partial[(size_t)b * NAVG + k];
    const double s = block_sum256(v, sh);
    if (threadIdx.x == 0) result[k] = s;
  }
}

// Gather of sampled field values: bases = {P, tauXX, tauYY, Ux[0], Uy[0],
// Ux[1], Uy[1]}; displacement entries (ids 3, 4) are redirected to the
// current ping-pong buffer.
__global__ void k_gather(const double* const* __restrict__ bases, const int2* __restrict__ table,
                         int n, int ucur, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    int f = table[k].x;
    if (f >= 3) f += 2 * ucur;
    out[k] = bases[f][table[k].y];
  }
}

// Plain streaming copy (bandwidth calibration).
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src,
                                              double2* __restrict__ dst, size_t n) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}

inline dim3 grid2(int nx, int ny, dim3 blk) {
  return dim3((nx + blk.x - 1) / blk.x, (ny + blk.y - 1) / blk.y, 1);
}

inline int tiles_x(int nX) { return (nX + TX - 1) / TX; }
inline int tiles_y(int nY) { return (nY + TY - 1) / TY; }

}  // namespace

// ---- launchers --------------------------------------------------------------
void launch_build_mask(unsigned char* mask, int nX, int nY, double dX, double dY, double rad,
                       double nPores, hipStream_t s) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(k_build_mask, grid2(nX, nY, blk), blk, 0, s, mask, nX, nY, dX, dY, rad,
                     nPores);
  EP2D_HIP(hipGetLastError());
}

void launch_fill(double* p, size_t n, double v, hipStream_t s) {
  if (!n) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, n, v);
  EP2D_HIP(hipGetLastError());
}

void launch_fill_f32(float* p, size_t n, float v, hipStream_t s) {
  if (!n) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_f32, dim3(blocks), dim3(256), 0, s, p, n, v);
  EP2D_HIP(hipGetLastError());
}

void launch_f32_to_f64(const float* src, double* dst, size_t n, hipStream_t s) {
  if (!n) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_f32_to_f64, dim3(blocks), dim3(256), 0, s, src, dst, n);
  EP2D_HIP(hipGetLastError());
}

void launch_f64_to_f32(const double* src, float* dst, size_t n, hipStream_t s) {
  if (!n) return;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks), dim3(256), 0, s, src, dst, n);
  EP2D_HIP(hipGetLastError());
}

void launch_affine(double* Ux, double* Uy, int nX, int nY, double dX, double dY, double dUxdx,
                   double dUydy, double dUxdy, double dUydx, hipStream_t s) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(k_affine, grid2(nX + 1, nY + 1, blk), blk, 0, s, Ux, Uy, nX, nY, dX, dY,
                     dUxdx, dUydy, dUxdy, dUydx);
  EP2D_HIP(hipGetLastError());
}

void launch_plain_stress(const RveDev& r, const double* Ux, const double* Uy, const StressDev& f,
                         hipStream_t s) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(k_plain_stress, grid2(r.nX, r.nY, blk), blk, 0, s, r, Ux, Uy, f);
  EP2D_HIP(hipGetLastError());
}

void launch_plain_j2(const RveDev& r, const StressDev& f, hipStream_t s) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(k_plain_j2, grid2(r.nX, r.nY, blk), blk, 0, s, r, f);
  EP2D_HIP(hipGetLastError());
}

void launch_plain_plasticity(const RveDev& r, const StressDev& f, hipStream_t s) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(k_plain_plasticity, grid2(r.nX, r.nY, blk), blk, 0, s, r, f);
  EP2D_HIP(hipGetLastError());
}

void launch_plain_disp(const RveDev& r, double* Ux, double* Uy, const StressDev& f,
                       hipStream_t s) {
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(k_plain_disp, grid2(r.nX, r.nY, blk), blk, 0, s, r, Ux, Uy, f);
  EP2D_HIP(hipGetLastError());
}

void launch_fused_fp64(const RveDev& r, int cur, bool plastic, hipStream_t s) {
  const int ntx = tiles_x(r.nX), nb = ntx * tiles_y(r.nY);
  if (plastic)
    hipLaunchKernelGGL((k_fused<double, double, true>), dim3(nb), dim3(NT), 0, s, r, cur, ntx, nb);
  else
    hipLaunchKernelGGL((k_fused<double, double, false>), dim3(nb), dim3(NT), 0, s, r, cur, ntx, nb);
  EP2D_HIP(hipGetLastError());
}

void launch_fused_split(const RveDev& r, int cur, bool plastic, hipStream_t s) {
  const int ntx = tiles_x(r.nX), nb = ntx * tiles_y(r.nY);
  if (plastic)
    hipLaunchKernelGGL((k_fused<float, float, true>), dim3(nb), dim3(NT), 0, s, r, cur, ntx, nb);
  else
    hipLaunchKernelGGL((k_fused<float, float, false>), dim3(nb), dim3(NT), 0, s, r, cur, ntx, nb);
  EP2D_HIP(hipGetLastError());
}

const char* fused_kernel_name(int mode) {
  switch (mode) {
    case 0: return "k_plain_disp";
    case 1: return "k_fused<double, double, true>";
    default: return "k_fused<float, float, true>";
  }
}

void launch_maxabs(const RveDev& r, bool f32, unsigned long long* out, hipStream_t s) {
  const size_t nx = (size_t)(r.nX + 1) * r.nY, ny = (size_t)r.nX * (r.nY + 1);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, out, 3);
  const int blocks = (int)std::min<size_t>((std::max(nx, ny) + 255) / 256, 2048);
  if (f32)
    hipLaunchKernelGGL((k_maxabs<float>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const float*>(r.Vx), nx, static_cast<const float*>(r.Vy), ny, out);
  else
    hipLaunchKernelGGL((k_maxabs<double>), dim3(blocks), dim3(256), 0, s,
                       static_cast<const double*>(r.Vx), nx, static_cast<const double*>(r.Vy), ny, out);
  EP2D_HIP(hipGetLastError());
}

int avg_partial_blocks(int nX, int nY) {
  return ((nX + AVG_TX - 1) / AVG_TX) * ((nY + AVG_TY - 1) / AVG_TY);
}

void launch_averages(const RveDev& r, const StressDev& f, const double* Ux, const double* Uy,
                     double nu0, int perOffX, int perOffY, double* partial, double* result,
                     hipStream_t s) {
  const int ntx = (r.nX + AVG_TX - 1) / AVG_TX;
  const int nb = avg_partial_blocks(r.nX, r.nY);
  hipLaunchKernelGGL(k_avg_partial, dim3(nb), dim3(256), 0, s, r, f, Ux, Uy, nu0, perOffX, perOffY,
                     ntx, partial);
  hipLaunchKernelGGL(k_avg_final, dim3(1), dim3(256), 0, s, partial, nb, result);
  EP2D_HIP(hipGetLastError());
}

void launch_gather(const double* const* bases, const int2* table, int n, int ucur, double* out,
                   hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, bases, table, n, ucur, out);
  EP2D_HIP(hipGetLastError());
}

void launch_copy(const double* src, double* dst, size_t n, hipStream_t s) {
  const size_t n2 = n / 2;
  const int blocks = (int)std::min<size_t>((n2 + 255) / 256, 8192);
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, s, (const double2*)src, (double2*)dst, n2);
  EP2D_HIP(hipGetLastError());
}

}  // namespace ep2d
