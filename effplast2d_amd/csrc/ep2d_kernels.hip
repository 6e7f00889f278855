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

#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <utility>

namespace ep2d {

namespace {


__device__ __forceinline__ double kmat(unsigned char m, double v0, double vp) {
  return (m & M_CELL_MAT) ? vp : v0;
}

// ---------------------------------------------------------------------------
// Pore masks.  SetMaterials (EffPlast2D.cu:708-728) and the two pore tests of
// ComputeStress (EffPlast2D.cu:77-97) evaluated once, with the reference's
// expressions, instead of in every iteration.
__device__ bool mat_pore(int i, int j, int nX, int nY, double dX, double dY, double rad,
                         double N) {
  const double x = -0.5 * dX * (nX - 1) + dX * i;
  const double y = -0.5 * dY * (nY - 1) + dY * j;
  const double Lx = dX * (nX - 1);
  const double Ly = dY * (nY - 1);
  bool in = false;
  for (int k = 0; k < N; k++) {
    for (int l = 0; l < N; l++) {
      const double ex = x - 0.5 * Lx * (1.0 - 1.0 / N) + (Lx / N) * k;
      const double ey = y - 0.5 * Ly * (1.0 - 1.0 / N) + (Ly / N) * l;
      if (sqrt(ex * ex + ey * ey) < rad) in = true;
    }
  }
  return in;
}

__global__ void k_build_mask(unsigned char* __restrict__ mask, int nX, int nY, double dX,
                             double dY, double rad, double nP) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i >= nX || j >= nY) return;
  const double N = nP;
  unsigned char m = mat_pore(i, j, nX, nY, dX, dY, rad, N) ? M_CELL_MAT : 0;
  for (int k = 0; k < N; k++) {
    for (int l = 0; l < N; l++) {
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
  if (i < nX - 1 && j < nY - 1) {
    const int c = (m & M_CELL_MAT ? 1 : 0) | (mat_pore(i + 1, j, nX, nY, dX, dY, rad, N) ? 2 : 0) |
                  (mat_pore(i, j + 1, nX, nY, dX, dY, rad, N) ? 4 : 0) |
                  (mat_pore(i + 1, j + 1, nX, nY, dX, dY, rad, N) ? 8 : 0);
    m |= (unsigned char)(c << M_VCLASS_SHIFT);
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
// Quotient by a uniform divisor d, given rd = RN(1/d): one product and one
// FMA correction (Markstein's method, the same correction step the AMDGPU and
// IA-64 division sequences end with).  For the operand ranges of this solver
// (no overflow, underflow or subnormal quotients) it returns the IEEE quotient
// a / d, so the fused FP64 kernel keeps the reference's arithmetic.
__device__ __forceinline__ double qdiv(double a, double d, double rd) {
  const double q = a * rd;
  return fma(fma(-q, d, a), rd, q);
}

// Global-memory access at a 32-bit element index: the base stays in SGPRs
// (uniform per wave) and the byte offset in one VGPR (saddr form).
template <typename T>
__device__ __forceinline__ T gld(const T* base, unsigned idx) {
  using GT = const __attribute__((address_space(1))) T;
  using GC = const __attribute__((address_space(1))) char;
  return *(GT*)((GC*)base + idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void gst(T* base, unsigned idx, T v) {
  using GT = __attribute__((address_space(1))) T;
  using GC = __attribute__((address_space(1))) char;
  *(GT*)((GC*)base + idx * (unsigned)sizeof(T)) = v;
}

// Raw buffer access for row marches: a resource (base in SGPRs; no range
// limit, every access is in bounds by construction), a per-lane byte offset
// fixed along the march and the row's byte offset, uniform, in an SGPR, so
// stepping to the next row costs no vector instruction.  readfirstlane states
// the row offset's uniformity (a value the compiler kept in a VGPR then costs
// one read, not a loop over lanes).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void* p) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, -1, 0x00020000);
}
// A resource bounded to the field's bytes (under 2 GiB): an access whose lane
// offset is kOob or more lies past the bound, so a load returns 0 and a store
// is dropped by the hardware.  A bound of 0 (bnull) does that for every lane.
constexpr unsigned kOob = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc_n(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes,
                                           0x00020000);
}
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ T bld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
  soff = __builtin_amdgcn_readfirstlane(soff);
  if constexpr (sizeof(T) == 8)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
  else if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, T v) {
  soff = __builtin_amdgcn_readfirstlane(soff);
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, voff, soff, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

// Return-mapping scale (ComputeJ2 / ComputePlasticity, EffPlast2D.cu:109-176):
// a stress whose squared invariant s2 lies outside the yield radius coh is
// scaled by sc = coh / sqrt(s2); returns whether it is.  FP64 keeps the
// reference's test and quotient (sqrt, compare, divide), skipped when s2 <= c2
// = coh2lo, below which sqrt(s2) <= coh exactly.  The FP32 split scheme tests
// s2 against c2 = coh^2 (rounded to FP32) and scales by coh * rsq(s2), the
// hardware reciprocal square root (about one ulp): the scaled stress lies on
// the yield surface to FP32 accuracy either way.
template <typename CT>
__device__ __forceinline__ bool rm_scale(CT s2, CT coh, CT c2, CT& sc) {
  if constexpr (sizeof(CT) == 8) {
    if (s2 > c2) {
      const CT J = sqrt(s2);
      if (J > coh) {
        sc = coh / J;
        return true;
      }
    }
    return false;
  } else {
    if (s2 > c2) {
      sc = coh * __builtin_amdgcn_rsqf(s2);
      return true;
    }
    return false;
  }
}

// The split scheme's FP32 arithmetic.  Every kernel computes these quantities
// through the functions below, so the split kernels agree with each other bit
// for bit.  The terms follow the reference's FP64 expressions
// (EffPlast2D.cu:16-33, 54-75, 109-176), which the FP64 scheme keeps as they
// are; here each product that ends in a sum is fused into it (one rounding
// instead of two, and one instruction fewer).  With residual checks every 500
// iterations, the 1024^2 RVE of three load steps stops at the same checks as
// the FP64 solve in every step, as it does with separate roundings.
//
// trial cell stress from the strain pair: P0 - K (exx + eyy) and
// 2G (e - (exx + eyy) / 3), with g2 = 2G
__device__ __forceinline__ void split_cell_trial(float exx, float eyy, float K, float g2, float P0,
                                                 float& pp, float& txx, float& tyy) {
  const float s = exx + eyy;
  pp = __builtin_fmaf(-K, s, P0);
  txx = g2 * __builtin_fmaf(s, -(1.0f / 3.0f), exx);
  tyy = g2 * __builtin_fmaf(s, -(1.0f / 3.0f), eyy);
}
// trial vertex shear: g (a / dY + b / dX) from the differences a of Ux in y
// and b of Uy in x
__device__ __forceinline__ float split_vertex_trial(float g, float a, float rdY, float b,
                                                   float rdX) {
  return g * __builtin_fmaf(a, rdY, b * rdX);
}
// squared invariants the return mapping tests: vertex, from the four-cell
// sums of the normal stresses; cell, from the four-vertex mean av
__device__ __forceinline__ float split_s2_vertex(float sxx, float syy, float txy) {
  return __builtin_fmaf(0.0625f * sxx, sxx, __builtin_fmaf(0.0625f * syy, syy, 2.0f * txy * txy));
}
__device__ __forceinline__ float split_s2_cell(float txx, float tyy, float av) {
  return __builtin_fmaf(txx, txx, __builtin_fmaf(tyy, tyy, 2.0f * av * av));
}
// damped velocity: v damp + dTrho (a ra + b rb), ra and rb the reciprocal
// grid steps of the two stress differences a and b
__device__ __forceinline__ float split_velocity(float v, float damp, float dTrho, float a,
                                               float ra, float b, float rb) {
  return __builtin_fmaf(v, damp, dTrho * __builtin_fmaf(a, ra, b * rb));
}
// displacement, FP64: u + v dT with one rounding (a fused multiply-add; the
// FP64 scheme keeps the reference's product and sum)
__device__ __forceinline__ double split_disp(double u, float v, double dT) {
  return __builtin_fma((double)v, dT, u);
}

// Workgroup barrier that waits for LDS traffic only: global loads issued
// before it (the next tile's prefetch) stay in flight.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// ---------------------------------------------------------------------------
// Fused iteration.  A workgroup of NW wave64s computes one tile of TX = 62 x TY
// points (i0 <= i < i0+62, j0 <= j < j0+TY) for Vx, Vy, Ux, Uy.  Lane a of a
// wave owns column a of every per-tile array, waves take rows in turn:
//
//   stage     U^n box [i0-2, i0+64) x [j0-2, j0+TY+2) (66 wide) and the pore
//             mask -> LDS; this thread's V^n and U^n -> registers
//   phase 1   trial stress (ComputeStress, EffPlast2D.cu:37-98): cells
//             i0-1+a and vertices i0-2+a, a = lane (64 wide)
//   phase 2a  vertex return mapping (ComputeJ2 / ComputePlasticity,
//             EffPlast2D.cu:138-144, 172-176), 63 wide, over the dead U box
//   phase 2b  cell return mapping (:109-137, :165-170), 63 wide, in place
//   phase 3   damped velocity and displacement (ComputeDisp, :3-35) -> HBM
//
// U is read from U[cur] and written to U[cur^1] (a neighbour tile may still
// read U^n); V is updated in place (a point reads only its own V).  The tile
// holding i = nX-1 also copies column nX of Ux and the tile holding j = nY-1
// copies row nY of Uy: ComputeDisp never updates those nodes.
//
// Tiles whose whole dependency box lies inside the grid run without bounds
// checks, edge clamps or boundary conditions (EDGE = false).  Return mappings
// go through rm_scale (the reference's test and quotient in FP64).
//
// CT is the arithmetic type of stress and velocity, VT the storage type of V.
// With CT = double every expression is the plain kernels' expression, in the
// same order, so the fused FP64 scheme reproduces the plain scheme bit for bit.
// One thread's global reads of a tile (the same for every Fused variant of a
// tile shape, so the persistent kernel can fetch before it knows whether the
// tile's RVE is plastic).
template <typename VT, int KB, int KP>
struct TileReads {
  double bux[KB], buy[KB];
  unsigned int bm[KB];
  VT vxr[KP], vyr[KP];
};

template <typename CT, typename VT, bool PLASTIC, int TY, int NW>
struct Fused {
  static constexpr int S = 64;                 // row stride of the per-tile arrays
  static constexpr int TX = S - 2;             // output columns
  static constexpr int NT = 64 * NW;
  static constexpr int BX = TX + 4, BY = TY + 4, NBOX = BX * BY;  // U box
  static constexpr int CY = TY + 2;            // trial rows
  static constexpr int QY = TY + 1;            // corrected rows
  static constexpr int KP = (TY + NW - 1) / NW;  // output rows per wave
  static constexpr size_t OFF_BOX = 128;       // after the 16-entry G table
  static constexpr size_t OFF_C = OFF_BOX + 16 * NBOX;
  static constexpr size_t OFF_M = OFF_C + 4 * sizeof(CT) * CY * S;
  static constexpr size_t SMEM = (OFF_M + NBOX + 15) / 16 * 16;
  static_assert(sizeof(CT) * QY * S <= 16 * NBOX, "corrected vertex stress must fit the U box");

  static __device__ __forceinline__ CT dx(double a, const RveDev& r) {
    if constexpr (sizeof(CT) == 8) return qdiv(a, r.dX, r.rdX);
    else return (CT)a * (CT)r.rdX;
  }
  static __device__ __forceinline__ CT dy(double a, const RveDev& r) {
    if constexpr (sizeof(CT) == 8) return qdiv(a, r.dY, r.rdY);
    else return (CT)a * (CT)r.rdY;
  }
  static __device__ __forceinline__ CT third(CT a) {
    if constexpr (sizeof(CT) == 8) return qdiv(a, 3.0, 1.0 / 3.0);
    else return a * (CT)(1.0 / 3.0);
  }

  // Global reads of one tile: the U box and mask (row-major, element
  // f = threadIdx.x + NT*q at (f % BX, f / BX), so every lane of every load is
  // used) and this thread's velocities.
  static constexpr int KB = (NBOX + NT - 1) / NT;   // box elements per thread
  using Pre = TileReads<VT, KB, KP>;
  template <bool EDGE>
  static __device__ __forceinline__ void fetch(const RveDev& r, int cur, int i0, int j0, Pre& p) {
    const int nX = r.nX, nY = r.nY;
    // the wave index is uniform: row loops and row addresses stay scalar
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double* __restrict__ Uxi = r.Ux[cur];
    const double* __restrict__ Uyi = r.Uy[cur];
    const unsigned char* __restrict__ msk = r.mask;
    const int sx = nX + 1;  // Ux row length
    // ---- stage: issue every load of this thread first, then fill LDS ----
    // The box is read in row-major order, element f = threadIdx.x + NT*q at
    // (f % BX, f / BX), so every lane of every load instruction is used.
#pragma unroll
    for (int q = 0; q < KB; q++) {
      const int f = threadIdx.x + NT * q;
      const int b = f / BX, a = f - (f / BX) * BX;
      const int gj = j0 - 2 + b, gi = i0 - 2 + a;
      double ux = 0.0, uy = 0.0;
      unsigned int m = 0;
      if (f < NBOX) {
        if (!EDGE) {
          ux = gld(Uxi, (unsigned)(gj * sx + gi));
          uy = gld(Uyi, (unsigned)(gj * nX + gi));
          m = gld(msk, (unsigned)(gj * nX + gi));
        } else if (gi >= 0 && gj >= 0) {
          if (gi <= nX && gj < nY) ux = gld(Uxi, (unsigned)(gj * sx + gi));
          if (gi < nX && gj <= nY) uy = gld(Uyi, (unsigned)(gj * nX + gi));
          if (gi < nX && gj < nY) m = gld(msk, (unsigned)(gj * nX + gi));
        }
      }
      p.bux[q] = ux;
      p.buy[q] = uy;
      p.bm[q] = m;
    }
    // this thread's velocities (consumed in phase 3: the latency overlaps phases 1-2)
    const VT* __restrict__ Vxi = static_cast<const VT*>(r.Vx);
    const VT* __restrict__ Vyi = static_cast<const VT*>(r.Vy);
#pragma unroll
    for (int k = 0; k < KP; k++) {
      const int b = wave + NW * k, i = i0 + lane, j = j0 + b;
      p.vxr[k] = 0;
      p.vyr[k] = 0;
      if (b < TY && lane < TX && (!EDGE || (i < nX && j < nY))) {
        p.vxr[k] = gld(Vxi + (size_t)j * sx, (unsigned)i);
        p.vyr[k] = gld(Vyi + (size_t)j * nX, (unsigned)i);
      }
    }
  }

  template <bool EDGE>
  static __device__ __forceinline__ void run(const RveDev& r, int cur, int i0, int j0, char* smem) {
    Pre p;
    fetch<EDGE>(r, cur, i0, j0, p);
    compute<EDGE>(r, cur, i0, j0, smem, p);
  }

  template <bool EDGE>
  static __device__ __forceinline__ void compute(const RveDev& r, int cur, int i0, int j0, char* smem,
                                                 const Pre& p) {
    const int nX = r.nX, nY = r.nY;
    // the wave index is uniform: row loops and row addresses stay scalar
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    CT* gt = reinterpret_cast<CT*>(smem);  // vertex G by class, in the arithmetic type
    double* sUx = reinterpret_cast<double*>(smem + OFF_BOX);
    double* sUy = sUx + NBOX;
    CT* sP = reinterpret_cast<CT*>(smem + OFF_C);
    CT* sTxx = sP + CY * S;
    CT* sTyy = sTxx + CY * S;
    CT* sTxy = sTyy + CY * S;
    CT* cTxy = reinterpret_cast<CT*>(smem + OFF_BOX);  // written after phase 1
    unsigned char* sM = reinterpret_cast<unsigned char*>(smem + OFF_M);
    const int sx = nX + 1;  // Ux row length

    // ---- stage: LDS fill (the global reads were issued by fetch) ----
    if (threadIdx.x == 0) {
#pragma unroll
      for (int c = 0; c < 16; c++) gt[c] = (CT)r.gtab[c];
    }
#pragma unroll
    for (int q = 0; q < KB; q++) {
      const int f = threadIdx.x + NT * q;
      if (f < NBOX) {
        sUx[f] = p.bux[q];
        sUy[f] = p.buy[q];
        sM[f] = (unsigned char)p.bm[q];
      }
    }
    lds_barrier();

    double uxo[KP], uyo[KP], uxe[KP], uye[KP];
#pragma unroll
    for (int k = 0; k < KP; k++) {
      const int b = wave + NW * k;
      const int bb = b < TY ? b : 0;
      const int a = lane < TX ? lane : 0;
      uxo[k] = sUx[(bb + 2) * BX + a + 2];
      uyo[k] = sUy[(bb + 2) * BX + a + 2];
      uxe[k] = EDGE ? sUx[(bb + 2) * BX + a + 3] : 0.0;
      uye[k] = EDGE ? sUy[(bb + 3) * BX + a + 2] : 0.0;
    }

    // the two values of K and G (pore or matrix), in the arithmetic type
    const CT K0 = (CT)r.K0, Kp = (CT)r.Kp, G0 = (CT)r.G0, Gp = (CT)r.Gp;

    // ---- phase 1: trial stress ----
    constexpr int K1 = (CY + NW - 1) / NW;
#pragma unroll
    for (int k1 = 0; k1 < K1; k1++) {
      const int b = wave + NW * k1;
      if (b >= CY) break;
      const int a = lane;
      {  // cell (i0-1+a, j0-1+b) <-> box (a+1, b+1)
        const int ci = i0 - 1 + a, cj = j0 - 1 + b;
        CT pp = 0, txx = 0, tyy = 0;
        if (!EDGE || (ci >= 0 && ci < nX && cj >= 0 && cj < nY)) {
          const unsigned char m = sM[(b + 1) * BX + a + 1];
          const CT K = (m & M_CELL_MAT) ? Kp : K0;
          const CT G = (m & M_CELL_MAT) ? Gp : G0;
          const CT exx = dx(sUx[(b + 1) * BX + a + 2] - sUx[(b + 1) * BX + a + 1], r);
          const CT eyy = dy(sUy[(b + 2) * BX + a + 1] - sUy[(b + 1) * BX + a + 1], r);
          if constexpr (sizeof(CT) == 8) {
            pp = (CT)r.P0 - K * (exx + eyy);
            txx = (CT)2.0 * G * (exx - third(exx + eyy));
            tyy = (CT)2.0 * G * (eyy - third(exx + eyy));
          } else {
            split_cell_trial(exx, eyy, K, (CT)2.0 * G, (CT)r.P0, pp, txx, tyy);
          }
          if (m & M_CELL_ZERO) {
            pp = 0;
            txx = 0;
            tyy = 0;
          }
        }
        sP[b * S + a] = pp;
        sTxx[b * S + a] = txx;
        sTyy[b * S + a] = tyy;
      }
      {  // vertex (i0-2+a, j0-2+b) <-> box (a, b)
        const int vi = i0 - 2 + a, vj = j0 - 2 + b;
        CT txy = 0;
        if (!EDGE || (vi >= 0 && vi < nX - 1 && vj >= 0 && vj < nY - 1)) {
          const unsigned char m00 = sM[b * BX + a];
          const CT g = gt[(m00 >> M_VCLASS_SHIFT) & 15];
          const double du = sUx[(b + 1) * BX + a + 1] - sUx[b * BX + a + 1];
          const double dv = sUy[(b + 1) * BX + a + 1] - sUy[(b + 1) * BX + a];
          if constexpr (sizeof(CT) == 8) txy = g * (dy(du, r) + dx(dv, r));
          else txy = split_vertex_trial(g, (CT)du, (CT)r.rdY, (CT)dv, (CT)r.rdX);
          if (m00 & M_VERT_ZERO) txy = 0;
        }
        sTxy[b * S + a] = txy;
      }
    }
    lds_barrier();

    if constexpr (PLASTIC) {
      const CT coh = (CT)r.coh;
      CT c2lo;
      if constexpr (sizeof(CT) == 8) c2lo = r.coh2lo;
      else c2lo = r.coh2lof;
      // phase 2a: vertex (i0-1+a, j0-1+b) <-> trial vertex (a+1, b+1), trial cells (a..a+1, b..b+1)
      constexpr int K2 = (QY + NW - 1) / NW;
#pragma unroll
      for (int k2 = 0; k2 < K2; k2++) {
        const int b = wave + NW * k2;
        if (b >= QY) break;
        const int a = lane;
        if (a < S - 1) {
          const int vi = i0 - 1 + a, vj = j0 - 1 + b;
          CT txy = sTxy[(b + 1) * S + a + 1];
          if (!EDGE || (vi >= 0 && vi < nX - 1 && vj >= 0 && vj < nY - 1)) {
            const CT sxx = sTxx[b * S + a] + sTxx[b * S + a + 1] + sTxx[(b + 1) * S + a] + sTxx[(b + 1) * S + a + 1];
            const CT syy = sTyy[b * S + a] + sTyy[b * S + a + 1] + sTyy[(b + 1) * S + a] + sTyy[(b + 1) * S + a + 1];
            CT s2;
            if constexpr (sizeof(CT) == 8)
              s2 = (CT)0.0625 * sxx * sxx + (CT)0.0625 * syy * syy + (CT)2.0 * txy * txy;
            else
              s2 = split_s2_vertex(sxx, syy, txy);
            CT sc;
            if (rm_scale(s2, coh, c2lo, sc)) txy *= sc;
          }
          cTxy[b * S + a] = txy;
        }
      }
      lds_barrier();
      // phase 2b: cell (i0-1+a, j0-1+b) <-> trial cell (a, b), corrected in place
#pragma unroll
      for (int k2 = 0; k2 < K2; k2++) {
        const int b = wave + NW * k2;
        if (b >= QY) break;
        const int a = lane;
        if (a < S - 1) {
          const int ci = i0 - 1 + a, cj = j0 - 1 + b;
          if (!EDGE || (ci >= 0 && ci < nX && cj >= 0 && cj < nY)) {
            const int va = EDGE ? min(max(ci - 1, 0), nX - 3) - (i0 - 2) : a;
            const int vb = EDGE ? min(max(cj - 1, 0), nY - 3) - (j0 - 2) : b;
            const CT av = (CT)0.25 * (sTxy[vb * S + va] + sTxy[vb * S + va + 1] +
                                      sTxy[(vb + 1) * S + va] + sTxy[(vb + 1) * S + va + 1]);
            const CT txx = sTxx[b * S + a], tyy = sTyy[b * S + a];
            CT s2;
            if constexpr (sizeof(CT) == 8) s2 = txx * txx + tyy * tyy + (CT)2.0 * av * av;
            else s2 = split_s2_cell(txx, tyy, av);
            CT sc;
            if (rm_scale(s2, coh, c2lo, sc)) {
              sTxx[b * S + a] = txx * sc;
              sTyy[b * S + a] = tyy * sc;
            }
          }
        }
      }
      lds_barrier();
    }

    // ---- phase 3: corrected cell (i, j) <-> (a+1, b+1); corrected vertex (i-1, j-1) <-> (a, b) ----
    VT* __restrict__ Vx = static_cast<VT*>(r.Vx);
    VT* __restrict__ Vy = static_cast<VT*>(r.Vy);
    double* __restrict__ Uxo = r.Ux[cur ^ 1];
    double* __restrict__ Uyo = r.Uy[cur ^ 1];
    const CT vdampX = (CT)r.vdampX, vdampY = (CT)r.vdampY, dTrho = (CT)r.dTrho;
    auto Txy = [&](int x, int y) -> CT {
      if constexpr (PLASTIC) return cTxy[y * S + x];
      else return sTxy[(y + 1) * S + x + 1];
    };
#pragma unroll
    for (int k = 0; k < KP; k++) {
      const int b = wave + NW * k, a = lane;
      const int i = i0 + a, j = j0 + b;
      if (b >= TY || a >= TX) continue;
      if (EDGE && (i >= nX || j >= nY)) continue;
      CT vx = (CT)p.vxr[k];
      CT vy = (CT)p.vyr[k];
      if (!EDGE || (i > 0 && j > 0 && j < nY - 1)) {
        const CT fa = -sP[(b + 1) * S + a + 1] + sP[(b + 1) * S + a] + sTxx[(b + 1) * S + a + 1] - sTxx[(b + 1) * S + a];
        const CT fb = Txy(a, b + 1) - Txy(a, b);
        if constexpr (sizeof(CT) == 8) vx = vx * vdampX + dTrho * (dx(fa, r) + dy(fb, r));
        else vx = split_velocity(vx, vdampX, dTrho, fa, (CT)r.rdX, fb, (CT)r.rdY);
        gst(Vx + (size_t)j * sx, (unsigned)i, (VT)vx);
      }
      if (!EDGE || (i > 0 && i < nX - 1 && j > 0)) {
        const CT fa = -sP[(b + 1) * S + a + 1] + sP[b * S + a + 1] + sTyy[(b + 1) * S + a + 1] - sTyy[b * S + a + 1];
        const CT fb = Txy(a + 1, b) - Txy(a, b);
        if constexpr (sizeof(CT) == 8) vy = vy * vdampY + dTrho * (dy(fa, r) + dx(fb, r));
        else vy = split_velocity(vy, vdampY, dTrho, fa, (CT)r.rdY, fb, (CT)r.rdX);
        gst(Vy + (size_t)j * nX, (unsigned)i, (VT)vy);
      }
      if constexpr (sizeof(CT) == 8) {
        gst(Uxo + (size_t)j * sx, (unsigned)i, uxo[k] + (double)vx * r.dT);
        gst(Uyo + (size_t)j * nX, (unsigned)i, uyo[k] + (double)vy * r.dT);
      } else {
        gst(Uxo + (size_t)j * sx, (unsigned)i, split_disp(uxo[k], vx, r.dT));
        gst(Uyo + (size_t)j * nX, (unsigned)i, split_disp(uyo[k], vy, r.dT));
      }
      if (EDGE) {
        if (i == nX - 1) gst(Uxo, (unsigned)(j * sx + i + 1), uxe[k]);
        if (j == nY - 1) gst(Uyo, (unsigned)((j + 1) * nX + i), uye[k]);
      }
    }
  }
};

// Block -> tile map.  Workgroups are dispatched round-robin over the 8 XCDs
// (block b to XCD b % 8), so XCD x receives its blocks as a contiguous run of
// the tile list: x*q + min(x, r) + b/8 with q = nb/8, r = nb%8 (the first r
// XCDs take one more).  A tile and its neighbours, which share halo rows,
// then mostly run on the same XCD and share its L2.
__device__ __forceinline__ int tile_of_block(int b, int nb) {
  const int x = b & 7, q = nb >> 3, r = nb & 7;
  return x * q + min(x, r) + (b >> 3);
}

template <typename CT, typename VT, bool PLASTIC, int TY, int NW>
__global__ __launch_bounds__(64 * NW) void k_fused(RveDev r, int cur, int ntx, int ntiles) {
  using F = Fused<CT, VT, PLASTIC, TY, NW>;
  __shared__ __attribute__((aligned(16))) char smem[F::SMEM];
  const int t = tile_of_block(blockIdx.x, ntiles);
  const int i0 = (t % ntx) * F::TX, j0 = (t / ntx) * TY;
  if (i0 >= 2 && i0 + F::TX + 2 <= r.nX && j0 >= 2 && j0 + TY + 2 <= r.nY)
    F::template run<false>(r, cur, i0, j0, smem);
  else
    F::template run<true>(r, cur, i0, j0, smem);
}

// ---------------------------------------------------------------------------
// Residual (EffPlast2D.cu:407-412, FindMaxAbs :805-813): max |Vx|, max |Vy|
// and a count of non-finite values.  Wave64 shuffle reduction, then one
// 64-bit atomicMax per block on the bit pattern (monotone for values >= 0).
// NaN is skipped as in FindMaxAbs, and counted.
template <typename VT, bool VEC> struct Vec16 { using T = VT; static constexpr int N = 1; };
template <> struct Vec16<float, true> { using T = float4; static constexpr int N = 4; };
template <> struct Vec16<double, true> { using T = double2; static constexpr int N = 2; };

__device__ __forceinline__ void maxabs_acc(double a, double& m, unsigned long long& bad) {
  const double v = fabs(a);
  if (v > m) m = v;                               // NaN never replaces the max
  if (!(v <= 1.7976931348623157e308)) bad++;      // NaN and infinity
}

template <typename VT, int N, typename V>
__device__ __forceinline__ void maxabs_vec(const V& v, double& m, unsigned long long& bad) {
  const VT* e = reinterpret_cast<const VT*>(&v);
#pragma unroll
  for (int q = 0; q < N; q++) maxabs_acc((double)e[q], m, bad);
}

// One block's share of the reduction: blocks b of nb (256 threads) stride
// over both arrays in one loop, so each thread has loads of both in flight:
// as 16-byte vectors when both start 16-byte aligned (VEC), then the few
// elements past the last whole vector.
template <typename VT, bool VEC>
__device__ __forceinline__ void maxabs_block(const VT* __restrict__ Vx, size_t nx,
                                             const VT* __restrict__ Vy, size_t ny,
                                             unsigned long long* out, unsigned b, unsigned nb) {
  using V = typename Vec16<VT, VEC>::T;
  constexpr int N = Vec16<VT, VEC>::N;
  double mx = 0.0, my = 0.0;
  unsigned long long bad = 0;
  const size_t stride = (size_t)nb * blockDim.x;
  const size_t t0 = b * (size_t)blockDim.x + threadIdx.x;
  const size_t nxv = nx / N, nyv = ny / N, nv = nxv > nyv ? nxv : nyv;
  const V* __restrict__ X = reinterpret_cast<const V*>(Vx);
  const V* __restrict__ Y = reinterpret_cast<const V*>(Vy);
#pragma unroll 2
  for (size_t k = t0; k < nv; k += stride) {
    if (k < nxv) maxabs_vec<VT, N>(X[k], mx, bad);
    if (k < nyv) maxabs_vec<VT, N>(Y[k], my, bad);
  }
  for (size_t k = nxv * N + t0; k < nx; k += stride) maxabs_acc((double)Vx[k], mx, bad);
  for (size_t k = nyv * N + t0; k < ny; k += stride) maxabs_acc((double)Vy[k], my, bad);
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

template <typename VT, bool VEC>
__global__ __launch_bounds__(256) void k_maxabs(const VT* __restrict__ Vx, size_t nx,
                                                const VT* __restrict__ Vy, size_t ny,
                                                unsigned long long* out) {
  maxabs_block<VT, VEC>(Vx, nx, Vy, ny, out, blockIdx.x, gridDim.x);
}

// The residuals of many RVEs in one launch: blocks (b, k) reduce job k's
// velocity fields into out[3k .. 3k+2] (gridDim.x blocks per job).
template <typename VT, bool VEC>
__global__ __launch_bounds__(256) void k_maxabs_multi(const MaxabsJob* __restrict__ jobs,
                                                      unsigned long long* out) {
  const MaxabsJob jb = jobs[blockIdx.y];
  maxabs_block<VT, VEC>(static_cast<const VT*>(jb.vx), (size_t)jb.nx,
                        static_cast<const VT*>(jb.vy), (size_t)jb.ny, out + 3 * (size_t)blockIdx.y,
                        blockIdx.x, gridDim.x);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) p[k] = 0ull;
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
  const size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[k];
}

inline dim3 grid2(int nx, int ny, dim3 blk) {
  return dim3((nx + blk.x - 1) / blk.x, (ny + blk.y - 1) / blk.y, 1);
}

// Many RVEs in one launch: the tile list names the RVE, the tile origin and
// the RVE's ping-pong offset; the RVE's parameters come from a device table.
template <typename CT, typename VT, int TY, int NW>
__global__ __launch_bounds__(64 * NW) void k_fused_batch(const RveDev* __restrict__ rves,
                                                         const int4* __restrict__ tiles,
                                                         int ntiles, int parity) {
  using FP = Fused<CT, VT, true, TY, NW>;
  using FE = Fused<CT, VT, false, TY, NW>;
  static_assert(FP::SMEM == FE::SMEM, "same LDS layout");
  __shared__ __attribute__((aligned(16))) char smem[FP::SMEM];
  const int4 t = tiles[tile_of_block(blockIdx.x, ntiles)];
  const RveDev& r = rves[t.x];
  const int cur = parity ^ r.cur;
  const int i0 = t.y, j0 = t.z;
  const bool interior = i0 >= 2 && i0 + FP::TX + 2 <= r.nX && j0 >= 2 && j0 + TY + 2 <= r.nY;
  if (r.plastic) {
    if (interior) FP::template run<false>(r, cur, i0, j0, smem);
    else FP::template run<true>(r, cur, i0, j0, smem);
  } else {
    if (interior) FE::template run<false>(r, cur, i0, j0, smem);
    else FE::template run<true>(r, cur, i0, j0, smem);
  }
}

// Persistent form of the batched kernel: a grid of about one wave of
// resident blocks, each walking a run of the tile list and fetching the next
// tile's U box, mask and velocities into registers while it computes the
// current one, so the global-read latency hides behind arithmetic.  Blocks
// of XCD x (blockIdx % 8) share the x-th eighth of the list (neighbouring
// tiles, shared halos in that XCD's L2).  Two register sets alternate, so no
// copy has to wait for a load.
template <typename CT, typename VT, int TY, int NW>
__global__ __launch_bounds__(64 * NW) void k_fused_persist(const RveDev* __restrict__ rves,
                                                          const int4* __restrict__ tiles,
                                                          int ntiles, int parity) {
  using FP = Fused<CT, VT, true, TY, NW>;
  using FE = Fused<CT, VT, false, TY, NW>;
  using Pre = typename FP::Pre;
  static_assert(FP::SMEM == FE::SMEM, "same LDS layout");
  __shared__ __attribute__((aligned(16))) char smem[FP::SMEM];
  const int nb = gridDim.x, b = blockIdx.x;
  const int x = b & 7, nbx = (nb >> 3) + ((b & 7) < (nb & 7) ? 1 : 0);
  const int lo = (int)((long)ntiles * x / 8), hi = (int)((long)ntiles * (x + 1) / 8);
  int t = lo + (b >> 3);
  if (t >= hi) return;
  auto fetch = [&](int tt, Pre& p) {
    const int4 q = tiles[tt];
    const RveDev& r = rves[q.x];
    const int cur = parity ^ r.cur;
    const bool interior = q.y >= 2 && q.y + FP::TX + 2 <= r.nX && q.z >= 2 && q.z + TY + 2 <= r.nY;
    if (interior) FP::template fetch<false>(r, cur, q.y, q.z, p);
    else FP::template fetch<true>(r, cur, q.y, q.z, p);
  };
  auto compute = [&](int tt, const Pre& p) {
    const int4 q = tiles[tt];
    const RveDev& r = rves[q.x];
    const int cur = parity ^ r.cur;
    const bool interior = q.y >= 2 && q.y + FP::TX + 2 <= r.nX && q.z >= 2 && q.z + TY + 2 <= r.nY;
    if (r.plastic) {
      if (interior) FP::template compute<false>(r, cur, q.y, q.z, smem, p);
      else FP::template compute<true>(r, cur, q.y, q.z, smem, p);
    } else {
      if (interior) FE::template compute<false>(r, cur, q.y, q.z, smem, p);
      else FE::template compute<true>(r, cur, q.y, q.z, smem, p);
    }
    lds_barrier();  // every wave is done with this tile's LDS
  };
  Pre A, B;
  fetch(t, A);
  for (;;) {
    const int t1 = t + nbx;
    if (t1 < hi) fetch(t1, B);
    compute(t, A);
    if (t1 >= hi) break;
    const int t2 = t1 + nbx;
    if (t2 < hi) fetch(t2, A);
    compute(t1, B);
    if (t2 >= hi) break;
    t = t2;
  }
}

// Tile heights and their workgroup sizes (waves per block).
template <int TY> struct TileNW;
template <> struct TileNW<6> { static constexpr int value = 8; };
template <> struct TileNW<8> { static constexpr int value = 4; };
template <> struct TileNW<14> { static constexpr int value = 8; };
template <> struct TileNW<30> { static constexpr int value = 16; };

template <typename CT, typename VT, bool P, int TY>
struct LaunchFn {
  static constexpr int NW = TileNW<TY>::value;
  static int call(const RveDev& r, int cur, const FusedPlan& pl, hipStream_t s, hipEvent_t e0,
                  hipEvent_t e1) {
    if (e0)
      hipExtLaunchKernelGGL((k_fused<CT, VT, P, TY, NW>), dim3(pl.grid), dim3(64 * NW), 0, s, e0,
                            e1, 0, r, cur, pl.ntx, pl.ntiles);
    else
      hipLaunchKernelGGL((k_fused<CT, VT, P, TY, NW>), dim3(pl.grid), dim3(64 * NW), 0, s, r,
                         cur, pl.ntx, pl.ntiles);
    return 0;
  }
};

template <typename CT, typename VT, bool P, int TY>
struct BatchFn {
  static constexpr int NW = TileNW<TY>::value;
  static int call(const RveDev* rves, const int4* tiles, int ntiles, int parity, hipStream_t s,
                  hipEvent_t e0, hipEvent_t e1) {
    if (e0)
      hipExtLaunchKernelGGL((k_fused_batch<CT, VT, TY, NW>), dim3(ntiles), dim3(64 * NW), 0, s,
                            e0, e1, 0, rves, tiles, ntiles, parity);
    else
      hipLaunchKernelGGL((k_fused_batch<CT, VT, TY, NW>), dim3(ntiles), dim3(64 * NW), 0, s,
                         rves, tiles, ntiles, parity);
    return 0;
  }
};

template <typename CT, typename VT, bool P, int TY>
struct PersistFn {
  static constexpr int NW = TileNW<TY>::value;
  static int call(const RveDev* rves, const int4* tiles, int ntiles, int parity, hipStream_t s,
                  hipEvent_t e0, hipEvent_t e1) {
    static int grid = 0;  // resident blocks on this device: CUs x blocks per CU, a multiple of 8
    if (grid == 0) {
      int nb = 0, dev = 0, cus = 0;
      EP2D_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &nb, (const void*)k_fused_persist<CT, VT, TY, NW>, 64 * NW, 0));
      EP2D_HIP(hipGetDevice(&dev));
      EP2D_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      grid = std::max(8, (std::max(nb, 1) * cus) / 8 * 8);
    }
    const int g = std::min(grid, (ntiles + 7) / 8 * 8);
    if (e0)
      hipExtLaunchKernelGGL((k_fused_persist<CT, VT, TY, NW>), dim3(g), dim3(64 * NW), 0, s, e0,
                            e1, 0, rves, tiles, ntiles, parity);
    else
      hipLaunchKernelGGL((k_fused_persist<CT, VT, TY, NW>), dim3(g), dim3(64 * NW), 0, s, rves,
                         tiles, ntiles, parity);
    return 0;
  }
};

template <typename CT, typename VT, bool P, int TY>
struct OccFn {
  static constexpr int NW = TileNW<TY>::value;
  static int call() {
    int nb = 0;
    EP2D_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, (const void*)k_fused<CT, VT, P, TY, NW>, 64 * NW, 0));
    return nb;
  }
};

template <typename CT, typename VT, bool P, int TY>
struct SmemFn {
  static int call() { return (int)Fused<CT, VT, P, TY, TileNW<TY>::value>::SMEM; }
};

template <typename CT, typename VT, bool P, int TY>
struct NwFn {
  static int call() { return TileNW<TY>::value; }
};

// Dispatch over scheme (double / float arithmetic), plasticity and tile height.
template <template <typename, typename, bool, int> class Fn, typename... A>
int dispatch_fused(int mode, bool plastic, int ty, A&&... a) {
#define EP2D_TY_CASES(CT, VT, P)                                    \
  switch (ty) {                                                     \
    case 6: return Fn<CT, VT, P, 6>::call(std::forward<A>(a)...);   \
    case 14: return Fn<CT, VT, P, 14>::call(std::forward<A>(a)...); \
    case 30: return Fn<CT, VT, P, 30>::call(std::forward<A>(a)...); \
    default: return Fn<CT, VT, P, 8>::call(std::forward<A>(a)...);  \
  }
  if (mode == 2) {
    if (plastic) { EP2D_TY_CASES(float, float, true) } else { EP2D_TY_CASES(float, float, false) }
  } else if (mode == kModeSplitStress64) {
    // the split scheme's FP32 velocity storage with the FP64 arithmetic
    // (tiles of 8 rows only)
    if (plastic) return Fn<double, float, true, 8>::call(std::forward<A>(a)...);
    return Fn<double, float, false, 8>::call(std::forward<A>(a)...);
  } else {
    if (plastic) { EP2D_TY_CASES(double, double, true) } else { EP2D_TY_CASES(double, double, false) }
  }
#undef EP2D_TY_CASES
}

// ---------------------------------------------------------------------------
// Row-marching fused iteration.  One wave64 owns a strip of 64 columns
// (lane L <-> column i = s + L) and walks down rows j0-2 .. j1-1 of one RVE.
// Step j loads row j+1 of Ux and the mask, row j+2 of Uy and row j of V, and
// produces, all in registers:
//
//   trial cell row j+1 and trial vertex row j      (ComputeStress, :37-98)
//   return-mapped vertex row j and cell row j       (ComputeJ2 / Plasticity,
//                                                     :100-177)
//   V and U of row j, lanes 2..61                   (ComputeDisp, :3-35)
//
// Neighbours in x come from the adjacent lanes through DPP wave shifts;
// neighbours in y are the rows carried from the previous steps, so nothing
// is staged in LDS and no barrier is needed.  The loads of step j+1 are
// issued before step j computes.  Steps j0-2 and j0-1 only fill the carried
// rows (the halo of the segment); outputs start at row j0.
//
// Every expression is the fused tile kernel's (and so the reference's) in the
// same order: the two kernels give identical results.  EDGE segments (strips
// or rows touching the RVE boundary) keep the boundary conditions, zero the
// stress outside the grid and use the clamped vertex block of the reference's
// ComputeJ2 edge branches for cells on the boundary.

// value of lane L+1 (0 in lane 63) and of lane L-1 (0 in lane 0)
__device__ __forceinline__ int lane_next_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xf, 0xf, true);  // wave_shl:1, bound_ctrl:0
}
__device__ __forceinline__ int lane_prev_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xf, 0xf, true);  // wave_shr:1, bound_ctrl:0
}
__device__ __forceinline__ float lane_next(float v) {
  return __builtin_bit_cast(float, lane_next_i(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float lane_prev(float v) {
  return __builtin_bit_cast(float, lane_prev_i(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double lane_next(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)lane_next_i((int)(unsigned)b);
  const unsigned hi = (unsigned)lane_next_i((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double lane_prev(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)lane_prev_i((int)(unsigned)b);
  const unsigned hi = (unsigned)lane_prev_i((int)(unsigned)(b >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

template <typename CT, typename VT>
struct March {
  static __device__ __forceinline__ CT dx(double a, const RveDev& r) {
    if constexpr (sizeof(CT) == 8) return qdiv(a, r.dX, r.rdX);
    else return (CT)a * (CT)r.rdX;
  }
  static __device__ __forceinline__ CT dy(double a, const RveDev& r) {
    if constexpr (sizeof(CT) == 8) return qdiv(a, r.dY, r.rdY);
    else return (CT)a * (CT)r.rdY;
  }
  static __device__ __forceinline__ CT third(CT a) {
    if constexpr (sizeof(CT) == 8) return qdiv(a, 3.0, 1.0 / 3.0);
    else return a * (CT)(1.0 / 3.0);
  }

  struct Ld {
    double ux, uy;   // Ux row j+1, Uy row j+2
    unsigned m;      // mask row j+1
    VT vx, vy;       // V row j
  };

  template <bool EDGE>
  static __device__ __forceinline__ Ld load(const RveDev& r, const double* __restrict__ Ux,
                                            const double* __restrict__ Uy, const VT* Vx,
                                            const VT* Vy, int i, int j) {
    const int nX = r.nX, nY = r.nY, sx = nX + 1;
    Ld d;
    if (!EDGE) {
      // row bases are uniform (scalar arithmetic), the lane offset i is fixed
      d.ux = gld(Ux + (size_t)(j + 1) * sx, (unsigned)i);
      d.uy = gld(Uy + (size_t)(j + 2) * nX, (unsigned)i);
      d.m = gld(r.mask + (size_t)(j + 1) * nX, (unsigned)i);
      d.vx = gld(Vx + (size_t)j * sx, (unsigned)i);
      d.vy = gld(Vy + (size_t)j * nX, (unsigned)i);
    } else {
      const bool ci = i >= 0 && i < nX;
      d.ux = (i >= 0 && i <= nX && j + 1 >= 0 && j + 1 < nY) ? gld(Ux, (unsigned)((j + 1) * sx + i)) : 0.0;
      d.uy = (ci && j + 2 >= 0 && j + 2 <= nY) ? gld(Uy, (unsigned)((j + 2) * nX + i)) : 0.0;
      d.m = (ci && j + 1 >= 0 && j + 1 < nY) ? (unsigned)gld(r.mask, (unsigned)((j + 1) * nX + i)) : 0u;
      const bool vr = ci && j >= 0 && j < nY;
      d.vx = vr ? gld(Vx, (unsigned)(j * sx + i)) : (VT)0;
      d.vy = vr ? gld(Vy, (unsigned)(j * nX + i)) : (VT)0;
    }
    return d;
  }

  // Return mapping of a vertex from its trial shear stress and the four trial
  // cells around it, summed in the reference's order.
  static __device__ __forceinline__ CT vertex_rm(CT txy, CT sxx, CT syy, CT coh, CT c2lo) {
    CT s2;
    if constexpr (sizeof(CT) == 8)
      s2 = (CT)0.0625 * sxx * sxx + (CT)0.0625 * syy * syy + (CT)2.0 * txy * txy;
    else
      s2 = split_s2_vertex(sxx, syy, txy);
    CT sc;
    if (rm_scale(s2, coh, c2lo, sc)) txy *= sc;
    return txy;
  }
  static __device__ __forceinline__ void cell_rm(CT& txx, CT& tyy, CT av, CT coh, CT c2lo) {
    CT s2;
    if constexpr (sizeof(CT) == 8) s2 = txx * txx + tyy * tyy + (CT)2.0 * av * av;
    else s2 = split_s2_cell(txx, tyy, av);
    CT sc;
    if (rm_scale(s2, coh, c2lo, sc)) {
      txx = txx * sc;
      tyy = tyy * sc;
    }
  }

  template <bool EDGE>
  static __device__ void run(const RveDev& r, int cur, int vpar, int s, int j0, int j1,
                             const CT* gt) {
    const int nX = r.nX, nY = r.nY, sx = nX + 1;
    const int L = threadIdx.x & 63;
    const int i = s + L;
    const double* __restrict__ Uxi = r.Ux[cur];
    const double* __restrict__ Uyi = r.Uy[cur];
    double* __restrict__ Uxo = r.Ux[cur ^ 1];
    double* __restrict__ Uyo = r.Uy[cur ^ 1];
    // V updated in place in the current buffer (Vx2/Vy2 when vpar)
    VT* Vx = static_cast<VT*>(vpar ? r.Vx2 : r.Vx);
    VT* Vy = static_cast<VT*>(vpar ? r.Vy2 : r.Vy);
    const bool plastic = r.plastic != 0;
    const CT coh = (CT)r.coh;
    CT c2lo;
    if constexpr (sizeof(CT) == 8) c2lo = r.coh2lo;
    else c2lo = r.coh2lof;
    const CT K0 = (CT)r.K0, Kp = (CT)r.Kp, G0 = (CT)r.G0, Gp = (CT)r.Gp, P0 = (CT)r.P0;
    const CT vdampX = (CT)r.vdampX, vdampY = (CT)r.vdampY, dTrho = (CT)r.dTrho;
    const double dT = r.dT;
    const bool outLane = L >= 2 && L < 62 && (!EDGE || i < nX);
    const bool colC = !EDGE || (i >= 0 && i < nX);
    const bool colV = !EDGE || (i >= 0 && i < nX - 1);
    const bool leftEdge = EDGE && i == 0, rightEdge = EDGE && i == nX - 1;

    // carried rows (see the step comments below)
    const int jb = j0 - 2;
    double ux_c = 0.0, uy_c = 0.0, uy_n = 0.0;
    unsigned m_c = 0;
    if (!EDGE) {
      ux_c = gld(Uxi, (unsigned)(jb * sx + i));
      uy_n = gld(Uyi, (unsigned)((jb + 1) * nX + i));
      m_c = gld(r.mask, (unsigned)(jb * nX + i));
    } else {
      if (i >= 0 && i <= nX && jb >= 0 && jb < nY) ux_c = gld(Uxi, (unsigned)(jb * sx + i));
      if (colC && jb + 1 >= 0 && jb + 1 <= nY) uy_n = gld(Uyi, (unsigned)((jb + 1) * nX + i));
      if (colC && jb >= 0 && jb < nY) m_c = gld(r.mask, (unsigned)(jb * nX + i));
    }
    double uxs_c = lane_next(ux_c);
    CT tP_c = 0, tXX_c = 0, tYY_c = 0, sXX_c = 0, sYY_c = 0;   // trial cell row j
    CT tP_p = 0, tXX_p = 0, tYY_p = 0;                         // trial cell row j-1 (EDGE)
    CT tXY_p = 0, tXY_pp = 0;                                  // trial vertex rows j-1, j-2
    CT cP_p = 0, cYY_p = 0, cXY_p = 0;                         // corrected row j-1

    // the loads are unconditional (past the last step they re-read rows in
    // range: interior segments end at least one row above nY-1, edge loads
    // are guarded), so no branch joins in front of a wait
    auto fetch = [&](int j) {
      return load<EDGE>(r, Uxi, Uyi, Vx, Vy, i, EDGE ? j : min(j, j1 - 1));
    };
    auto step = [&](const int j, const Ld& ld) {

      // trial cell row j+1 (cell i, row j+1)
      const double ux_n = ld.ux, uy_nn = ld.uy;
      const unsigned m_n = ld.m;
      const double uxs_n = lane_next(ux_n);
      CT tP_n = 0, tXX_n = 0, tYY_n = 0;
      {
        const CT K = (m_n & M_CELL_MAT) ? Kp : K0;
        const CT G = (m_n & M_CELL_MAT) ? Gp : G0;
        const CT exx = dx(uxs_n - ux_n, r);
        const CT eyy = dy(uy_nn - uy_n, r);
        CT pp, txx, tyy;
        if constexpr (sizeof(CT) == 8) {
          pp = P0 - K * (exx + eyy);
          txx = (CT)2.0 * G * (exx - third(exx + eyy));
          tyy = (CT)2.0 * G * (eyy - third(exx + eyy));
        } else {
          split_cell_trial(exx, eyy, K, (CT)2.0 * G, P0, pp, txx, tyy);
        }
        const bool ok = !EDGE || (colC && j + 1 >= 0 && j + 1 < nY);
        if ((m_n & M_CELL_ZERO) || !ok) {
          pp = 0;
          txx = 0;
          tyy = 0;
        }
        tP_n = pp;
        tXX_n = txx;
        tYY_n = tyy;
      }
      const CT sXX_n = lane_next(tXX_n), sYY_n = lane_next(tYY_n);

      // trial vertex row j (vertex i, row j)
      const double uys_n = lane_next(uy_n);
      CT tXY_c;
      {
        const CT g = gt[(m_c >> M_VCLASS_SHIFT) & 15];
        CT txy;
        if constexpr (sizeof(CT) == 8) txy = g * (dy(uxs_n - uxs_c, r) + dx(uys_n - uy_n, r));
        else txy = split_vertex_trial(g, (CT)(uxs_n - uxs_c), (CT)r.rdY, (CT)(uys_n - uy_n),
                                      (CT)r.rdX);
        const bool ok = !EDGE || (colV && j >= 0 && j < nY - 1);
        if ((m_c & M_VERT_ZERO) || !ok) txy = 0;
        tXY_c = txy;
      }

      // return mapping: vertex row j, cell row j (and, on the top boundary,
      // cell row 0 once vertex row 1 exists)
      CT cXY_c = tXY_c, cP_c = tP_c, cXX_c = tXX_c, cYY_c = tYY_c;
      if (plastic) {
        const bool vok = !EDGE || (colV && j >= 0 && j < nY - 1);
        const CT sxx = tXX_c + sXX_c + tXX_n + sXX_n;
        const CT syy = tYY_c + sYY_c + tYY_n + sYY_n;
        if (vok) cXY_c = vertex_rm(tXY_c, sxx, syy, coh, c2lo);

        const CT pr_p = lane_prev(tXY_p), pr_c = lane_prev(tXY_c);
        if (!EDGE) {
          const CT av = (CT)0.25 * (pr_p + tXY_p + pr_c + tXY_c);
          cell_rm(cXX_c, cYY_c, av, coh, c2lo);
        } else {
          // vertex pair of this lane's cell in a vertex row: (first, second)
          const CT nx_p = lane_next(tXY_p), nx_c = lane_next(tXY_c), nx_pp = lane_next(tXY_pp);
          const CT pr_pp = lane_prev(tXY_pp);
          const CT pp_p = lane_prev(pr_p), pp_c = lane_prev(pr_c), pp_pp = lane_prev(pr_pp);
          auto first = [&](CT t, CT pr, CT pp) { return leftEdge ? t : (rightEdge ? pp : pr); };
          auto second = [&](CT t, CT pr, CT nxv) { return leftEdge ? nxv : (rightEdge ? pr : t); };
          if (colC && j >= 1 && j < nY - 1) {
            const CT av = (CT)0.25 * (first(tXY_p, pr_p, pp_p) + second(tXY_p, pr_p, nx_p) +
                                      first(tXY_c, pr_c, pp_c) + second(tXY_c, pr_c, nx_c));
            cell_rm(cXX_c, cYY_c, av, coh, c2lo);
          } else if (colC && j == nY - 1 && j >= 1) {  // rows nY-3, nY-2
            const CT av = (CT)0.25 * (first(tXY_pp, pr_pp, pp_pp) + second(tXY_pp, pr_pp, nx_pp) +
                                      first(tXY_p, pr_p, pp_p) + second(tXY_p, pr_p, nx_p));
            cell_rm(cXX_c, cYY_c, av, coh, c2lo);
          }
          if (j == 1) {  // cell row 0 from vertex rows 0 and 1
            CT xx = tXX_p, yy = tYY_p;
            if (colC) {
              const CT av = (CT)0.25 * (first(tXY_p, pr_p, pp_p) + second(tXY_p, pr_p, nx_p) +
                                        first(tXY_c, pr_c, pp_c) + second(tXY_c, pr_c, nx_c));
              cell_rm(xx, yy, av, coh, c2lo);
            }
            cP_p = tP_p;
            cYY_p = yy;
          }
        }
      }

      // velocity and displacement of row j
      const CT rP = lane_prev(cP_c), rXX = lane_prev(cXX_c);
      const CT rXY_c = lane_prev(cXY_c), rXY_p = lane_prev(cXY_p);
      if (j >= j0) {
        CT vx = (CT)ld.vx, vy = (CT)ld.vy;
        if (!EDGE || (i > 0 && j > 0 && j < nY - 1)) {
          if constexpr (sizeof(CT) == 8)
            vx = vx * vdampX + dTrho * (dx(-cP_c + rP + cXX_c - rXX, r) + dy(rXY_c - rXY_p, r));
          else
            vx = split_velocity(vx, vdampX, dTrho, -cP_c + rP + cXX_c - rXX, (CT)r.rdX,
                                rXY_c - rXY_p, (CT)r.rdY);
          if (outLane) gst(Vx + (size_t)j * sx, (unsigned)i, (VT)vx);
        }
        if (!EDGE || (i > 0 && i < nX - 1 && j > 0)) {
          if constexpr (sizeof(CT) == 8)
            vy = vy * vdampY + dTrho * (dy(-cP_c + cP_p + cYY_c - cYY_p, r) + dx(cXY_p - rXY_p, r));
          else
            vy = split_velocity(vy, vdampY, dTrho, -cP_c + cP_p + cYY_c - cYY_p, (CT)r.rdY,
                                cXY_p - rXY_p, (CT)r.rdX);
          if (outLane) gst(Vy + (size_t)j * nX, (unsigned)i, (VT)vy);
        }
        if (outLane && (!EDGE || j < nY)) {
          if constexpr (sizeof(CT) == 8) {
            gst(Uxo + (size_t)j * sx, (unsigned)i, ux_c + (double)vx * dT);
            gst(Uyo + (size_t)j * nX, (unsigned)i, uy_c + (double)vy * dT);
          } else {
            gst(Uxo + (size_t)j * sx, (unsigned)i, split_disp(ux_c, vx, dT));
            gst(Uyo + (size_t)j * nX, (unsigned)i, split_disp(uy_c, vy, dT));
          }
          if (EDGE) {
            if (i == nX - 1) gst(Uxo, (unsigned)(j * sx + i + 1), uxs_c);
            if (j == nY - 1) gst(Uyo, (unsigned)((j + 1) * nX + i), uy_n);
          }
        }
      }

      // rotate
      ux_c = ux_n;
      uxs_c = uxs_n;
      uy_c = uy_n;
      uy_n = uy_nn;
      m_c = m_n;
      tP_p = tP_c;
      tXX_p = tXX_c;
      tYY_p = tYY_c;
      tP_c = tP_n;
      tXX_c = tXX_n;
      tYY_c = tYY_n;
      sXX_c = sXX_n;
      sYY_c = sYY_n;
      tXY_pp = tXY_p;
      tXY_p = tXY_c;
      cP_p = cP_c;
      cYY_p = cYY_c;
      cXY_p = cXY_c;
    };
    // loads run one step ahead of the arithmetic: the wait for a step's rows
    // leaves the next step's loads in flight
    // (two steps per trip, each buffer refilled right after the step that
    // reads it, so no register copy has to wait for a load)
    Ld A = fetch(jb);
    for (int j = jb; j < j1; j += 2) {
      const Ld B = fetch(j + 1);
      step(j, A);
      if (j + 1 >= j1) break;
      A = fetch(j + 2);
      step(j + 1, B);
    }
  }
};

// One wave per segment (rve, s, j0, j1); four independent waves per block.
// Blocks are dealt to the XCDs in contiguous runs of the segment list, which
// is ordered RVE by RVE and band by band, so neighbouring segments (sharing
// halo columns) run together and meet in the same L2.
template <typename CT, typename VT>
__global__ __launch_bounds__(256) void k_march(const RveDev* __restrict__ rves,
                                               const int4* __restrict__ segs, int nsegs,
                                               int parity, int vpar) {
  __shared__ CT gt[4][16];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = tile_of_block(blockIdx.x, gridDim.x);
  const int k = b * 4 + wave;
  if (k >= nsegs) return;
  const int4 sg = segs[k];
  const RveDev& r = rves[sg.x];
  if (lane < 16) gt[wave][lane] = (CT)r.gtab[lane];
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's table is in LDS
  __builtin_amdgcn_wave_barrier();
  const int cur = parity ^ r.cur;
  const int s = sg.y, j0 = sg.z, j1 = sg.w;
  const bool interior = s >= 1 && s + 63 <= r.nX - 2 && j0 >= 2 && j1 <= r.nY - 1;
  if (interior) March<CT, VT>::template run<false>(r, cur, vpar, s, j0, j1, gt[wave]);
  else March<CT, VT>::template run<true>(r, cur, vpar, s, j0, j1, gt[wave]);
}

// ---------------------------------------------------------------------------
// Two iterations per launch (temporal blocking of the row-marching kernel).
// One iteration is written as a row stream (MarchStage): step j takes row j+1
// of Ux and the mask, row j+2 of Uy and row j of V of its input state and
// gives row j of its output state -- V, and U at every node of the row
// (updated where ComputeDisp updates it, unchanged where it does not, zero
// outside the grid) -- so a second stage can take it as input.  In k_march2
// stage A reads the state from HBM and stage B, two rows behind, reads A's
// rows from registers; only B's rows are written.  U and V are read from one
// buffer and written to the other (a neighbouring segment's halo needs the
// old velocities).  Valid outputs shrink by two lanes per stage: lanes 4..59,
// strips of 56 columns.  The arithmetic of each stage is the single-iteration
// kernels', so two stages give the same bits as two launches.
//
// EX stages run on strips that touch the left or right boundary (per-column
// tests), EY stages on segments whose rows reach the top or bottom boundary
// (per-row tests, uniform over the wave); EDGE is either.  CLEAN stages
// (segments whose mask bytes are all 0, chain kernels) read no mask: matrix
// K and G, no pore zeroing and the vertex modulus of four matrix cells, which
// is what the mask path computes from zero mask bytes.
template <typename CT, typename VT, bool EX, bool EY = EX, bool CLEAN = false>
struct MarchStage {
  static constexpr bool EDGE = EX || EY;
  using M = March<CT, VT>;
  const RveDev* r;
  const CT* gt;
  int nX, nY, i;
  bool colC, colV, leftEdge, rightEdge, inX, plastic;
  CT coh, c2lo, K0, Kp, G2_0, G2_p, P0, vdampX, vdampY, dTrho, g0;
  double dT;
  // grid steps held in registers: the stores of a step may alias the RVE
  // table, so values read from it would be read again after every store
  double dXd, dYd, rdXd, rdYd;
  CT rdXc, rdYc;
  // carried rows: U rows j and j+1, mask row j, trial cells j (and j-1),
  // trial vertices j-1 and j-2, corrected row j-1
  double ux_c, uxs_c, uy_c, uy_n;
  unsigned m_c;
  CT tP_c, tXX_c, tYY_c, sXX_c, sYY_c, tP_p, tXX_p, tYY_p, tXY_p, tXY_pp, cP_p, cYY_p, cXY_p;

  struct Out {
    double ux, uy;   // output U of row j at this lane's column
    CT vx, vy;       // output V of row j (rounded to the storage type)
    bool updX, updY; // V was updated here (else it is the input)
  };

  __device__ __forceinline__ void setup(const RveDev& rr, const CT* g, int col) {
    r = &rr;
    gt = g;
    nX = rr.nX;
    nY = rr.nY;
    i = col;
    colC = !EX || (i >= 0 && i < nX);
    colV = !EX || (i >= 0 && i < nX - 1);
    inX = colC;
    leftEdge = EX && i == 0;
    rightEdge = EX && i == nX - 1;
    plastic = rr.plastic != 0;
    coh = (CT)rr.coh;
    if constexpr (sizeof(CT) == 8) c2lo = rr.coh2lo;
    else c2lo = rr.coh2lof;
    K0 = (CT)rr.K0;
    Kp = (CT)rr.Kp;
    G2_0 = (CT)2.0 * (CT)rr.G0;  // 2G, exact
    g0 = (CT)rr.gtab[0];
    G2_p = (CT)2.0 * (CT)rr.Gp;
    P0 = (CT)rr.P0;
    dXd = rr.dX;
    dYd = rr.dY;
    rdXd = rr.rdX;
    rdYd = rr.rdY;
    rdXc = (CT)rr.rdX;
    rdYc = (CT)rr.rdY;
    vdampX = (CT)rr.vdampX;
    vdampY = (CT)rr.vdampY;
    dTrho = (CT)rr.dTrho;
    dT = rr.dT;
  }

  // difference quotients: March::dx / March::dy with the steps in registers
  __device__ __forceinline__ CT dxs(double a) const {
    if constexpr (sizeof(CT) == 8) return qdiv(a, dXd, rdXd);
    else return (CT)a * rdXc;
  }
  __device__ __forceinline__ CT dys(double a) const {
    if constexpr (sizeof(CT) == 8) return qdiv(a, dYd, rdYd);
    else return (CT)a * rdYc;
  }

  // Ux row jb, Uy row jb+1, mask row jb (jb = first step)
  __device__ __forceinline__ void start(double ux, double uy1, unsigned m) {
    startS(ux, lane_next(ux), uy1, m);
  }
  // the same with Ux(i+1, jb) given (read from a ring row instead of shifted)
  __device__ __forceinline__ void startS(double ux, double uxs, double uy1, unsigned m) {
    ux_c = ux;
    uxs_c = uxs;
    uy_c = 0.0;
    uy_n = uy1;
    m_c = m;
    tP_c = tXX_c = tYY_c = sXX_c = sYY_c = 0;
    tP_p = tXX_p = tYY_p = 0;
    tXY_p = tXY_pp = 0;
    cP_p = cYY_p = cXY_p = 0;
  }

  __device__ __forceinline__ Out step(const int j, const double ux_n, const double uy_nn,
                                      const unsigned m_n, const CT vx_in, const CT vy_in) {
    return stepS(j, ux_n, lane_next(ux_n), lane_next(uy_n), uy_nn, m_n, vx_in, vy_in);
  }
  // the step with the right neighbours uxs_n = Ux(i+1, j+1) and uys_n =
  // Uy(i+1, j+1) given: a stage reading its rows from an LDS ring reads them
  // at the next position instead of shifting lanes (the same values in every
  // lane but the last, whose outputs no stage uses)
  __device__ __forceinline__ Out stepS(const int j, const double ux_n, const double uxs_n,
                                       const double uys_n, const double uy_nn,
                                       const unsigned m_n, const CT vx_in, const CT vy_in) {
    // trial cell row j+1
    CT tP_n, tXX_n, tYY_n;
    {
      const CT K = !CLEAN && (m_n & M_CELL_MAT) ? Kp : K0;
      const CT G2 = !CLEAN && (m_n & M_CELL_MAT) ? G2_p : G2_0;
      const CT exx = dxs(uxs_n - ux_n);
      const CT eyy = dys(uy_nn - uy_n);
      CT pp, txx, tyy;
      if constexpr (sizeof(CT) == 8) {
        pp = P0 - K * (exx + eyy);
        txx = G2 * (exx - M::third(exx + eyy));
        tyy = G2 * (eyy - M::third(exx + eyy));
      } else {
        split_cell_trial(exx, eyy, K, G2, P0, pp, txx, tyy);
      }
      const bool ok = (!EX || colC) && (!EY || (j + 1 >= 0 && j + 1 < nY));
      if ((!CLEAN && (m_n & M_CELL_ZERO)) || !ok) {
        pp = 0;
        txx = 0;
        tyy = 0;
      }
      tP_n = pp;
      tXX_n = txx;
      tYY_n = tyy;
    }
    const CT sXX_n = lane_next(tXX_n), sYY_n = lane_next(tYY_n);
    // trial vertex row j
    CT tXY_c;
    {
      const CT g = CLEAN ? g0 : gt[(m_c >> M_VCLASS_SHIFT) & 15];
      CT txy;
      if constexpr (sizeof(CT) == 8) txy = g * (dys(uxs_n - uxs_c) + dxs(uys_n - uy_n));
      else txy = split_vertex_trial(g, (CT)(uxs_n - uxs_c), rdYc, (CT)(uys_n - uy_n), rdXc);
      const bool ok = (!EX || colV) && (!EY || (j >= 0 && j < nY - 1));
      if ((!CLEAN && (m_c & M_VERT_ZERO)) || !ok) txy = 0;
      tXY_c = txy;
    }
    // return mapping
    CT cXY_c = tXY_c, cP_c = tP_c, cXX_c = tXX_c, cYY_c = tYY_c;
    if (plastic && !EDGE) {
      // M::vertex_rm and M::cell_rm, with the scaling skipped where no lane
      // of the wave yields (the squared invariants decide, as there)
      const CT sxx = tXX_c + sXX_c + tXX_n + sXX_n;
      const CT syy = tYY_c + sYY_c + tYY_n + sYY_n;
      CT s2v, s2c;
      if constexpr (sizeof(CT) == 8)
        s2v = (CT)0.0625 * sxx * sxx + (CT)0.0625 * syy * syy + (CT)2.0 * tXY_c * tXY_c;
      else
        s2v = split_s2_vertex(sxx, syy, tXY_c);
      const CT pr_p = lane_prev(tXY_p), pr_c = lane_prev(tXY_c);
      const CT av = (CT)0.25 * (pr_p + tXY_p + pr_c + tXY_c);
      if constexpr (sizeof(CT) == 8) s2c = cXX_c * cXX_c + cYY_c * cYY_c + (CT)2.0 * av * av;
      else s2c = split_s2_cell(cXX_c, cYY_c, av);
      if (__any(s2v > c2lo || s2c > c2lo)) {
        CT sc;
        if (rm_scale(s2v, coh, c2lo, sc)) cXY_c *= sc;
        if (rm_scale(s2c, coh, c2lo, sc)) {
          cXX_c = cXX_c * sc;
          cYY_c = cYY_c * sc;
        }
      }
    } else if (plastic) {
      const bool vok = (!EX || colV) && (!EY || (j >= 0 && j < nY - 1));
      const CT sxx = tXX_c + sXX_c + tXX_n + sXX_n;
      const CT syy = tYY_c + sYY_c + tYY_n + sYY_n;
      if (vok) cXY_c = M::vertex_rm(tXY_c, sxx, syy, coh, c2lo);
      const CT pr_p = lane_prev(tXY_p), pr_c = lane_prev(tXY_c);
      {
        // the clamped vertex blocks of the boundary columns (EX only): one
        // column right and two left
        CT nx_p = 0, nx_c = 0, nx_pp = 0, pp_p = 0, pp_c = 0, pp_pp = 0;
        const CT pr_pp = lane_prev(tXY_pp);
        if constexpr (EX) {
          nx_p = lane_next(tXY_p);
          nx_c = lane_next(tXY_c);
          nx_pp = lane_next(tXY_pp);
          pp_p = lane_prev(pr_p);
          pp_c = lane_prev(pr_c);
          pp_pp = lane_prev(pr_pp);
        }
        auto first = [&](CT t, CT pr, CT pp) { return leftEdge ? t : (rightEdge ? pp : pr); };
        auto second = [&](CT t, CT pr, CT nxv) { return leftEdge ? nxv : (rightEdge ? pr : t); };
        if (colC && (!EY || (j >= 1 && j < nY - 1))) {
          const CT av = (CT)0.25 * (first(tXY_p, pr_p, pp_p) + second(tXY_p, pr_p, nx_p) +
                                    first(tXY_c, pr_c, pp_c) + second(tXY_c, pr_c, nx_c));
          M::cell_rm(cXX_c, cYY_c, av, coh, c2lo);
        } else if (EY && colC && j == nY - 1 && j >= 1) {
          const CT av = (CT)0.25 * (first(tXY_pp, pr_pp, pp_pp) + second(tXY_pp, pr_pp, nx_pp) +
                                    first(tXY_p, pr_p, pp_p) + second(tXY_p, pr_p, nx_p));
          M::cell_rm(cXX_c, cYY_c, av, coh, c2lo);
        }
        if (EY && j == 1) {
          CT xx = tXX_p, yy = tYY_p;
          if (colC) {
            const CT av = (CT)0.25 * (first(tXY_p, pr_p, pp_p) + second(tXY_p, pr_p, nx_p) +
                                      first(tXY_c, pr_c, pp_c) + second(tXY_c, pr_c, nx_c));
            M::cell_rm(xx, yy, av, coh, c2lo);
          }
          cP_p = tP_p;
          cYY_p = yy;
        }
      }
    }
    // velocity and displacement of row j
    const CT rP = lane_prev(cP_c), rXX = lane_prev(cXX_c);
    const CT rXY_c = lane_prev(cXY_c), rXY_p = lane_prev(cXY_p);
    Out o;
    CT vx = vx_in, vy = vy_in;
    o.updX = (!EX || (i > 0 && i < nX)) && (!EY || (j > 0 && j < nY - 1));
    o.updY = (!EX || (i > 0 && i < nX - 1)) && (!EY || (j > 0 && j < nY));
    if constexpr (sizeof(CT) == 8) {
      // U advances with the velocity before its rounding to the storage type
      // (as the tile and march kernels do); the next stage gets the stored value
      if (o.updX)
        vx = vx * vdampX + dTrho * (dxs(-cP_c + rP + cXX_c - rXX) + dys(rXY_c - rXY_p));
      if (o.updY)
        vy = vy * vdampY + dTrho * (dys(-cP_c + cP_p + cYY_c - cYY_p) + dxs(cXY_p - rXY_p));
    } else {
      if (o.updX)
        vx = (CT)(VT)split_velocity(vx, vdampX, dTrho, -cP_c + rP + cXX_c - rXX, rdXc,
                                    rXY_c - rXY_p, rdYc);
      if (o.updY)
        vy = (CT)(VT)split_velocity(vy, vdampY, dTrho, -cP_c + cP_p + cYY_c - cYY_p, rdYc,
                                    cXY_p - rXY_p, rdXc);
    }
    const bool in = (!EX || inX) && (!EY || (j >= 0 && j < nY));
    o.vx = (CT)(VT)vx;
    o.vy = (CT)(VT)vy;
    if constexpr (sizeof(CT) == 8) {
      o.ux = in ? ux_c + (double)vx * dT : ux_c;
      o.uy = in ? uy_c + (double)vy * dT : uy_c;
    } else {
      o.ux = in ? split_disp(ux_c, vx, dT) : ux_c;
      o.uy = in ? split_disp(uy_c, vy, dT) : uy_c;
    }
    // rotate
    ux_c = ux_n;
    uxs_c = uxs_n;
    uy_c = uy_n;
    uy_n = uy_nn;
    m_c = m_n;
    tP_p = tP_c;
    tXX_p = tXX_c;
    tYY_p = tYY_c;
    tP_c = tP_n;
    tXX_c = tXX_n;
    tYY_c = tYY_n;
    sXX_c = sXX_n;
    sYY_c = sYY_n;
    tXY_pp = tXY_p;
    tXY_p = tXY_c;
    cP_p = cP_c;
    cYY_p = cYY_c;
    cXY_p = cXY_c;
    return o;
  }
};

// Stage A's inputs in a two-iteration segment: the first rows (start) and,
// for step t, Ux row t+1, Uy row t+2, mask row t+1 and V row t of the state in
// HBM, through buffer resources (the lane's byte offset fixed, the row's in an
// SGPR).  Interior segments clamp rows to the segment's last step tEnd.  Edge
// segments bound each resource to its field: lanes outside the field's
// columns carry the offset kOob, and a row outside the field reads through a
// resource of bound 0, so every such load returns 0 without a branch.
template <typename CT, typename VT, bool EX, bool EY = EX, bool CLEAN = false>
struct StageInput {
  static constexpr bool EDGE = EX || EY;
  using Ld = typename March<CT, VT>::Ld;
  const double* Uxi;
  const double* Uyi;
  const VT* Vxi;
  const VT* Vyi;
  const unsigned char* mask;
  int nX, nY, sx, i, tEnd;
  unsigned o8, oV;
  __amdgpu_buffer_rsrc_t bUx, bUy, bVx, bVy, bM;
  // edge segments: lane offsets with the column test folded in (Ux columns
  // 0..nX, the others 0..nX-1; V is read for columns 0..nX-1 only)
  unsigned oUx, oC8, oM, oVc;

  __device__ __forceinline__ StageInput(const RveDev& r, int cur, int vpar, int col, int tEnd_)
      : Uxi(r.Ux[cur]), Uyi(r.Uy[cur]),
        Vxi(static_cast<const VT*>(vpar ? r.Vx2 : r.Vx)),
        Vyi(static_cast<const VT*>(vpar ? r.Vy2 : r.Vy)), mask(r.mask), nX(r.nX), nY(r.nY),
        sx(r.nX + 1), i(col), tEnd(tEnd_), o8((unsigned)col * 8u),
        oV((unsigned)col * (unsigned)sizeof(VT)) {
    if (!EDGE) {
      bUx = brsrc(Uxi);
      bUy = brsrc(Uyi);
      bVx = brsrc(Vxi);
      bVy = brsrc(Vyi);
      bM = brsrc(mask);
    } else {
      const unsigned ux = (unsigned)nY * (unsigned)sx, uy = (unsigned)(nY + 1) * (unsigned)nX;
      bUx = brsrc_n(Uxi, ux * 8u);
      bUy = brsrc_n(Uyi, uy * 8u);
      bVx = brsrc_n(Vxi, ux * (unsigned)sizeof(VT));
      bVy = brsrc_n(Vyi, uy * (unsigned)sizeof(VT));
      bM = brsrc_n(mask, (unsigned)nY * (unsigned)nX);
      const bool cx = col >= 0 && col <= nX, cc = col >= 0 && col < nX;
      oUx = cx ? o8 : kOob;
      oC8 = cc ? o8 : kOob;
      oM = cc ? (unsigned)col : kOob;
      oVc = cc ? oV : kOob;
    }
  }

  // Ux row jb, Uy row jb+1, mask row jb into the stage
  __device__ __forceinline__ void start(MarchStage<CT, VT, EX, EY, CLEAN>& st, int jb) const {
    double ux = 0.0, uy = 0.0;
    unsigned m = 0;
    const bool colC = !EDGE || (i >= 0 && i < nX);
    if (!EDGE) {
      ux = gld(Uxi, (unsigned)(jb * sx + i));
      uy = gld(Uyi, (unsigned)((jb + 1) * nX + i));
      if (!CLEAN) m = gld(mask, (unsigned)(jb * nX + i));
    } else {
      if (i >= 0 && i <= nX && jb >= 0 && jb < nY) ux = gld(Uxi, (unsigned)(jb * sx + i));
      if (colC && jb + 1 >= 0 && jb + 1 <= nY) uy = gld(Uyi, (unsigned)((jb + 1) * nX + i));
      if (!CLEAN && colC && jb >= 0 && jb < nY) m = gld(mask, (unsigned)(jb * nX + i));
    }
    st.start(ux, uy, m);
  }

  __device__ __forceinline__ Ld fetch(int t0) const {
    const int t = EDGE ? t0 : min(t0, tEnd);
    Ld d;
    if (!EDGE) {
      d.ux = bld<double>(bUx, o8, (unsigned)(t + 1) * (unsigned)sx * 8u);
      d.uy = bld<double>(bUy, o8, (unsigned)(t + 2) * (unsigned)nX * 8u);
      d.m = CLEAN ? 0u : bld<unsigned char>(bM, (unsigned)i, (unsigned)(t + 1) * (unsigned)nX);
      d.vx = bld<VT>(bVx, oV, (unsigned)t * (unsigned)sx * (unsigned)sizeof(VT));
      d.vy = bld<VT>(bVy, oV, (unsigned)t * (unsigned)nX * (unsigned)sizeof(VT));
    } else {
      // rows outside a field read through a resource of bound 0 (the row
      // tests are uniform: scalar selects); the row offset is then 0
      const __amdgpu_buffer_rsrc_t none = brsrc_n(Uxi, 0u);
      const bool r1 = t + 1 >= 0 && t + 1 < nY, r2 = t + 2 >= 0 && t + 2 <= nY,
                 r0 = t >= 0 && t < nY;
      const unsigned s1 = r1 ? (unsigned)(t + 1) : 0u, s2 = r2 ? (unsigned)(t + 2) : 0u,
                     s0 = r0 ? (unsigned)t : 0u;
      d.ux = bld<double>(r1 ? bUx : none, oUx, s1 * (unsigned)sx * 8u);
      d.uy = bld<double>(r2 ? bUy : none, oC8, s2 * (unsigned)nX * 8u);
      d.m = CLEAN ? 0u : bld<unsigned char>(r1 ? bM : none, oM, s1 * (unsigned)nX);
      d.vx = bld<VT>(r0 ? bVx : none, oVc, s0 * (unsigned)sx * (unsigned)sizeof(VT));
      d.vy = bld<VT>(r0 ? bVy : none, oVc, s0 * (unsigned)nX * (unsigned)sizeof(VT));
    }
    return d;
  }
};

template <typename CT, typename VT, bool EDGE>
__device__ __forceinline__ void march2_run(const RveDev& r, int cur, int vpar, int s, int j0,
                                           int j1, const CT* gt) {
  using M = March<CT, VT>;
  using Ld = typename M::Ld;
  const int nX = r.nX, nY = r.nY, sx = nX + 1;
  const int L = threadIdx.x & 63;
  const int i = s + L;
  double* __restrict__ Uxo = r.Ux[cur ^ 1];
  double* __restrict__ Uyo = r.Uy[cur ^ 1];
  // velocities are read from one buffer and written to the other
  VT* Vxo = static_cast<VT*>(vpar ? r.Vx : r.Vx2);
  VT* Vyo = static_cast<VT*>(vpar ? r.Vy : r.Vy2);
  const bool outLane = L >= 4 && L < 60 && (!EDGE || i < nX);

  MarchStage<CT, VT, EDGE> A, B;
  A.setup(r, gt, i);
  B.setup(r, gt, i);
  const int jbA = j0 - 4, jbB = j0 - 2;
  const int tEnd = j1 + 1;  // A's last step: B's last step j1-1 needs A's row j1+1
  const StageInput<CT, VT, EDGE> in(r, cur, vpar, i, tEnd);
  in.start(A, jbA);
  // A's outputs of rows t-1 (Ux) and t-1, t-2 (V); A's input mask rows t, t-1
  double h1ux = 0.0;
  CT h1vx = 0, h1vy = 0, h2vx = 0, h2vy = 0;
  unsigned mh1 = 0, mh2 = 0;
  auto fetch = [&](int t) { return in.fetch(t); };
  auto body = [&](const int t, const Ld& ld) {
    const auto oA = A.step(t, ld.ux, ld.uy, ld.m, (CT)ld.vx, (CT)ld.vy);
    if (t == jbB + 1) B.start(h1ux, oA.uy, mh2);  // Ux' row jbB, Uy' row jbB+1, mask row jbB
    if (t >= jbB + 2) {
      const int j = t - 2;
      const double uxsB = B.uxs_c;  // Ux(i+1, j) of B's input, before the step
      const auto oB = B.step(j, h1ux, oA.uy, mh2, h2vx, h2vy);
      if (j >= j0 && outLane) {
        if (oB.updX) gst(Vxo + (size_t)j * sx, (unsigned)i, (VT)oB.vx);
        if (oB.updY) gst(Vyo + (size_t)j * nX, (unsigned)i, (VT)oB.vy);
        gst(Uxo + (size_t)j * sx, (unsigned)i, oB.ux);
        gst(Uyo + (size_t)j * nX, (unsigned)i, oB.uy);
        if (EDGE) {
          // nodes ComputeDisp never updates: Ux column nX, Uy row nY
          if (i == nX - 1) gst(Uxo, (unsigned)(j * sx + i + 1), uxsB);
          if (j == nY - 1) gst(Uyo, (unsigned)((j + 1) * nX + i), B.uy_c);  // row j+1 after the step
        }
      }
    }
    h2vx = h1vx;
    h2vy = h1vy;
    h1ux = oA.ux;
    h1vx = oA.vx;
    h1vy = oA.vy;
    mh2 = mh1;
    mh1 = ld.m;
  };
  Ld P = fetch(jbA);
  for (int t = jbA; t <= tEnd; t += 2) {
    const Ld Q = fetch(t + 1);
    body(t, P);
    if (t + 1 > tEnd) break;
    P = fetch(t + 2);
    body(t + 1, Q);
  }
}

template <typename CT, typename VT>
__global__ __launch_bounds__(256) void k_march2(const RveDev* __restrict__ rves,
                                                const int4* __restrict__ segs, int nsegs,
                                                int parity, int vpar) {
  __shared__ CT gt[4][16];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = tile_of_block(blockIdx.x, gridDim.x);
  const int k = b * 4 + wave;
  if (k >= nsegs) return;
  const int4 sg = segs[k];
  const RveDev& r = rves[sg.x];
  if (lane < 16) gt[wave][lane] = (CT)r.gtab[lane];
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's table is in LDS
  __builtin_amdgcn_wave_barrier();
  const int cur = parity ^ r.cur;
  const int s = sg.y, j0 = sg.z, j1 = sg.w;
  const bool interior = s >= 1 && s + 63 <= r.nX - 2 && j0 >= 4 && j1 <= r.nY - 3;
  if (interior) march2_run<CT, VT, false>(r, cur, vpar, s, j0, j1, gt[wave]);
  else march2_run<CT, VT, true>(r, cur, vpar, s, j0, j1, gt[wave]);
}

// ---------------------------------------------------------------------------
// Two iterations per launch on a pair of waves.  Of each pair of waves the
// first runs stage A (iteration n+1 from the state in HBM) and the second
// stage B (iteration n+2) three rows behind; A hands every row's outputs (Ux,
// Uy, V, and the mask row it read) to B through a four-row LDS ring, and the
// workgroup advances one row per barrier.  A wave holds one stage's state
// instead of two, so about twice as many waves fit on a SIMD to hide the
// memory latency.  Segments, arithmetic and outputs are those of k_march2.
//
// RowRing: the four-row LDS ring through which a stage hands its output rows
// to the next (also the ring of the wave chains below).
// W waves side by side per stage (the chains): rows of 64 W entries.
template <typename CT, int W = 1>
struct RowRing {
  double ux[4][64 * W];
  double uy[4][64 * W];
  CT vx[4][64 * W];
  CT vy[4][64 * W];
  unsigned m[4][64 * W];
};

// Lockstep step q (0 <= q < nSteps) is row t = j0-4+q of stage A and row
// t-3 of stage B.  A at row t writes ring slot q&3; B at row j reads Ux' of
// row j+1 (slot (q-2)&3), Uy' of row j+2 (slot (q-1)&3), and V' of row j and
// mask row j+1 (slot (q-3)&3), so reads and the write of one step never share
// a slot and the barrier after each step orders a write before its reads.
template <typename CT, typename VT, bool EDGE>
__device__ __forceinline__ void march2p_run(const RveDev& r, int cur, int vpar, int s, int j0,
                                            int j1, const CT* gt, bool stageB,
                                            RowRing<CT>& ring, int nSteps) {
  using M = March<CT, VT>;
  using Ld = typename M::Ld;
  const int nX = r.nX, nY = r.nY, sx = nX + 1;
  const int L = threadIdx.x & 63;
  const int i = s + L;
  const int jbA = j0 - 4, jbB = j0 - 2, tEnd = j1 + 1;
  MarchStage<CT, VT, EDGE> st;
  st.setup(r, gt, i);
  if (!stageB) {
    const StageInput<CT, VT, EDGE> in(r, cur, vpar, i, tEnd);
    in.start(st, jbA);
    auto fetch = [&](int t) { return in.fetch(t); };
    auto stepA = [&](int q, const Ld& ld, int sl) {
      const int t = jbA + q;
      if (t <= tEnd) {
        const auto o = st.step(t, ld.ux, ld.uy, ld.m, (CT)ld.vx, (CT)ld.vy);
        ring.ux[sl][L] = o.ux;
        ring.uy[sl][L] = o.uy;
        ring.vx[sl][L] = o.vx;
        ring.vy[sl][L] = o.vy;
        ring.m[sl][L] = ld.m;
      }
      lds_barrier();
    };
    // two row buffers (an unroll by four, for constant ring slots, would
    // keep more registers live than the immediate offsets save)
    Ld P = fetch(jbA);
    for (int q = 0; q < nSteps; q += 2) {
      const Ld Q = fetch(jbA + q + 1);
      stepA(q, P, q & 3);
      if (q + 1 >= nSteps) break;
      P = fetch(jbA + q + 2);
      stepA(q + 1, Q, (q + 1) & 3);
    }
  } else {
    double* __restrict__ Uxo = r.Ux[cur ^ 1];
    double* __restrict__ Uyo = r.Uy[cur ^ 1];
    VT* Vxo = static_cast<VT*>(vpar ? r.Vx : r.Vx2);
    VT* Vyo = static_cast<VT*>(vpar ? r.Vy : r.Vy2);
    const bool outLane = L >= 4 && L < 60 && (!EDGE || i < nX);
    const unsigned o8 = (unsigned)i * 8u, oV = (unsigned)i * (unsigned)sizeof(VT);
    // edge segments bound the V resources to their fields: a lane whose V
    // ComputeDisp leaves alone stores at kOob, which the hardware drops
    const unsigned vxb = (unsigned)nY * (unsigned)sx * (unsigned)sizeof(VT),
                   vyb = (unsigned)(nY + 1) * (unsigned)nX * (unsigned)sizeof(VT);
    const __amdgpu_buffer_rsrc_t bUx = brsrc(Uxo), bUy = brsrc(Uyo),
                                 bVx = EDGE ? brsrc_n(Vxo, vxb) : brsrc(Vxo),
                                 bVy = EDGE ? brsrc_n(Vyo, vyb) : brsrc(Vyo);
    // slot phase ph = q & 3 known at compile time (the loop is unrolled by
    // four): ring addresses are immediate offsets
    auto stepB = [&](int q, auto ph) {
      constexpr int s1 = (ph + 3) & 3, s2 = (ph + 2) & 3, s3 = (ph + 1) & 3;  // q-1, q-2, q-3
      const int t = jbA + q;
      if (t == jbB + 2) {
        st.start(ring.ux[s2][L], ring.uy[s1][L], ring.m[s3][L]);
      } else if (t >= jbB + 3 && t <= tEnd + 1) {
        const int j = t - 3;
        const double uxn = ring.ux[s2][L], uynn = ring.uy[s1][L];
        const unsigned mn = ring.m[s3][L];
        const CT vxi = ring.vx[s3][L], vyi = ring.vy[s3][L];
        const double uxsB = st.uxs_c;  // Ux(i+1, j) of B's input, before the step
        const auto o = st.step(j, uxn, uynn, mn, vxi, vyi);
        if (j >= j0 && outLane) {
          // interior: every lane's V is updated
          const unsigned ovx = EDGE && !o.updX ? kOob : oV, ovy = EDGE && !o.updY ? kOob : oV;
          bst(bVx, ovx, (unsigned)j * (unsigned)sx * (unsigned)sizeof(VT), (VT)o.vx);
          bst(bVy, ovy, (unsigned)j * (unsigned)nX * (unsigned)sizeof(VT), (VT)o.vy);
          bst(bUx, o8, (unsigned)j * (unsigned)sx * 8u, o.ux);
          bst(bUy, o8, (unsigned)j * (unsigned)nX * 8u, o.uy);
          if (EDGE) {
            // nodes ComputeDisp never updates: Ux column nX, Uy row nY
            if (i == nX - 1) gst(Uxo, (unsigned)(j * sx + i + 1), uxsB);
            if (j == nY - 1) gst(Uyo, (unsigned)((j + 1) * nX + i), st.uy_c);  // row j+1 after the step
          }
        }
      }
      lds_barrier();
    };
    for (int q = 0; q < nSteps; q += 4) {
      stepB(q, std::integral_constant<int, 0>{});
      if (q + 1 >= nSteps) break;
      stepB(q + 1, std::integral_constant<int, 1>{});
      if (q + 2 >= nSteps) break;
      stepB(q + 2, std::integral_constant<int, 2>{});
      if (q + 3 >= nSteps) break;
      stepB(q + 3, std::integral_constant<int, 3>{});
    }
  }
}

template <typename CT, typename VT>
__global__ __launch_bounds__(256) void k_march2p(const RveDev* __restrict__ rves,
                                                 const int4* __restrict__ segs, int nsegs,
                                                 int parity, int vpar) {
  __shared__ CT gt[4][16];
  __shared__ RowRing<CT> ring[2];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = tile_of_block(blockIdx.x, gridDim.x);
  const int k0 = 2 * b;
  if (k0 >= nsegs) return;  // the whole workgroup
  // the workgroup marches the longer of its two segments, one barrier a step
  const int4 s0 = segs[k0];
  int nSteps = s0.w - s0.z + 7;
  if (k0 + 1 < nsegs) {
    const int4 s1 = segs[k0 + 1];
    nSteps = max(nSteps, s1.w - s1.z + 7);
  }
  const int k = k0 + (wave >> 1);
  if (k >= nsegs) {  // a pair without a segment keeps the barrier count
    for (int q = 0; q < nSteps; q++) lds_barrier();
    return;
  }
  const int4 sg = segs[k];
  const RveDev& r = rves[sg.x];
  if (lane < 16) gt[wave][lane] = (CT)r.gtab[lane];
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's table is in LDS
  __builtin_amdgcn_wave_barrier();
  const int cur = parity ^ r.cur;
  const int s = sg.y, j0 = sg.z, j1 = sg.w;
  const bool stageB = (wave & 1) != 0;
  const bool interior = s >= 1 && s + 63 <= r.nX - 2 && j0 >= 4 && j1 <= r.nY - 3;
  if (interior)
    march2p_run<CT, VT, false>(r, cur, vpar, s, j0, j1, gt[wave], stageB, ring[wave >> 1], nSteps);
  else
    march2p_run<CT, VT, true>(r, cur, vpar, s, j0, j1, gt[wave], stageB, ring[wave >> 1], nSteps);
}

// Block map of the mask: one byte per kMapRows x kMapCols block of cells,
// nonzero when a mask byte of the block is.
__global__ void k_mask_map(const unsigned char* __restrict__ mask, int nX, int nY,
                           unsigned char* __restrict__ map, int nbx, int nby) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbx * nby) return;
  const int bx = b % nbx, by = b / nbx;
  const int i0 = bx * kMapCols, j0 = by * kMapRows;
  const int i1 = min(i0 + kMapCols, nX), j1 = min(j0 + kMapRows, nY);
  unsigned any = 0;
  for (int j = j0; j < j1; j++)
    for (int i = i0; i < i1; i++) any |= mask[(size_t)j * nX + i];
  map[b] = any ? 1 : 0;
}

// Whether the mask is 0 over rows [jlo, jhi] and columns [ilo, ihi] (clipped
// to the grid), from the block map; the same answer in every lane.
__device__ __forceinline__ bool mask_clear(const RveDev& r, int ilo, int ihi, int jlo, int jhi) {
  ilo = max(ilo, 0);
  jlo = max(jlo, 0);
  ihi = min(ihi, r.nX - 1);
  jhi = min(jhi, r.nY - 1);
  if (ilo > ihi || jlo > jhi) return true;
  const int bx0 = ilo / kMapCols, bx1 = ihi / kMapCols;
  const int by0 = jlo / kMapRows, by1 = jhi / kMapRows;
  const int nbx = (r.nX + kMapCols - 1) / kMapCols;
  const int w = bx1 - bx0 + 1, n = w * (by1 - by0 + 1);
  unsigned any = 0;
  for (int q = (int)(threadIdx.x & 63); q < n; q += 64)
    any |= r.mmap[(by0 + q / w) * nbx + bx0 + q % w];
  return !__any(any != 0);
}

// NS iterations per launch on a chain of NS waves (one workgroup).  Stage 0
// reads the state from HBM and runs the first iteration, stage k the (k+1)-th
// from the rows of stage k-1, and the last stage writes the result.  Each
// stage hands its rows on through a four-row LDS ring, and each lags the one
// before by three rows, as stage B lags stage A in march2p_run: lockstep step
// q is row t = j0-2NS+q of stage 0 and row t-3k of stage k.  Every stage loses
// two columns and two rows of halo, so strips carry 64-4NS output columns
// (lanes 2NS..63-2NS) and stage 0 starts 2NS rows above the segment.  A launch
// reads and writes every field once per NS iterations.
//
// Stage k >= 1 marches rows jb = j0-2(NS-k) .. j1-1+2(NS-1-k): it takes its
// first rows at t = jb+3k-1 and steps row j = t-3k for t >= jb+3k, reading
// Ux' of row j+1 (slot q-2), Uy' of row j+2 (slot q-1), V' of row j and mask
// row j+1 (slot q-3) of its ring; a middle stage writes its row and that mask
// row to slot q of the next ring.  Reads and the write of one step never share
// a slot, and the barrier after each step orders a write before its reads.
//
// With W > 1 waves per stage, wave c of a stage takes columns s+60c .. s+60c+63
// (neighbouring waves overlap by four columns) and the stage's rows span the
// W waves' union of 60W+4 columns: a wave reads its 64 entries of the ring row
// at position 60c+L, so its input halo comes from its neighbours' outputs,
// and writes the entries it computes exactly (lanes 2..61, the union's outer
// lanes 0, 1 and 62, 63 included), each position once.  Only the union's
// outer edges lose two columns per stage: strips of 60W+4-4NS outputs.
template <typename CT, typename VT, bool EX, bool EY, bool CLEAN, int NS, int W>
__device__ __forceinline__ void chain_run(const RveDev& r, int cur, int vpar, int s, int j0,
                                          int j1, const CT* gt, int role, int c,
                                          RowRing<CT, W>* ring, int nSteps) {
  constexpr bool EDGE = EX || EY;
  constexpr int UW = 60 * W + 4;  // columns of the union of a stage's waves
  using M = March<CT, VT>;
  using Ld = typename M::Ld;
  using Stage = MarchStage<CT, VT, EX, EY, CLEAN>;
  const int nX = r.nX, nY = r.nY, sx = nX + 1;
  const int L = threadIdx.x & 63;
  const int P = W == 1 ? L : 60 * c + L;  // position in the union (and the ring row)
  const int i = s + P;
  // the lanes whose outputs this wave hands on (each ring position once)
  const bool wl = W == 1 || ((c == 0 || L >= 2) && (c == W - 1 || L < 62));
  const int jbA = j0 - 2 * NS, tEnd = j1 + 2 * NS - 3;
  Stage st;
  st.setup(r, gt, i);
  // Each stage runs its lockstep steps in three loops -- idle steps before
  // its first row, its rows, idle steps after its last -- so the loop over
  // its rows carries only the stage's state, and the ring slot of a step
  // (q & 3) is known at compile time in a loop unrolled by four.
  auto idle = [&](int q0, int q1) {
    for (int q = q0; q < q1; q++) lds_barrier();
  };
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  if (role == 0) {
    // stage 0: rows jbA..tEnd (steps 0 .. tEnd-jbA) from HBM into ring 0
    RowRing<CT, W>& out = ring[0];
    const StageInput<CT, VT, EX, EY, CLEAN> in(r, cur, vpar, i, tEnd);
    in.start(st, jbA);
    const int nAct = tEnd - jbA + 1;
    auto stepA = [&](int q, const Ld& ld, auto sl) {
      const auto o = st.step(jbA + q, ld.ux, ld.uy, ld.m, (CT)ld.vx, (CT)ld.vy);
      if (wl) {
        out.ux[sl][P] = o.ux;
        out.uy[sl][P] = o.uy;
        out.vx[sl][P] = o.vx;
        out.vy[sl][P] = o.vy;
        if (!CLEAN) out.m[sl][P] = ld.m;
      }
      lds_barrier();
    };
    Ld P = in.fetch(jbA), Q;
    int q = 0;
    for (; q + 4 <= nAct; q += 4) {
      Q = in.fetch(jbA + q + 1);
      stepA(q, P, I0{});
      P = in.fetch(jbA + q + 2);
      stepA(q + 1, Q, I1{});
      Q = in.fetch(jbA + q + 3);
      stepA(q + 2, P, I2{});
      P = in.fetch(jbA + q + 4);
      stepA(q + 3, Q, I3{});
    }
    if (q < nAct) {
      Q = in.fetch(jbA + q + 1);
      stepA(q, P, I0{});
      if (q + 1 < nAct) {
        P = in.fetch(jbA + q + 2);
        stepA(q + 1, Q, I1{});
        if (q + 2 < nAct) stepA(q + 2, P, I2{});
      }
    }
    idle(nAct, nSteps);
    return;
  }
  // every later stage is its own loop with its stage number K known at
  // compile time: ring addresses are constants and no stage keeps another's
  // values live
  auto march = [&](auto K) {
    constexpr int k = decltype(K)::value, lag = 3 * k;
    constexpr bool last = k == NS - 1;
    // first step (rows in from the ring) and last step of this stage
    constexpr int qStart = 5 * k - 1;
    const int qLast = tEnd + k - jbA;
    RowRing<CT, W>& in = ring[k - 1];
    RowRing<CT, W>& out = ring[last ? k - 1 : k];  // not written by the last stage
    double* __restrict__ Uxo = r.Ux[cur ^ 1];
    double* __restrict__ Uyo = r.Uy[cur ^ 1];
    VT* Vxo = static_cast<VT*>(vpar ? r.Vx : r.Vx2);
    VT* Vyo = static_cast<VT*>(vpar ? r.Vy : r.Vy2);
    const bool outLane = wl && P >= 2 * NS && P < UW - 2 * NS && (!EX || i < nX);
    const unsigned o8 = (unsigned)i * 8u, oV = (unsigned)i * (unsigned)sizeof(VT);
    // edge segments bound the V resources to their fields: a lane whose V
    // ComputeDisp leaves alone stores at kOob, which the hardware drops
    const unsigned vxb = (unsigned)nY * (unsigned)sx * (unsigned)sizeof(VT),
                   vyb = (unsigned)(nY + 1) * (unsigned)nX * (unsigned)sizeof(VT);
    const __amdgpu_buffer_rsrc_t bUx = brsrc(Uxo), bUy = brsrc(Uyo),
                                 bVx = EDGE ? brsrc_n(Vxo, vxb) : brsrc(Vxo),
                                 bVy = EDGE ? brsrc_n(Vyo, vyb) : brsrc(Vyo);
    // step q reads Ux' of row j+1 (slot q-2), Uy' of row j+2 (slot q-1), V'
    // of row j and mask row j+1 (slot q-3); a middle stage writes slot q
    auto stepK = [&](int q, auto ph) {
      constexpr int s1 = (ph + 3) & 3, s2 = (ph + 2) & 3, s3 = (ph + 1) & 3;  // q-1, q-2, q-3
      const int j = jbA + q - lag;
      const double uxn = in.ux[s2][P], uynn = in.uy[s1][P];
      const unsigned mn = CLEAN ? 0u : in.m[s3][P];
      const CT vxi = in.vx[s3][P], vyi = in.vy[s3][P];
      const double uxsK = st.uxs_c;  // Ux(i+1, j) of this stage's input, before the step
      // with two waves per stage the ring rows are 128 entries wide and the
      // right neighbours are read at P+1 (slot q-2 holds Ux' and Uy' of row
      // j+1); one wave per stage shifts lanes
      const auto o = W > 1 ? st.stepS(j, uxn, in.ux[s2][P + 1], in.uy[s2][P + 1], uynn, mn,
                                      vxi, vyi)
                           : st.step(j, uxn, uynn, mn, vxi, vyi);
      if constexpr (!last) {
        if (wl) {
          out.ux[ph][P] = o.ux;
          out.uy[ph][P] = o.uy;
          out.vx[ph][P] = o.vx;
          out.vy[ph][P] = o.vy;
          if (!CLEAN) out.m[ph][P] = mn;
        }
      } else if (j >= j0 && outLane) {
        const unsigned ovx = EDGE && !o.updX ? kOob : oV, ovy = EDGE && !o.updY ? kOob : oV;
        bst(bVx, ovx, (unsigned)j * (unsigned)sx * (unsigned)sizeof(VT), (VT)o.vx);
        bst(bVy, ovy, (unsigned)j * (unsigned)nX * (unsigned)sizeof(VT), (VT)o.vy);
        bst(bUx, o8, (unsigned)j * (unsigned)sx * 8u, o.ux);
        bst(bUy, o8, (unsigned)j * (unsigned)nX * 8u, o.uy);
        // nodes ComputeDisp never updates: Ux column nX, Uy row nY
        if (EX && i == nX - 1) gst(Uxo, (unsigned)(j * sx + i + 1), uxsK);
        if (EY && j == nY - 1) gst(Uyo, (unsigned)((j + 1) * nX + i), st.uy_c);  // row j+1 after the step
      }
      lds_barrier();
    };
    idle(0, qStart);
    {
      // first rows: Ux' row jb (slot qStart-2), Uy' row jb+1 (slot
      // qStart-1), mask row jb (slot qStart-3)
      constexpr int ph = qStart & 3;
      constexpr int s1 = (ph + 3) & 3, s2 = (ph + 2) & 3, s3 = (ph + 1) & 3;
      if constexpr (W > 1)
        st.startS(in.ux[s2][P], in.ux[s2][P + 1], in.uy[s1][P], CLEAN ? 0u : in.m[s3][P]);
      else
        st.start(in.ux[s2][P], in.uy[s1][P], CLEAN ? 0u : in.m[s3][P]);
      lds_barrier();
    }
    constexpr int P0 = (qStart + 1) & 3;
    using Q0 = std::integral_constant<int, P0>;
    using Q1 = std::integral_constant<int, (P0 + 1) & 3>;
    using Q2 = std::integral_constant<int, (P0 + 2) & 3>;
    using Q3 = std::integral_constant<int, (P0 + 3) & 3>;
    int q = qStart + 1;
    for (; q + 3 <= qLast; q += 4) {
      stepK(q, Q0{});
      stepK(q + 1, Q1{});
      stepK(q + 2, Q2{});
      stepK(q + 3, Q3{});
    }
    if (q <= qLast) {
      stepK(q, Q0{});
      if (q + 1 <= qLast) {
        stepK(q + 1, Q1{});
        if (q + 2 <= qLast) stepK(q + 2, Q2{});
      }
    }
    idle(qLast + 1, nSteps);
  };
  static_assert(NS == 3 || NS == 4, "chains of three or four stages");
  if (role == 1) {
    march(std::integral_constant<int, 1>{});
  } else if (role == 2) {
    march(std::integral_constant<int, 2>{});
  } else {
    if constexpr (NS == 4) march(std::integral_constant<int, 3>{});
  }
}

// One segment per workgroup of NS x W waves (stage k on waves kW .. kW+W-1,
// wave c of a stage on columns s+60c ..).  The split scheme's register
// budget holds seven waves per SIMD (six with two waves per stage, where the
// LDS of three workgroups of eight waves fills a CU); FP64 chains of one
// wave per stage are held to five (a few registers spilled, 2-4% faster than
// four), with two waves per stage they run four.
template <typename CT, typename VT, int NS, int W>
__global__ __launch_bounds__(64 * NS * W)
__attribute__((amdgpu_waves_per_eu(sizeof(CT) == 4 ? (W == 1 ? 7 : 6) : (W == 1 ? 5 : 1)))) void k_chain(const RveDev* __restrict__ rves,
                                                   const int4* __restrict__ segs, int nsegs,
                                                   int parity, int vpar) {
  constexpr int UW = 60 * W + 4;
  __shared__ CT gt[NS * W][16];
  __shared__ RowRing<CT, W> ring[NS - 1];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int role = wave / W, c = wave % W;
  const int k = tile_of_block(blockIdx.x, gridDim.x);
  if (k >= nsegs) return;  // the whole workgroup
  const int4 sg = segs[k];
  const RveDev& r = rves[sg.x];
  if (lane < 16) gt[wave][lane] = (CT)r.gtab[lane];
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's table is in LDS
  __builtin_amdgcn_wave_barrier();
  const int cur = parity ^ r.cur;
  const int s = sg.y, j0 = sg.z, j1 = sg.w;
  const int nSteps = j1 - j0 + 5 * NS - 3;
  // column edge (this wave's 64 columns touch the left or right boundary;
  // the other waves of the stage take the interior path, which computes the
  // same values there), row edge (the segment's rows reach the top or
  // bottom), and whether every mask byte the segment reads (rows j0-2NS ..
  // j1+2NS-2, columns s .. s+UW-1) is 0, decided for the whole workgroup: a
  // mask-free stage hands no mask rows on
  const int sw = s + 60 * c;
  const bool ex = !(sw >= 1 && sw + 63 <= r.nX - 2);
  const bool exAny = !(s >= 1 && s + UW - 1 <= r.nX - 2);
  const bool ey = !(j0 >= 2 * NS && j1 <= r.nY - (2 * NS - 1));
  const bool clean = !exAny && mask_clear(r, s, s + UW - 1, j0 - 2 * NS, j1 + 2 * NS - 2);
  const CT* g = gt[wave];
  if (ex)
    chain_run<CT, VT, true, true, false, NS, W>(r, cur, vpar, s, j0, j1, g, role, c, ring,
                                                nSteps);
  else if (ey && clean)
    chain_run<CT, VT, false, true, true, NS, W>(r, cur, vpar, s, j0, j1, g, role, c, ring,
                                                nSteps);
  else if (ey)
    chain_run<CT, VT, false, true, false, NS, W>(r, cur, vpar, s, j0, j1, g, role, c, ring,
                                                 nSteps);
  else if (clean)
    chain_run<CT, VT, false, false, true, NS, W>(r, cur, vpar, s, j0, j1, g, role, c, ring,
                                                 nSteps);
  else
    chain_run<CT, VT, false, false, false, NS, W>(r, cur, vpar, s, j0, j1, g, role, c, ring,
                                                  nSteps);
}

// ---------------------------------------------------------------------------
// Wide chains (split scheme): the chain kernels with two columns per lane.
// Lane L holds columns i0 = s+2L (element 0) and i0+1 (element 1) of every
// row, so one wave spans 128 columns: a stage's four halo columns take 4 of
// 128 lanes' columns instead of 4 of 64, and of the two x-neighbours of a
// lane's columns one is the lane's other column, so one lane shift serves
// both.  Strips carry 128 - 4 NS output columns (lanes NS .. 63-NS).  Rows,
// stage lags, the order of ring slots and the U and V buffers are those of
// the chain kernels, and each element is the split scheme's arithmetic in the
// same order, so the results equal the chain and tile kernels' bit for bit.
//
// Segments are classified as a workgroup starts.  EX: the strip touches the
// left or right boundary (per-column boundary tests); EY: the segment's rows
// reach the top or bottom boundary (per-row tests, uniform over the wave).
// CLEAN: every mask byte the segment reads is 0 (from the RVE's block map of
// the mask), so the stages read no mask and use the matrix K and G, no pore
// zeroing and the vertex modulus of four matrix cells -- exactly what the
// mask path computes from zero mask bytes.

// two doubles / two floats at consecutive columns through a buffer resource,
// one access per element: a row of Ux or Vx (nX+1 elements) starts at an odd
// element on every other row, so a pair is aligned to one element only
__device__ __forceinline__ void bld2d(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                      double (&v)[2]) {
  v[0] = bld<double>(r, voff, soff);
  v[1] = bld<double>(r, voff + 8u, soff);
}
__device__ __forceinline__ void bld2f(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                      float (&v)[2]) {
  v[0] = bld<float>(r, voff, soff);
  v[1] = bld<float>(r, voff + 4u, soff);
}
__device__ __forceinline__ void bst2d(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                      const double (&v)[2]) {
  bst(r, voff, soff, v[0]);
  bst(r, voff + 8u, soff, v[1]);
}
__device__ __forceinline__ void bst2f(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff,
                                      const float (&v)[2]) {
  bst(r, voff, soff, v[0]);
  bst(r, voff + 4u, soff, v[1]);
}

// One wide stage: MarchStage<float, float, EDGE> on two columns per lane.
template <bool EX, bool EY, bool CLEAN>
struct WideStage {
  static constexpr bool EDGE = EX || EY;
  int nX, nY, i0;
  bool colC[2], colV[2], leftE[2], rightE[2];
  bool plastic;
  float coh, c2lo, K0, Kp, G2_0, G2_p, P0, vdampX, vdampY, dTrho, rdX, rdY, g0;
  double dT;
  const float* gt;
  // carried rows, element c at column i0 + c: U rows j and j+1, mask row j,
  // trial cells j (and j-1), trial vertices j-1 and j-2, corrected row j-1;
  // the values at column i0+2 (next lane) and i0-1 (previous lane) that a
  // step needs again are carried too
  double ux_c[2], uy_c[2], uy_n[2], uxs_c1;
  unsigned m_c[2];
  float tP_c[2], tXX_c[2], tYY_c[2], sXX_c1, sYY_c1;
  float tP_p[2], tXX_p[2], tYY_p[2];
  float tXY_p[2], tXY_pp[2], prXY_p0;
  float cP_p[2], cYY_p[2], cXY_p[2], rXY_p0;

  struct Out {
    double ux[2], uy[2];
    float vx[2], vy[2];
    bool updX[2], updY[2];
  };

  __device__ __forceinline__ void setup(const RveDev& rr, const float* g, int col0) {
    gt = g;
    nX = rr.nX;
    nY = rr.nY;
    i0 = col0;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int i = col0 + c;
      colC[c] = !EX || (i >= 0 && i < nX);
      colV[c] = !EX || (i >= 0 && i < nX - 1);
      leftE[c] = EX && i == 0;
      rightE[c] = EX && i == nX - 1;
    }
    plastic = rr.plastic != 0;
    coh = (float)rr.coh;
    c2lo = rr.coh2lof;
    K0 = (float)rr.K0;
    Kp = (float)rr.Kp;
    G2_0 = 2.0f * (float)rr.G0;
    G2_p = 2.0f * (float)rr.Gp;
    P0 = (float)rr.P0;
    rdX = (float)rr.rdX;
    rdY = (float)rr.rdY;
    vdampX = (float)rr.vdampX;
    vdampY = (float)rr.vdampY;
    dTrho = (float)rr.dTrho;
    dT = rr.dT;
    g0 = (float)rr.gtab[0];
  }

  // Ux row jb, Uy row jb+1, mask row jb (jb = first step)
  __device__ __forceinline__ void start(const double (&ux)[2], const double (&uy1)[2],
                                        const unsigned (&m)[2]) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      ux_c[c] = ux[c];
      uy_c[c] = 0.0;
      uy_n[c] = uy1[c];
      m_c[c] = m[c];
      tP_c[c] = tXX_c[c] = tYY_c[c] = 0;
      tP_p[c] = tXX_p[c] = tYY_p[c] = 0;
      tXY_p[c] = tXY_pp[c] = 0;
      cP_p[c] = cYY_p[c] = cXY_p[c] = 0;
    }
    uxs_c1 = lane_next(ux[0]);
    sXX_c1 = sYY_c1 = 0;
    prXY_p0 = rXY_p0 = 0;
  }

  __device__ __forceinline__ Out step(const int j, const double (&ux_n)[2],
                                      const double (&uy_nn)[2], const unsigned (&m_n)[2],
                                      const float (&vx_in)[2], const float (&vy_in)[2]) {
    using M = March<float, float>;
    // trial cell row j+1
    const double uxs_n[2] = {ux_n[1], lane_next(ux_n[0])};
    float tP_n[2], tXX_n[2], tYY_n[2];
    const bool rowC = !EY || (j + 1 >= 0 && j + 1 < nY);
#pragma unroll
    for (int c = 0; c < 2; c++) {
      float K = K0, G2 = G2_0;
      if constexpr (!CLEAN) {
        K = (m_n[c] & M_CELL_MAT) ? Kp : K0;
        G2 = (m_n[c] & M_CELL_MAT) ? G2_p : G2_0;
      }
      const float exx = (float)(uxs_n[c] - ux_n[c]) * rdX;
      const float eyy = (float)(uy_nn[c] - uy_n[c]) * rdY;
      float pp, txx, tyy;
      split_cell_trial(exx, eyy, K, G2, P0, pp, txx, tyy);
      const bool ok = rowC && colC[c];
      if ((!CLEAN && (m_n[c] & M_CELL_ZERO)) || !ok) {
        pp = 0;
        txx = 0;
        tyy = 0;
      }
      tP_n[c] = pp;
      tXX_n[c] = txx;
      tYY_n[c] = tyy;
    }
    const float sXX_n[2] = {tXX_n[1], lane_next(tXX_n[0])};
    const float sYY_n[2] = {tYY_n[1], lane_next(tYY_n[0])};
    // trial vertex row j
    const double uys_n[2] = {uy_n[1], lane_next(uy_n[0])};
    const double uxs_c[2] = {ux_c[1], uxs_c1};
    const bool rowV = !EY || (j >= 0 && j < nY - 1);
    float tXY_c[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      float g = g0;
      if constexpr (!CLEAN) g = gt[(m_c[c] >> M_VCLASS_SHIFT) & 15];
      float txy = split_vertex_trial(g, (float)(uxs_n[c] - uxs_c[c]), rdY,
                                     (float)(uys_n[c] - uy_n[c]), rdX);
      const bool ok = rowV && colV[c];
      if ((!CLEAN && (m_c[c] & M_VERT_ZERO)) || !ok) txy = 0;
      tXY_c[c] = txy;
    }
    // return mapping
    float cXY_c[2], cP_c[2], cXX_c[2], cYY_c[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      cXY_c[c] = tXY_c[c];
      cP_c[c] = tP_c[c];
      cXX_c[c] = tXX_c[c];
      cYY_c[c] = tYY_c[c];
    }
    float prc0 = 0;
    if (plastic) {
      const float sXX_c[2] = {tXX_c[1], sXX_c1}, sYY_c[2] = {tYY_c[1], sYY_c1};
      prc0 = lane_prev(tXY_c[1]);
      const float pr_p[2] = {prXY_p0, tXY_p[0]}, pr_c[2] = {prc0, tXY_c[0]};
      if constexpr (!EDGE) {
        // M::vertex_rm and M::cell_rm, with the scaling skipped where no lane
        // of the wave yields (the squared invariants decide, as there)
        float s2v[2], s2c[2];
        bool yield = false;
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const float sxx = tXX_c[c] + sXX_c[c] + tXX_n[c] + sXX_n[c];
          const float syy = tYY_c[c] + sYY_c[c] + tYY_n[c] + sYY_n[c];
          s2v[c] = split_s2_vertex(sxx, syy, tXY_c[c]);
          const float av = 0.25f * (pr_p[c] + tXY_p[c] + pr_c[c] + tXY_c[c]);
          s2c[c] = split_s2_cell(cXX_c[c], cYY_c[c], av);
          yield = yield || s2v[c] > c2lo || s2c[c] > c2lo;
        }
        if (__any(yield)) {
#pragma unroll
          for (int c = 0; c < 2; c++) {
            float sc;
            if (rm_scale(s2v[c], coh, c2lo, sc)) cXY_c[c] *= sc;
            if (rm_scale(s2c[c], coh, c2lo, sc)) {
              cXX_c[c] = cXX_c[c] * sc;
              cYY_c[c] = cYY_c[c] * sc;
            }
          }
        }
      } else {
#pragma unroll
        for (int c = 0; c < 2; c++) {
          const float sxx = tXX_c[c] + sXX_c[c] + tXX_n[c] + sXX_n[c];
          const float syy = tYY_c[c] + sYY_c[c] + tYY_n[c] + sYY_n[c];
          if (rowV && colV[c]) cXY_c[c] = M::vertex_rm(tXY_c[c], sxx, syy, coh, c2lo);
        }
        // the clamped vertex blocks of the edge cells (first/second as in
        // MarchStage): one column right and two left of each element
        float nx_p[2], nx_c[2], nx_pp[2], pr_pp[2], pp_p[2], pp_c[2], pp_pp[2];
        if constexpr (EX) {
          nx_p[0] = tXY_p[1];
          nx_p[1] = lane_next(tXY_p[0]);
          nx_c[0] = tXY_c[1];
          nx_c[1] = lane_next(tXY_c[0]);
          nx_pp[0] = tXY_pp[1];
          nx_pp[1] = lane_next(tXY_pp[0]);
          pp_p[0] = lane_prev(tXY_p[0]);
          pp_p[1] = pr_p[0];
          pp_c[0] = lane_prev(tXY_c[0]);
          pp_c[1] = prc0;
          pr_pp[0] = lane_prev(tXY_pp[1]);
          pr_pp[1] = tXY_pp[0];
          pp_pp[0] = lane_prev(tXY_pp[0]);
          pp_pp[1] = pr_pp[0];
        } else {
#pragma unroll
          for (int c = 0; c < 2; c++) nx_p[c] = nx_c[c] = nx_pp[c] = pp_p[c] = pp_c[c] = pp_pp[c] = 0;
          pr_pp[0] = lane_prev(tXY_pp[1]);
          pr_pp[1] = tXY_pp[0];
        }
#pragma unroll
        for (int c = 0; c < 2; c++) {
          auto first = [&](float t, float pr, float pp) {
            return leftE[c] ? t : (rightE[c] ? pp : pr);
          };
          auto second = [&](float t, float pr, float nxv) {
            return leftE[c] ? nxv : (rightE[c] ? pr : t);
          };
          const bool cc = colC[c];
          if (cc && (!EY || (j >= 1 && j < nY - 1))) {
            const float av = 0.25f * (first(tXY_p[c], pr_p[c], pp_p[c]) +
                                      second(tXY_p[c], pr_p[c], nx_p[c]) +
                                      first(tXY_c[c], pr_c[c], pp_c[c]) +
                                      second(tXY_c[c], pr_c[c], nx_c[c]));
            M::cell_rm(cXX_c[c], cYY_c[c], av, coh, c2lo);
          } else if (EY && cc && j == nY - 1 && j >= 1) {
            const float av = 0.25f * (first(tXY_pp[c], pr_pp[c], pp_pp[c]) +
                                      second(tXY_pp[c], pr_pp[c], nx_pp[c]) +
                                      first(tXY_p[c], pr_p[c], pp_p[c]) +
                                      second(tXY_p[c], pr_p[c], nx_p[c]));
            M::cell_rm(cXX_c[c], cYY_c[c], av, coh, c2lo);
          }
          if (EY && j == 1) {
            float xx = tXX_p[c], yy = tYY_p[c];
            if (cc) {
              const float av = 0.25f * (first(tXY_p[c], pr_p[c], pp_p[c]) +
                                        second(tXY_p[c], pr_p[c], nx_p[c]) +
                                        first(tXY_c[c], pr_c[c], pp_c[c]) +
                                        second(tXY_c[c], pr_c[c], nx_c[c]));
              M::cell_rm(xx, yy, av, coh, c2lo);
            }
            cP_p[c] = tP_p[c];
            cYY_p[c] = yy;
          }
        }
      }
    }
    // velocity and displacement of row j
    const float rP[2] = {lane_prev(cP_c[1]), cP_c[0]};
    const float rXX[2] = {lane_prev(cXX_c[1]), cXX_c[0]};
    const float rxc0 = lane_prev(cXY_c[1]);
    const float rXY_c[2] = {rxc0, cXY_c[0]}, rXY_p[2] = {rXY_p0, cXY_p[0]};
    Out o;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int i = i0 + c;
      float vx = vx_in[c], vy = vy_in[c];
      o.updX[c] = (!EX || (i > 0 && i < nX)) && (!EY || (j > 0 && j < nY - 1));
      o.updY[c] = (!EX || (i > 0 && i < nX - 1)) && (!EY || (j > 0 && j < nY));
      if (o.updX[c])
        vx = split_velocity(vx, vdampX, dTrho, -cP_c[c] + rP[c] + cXX_c[c] - rXX[c], rdX,
                            rXY_c[c] - rXY_p[c], rdY);
      if (o.updY[c])
        vy = split_velocity(vy, vdampY, dTrho, -cP_c[c] + cP_p[c] + cYY_c[c] - cYY_p[c], rdY,
                            cXY_p[c] - rXY_p[c], rdX);
      const bool in = colC[c] && (!EY || (j >= 0 && j < nY));
      o.vx[c] = vx;
      o.vy[c] = vy;
      o.ux[c] = in ? split_disp(ux_c[c], vx, dT) : ux_c[c];
      o.uy[c] = in ? split_disp(uy_c[c], vy, dT) : uy_c[c];
    }
    // rotate
#pragma unroll
    for (int c = 0; c < 2; c++) {
      ux_c[c] = ux_n[c];
      uy_c[c] = uy_n[c];
      uy_n[c] = uy_nn[c];
      m_c[c] = m_n[c];
      tP_p[c] = tP_c[c];
      tXX_p[c] = tXX_c[c];
      tYY_p[c] = tYY_c[c];
      tP_c[c] = tP_n[c];
      tXX_c[c] = tXX_n[c];
      tYY_c[c] = tYY_n[c];
      tXY_pp[c] = tXY_p[c];
      tXY_p[c] = tXY_c[c];
      cP_p[c] = cP_c[c];
      cYY_p[c] = cYY_c[c];
      cXY_p[c] = cXY_c[c];
    }
    uxs_c1 = uxs_n[1];
    sXX_c1 = sXX_n[1];
    sYY_c1 = sYY_n[1];
    prXY_p0 = prc0;
    rXY_p0 = rxc0;
    return o;
  }
};

// Stage 0's inputs of a wide segment: StageInput on two columns per lane.
// Interior strips (not EX) read the lane's two columns at consecutive
// offsets; EX strips read each column through a resource bounded to its
// field (lane offsets of columns outside the field are kOob, loads return 0).
template <bool EX, bool EY, bool CLEAN>
struct WideInput {
  struct Ld {
    double ux[2], uy[2];   // Ux row t+1, Uy row t+2
    unsigned m[2];         // mask row t+1
    float vx[2], vy[2];    // V row t
  };
  int nX, nY, sx, i0, tEnd;
  unsigned o8, o4;
  __amdgpu_buffer_rsrc_t bUx, bUy, bVx, bVy, bM, none;
  unsigned oUx[2], oC8[2], oM[2], oVc[2];

  __device__ __forceinline__ WideInput(const RveDev& r, int cur, int vpar, int col0, int tEnd_)
      : nX(r.nX), nY(r.nY), sx(r.nX + 1), i0(col0), tEnd(tEnd_), o8((unsigned)col0 * 8u),
        o4((unsigned)col0 * 4u) {
    const double* Ux = r.Ux[cur];
    const double* Uy = r.Uy[cur];
    const float* Vx = static_cast<const float*>(vpar ? r.Vx2 : r.Vx);
    const float* Vy = static_cast<const float*>(vpar ? r.Vy2 : r.Vy);
    if (!EDGE()) {
      bUx = brsrc(Ux);
      bUy = brsrc(Uy);
      bVx = brsrc(Vx);
      bVy = brsrc(Vy);
      bM = brsrc(r.mask);
      none = bUx;
    } else {
      const unsigned ux = (unsigned)nY * (unsigned)sx, uy = (unsigned)(nY + 1) * (unsigned)nX;
      bUx = brsrc_n(Ux, ux * 8u);
      bUy = brsrc_n(Uy, uy * 8u);
      bVx = brsrc_n(Vx, ux * 4u);
      bVy = brsrc_n(Vy, uy * 4u);
      bM = brsrc_n(r.mask, (unsigned)nY * (unsigned)nX);
      none = brsrc_n(Ux, 0u);
#pragma unroll
      for (int c = 0; c < 2; c++) {
        const int i = col0 + c;
        const bool cx = i >= 0 && i <= nX, cc = i >= 0 && i < nX;
        oUx[c] = cx ? (unsigned)i * 8u : kOob;
        oC8[c] = cc ? (unsigned)i * 8u : kOob;
        oM[c] = cc ? (unsigned)i : kOob;
        oVc[c] = cc ? (unsigned)i * 4u : kOob;
      }
    }
  }
  static constexpr bool EDGE() { return EX || EY; }

  // one row of a field: two columns at element offsets in (EX: each through
  // its own offset)
  __device__ __forceinline__ void rowd(__amdgpu_buffer_rsrc_t b, const unsigned (&oc)[2],
                                       unsigned soff, double (&v)[2]) const {
    if constexpr (!EX) {
      bld2d(b, o8, soff, v);
    } else {
      v[0] = bld<double>(b, oc[0], soff);
      v[1] = bld<double>(b, oc[1], soff);
    }
  }
  __device__ __forceinline__ void rowf(__amdgpu_buffer_rsrc_t b, unsigned soff,
                                       float (&v)[2]) const {
    if constexpr (!EX) {
      bld2f(b, o4, soff, v);
    } else {
      v[0] = bld<float>(b, oVc[0], soff);
      v[1] = bld<float>(b, oVc[1], soff);
    }
  }
  __device__ __forceinline__ void rowm(__amdgpu_buffer_rsrc_t b, unsigned soff,
                                       unsigned (&m)[2]) const {
    if constexpr (CLEAN) {
      m[0] = m[1] = 0;
    } else if constexpr (!EX) {
      m[0] = bld<unsigned char>(b, (unsigned)i0, soff);
      m[1] = bld<unsigned char>(b, (unsigned)i0 + 1u, soff);
    } else {
      m[0] = bld<unsigned char>(b, oM[0], soff);
      m[1] = bld<unsigned char>(b, oM[1], soff);
    }
  }

  // Ux row jb, Uy row jb+1, mask row jb into the stage
  __device__ __forceinline__ void start(WideStage<EX, EY, CLEAN>& st, int jb) const {
    double ux[2], uy[2];
    unsigned m[2];
    const bool rx = !EY || (jb >= 0 && jb < nY), ry = !EY || (jb + 1 >= 0 && jb + 1 <= nY);
    const unsigned sx0 = rx ? (unsigned)jb : 0u, sy1 = ry ? (unsigned)(jb + 1) : 0u;
    rowd(rx ? bUx : none, oUx, sx0 * (unsigned)sx * 8u, ux);
    rowd(ry ? bUy : none, oC8, sy1 * (unsigned)nX * 8u, uy);
    rowm(rx ? bM : none, sx0 * (unsigned)nX, m);
    st.start(ux, uy, m);
  }

  __device__ __forceinline__ Ld fetch(int t0) const {
    const int t = EY ? t0 : min(t0, tEnd);
    Ld d;
    // rows outside a field read through a resource of bound 0 (EY only; the
    // row tests are uniform: scalar selects)
    const bool r1 = !EY || (t + 1 >= 0 && t + 1 < nY), r2 = !EY || (t + 2 >= 0 && t + 2 <= nY),
               r0 = !EY || (t >= 0 && t < nY);
    const unsigned s1 = r1 ? (unsigned)(t + 1) : 0u, s2 = r2 ? (unsigned)(t + 2) : 0u,
                   s0 = r0 ? (unsigned)t : 0u;
    rowd(r1 ? bUx : none, oUx, s1 * (unsigned)sx * 8u, d.ux);
    rowd(r2 ? bUy : none, oC8, s2 * (unsigned)nX * 8u, d.uy);
    rowm(r1 ? bM : none, s1 * (unsigned)nX, d.m);
    rowf(r0 ? bVx : none, s0 * (unsigned)sx * 4u, d.vx);
    rowf(r0 ? bVy : none, s0 * (unsigned)nX * 4u, d.vy);
    return d;
  }
};

// The four-row LDS ring of the wide chains: U and V of both columns of every
// lane, and the two mask bytes (dirty segments).
struct WideRing {
  double ux[4][128];
  double uy[4][128];
  float vx[4][128];
  float vy[4][128];
  unsigned m[4][64];
};

template <bool EX, bool EY, bool CLEAN, int NS>
__device__ __forceinline__ void wide_run(const RveDev& r, int cur, int vpar, int s, int j0,
                                         int j1, const float* gt, int role, WideRing* ring,
                                         int nSteps) {
  using Stage = WideStage<EX, EY, CLEAN>;
  using In = WideInput<EX, EY, CLEAN>;
  using Ld = typename In::Ld;
  const int nX = r.nX, nY = r.nY, sx = nX + 1;
  const int L = threadIdx.x & 63;
  const int i0 = s + 2 * L;
  const int jbA = j0 - 2 * NS, tEnd = j1 + 2 * NS - 3;
  Stage st;
  st.setup(r, gt, i0);
  auto put = [&](WideRing& w, int sl, const typename Stage::Out& o, const unsigned (&m)[2]) {
    *reinterpret_cast<double2*>(&w.ux[sl][2 * L]) = make_double2(o.ux[0], o.ux[1]);
    *reinterpret_cast<double2*>(&w.uy[sl][2 * L]) = make_double2(o.uy[0], o.uy[1]);
    *reinterpret_cast<float2*>(&w.vx[sl][2 * L]) = make_float2(o.vx[0], o.vx[1]);
    *reinterpret_cast<float2*>(&w.vy[sl][2 * L]) = make_float2(o.vy[0], o.vy[1]);
    if constexpr (!CLEAN) w.m[sl][L] = m[0] | (m[1] << 8);
  };
  // Lockstep step q (0 <= q < nSteps, one barrier each) is row jbA+q of
  // stage 0 and row jbA+q-3k of stage k.  Each stage runs its steps in three
  // loops -- idle steps before its first row, its rows, idle steps after its
  // last -- so the loop over its rows carries only the stage's state, and
  // the slot of a step (q & 3) is known at compile time in a loop unrolled
  // by four.
  auto idle = [&](int q0, int q1) {
    for (int q = q0; q < q1; q++) lds_barrier();
  };
  if (role == 0) {
    // stage 0: rows jbA..tEnd (steps 0 .. tEnd-jbA) from HBM into ring 0
    WideRing& out = ring[0];
    const In in(r, cur, vpar, i0, tEnd);
    in.start(st, jbA);
    const int nAct = tEnd - jbA + 1;
    auto stepA = [&](int q, const Ld& ld, auto sl) {
      const auto o = st.step(jbA + q, ld.ux, ld.uy, ld.m, ld.vx, ld.vy);
      put(out, sl, o, ld.m);
      lds_barrier();
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    Ld P = in.fetch(jbA), Q;
    int q = 0;
    for (; q + 4 <= nAct; q += 4) {
      Q = in.fetch(jbA + q + 1);
      stepA(q, P, I0{});
      P = in.fetch(jbA + q + 2);
      stepA(q + 1, Q, I1{});
      Q = in.fetch(jbA + q + 3);
      stepA(q + 2, P, I2{});
      P = in.fetch(jbA + q + 4);
      stepA(q + 3, Q, I3{});
    }
    if (q < nAct) {
      Q = in.fetch(jbA + q + 1);
      stepA(q, P, I0{});
      if (q + 1 < nAct) {
        P = in.fetch(jbA + q + 2);
        stepA(q + 1, Q, I1{});
        if (q + 2 < nAct) stepA(q + 2, P, I2{});
      }
    }
    idle(nAct, nSteps);
    return;
  }
  auto march = [&](auto K) {
    constexpr int k = decltype(K)::value, lag = 3 * k;
    constexpr bool last = k == NS - 1;
    // first step (rows in from the ring) and last step of this stage
    constexpr int qStart = 5 * k - 1;
    const int qLast = tEnd + k - jbA;
    WideRing& in = ring[k - 1];
    WideRing& out = ring[last ? k - 1 : k];  // not written by the last stage
    double* __restrict__ Uxo = r.Ux[cur ^ 1];
    double* __restrict__ Uyo = r.Uy[cur ^ 1];
    float* Vxo = static_cast<float*>(vpar ? r.Vx : r.Vx2);
    float* Vyo = static_cast<float*>(vpar ? r.Vy : r.Vy2);
    const bool outLane = L >= NS && L < 64 - NS;
    const unsigned o8 = (unsigned)i0 * 8u, o4 = (unsigned)i0 * 4u;
    const unsigned vxb = (unsigned)nY * (unsigned)sx * 4u,
                   vyb = (unsigned)(nY + 1) * (unsigned)nX * 4u;
    const __amdgpu_buffer_rsrc_t bUx = brsrc(Uxo), bUy = brsrc(Uyo),
                                 bVx = EX ? brsrc_n(Vxo, vxb) : brsrc(Vxo),
                                 bVy = EX ? brsrc_n(Vyo, vyb) : brsrc(Vyo);
    // step q reads Ux' of row j+1 (slot q-2), Uy' of row j+2 (slot q-1), V'
    // of row j and mask row j+1 (slot q-3); a middle stage writes slot q
    auto stepK = [&](int q, auto ph) {
      constexpr int s1 = (ph + 3) & 3, s2 = (ph + 2) & 3, s3 = (ph + 1) & 3;  // q-1, q-2, q-3
      const int j = jbA + q - lag;
      double uxn[2], uynn[2];
      unsigned mn[2];
      float vxi[2], vyi[2];
      {
        const double2 a = *reinterpret_cast<const double2*>(&in.ux[s2][2 * L]);
        const double2 b = *reinterpret_cast<const double2*>(&in.uy[s1][2 * L]);
        uxn[0] = a.x;
        uxn[1] = a.y;
        uynn[0] = b.x;
        uynn[1] = b.y;
        if constexpr (CLEAN) {
          mn[0] = mn[1] = 0;
        } else {
          const unsigned mm = in.m[s3][L];
          mn[0] = mm & 0xffu;
          mn[1] = mm >> 8;
        }
        const float2 x = *reinterpret_cast<const float2*>(&in.vx[s3][2 * L]);
        const float2 y = *reinterpret_cast<const float2*>(&in.vy[s3][2 * L]);
        vxi[0] = x.x;
        vxi[1] = x.y;
        vyi[0] = y.x;
        vyi[1] = y.y;
      }
      // Ux(i+1, j) of this stage's input, before the step
      const double uxsK[2] = {st.ux_c[1], st.uxs_c1};
      const auto o = st.step(j, uxn, uynn, mn, vxi, vyi);
      if constexpr (!last) {
        put(out, ph, o, mn);
      } else if (j >= j0 && outLane) {
        const unsigned rowx = (unsigned)j * (unsigned)sx, rowy = (unsigned)j * (unsigned)nX;
        if constexpr (!EX) {
          if (!EY || o.updX[0]) bst2f(bVx, o4, rowx * 4u, o.vx);
          if (!EY || o.updY[0]) bst2f(bVy, o4, rowy * 4u, o.vy);
          bst2d(bUx, o8, rowx * 8u, o.ux);
          bst2d(bUy, o8, rowy * 8u, o.uy);
          if (EY && j == nY - 1) {
            // Uy row nY: ComputeDisp never updates it
            const double uyl[2] = {st.uy_c[0], st.uy_c[1]};
            bst2d(bUy, o8, (rowy + (unsigned)nX) * 8u, uyl);
          }
        } else {
#pragma unroll
          for (int c = 0; c < 2; c++) {
            const int i = i0 + c;
            if (i >= nX) continue;
            const unsigned oi = (unsigned)i;
            bst(bVx, o.updX[c] ? oi * 4u : kOob, rowx * 4u, o.vx[c]);
            bst(bVy, o.updY[c] ? oi * 4u : kOob, rowy * 4u, o.vy[c]);
            bst(bUx, oi * 8u, rowx * 8u, o.ux[c]);
            bst(bUy, oi * 8u, rowy * 8u, o.uy[c]);
            // nodes ComputeDisp never updates: Ux column nX, Uy row nY
            if (i == nX - 1) gst(Uxo, (unsigned)(j * sx + i + 1), uxsK[c]);
            if (j == nY - 1) gst(Uyo, (unsigned)((j + 1) * nX + i), st.uy_c[c]);
          }
        }
      }
      lds_barrier();
    };
    idle(0, qStart);
    {
      // first rows: Ux' row jb (slot qStart-2), Uy' row jb+1 (slot
      // qStart-1), mask row jb (slot qStart-3)
      constexpr int ph = qStart & 3;
      constexpr int s1 = (ph + 3) & 3, s2 = (ph + 2) & 3, s3 = (ph + 1) & 3;
      const double2 a = *reinterpret_cast<const double2*>(&in.ux[s2][2 * L]);
      const double2 b = *reinterpret_cast<const double2*>(&in.uy[s1][2 * L]);
      const double ux[2] = {a.x, a.y}, uy[2] = {b.x, b.y};
      unsigned m[2] = {0u, 0u};
      if constexpr (!CLEAN) {
        const unsigned mm = in.m[s3][L];
        m[0] = mm & 0xffu;
        m[1] = mm >> 8;
      }
      st.start(ux, uy, m);
      lds_barrier();
    }
    constexpr int P0 = (qStart + 1) & 3;
    using Q0 = std::integral_constant<int, P0>;
    using Q1 = std::integral_constant<int, (P0 + 1) & 3>;
    using Q2 = std::integral_constant<int, (P0 + 2) & 3>;
    using Q3 = std::integral_constant<int, (P0 + 3) & 3>;
    int q = qStart + 1;
    for (; q + 3 <= qLast; q += 4) {
      stepK(q, Q0{});
      stepK(q + 1, Q1{});
      stepK(q + 2, Q2{});
      stepK(q + 3, Q3{});
    }
    if (q <= qLast) {
      stepK(q, Q0{});
      if (q + 1 <= qLast) {
        stepK(q + 1, Q1{});
        if (q + 2 <= qLast) stepK(q + 2, Q2{});
      }
    }
    idle(qLast + 1, nSteps);
  };
  static_assert(NS == 3 || NS == 4, "wide chains of three or four stages");
  if (role == 1) {
    march(std::integral_constant<int, 1>{});
  } else if (role == 2) {
    march(std::integral_constant<int, 2>{});
  } else {
    if constexpr (NS == 4) march(std::integral_constant<int, 3>{});
  }
}

// One wide segment per workgroup of NS waves (stage k on wave k).  The
// register budget holds four waves per SIMD (at most 128 VGPRs): the
// boundary-strip variants need 127 without spills, and one kernel for every
// variant keeps a launch a single dispatch.
template <int NS>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(4))) void k_wide(const RveDev* __restrict__ rves,
                                                  const int4* __restrict__ segs, int nsegs,
                                                  int parity, int vpar) {
  __shared__ float gt[NS][16];
  __shared__ WideRing ring[NS - 1];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int k = tile_of_block(blockIdx.x, gridDim.x);
  if (k >= nsegs) return;  // the whole workgroup
  const int4 sg = segs[k];
  const RveDev& r = rves[sg.x];
  if (lane < 16) gt[wave][lane] = (float)r.gtab[lane];
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's table is in LDS
  __builtin_amdgcn_wave_barrier();
  const int cur = parity ^ r.cur;
  const int s = sg.y, j0 = sg.z, j1 = sg.w;
  const int nSteps = j1 - j0 + 5 * NS - 3;
  const bool ex = !(s >= 1 && s + 127 <= r.nX - 2);
  const bool ey = !(j0 >= 2 * NS && j1 <= r.nY - (2 * NS - 1));
  // mask rows the segment reads: j0-2NS .. j1+2NS-2, columns s .. s+127
  const bool clean = mask_clear(r, s, s + 127, j0 - 2 * NS, j1 + 2 * NS - 2);
  const float* g = gt[wave];
  if (!ex && !ey) {
    if (clean) wide_run<false, false, true, NS>(r, cur, vpar, s, j0, j1, g, wave, ring, nSteps);
    else wide_run<false, false, false, NS>(r, cur, vpar, s, j0, j1, g, wave, ring, nSteps);
  } else if (!ex) {
    if (clean) wide_run<false, true, true, NS>(r, cur, vpar, s, j0, j1, g, wave, ring, nSteps);
    else wide_run<false, true, false, NS>(r, cur, vpar, s, j0, j1, g, wave, ring, nSteps);
  } else {
    if (clean) wide_run<true, true, true, NS>(r, cur, vpar, s, j0, j1, g, wave, ring, nSteps);
    else wide_run<true, true, false, NS>(r, cur, vpar, s, j0, j1, g, wave, ring, nSteps);
  }
}

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
                       hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  const dim3 blk(64, 4);
  if (e0)
    hipExtLaunchKernelGGL(k_plain_disp, grid2(r.nX, r.nY, blk), blk, 0, s, e0, e1, 0, r, Ux, Uy, f);
  else
    hipLaunchKernelGGL(k_plain_disp, grid2(r.nX, r.nY, blk), blk, 0, s, r, Ux, Uy, f);
  EP2D_HIP(hipGetLastError());
}

bool valid_tile_rows(int ty) { return ty == 6 || ty == 8 || ty == 14 || ty == 30; }

FusedPlan plan_fused(int mode, bool plastic, int ty, int nX, int nY, int device) {
  (void)device;
  if (!valid_tile_rows(ty)) ty = 8;
  FusedPlan pl;
  pl.ty = ty;
  pl.ntx = (nX + 61) / 62;
  pl.ntiles = pl.ntx * ((nY + ty - 1) / ty);
  pl.grid = pl.ntiles;
  pl.blocksPerCU = dispatch_fused<OccFn>(mode, plastic, ty);
  pl.smem = dispatch_fused<SmemFn>(mode, plastic, ty);
  pl.waves = dispatch_fused<NwFn>(mode, plastic, ty);
  return pl;
}

void launch_fused(const RveDev& r, int cur, int mode, bool plastic, const FusedPlan& pl,
                  hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  dispatch_fused<LaunchFn>(mode, plastic, pl.ty, r, cur, pl, s, e0, e1);
  EP2D_HIP(hipGetLastError());
}

void launch_fused_batch(const RveDev* rves, const int4* tiles, int ntiles, int mode, int ty,
                        int parity, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (ntiles <= 0) return;
  dispatch_fused<BatchFn>(mode, true, ty, rves, tiles, ntiles, parity, s, e0, e1);
  EP2D_HIP(hipGetLastError());
}

void launch_fused_persist(const RveDev* rves, const int4* tiles, int ntiles, int mode, int ty,
                          int parity, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (ntiles <= 0) return;
  dispatch_fused<PersistFn>(mode, true, ty, rves, tiles, ntiles, parity, s, e0, e1);
  EP2D_HIP(hipGetLastError());
}

int march_segments(int nX, int nY, int rows, int rve, int4* out) {
  if (rows < 1) rows = 1;
  const int ns = (nX + kMarchCols - 1) / kMarchCols;
  int n = 0;
  for (int j0 = 0; j0 < nY; j0 += rows)
    for (int k = 0; k < ns; k++) {
      if (out) out[n] = make_int4(rve, kMarchCols * k - 2, j0, std::min(j0 + rows, nY));
      n++;
    }
  return n;
}

void launch_march(const RveDev* rves, const int4* segs, int nsegs, int mode, int parity,
                  int vpar, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (nsegs <= 0) return;
  const dim3 grid((nsegs + 3) / 4), block(256);
  auto go = [&](auto kern) {
    if (e0)
      hipExtLaunchKernelGGL(kern, grid, block, 0, s, e0, e1, 0, rves, segs, nsegs, parity, vpar);
    else
      hipLaunchKernelGGL(kern, grid, block, 0, s, rves, segs, nsegs, parity, vpar);
  };
  if (mode == 2) go(k_march<float, float>);  // split: FP32 stress and velocity
  else if (mode == kModeSplitStress64) go(k_march<double, float>);  // split state, FP64 stress
  else go(k_march<double, double>);
  EP2D_HIP(hipGetLastError());
}

int march2_segments(int nX, int nY, int rows, int rve, int4* out) {
  if (rows < 1) rows = 1;
  const int ns = (nX + kMarch2Cols - 1) / kMarch2Cols;
  int n = 0;
  for (int j0 = 0; j0 < nY; j0 += rows)
    for (int k = 0; k < ns; k++) {
      if (out) out[n] = make_int4(rve, kMarch2Cols * k - 4, j0, std::min(j0 + rows, nY));
      n++;
    }
  return n;
}

// Order of a chain work list.  The chain kernel gives XCD x the contiguous
// index range [x q + min(x, r), ...) of an n-entry list (q = n / 8,
// r = n % 8, tile_of_block), and an XCD starts its workgroups in index
// order.  Two measures shorten the end of a launch, where few workgroups
// are left (20 x 1024^2, four stages on two waves):
//  * column-edge segments first in each XCD's share: they run about a third
//    longer than interior ones (123 against 90 us), and started first they
//    leave workgroups of more even length for the end (3% shorter launches;
//    moving the row-edge segments forward as well gains less);
//  * the last twelfth of each share cut into shorter segments, so the last
//    workgroups to start end sooner: halves, and quarters for the last
//    forty-eighth (2.6% and 0.8% more; cutting a sixteenth or an eighth
//    into halves gains less).
// Each XCD's share grows by the pieces added, so the shares stay where the
// kernel looks for them.  Segments write disjoint outputs (a piece with no
// rows is replaced by the whole segment: the same outputs twice), so the
// order and the cuts change no result.  A list that the GPU holds at
// once (n <= resident workgroups) has no last round and is not cut.
// Returns the entries of the ordered list; out may be the input (or null,
// for the count alone).
int chain_work_order(const int4* in, int n, int w, const int2* dims, int resident,
                     int4* out) {
  const int q = n >> 3, r = n & 7, sp = n > resident ? q / 12 : 0;
  const int s4 = n > resident ? q / 48 : 0, s2 = sp - s4, q2 = q + s2 + 3 * s4;
  const int uw = 60 * w + 4;
  if (!out) return n + 8 * (s2 + 3 * s4);
  const std::vector<int4> tmp(in, in + n);
  for (int x = 0; x < 8; x++) {
    const int lo = x * q + std::min(x, r), cnt = q + (x < r ? 1 : 0);
    std::vector<int4> share(tmp.begin() + lo, tmp.begin() + lo + cnt);
    std::stable_partition(share.begin(), share.end(), [&](const int4& a) {
      const int2 d = dims[a.x];
      return !(a.y >= 1 && a.y + uw - 1 <= d.x - 2);  // the column-edge path
    });
    int4* o = out + x * q2 + std::min(x, r);
    for (int k = 0; k < cnt - sp; k++) *o++ = share[(size_t)k];
    for (int k = cnt - sp; k < cnt; k++) {
      const int4 a = share[(size_t)k];
      const int parts = k < cnt - s4 ? 2 : 4;
      for (int p = 0; p < parts; p++) {
        const int b0 = a.z + (a.w - a.z) * p / parts, b1 = a.z + (a.w - a.z) * (p + 1) / parts;
        *o++ = b1 > b0 ? make_int4(a.x, a.y, b0, b1) : a;
      }
    }
  }
  return n + 8 * (s2 + 3 * s4);
}

// Workgroups of k_chain<., ., ns, w> the device holds at once.
int chain_resident(int ns, int w, bool f32) {
  int nb = 0, dev = 0, cus = 0;
  auto occ = [&](auto kern) {
    EP2D_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)kern, 64 * ns * w, 0));
  };
  if (w == 2 && ns == 4) f32 ? occ(k_chain<float, float, 4, 2>) : occ(k_chain<double, double, 4, 2>);
  else if (w == 2) f32 ? occ(k_chain<float, float, 3, 2>) : occ(k_chain<double, double, 3, 2>);
  else if (ns == 4) f32 ? occ(k_chain<float, float, 4, 1>) : occ(k_chain<double, double, 4, 1>);
  else f32 ? occ(k_chain<float, float, 3, 1>) : occ(k_chain<double, double, 3, 1>);
  EP2D_HIP(hipGetDevice(&dev));
  EP2D_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return std::max(nb, 1) * std::max(cus, 1);
}

int chain_segments(int ns_, int nX, int nY, int rows, int rve, int4* out, int w) {
  if (rows < 1) rows = 1;
  const int cols = chain_cols(ns_, w);
  const int ns = (nX + cols - 1) / cols;
  int n = 0;
  for (int j0 = 0; j0 < nY; j0 += rows)
    for (int k = 0; k < ns; k++) {
      if (out) out[n] = make_int4(rve, cols * k - 2 * ns_, j0, std::min(j0 + rows, nY));
      n++;
    }
  return n;
}

void launch_chain(int ns, const RveDev* rves, const int4* segs, int nsegs, int mode, int parity,
                  int vpar, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int w) {
  if (nsegs <= 0) return;
  const dim3 grid(nsegs), block(64 * ns * w);
  auto go = [&](auto kern) {
    if (e0)
      hipExtLaunchKernelGGL(kern, grid, block, 0, s, e0, e1, 0, rves, segs, nsegs, parity, vpar);
    else
      hipLaunchKernelGGL(kern, grid, block, 0, s, rves, segs, nsegs, parity, vpar);
  };
  const bool f32 = mode == 2;  // split: FP32 stress and velocity
  // the split state (FP32 V) advanced with the FP64 arithmetic: the
  // three-iteration chains of three waves and the four-iteration chains of
  // two waves per stage (the shapes the engine's FP64 finish runs)
  const bool s64 = mode == kModeSplitStress64;
  if (s64 && w == 1 && ns == 3) {
    go(k_chain<double, float, 3, 1>);
  } else if (s64 && w == 2 && ns == 4) {
    go(k_chain<double, float, 4, 2>);
  } else if (s64) {
    throw std::invalid_argument("the FP64 finish runs chains of three stages on one wave or "
                                "four stages on two waves");
  } else if (w == 1 && ns == 3) {
    if (f32) go(k_chain<float, float, 3, 1>);
    else go(k_chain<double, double, 3, 1>);
  } else if (w == 1 && ns == 4) {
    if (f32) go(k_chain<float, float, 4, 1>);
    else go(k_chain<double, double, 4, 1>);
  } else if (w == 2 && ns == 3) {
    if (f32) go(k_chain<float, float, 3, 2>);
    else go(k_chain<double, double, 3, 2>);
  } else if (w == 2 && ns == 4) {
    if (f32) go(k_chain<float, float, 4, 2>);
    else go(k_chain<double, double, 4, 2>);
  } else {
    throw std::invalid_argument("chain kernels run 3 or 4 iterations per launch on 1 or 2 "
                                "waves per iteration");
  }
  EP2D_HIP(hipGetLastError());
}

int wide_segments(int ns_, int nX, int nY, int rows, int rve, int4* out) {
  if (rows < 1) rows = 1;
  const int cols = wide_cols(ns_);
  const int ns = (nX + cols - 1) / cols;
  int n = 0;
  for (int j0 = 0; j0 < nY; j0 += rows)
    for (int k = 0; k < ns; k++) {
      if (out) out[n] = make_int4(rve, cols * k - 2 * ns_, j0, std::min(j0 + rows, nY));
      n++;
    }
  return n;
}

void launch_wide(int ns, const RveDev* rves, const int4* segs, int nsegs, int parity, int vpar,
                 hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  if (nsegs <= 0) return;
  const dim3 grid(nsegs), block(64 * ns);
  auto go = [&](auto kern) {
    if (e0)
      hipExtLaunchKernelGGL(kern, grid, block, 0, s, e0, e1, 0, rves, segs, nsegs, parity, vpar);
    else
      hipLaunchKernelGGL(kern, grid, block, 0, s, rves, segs, nsegs, parity, vpar);
  };
  if (ns == 3) go(k_wide<3>);
  else if (ns == 4) go(k_wide<4>);
  else throw std::invalid_argument("wide chains run 3 or 4 iterations per launch");
  EP2D_HIP(hipGetLastError());
}

void launch_mask_map(const unsigned char* mask, int nX, int nY, unsigned char* map,
                     hipStream_t s) {
  const int n = (int)mask_map_bytes(nX, nY);
  hipLaunchKernelGGL(k_mask_map, dim3((n + 255) / 256), dim3(256), 0, s, mask, nX, nY, map,
                     (nX + kMapCols - 1) / kMapCols, (nY + kMapRows - 1) / kMapRows);
  EP2D_HIP(hipGetLastError());
}

void launch_march2(const RveDev* rves, const int4* segs, int nsegs, int mode, int parity,
                   int vpar, hipStream_t s, hipEvent_t e0, hipEvent_t e1, bool pairs) {
  if (nsegs <= 0) return;
  const dim3 grid(pairs ? (nsegs + 1) / 2 : (nsegs + 3) / 4), block(256);
  auto go = [&](auto kern) {
    if (e0)
      hipExtLaunchKernelGGL(kern, grid, block, 0, s, e0, e1, 0, rves, segs, nsegs, parity, vpar);
    else
      hipLaunchKernelGGL(kern, grid, block, 0, s, rves, segs, nsegs, parity, vpar);
  };
  if (pairs) {
    if (mode == 2) go(k_march2p<float, float>);
    else if (mode == kModeSplitStress64) go(k_march2p<double, float>);
    else go(k_march2p<double, double>);
  } else if (mode == 2) {
    go(k_march2<float, float>);  // split: FP32 stress and velocity
  } else {
    go(k_march2<double, double>);
  }
  EP2D_HIP(hipGetLastError());
}

const char* fused_kernel_name(int mode) {
  switch (mode) {
    case 0: return "k_plain_stress + k_plain_j2 + k_plain_plasticity + k_plain_disp";
    case 1: return "k_fused<double, double, true, TY>";
    default: return "k_fused<float, float, true, TY>";
  }
}

void launch_maxabs(const RveDev& r, bool f32, unsigned long long* out, hipStream_t s) {
  const size_t nx = (size_t)(r.nX + 1) * r.nY, ny = (size_t)r.nX * (r.nY + 1);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, out, 3);
  // 16 elements of each array per thread, at most 1024 blocks: the blocks'
  // atomics on the three result words serialise, so fewer is faster
  const int blocks = (int)std::max<size_t>(1, std::min<size_t>((std::max(nx, ny) + 4095) / 4096, 1024));
  const bool vec = ((reinterpret_cast<uintptr_t>(r.Vx) | reinterpret_cast<uintptr_t>(r.Vy)) & 15u) == 0;
  auto go = [&](auto kern, auto* vx, auto* vy) {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, vx, nx, vy, ny, out);
  };
  if (f32) {
    const float* vx = static_cast<const float*>(r.Vx);
    const float* vy = static_cast<const float*>(r.Vy);
    if (vec) go(k_maxabs<float, true>, vx, vy);
    else go(k_maxabs<float, false>, vx, vy);
  } else {
    const double* vx = static_cast<const double*>(r.Vx);
    const double* vy = static_cast<const double*>(r.Vy);
    if (vec) go(k_maxabs<double, true>, vx, vy);
    else go(k_maxabs<double, false>, vx, vy);
  }
  EP2D_HIP(hipGetLastError());
}

void launch_maxabs_multi(const MaxabsJob* hJobs, const MaxabsJob* dJobs, int n, bool f32,
                         unsigned long long* out, hipStream_t s) {
  if (n <= 0) return;
  size_t most = 0;
  bool vec = true;
  for (int k = 0; k < n; k++) {
    most = std::max<size_t>(most, std::max(hJobs[k].nx, hJobs[k].ny));
    vec = vec && ((reinterpret_cast<uintptr_t>(hJobs[k].vx) |
                   reinterpret_cast<uintptr_t>(hJobs[k].vy)) & 15u) == 0;
  }
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(256), 0, s, out, 3 * n);
  // as launch_maxabs for the largest job; the blocks past a smaller job's
  // elements add nothing
  const unsigned bx = (unsigned)std::max<size_t>(1, std::min<size_t>((most + 4095) / 4096, 1024));
  // at most 65535 jobs per launch (the grid's y extent)
  for (int j0 = 0; j0 < n; j0 += 65535) {
    const dim3 grid(bx, (unsigned)std::min(n - j0, 65535));
    const MaxabsJob* jb = dJobs + j0;
    unsigned long long* o = out + 3 * (size_t)j0;
    if (f32) {
      if (vec) hipLaunchKernelGGL((k_maxabs_multi<float, true>), grid, dim3(256), 0, s, jb, o);
      else hipLaunchKernelGGL((k_maxabs_multi<float, false>), grid, dim3(256), 0, s, jb, o);
    } else {
      if (vec) hipLaunchKernelGGL((k_maxabs_multi<double, true>), grid, dim3(256), 0, s, jb, o);
      else hipLaunchKernelGGL((k_maxabs_multi<double, false>), grid, dim3(256), 0, s, jb, o);
    }
  }
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
  // one 16-byte element per thread, no grid-stride loop: the form that
  // reaches the highest copy bandwidth on MI355X
  const size_t blocks = (n2 + 255) / 256;
  if (blocks == 0 || blocks > 0x7fffffffu) return;  // callers bound the size
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const double2*)src,
                     (double2*)dst, n2);
  EP2D_HIP(hipGetLastError());
}

}  // namespace ep2d
