// ep2d_internal.h -- shared between the HIP kernels (ep2d_kernels.hip) and
// the host driver (ep2d_solver.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace ep2d {

// Pore mask bits, one byte per cell (index j*nX + i).
enum : unsigned char {
  M_CELL_MAT = 1,   // SetMaterials pore: K = 0.01 K0, G = 0.01 G0 (EffPlast2D.cu:708-728)
  M_CELL_ZERO = 2,  // ComputeStress cell test: P = tauXX = tauYY = 0 (EffPlast2D.cu:77-87)
  M_VERT_ZERO = 4   // ComputeStress vertex test at vertex (i, j): tauXY = 0 (EffPlast2D.cu:88-95)
};
// Bits 3..6: SetMaterials pore bits of the four cells around vertex (i, j) --
// (i,j), (i+1,j), (i,j+1), (i+1,j+1) -- which select the vertex shear modulus
// 0.25*(G + G + G + G) from a 16-entry table (EffPlast2D.cu:72).
constexpr int M_VCLASS_SHIFT = 3;

struct HipError : std::runtime_error {
  explicit HipError(const std::string& s) : std::runtime_error(s) {}
};

#define EP2D_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::ep2d::HipError(std::string(#call) + ": " + hipGetErrorString(e_) +    \
                             " (" __FILE__ ":" + std::to_string(__LINE__) + ")");   \
  } while (0)

// Everything an iteration kernel needs about one RVE, passed by value.
struct RveDev {
  double* Ux[2];            // displacement, ping-pong in the fused schemes
  double* Uy[2];
  void* Vx;                 // velocity: double (plain, FP64) or float (split)
  void* Vy;
  void* Vx2;                // second velocity buffer of the fused schemes (two-
  void* Vy2;                // iteration launches read one and write the other)
  const unsigned char* mask;
  int nX, nY;
  double dX, dY, dT;
  double rdX, rdY;          // 1/dX, 1/dY (split scheme only)
  double K0, Kp, G0, Gp;    // material outside / inside pores (Kp = 0.01*K0)
  double P0;                // initial pressure (SetInitPressure: 0)
  double coh;               // yield radius pa[8]
  double vdampX, vdampY;    // 1 - dT*dampX, 1 - dT*dampY
  double dTrho;             // dT / rho
  int plastic;              // 1: return mapping on (nTasks > 1, EffPlast2D.cu:392)
  int cur;      // batched path: buffer index of the current displacements
  double coh2lo;            // below coh^2: a smaller squared J2 cannot exceed coh
  float coh2lof;            // split scheme: coh^2 in FP32, the yield test of its FP32 stress
  double gtab[16];          // vertex shear modulus by vertex class (mask bits 3..6)
  const unsigned char* mmap;  // block map of the mask: one byte per kMapRows x kMapCols
                              // cells, nonzero where a mask byte of the block is
};

// Blocks of the mask map (RveDev::mmap), row-major, ceil(nX / kMapCols)
// blocks per row.
constexpr int kMapRows = 8;
constexpr int kMapCols = 32;
inline size_t mask_map_bytes(int nX, int nY) {
  return (size_t)((nX + kMapCols - 1) / kMapCols) * (size_t)((nY + kMapRows - 1) / kMapRows);
}

// Stress fields in HBM (plain scheme; also the finalised fields of the fused
// schemes).
struct StressDev {
  double* P;
  double* tauXX;
  double* tauYY;
  double* tauXY;
  double* tauXYav;
  double* J2;
  double* J2XY;
};

// Launch plan of a fused kernel for one grid: tiles are 62 points wide and
// ty (6, 8, 14 or 30) tall, one workgroup per tile.
struct FusedPlan {
  int ty = 8;
  int ntx = 0;
  int ntiles = 0;
  int grid = 0;
  int blocksPerCU = 0;
  int smem = 0;
  int waves = 0;
};
bool valid_tile_rows(int ty);

// Partial sums of the volume averages: 11 accumulators.
constexpr int NAVG = 11;
enum AvgIdx {
  A_S0 = 0, A_S1, A_S2, A_SP0, A_SP1, A_SP2, A_SP3, A_E0, A_E1, A_E2, A_SXY
};
constexpr int AVG_TX = 64;
constexpr int AVG_TY = 4;

// ---- launchers (ep2d_kernels.hip) ------------------------------------------
void launch_build_mask(unsigned char* mask, int nX, int nY, double dX, double dY,
                       double rad, double nPores, hipStream_t s);
void launch_mask_map(const unsigned char* mask, int nX, int nY, unsigned char* map,
                     hipStream_t s);
void launch_fill(double* p, size_t n, double v, hipStream_t s);
void launch_fill_f32(float* p, size_t n, float v, hipStream_t s);
void launch_affine(double* Ux, double* Uy, int nX, int nY, double dX, double dY,
                   double dUxdx, double dUydy, double dUxdy, double dUydx, hipStream_t s);
// plain scheme (also used to finalise the fused schemes' stress fields)
void launch_plain_stress(const RveDev& r, const double* Ux, const double* Uy,
                         const StressDev& f, hipStream_t s);
void launch_plain_j2(const RveDev& r, const StressDev& f, hipStream_t s);
void launch_plain_plasticity(const RveDev& r, const StressDev& f, hipStream_t s);
void launch_plain_disp(const RveDev& r, double* Ux, double* Uy, const StressDev& f,
                       hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// fused schemes (mode 1: FP64, mode 2: split): reads U[cur], writes U[cur ^ 1],
// updates V in place.  e0/e1 (optional) receive the dispatch's start/end.
// kModeSplitStress64: the split scheme's state (FP64 U, FP32 V) advanced
// with the FP64 arithmetic (tiles of 8 rows; the finish of a split solve
// whose FP32 stress cannot resolve the residual target).
constexpr int kModeSplitStress64 = 3;
FusedPlan plan_fused(int mode, bool plastic, int ty, int nX, int nY, int device);
void launch_fused(const RveDev& r, int cur, int mode, bool plastic, const FusedPlan& pl,
                  hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
const char* fused_kernel_name(int mode);

// Row-marching fused iteration: one wave64 per segment (rve, s, j0, j1), a
// strip of 64 columns from s (60 output columns from s+2) and rows [j0, j1).
constexpr int kMarchCols = 60;
// Segments of one RVE, `rows` rows each (the last one of a strip shorter),
// ordered band by band (all strips of rows [j0, j0+rows) together, so the
// cache lines neighbouring strips share are read while both run); returns the
// count and fills out when it is not null.
int march_segments(int nX, int nY, int rows, int rve, int4* out);
// V is updated in place in Vx/Vy (vpar = 0) or Vx2/Vy2 (vpar = 1).
void launch_march(const RveDev* rves, const int4* segs, int nsegs, int mode, int parity,
                  int vpar, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// Two iterations per launch: strips of 56 output columns; U read from U[cur]
// (cur = parity ^ rve.cur) and written to U[cur^1]; V read from Vx/Vy when
// vpar = 0 (else Vx2/Vy2) and written to the other pair.  pairs: each segment
// on two waves (one per iteration, rows handed over in LDS), else on one.
constexpr int kMarch2Cols = 56;
// The multi-iteration kernels (pairs, chains) address fields through buffer resources with
// 32-bit byte offsets, and edge lanes outside the grid read past a bound set
// to the field's size at offset 2^31: every field must stay under 2 GiB,
// nX, nY <= 16382 (the largest field is (nX+1)(nY+1) doubles at most).
inline bool multi_iter_fits(int nX, int nY) {
  return (long long)(nX + 1) * (long long)(nY + 1) * 8 < (1LL << 31);
}
int march2_segments(int nX, int nY, int rows, int rve, int4* out);
void launch_march2(const RveDev* rves, const int4* segs, int nsegs, int mode, int parity,
                   int vpar, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                   bool pairs = false);
// ns = 3 or 4 iterations per launch on chains of ns stages (one per
// iteration, rows handed over in LDS), each stage on w = 1 or 2 waves side
// by side: strips of 60 w + 4 - 4 ns output columns; U and V buffers as for
// launch_march2.
constexpr int chain_cols(int ns, int w = 1) { return 60 * w + 4 - 4 * ns; }
int chain_segments(int ns, int nX, int nY, int rows, int rve, int4* out, int w = 1);
// The launch order of a chain work list of n segments (stages of w waves;
// dims[rve] = (nX, nY) of the RVE a segment's .x names): column-edge
// segments first in each XCD's share, the last twelfth of each share cut
// into halves and quarters of their height unless the device holds all n at once
// (resident workgroups).  Writes the ordered list to out (which may be in)
// and returns its length; with out null, returns the length only.
int chain_work_order(const int4* in, int n, int w, const int2* dims, int resident, int4* out);
// workgroups of k_chain<., ., ns, w> (FP32 or FP64 stress) the device holds at once
int chain_resident(int ns, int w, bool f32);
void launch_chain(int ns, const RveDev* rves, const int4* segs, int nsegs, int mode, int parity,
                  int vpar, hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                  int w = 1);
// Wide chains (split scheme only): the chains with two columns per lane,
// strips of 128 - 4 ns output columns; same segments order, U and V buffers
// and results as launch_chain.
constexpr int wide_cols(int ns) { return 128 - 4 * ns; }
int wide_segments(int ns, int nX, int nY, int rows, int rve, int4* out);
void launch_wide(int ns, const RveDev* rves, const int4* segs, int nsegs, int parity, int vpar,
                 hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// batched fused iteration over many RVEs: tiles[k] = (rve, i0, j0, parity
// offset); RVE k reads U[parity ^ offset].  All RVEs share mode and tile height.
void launch_fused_batch(const RveDev* rves, const int4* tiles, int ntiles, int mode, int ty,
                        int parity, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr);
// The same tile list through the persistent kernel (one wave of resident
// blocks, next tile's reads in flight during the current tile's arithmetic).
void launch_fused_persist(const RveDev* rves, const int4* tiles, int ntiles, int mode, int ty,
                        int parity, hipStream_t s, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr);
// residual: out[0] = bits of max|Vx|, out[1] = bits of max|Vy|, out[2] = nonfinite count
void launch_maxabs(const RveDev& r, bool f32, unsigned long long* out, hipStream_t s);
// One residual reduction job: an RVE's current velocity fields.
struct MaxabsJob {
  const void* vx;
  const void* vy;
  unsigned long long nx, ny;  // element counts
};
// launch_maxabs for n RVEs in one launch (after zeroing the 3 n result
// words): hJobs is the host copy of the device table dJobs (for the launch
// shape); RVE k's three words land at out[3k].
void launch_maxabs_multi(const MaxabsJob* hJobs, const MaxabsJob* dJobs, int n, bool f32,
                         unsigned long long* out, hipStream_t s);
// volume averages: partial[nblocks*NAVG] then result[NAVG]
int avg_partial_blocks(int nX, int nY);
void launch_averages(const RveDev& r, const StressDev& f, const double* Ux,
                     const double* Uy, double nu0, int perOffX, int perOffY,
                     double* partial, double* result, hipStream_t s);
// gather: out[k] = bases[f][table[k].y], f = table[k].x (+2*ucur for ids 3, 4)
void launch_gather(const double* const* bases, const int2* table, int n, int ucur, double* out,
                   hipStream_t s);
void launch_f32_to_f64(const float* src, double* dst, size_t n, hipStream_t s);
void launch_copy(const double* src, double* dst, size_t n, hipStream_t s);
void launch_f64_to_f32(const double* src, float* dst, size_t n, hipStream_t s);

}  // namespace ep2d
