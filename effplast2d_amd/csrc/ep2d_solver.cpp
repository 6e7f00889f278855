// ep2d_solver.cpp -- host driver of one RVE on one MI355X.
//
// Mirrors EffPlast2D::ComputeEffModuli / ComputeEffParams (EffPlast2D.cu:179-698)
// with device-resident state: the iteration loop, the residual, the affine
// load increment and the FP64 volume averages run on the GPU; only scalars and
// a few hundred sampled values cross PCIe.  The per-step scalar arithmetic
// (boundary means, hole area, analytic and final moduli) is evaluated on the
// host in the reference's operation order (compiled with -ffp-contract=off).
#include "ep2d_solver.h"

#include "ep2d_analytic.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <sstream>

namespace ep2d {

namespace {
double std_max(double a, double b) { return (a < b) ? b : a; }  // std::max semantics

// Tile height of the fused kernels when the caller does not choose one.
// FP64 from 1024^2 on: tiles of 14 rows (16.5 against 19.3 us per
// iteration for 8 rows at 1024^2, one RVE); 8 rows otherwise.
int default_tile_rows(int mode, int nX) {
  return mode == EP2D_MODE_FP64 && nX >= 1024 ? 14 : 8;
}
}  // namespace

// get_sigma_2D.m:3-32 (MATLAB evaluates left to right).
void derive_params(const ep2d_physics& ph, double pa[EP2D_NPARS], int* nXout) {
  const int nX = ph.nX > 0 ? ph.nX : 32 * ph.nGrid;
  if (nX < 8) throw ArgError("grid too small: nX = 32*nGrid must be >= 8");
  if (!(ph.Lx > 0) || !(ph.Ly > 0)) throw ArgError("Lx and Ly must be positive");
  if (ph.nPores < 1) throw ArgError("nPores must be >= 1");
  if (!(ph.porosity >= 0) || !(ph.porosity < 1)) throw ArgError("porosity must be in [0, 1)");
  if (!(ph.K0 > 0) || !(ph.G0 > 0) || !(ph.rho0 > 0) || !(ph.CFL > 0))
    throw ArgError("K0, G0, rho0 and CFL must be positive");
  const double N = (double)ph.nPores;
  const double coh = ph.Y * std::sqrt(2.0);
  const double rad = std::sqrt(ph.porosity * ph.Lx * ph.Ly / (M_PI * N * N));
  const double dX = ph.Lx / (double)(nX - 1);
  const double dY = ph.Ly / (double)(nX - 1);
  const double mn = dX < dY ? dX : dY;
  const double dt = ph.CFL * mn / std::sqrt((ph.K0 + 4.0 * ph.G0 / 3.0) / ph.rho0);
  const double dampX = 4.0 / dt / (double)nX;
  const double dampY = 4.0 / dt / (double)nX;
  const double v[EP2D_NPARS] = {dX, dY, dt, ph.K0, ph.G0, ph.rho0, dampX, dampY, coh, rad, N};
  std::memcpy(pa, v, sizeof(v));
  if (nXout) *nXout = nX;
}

// ---------------------------------------------------------------------------
void* MemPool::take(size_t bytes, bool host) {
  auto& m = host ? host_ : dev_;
  const auto it = m.find(bytes);
  if (it == m.end()) return nullptr;
  void* p = it->second;
  m.erase(it);
  held_ -= bytes;
  return p;
}

void MemPool::give(void* p, size_t bytes, bool host) {
  (host ? host_ : dev_).emplace(bytes, p);
  held_ += bytes;
}

void MemPool::trim() {
  for (auto& kv : dev_) (void)hipFree(kv.second);
  for (auto& kv : host_) (void)hipHostFree(kv.second);
  dev_.clear();
  host_.clear();
  held_ = 0;
}

// Device buffer of `bytes` (at least 256): from the pool when it holds one of
// that size, else hipMalloc (after trimming the pool if memory runs short).
// Buffers are not cleared: every user initialises what it reads.
void* Solver::alloc(size_t bytes) {
  bytes = std::max<size_t>(bytes, 256);
  void* p = pool_ ? pool_->take(bytes, false) : nullptr;
  if (!p) {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess && pool_ && pool_->heldBytes() > 0) {
      (void)hipGetLastError();
      pool_->trim();
      e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw StatusError(EP2D_ERR_MEMORY, std::string("hipMalloc of ") + std::to_string(bytes) +
                                             " bytes failed: " + hipGetErrorString(e));
    }
  }
  allocs_.push_back({p, bytes});
  devBytes_ += bytes;
  return p;
}

// Pinned host buffer, pooled like the device buffers.
void* Solver::hostAlloc(size_t bytes) {
  void* p = pool_ ? pool_->take(bytes, true) : nullptr;
  if (!p) EP2D_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
  hostAllocs_.push_back({p, bytes});
  return p;
}

template <typename T>
void Solver::h2d(T* dst, const T* src, size_t n) {
  EP2D_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, stream_));
  EP2D_HIP(hipStreamSynchronize(stream_));
}

template <typename T>
void Solver::d2h(T* dst, const T* src, size_t n) {
  EP2D_HIP(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, stream_));
  EP2D_HIP(hipStreamSynchronize(stream_));
}

// EffPlast2D::EffPlast2D, EffPlast2D.cu:1541-1612.
Solver::Solver(const double pa[EP2D_NPARS], int nX, int nY, const ep2d_options& opt,
               hipStream_t stream, MemPool* pool)
    : nX_(nX), nY_(nY), opt_(opt), pool_(pool) {
  try {
    init(pa, stream);
  } catch (...) {
    release();
    throw;
  }
}

void Solver::init(const double pa[EP2D_NPARS], hipStream_t stream) {
  const int nX = nX_, nY = nY_;
  const ep2d_options opt = opt_;
  std::memcpy(pa_, pa, sizeof(pa_));
  if (opt.logPath) logPath_ = opt.logPath;
  opt_.logPath = nullptr;
  if (opt_.checkInterval <= 0) opt_.checkInterval = 10000;
  mode_ = opt_.mode;
  if (nX < 8 || nY < 8) throw ArgError("nX and nY must be >= 8");
  if (nX != nY) throw ArgError("the RVE grid must be square (nX == nY), as in the reference");
  // the fused kernels address a field with 32-bit byte offsets
  if ((long)(nX + 1) * (long)(nY + 1) > (1L << 29))
    throw ArgError("grid too large: every field must stay under 4 GiB (nX, nY <= 23169)");
  if (opt_.itersPerLaunch < 0 || opt_.itersPerLaunch > 4)
    throw ArgError("itersPerLaunch must be 1, 2, 3 or 4");
  if (opt_.itersPerLaunch >= 2 && opt_.marchRows <= 0)
    throw ArgError("itersPerLaunch 2 to 4 needs the row-marching kernel (marchRows > 0)");
  if (opt_.itersPerLaunch >= 2 && !multi_iter_fits(nX, nY))
    throw ArgError("itersPerLaunch 2 to 4 needs every field under 2 GiB (nX, nY <= 16382)");
  if (opt_.laneColumns < 0 || opt_.laneColumns > 2) throw ArgError("laneColumns must be 0, 1 or 2");
  if (opt_.stageWaves < 0 || opt_.stageWaves > 2) throw ArgError("stageWaves must be 0, 1 or 2");
  if (mode_ < EP2D_MODE_PLAIN || mode_ > EP2D_MODE_SPLIT) throw ArgError("unknown mode");
  if (opt_.nTasks < 1) throw ArgError("nTasks must be >= 1");
  if (opt_.nIter < 1) throw ArgError("nIter must be >= 1");
  if (opt_.nTimeSteps < 1) throw ArgError("nTimeSteps must be >= 1");
  dX_ = pa_[EP2D_PA_DX];
  dY_ = pa_[EP2D_PA_DY];
  dT_ = pa_[EP2D_PA_DT];
  K0_ = pa_[EP2D_PA_K0];
  G0_ = pa_[EP2D_PA_G0];
  rho_ = pa_[EP2D_PA_RHO];
  dampX_ = pa_[EP2D_PA_DAMPX];
  dampY_ = pa_[EP2D_PA_DAMPY];
  coh_ = pa_[EP2D_PA_COH];
  rad_ = pa_[EP2D_PA_RAD];
  nPores_ = pa_[EP2D_PA_NPORES];
  if (!(dX_ > 0) || !(dY_ > 0) || !(dT_ > 0) || !(rho_ > 0))
    throw ArgError("pa: dX, dY, dt and rho0 must be positive");
  if (!(nPores_ >= 1) || nPores_ != std::floor(nPores_) || nPores_ > 64)
    throw ArgError("pa: nPores must be an integer in [1, 64]");
  if (!(rad_ >= 0)) throw ArgError("pa: rad must be >= 0");
  E0_ = 9.0 * K0_ * G0_ / (3.0 * K0_ + G0_);                  // :1568
  nu0_ = (1.5 * K0_ - G0_) / (3.0 * K0_ + G0_);               // :1569
  Y_ = pa_[EP2D_PA_COH] / std::sqrt(2.0);                     // :1572
  const double lX = (nX_ - 1) * dX_, lY = (nY_ - 1) * dY_;    // :1609-1610
  porosity_ = 3.1415926 * rad_ * rad_ * nPores_ * nPores_ / (lX * lY);  // :1611

  device_ = opt_.device;
  int ndev = 0;
  EP2D_HIP(hipGetDeviceCount(&ndev));
  if (device_ < 0 || device_ >= ndev) throw ArgError("device index out of range");
  EP2D_HIP(hipSetDevice(device_));
  if (stream) {
    stream_ = stream;
    ownStream_ = false;
  } else {
    EP2D_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  }

  const size_t nC = (size_t)nX * nY, nV = (size_t)(nX - 1) * (nY - 1);
  const size_t nUx = (size_t)(nX + 1) * nY, nUy = (size_t)nX * (nY + 1);
  const bool fused = mode_ != EP2D_MODE_PLAIN;
  const size_t vbytes = mode_ == EP2D_MODE_SPLIT ? sizeof(float) : sizeof(double);
  r_.Ux[0] = (double*)alloc(nUx * 8);
  r_.Uy[0] = (double*)alloc(nUy * 8);
  r_.Ux[1] = fused ? (double*)alloc(nUx * 8) : r_.Ux[0];
  r_.Uy[1] = fused ? (double*)alloc(nUy * 8) : r_.Uy[0];
  r_.Vx = alloc(nUx * vbytes);
  r_.Vy = alloc(nUy * vbytes);
  // second velocity buffers: multi-iteration launches read one and write the
  // other; boundary velocities (never updated) are kept equal in both
  r_.Vx2 = fused ? alloc(nUx * vbytes) : r_.Vx;
  r_.Vy2 = fused ? alloc(nUy * vbytes) : r_.Vy;
  r_.mask = (unsigned char*)alloc(nC);
  r_.mmap = (unsigned char*)alloc(mask_map_bytes(nX, nY));
  f_.P = (double*)alloc(nC * 8);
  f_.tauXX = (double*)alloc(nC * 8);
  f_.tauYY = (double*)alloc(nC * 8);
  f_.tauXYav = (double*)alloc(nC * 8);
  f_.J2 = (double*)alloc(nC * 8);
  f_.tauXY = (double*)alloc(nV * 8);
  f_.J2XY = (double*)alloc(nV * 8);
  dMax_ = (unsigned long long*)alloc(4 * sizeof(unsigned long long));
  dAvgPartial_ = (double*)alloc((size_t)avg_partial_blocks(nX, nY) * NAVG * 8);
  dAvg_ = (double*)alloc(NAVG * 8);
  hMax_ = (unsigned long long*)hostAlloc(4 * sizeof(unsigned long long));
  hAvg_ = (double*)hostAlloc(NAVG * sizeof(double));

  r_.nX = nX;
  r_.nY = nY;
  r_.dX = dX_;
  r_.dY = dY_;
  r_.dT = dT_;
  r_.rdX = 1.0 / dX_;
  r_.rdY = 1.0 / dY_;
  r_.K0 = K0_;
  r_.Kp = 0.01 * K0_;                  // :721
  r_.G0 = G0_;
  r_.Gp = 0.01 * G0_;                  // :722
  r_.P0 = 0.0;                         // SetInitPressure, :770-783
  r_.coh = coh_;
  r_.plastic = opt_.nTasks > 1 ? 1 : 0;
  r_.cur = 0;
  r_.vdampX = 1.0 - dT_ * dampX_;      // :19
  r_.vdampY = 1.0 - dT_ * dampY_;      // :26
  r_.dTrho = dT_ / rho_;               // :19
  // coh2lo <= coh^2 (real): s <= coh2lo implies RN(sqrt(s)) <= coh, i.e. no yield
  r_.coh2lo = (coh_ * coh_) * (1.0 - 0x1p-50);
  // split scheme: yield when the FP32 squared invariant exceeds coh^2 in FP32
  {
    const float cf = (float)coh_;
    r_.coh2lof = cf * cf;
  }
  // vertex shear modulus 0.25*(G[i,j] + G[i+1,j] + G[i,j+1] + G[i+1,j+1]) (:72)
  // for every in-pore pattern of the four cells, summed in the reference order
  for (int c = 0; c < 16; c++) {
    auto g = [&](int bit) { return ((c >> bit) & 1) ? r_.Gp : r_.G0; };
    r_.gtab[c] = 0.25 * (g(0) + g(1) + g(2) + g(3));
  }

  // SetMaterials / SetMatrixZero (:708-794): masks, zero fields.
  launch_build_mask(const_cast<unsigned char*>(r_.mask), nX, nY, dX_, dY_, rad_, nPores_, stream_);
  launch_mask_map(r_.mask, nX, nY, const_cast<unsigned char*>(r_.mmap), stream_);
  for (int k = 0; k < 2; k++) {
    launch_fill(r_.Ux[k], nUx, 0.0, stream_);
    launch_fill(r_.Uy[k], nUy, 0.0, stream_);
  }
  for (int k = 0; k < (fused ? 2 : 1); k++) {
    void* vx = k ? r_.Vx2 : r_.Vx;
    void* vy = k ? r_.Vy2 : r_.Vy;
    if (mode_ == EP2D_MODE_SPLIT) {
      launch_fill_f32((float*)vx, nUx, 0.0f, stream_);
      launch_fill_f32((float*)vy, nUy, 0.0f, stream_);
    } else {
      launch_fill((double*)vx, nUx, 0.0, stream_);
      launch_fill((double*)vy, nUy, 0.0, stream_);
    }
  }
  for (double* p : {f_.P, f_.tauXX, f_.tauYY, f_.tauXYav, f_.J2}) launch_fill(p, nC, 0.0, stream_);
  for (double* p : {f_.tauXY, f_.J2XY}) launch_fill(p, nV, 0.0, stream_);
  buildGatherTable();
  if (fused && opt_.tileRows == 0 && opt_.marchRows == 0 && opt_.itersPerLaunch == 0) {
    // automatic shape for one RVE, from one-RVE measurements: both schemes
    // from 1024^2 on four-stage chains with two waves per stage (four
    // iterations per launch, in the launch order of chain_work_order): at
    // 1024^2 the shortest segments whose workgroups the device holds in one
    // round (14 rows for split, 21 for FP64 on 256 CUs: 7.4 and 15.6 us per
    // iteration; 16 split rows take 8.6, 20 FP64 rows 17.3 and the FP64
    // tile kernel 16.7), 48 rows at 2048^2 and 96 beyond (split 16.4 and
    // 57.6 us, FP64 43.0 and 124.3 against 47.6 and 154.7 on wave pairs);
    // the tile kernel below 1024^2.
    const int n = std::max(nX, nY);
    const bool chains = (mode_ == EP2D_MODE_SPLIT || mode_ == EP2D_MODE_FP64) && n >= 1024 &&
                        multi_iter_fits(nX, nY) &&
                        !(mode_ == EP2D_MODE_FP64 && opt_.laneColumns == 2);
    if (chains) {
      int rows = n <= 2048 ? 48 : 96;
      if (n <= 1024) {
        const int cols = chain_cols(4, 2);
        const int strips = (nX + cols - 1) / cols;
        const int bands = std::max(1, chain_resident(4, 2, mode_ == EP2D_MODE_SPLIT) / strips);
        rows = std::max(8, (nY + bands - 1) / bands);
      }
      opt_.marchRows = rows;
      opt_.itersPerLaunch = 4;
      if (opt_.laneColumns != 2 && opt_.stageWaves == 0) opt_.stageWaves = 2;
    }
  }
  if (fused) {
    int ty = opt_.tileRows;
    if (ty != 0 && !valid_tile_rows(ty)) throw ArgError("tileRows must be 0, 6, 8, 14 or 30");
    if (ty == 0) ty = default_tile_rows(mode_, nX);
    fplan_ = plan_fused(mode_, opt_.nTasks > 1, ty, nX, nY, device_);
    if (mode_ == EP2D_MODE_SPLIT)
      fplan64_ = plan_fused(kModeSplitStress64, opt_.nTasks > 1, 8, nX, nY, device_);
    if (opt_.marchRows < 0) throw ArgError("marchRows must be >= 0");
    if (opt_.itersPerLaunch < 0 || opt_.itersPerLaunch > 4)
      throw ArgError("itersPerLaunch must be 1, 2, 3 or 4");
    if (opt_.itersPerLaunch >= 2 && opt_.marchRows <= 0)
      throw ArgError("itersPerLaunch 2 to 4 needs the row-marching kernel (marchRows > 0)");
    ipl_ = opt_.itersPerLaunch >= 2 ? opt_.itersPerLaunch : 1;
    multiIter_ = ipl_ > 1;
    if (opt_.laneColumns == 2 && (mode_ != EP2D_MODE_SPLIT || ipl_ < 3))
      throw ArgError("laneColumns 2 (wide chains) needs the split scheme and itersPerLaunch 3 or 4");
    wide_ = opt_.laneColumns == 2;
    if (opt_.stageWaves == 2 && (ipl_ < 3 || wide_))
      throw ArgError("stageWaves 2 needs itersPerLaunch 3 or 4 with one column per lane");
    stageWaves_ = opt_.stageWaves == 2 ? 2 : 1;
    if (opt_.marchRows > 0) {
      // row-marching kernel: the parameter table and the segment list live
      // on the device (the kernel reads the current buffer as parity ^ cur)
      marchRows_ = opt_.marchRows;
      nSegs_ = march_segments(nX, nY, marchRows_, 0, nullptr);
      std::vector<int4> segs((size_t)nSegs_);
      march_segments(nX, nY, marchRows_, 0, segs.data());
      dRve_ = (RveDev*)alloc(sizeof(RveDev));
      dSegs_ = (int4*)alloc(segs.size() * sizeof(int4));
      RveDev rd = r_;
      rd.cur = 0;
      EP2D_HIP(hipMemcpyAsync(dRve_, &rd, sizeof(RveDev), hipMemcpyHostToDevice, stream_));
      EP2D_HIP(hipMemcpyAsync(dSegs_, segs.data(), segs.size() * sizeof(int4),
                              hipMemcpyHostToDevice, stream_));
      if (multiIter_) {
        auto segsOf = [&](int4* out) {
          if (ipl_ == 2) return march2_segments(nX, nY, marchRows_, 0, out);
          if (wide_) return wide_segments(ipl_, nX, nY, marchRows_, 0, out);
          return chain_segments(ipl_, nX, nY, marchRows_, 0, out, stageWaves_);
        };
        const int raw = segsOf(nullptr);
        const int2 dims = make_int2(nX, nY);
        const bool ordered = ipl_ >= 3 && !wide_;
        const int res = ordered ? chain_resident(ipl_, stageWaves_, mode_ == EP2D_MODE_SPLIT) : 0;
        nSegs2_ = ordered ? chain_work_order(nullptr, raw, stageWaves_, &dims, res, nullptr) : raw;
        std::vector<int4> s2((size_t)nSegs2_);
        segsOf(s2.data());
        if (ordered) chain_work_order(s2.data(), raw, stageWaves_, &dims, res, s2.data());
        dSegs2_ = (int4*)alloc(s2.size() * sizeof(int4));
        EP2D_HIP(hipMemcpyAsync(dSegs2_, s2.data(), s2.size() * sizeof(int4),
                                hipMemcpyHostToDevice, stream_));
        EP2D_HIP(hipStreamSynchronize(stream_));
      }
      if (ipl_ >= 3) {
        std::vector<int4> c3, c2((size_t)march2_segments(nX, nY, 26, 0, nullptr));
        march2_segments(nX, nY, 26, 0, c2.data());
        if (ipl_ == 4) {
          const int raw = chain_segments(3, nX, nY, 42, 0, nullptr, 1);
          const int2 dims = make_int2(nX, nY);
          const int res = chain_resident(3, 1, mode_ == EP2D_MODE_SPLIT);
          c3.resize((size_t)chain_work_order(nullptr, raw, 1, &dims, res, nullptr));
          chain_segments(3, nX, nY, 42, 0, c3.data(), 1);
          chain_work_order(c3.data(), raw, 1, &dims, res, c3.data());
        }
        nSegsC3_ = (int)c3.size();
        nSegsC2_ = (int)c2.size();
        if (nSegsC3_ > 0) {
          dSegsC3_ = (int4*)alloc(c3.size() * sizeof(int4));
          EP2D_HIP(hipMemcpyAsync(dSegsC3_, c3.data(), c3.size() * sizeof(int4),
                                  hipMemcpyHostToDevice, stream_));
        }
        dSegsC2_ = (int4*)alloc(c2.size() * sizeof(int4));
        EP2D_HIP(hipMemcpyAsync(dSegsC2_, c2.data(), c2.size() * sizeof(int4),
                                hipMemcpyHostToDevice, stream_));
      }
      EP2D_HIP(hipStreamSynchronize(stream_));
    }
  }
  EP2D_HIP(hipStreamSynchronize(stream_));

  if (!logPath_.empty()) log_.open(logPath_, std::ios_base::app);   // :1607
}

Solver::~Solver() { release(); }

void Solver::release() {
  // teardown: errors are ignored, there is no caller to report them to
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& e : pending_) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  for (auto& e : freeEvents_) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  pending_.clear();
  pendingIters_.clear();
  freeEvents_.clear();
  // the stream is idle: the blocks can go back to the pool
  for (auto& a : allocs_) {
    if (pool_) pool_->give(a.first, a.second, false);
    else (void)hipFree(a.first);
  }
  allocs_.clear();
  for (auto& a : hostAllocs_) {
    if (pool_) pool_->give(a.first, a.second, true);
    else (void)hipHostFree(a.first);
  }
  hostAllocs_.clear();
  hMax_ = nullptr;
  hAvg_ = nullptr;
  hGather_ = nullptr;
  if (stream_ && ownStream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
}

// Entries of the gather table buildGatherTable() makes (same loops, counted).
static size_t gather_entries(long nX, long nY, double nPores) {
  // init() rejects pore counts outside [1, 64]; count within that range
  if (!(nPores >= 1.0)) nPores = 1.0;
  if (nPores > 64.0) nPores = 64.0;
  size_t n = 3 * (size_t)(2 * (nX - 2) + 2 * (nY - 2));
  if (nPores > 2) {
    const long nP = (long)nPores;
    long c = 0;
    for (long i = nX / nP; i < nX * (nP - 1) / nP; i++) c++;
    n += 3 * 2 * (size_t)c;
    c = 0;
    for (long j = nY / nP; j < nY * (nP - 1) / nPores; j++) c++;
    n += 3 * 2 * (size_t)c;
  }
  return n + 20 * (size_t)nPores * (size_t)nPores;
}

// Device bytes init() allocates for an RVE on the one-iteration shapes the
// batched engine uses (no segment tables): the sum of alloc() sizes, each at
// least 256 bytes, in the same order.
size_t Solver::plannedDeviceBytes(int nX, int nY, int mode, double nPores) {
  auto a = [](size_t b) { return std::max<size_t>(b, 256); };
  const size_t nC = (size_t)nX * nY, nV = (size_t)(nX - 1) * (nY - 1);
  const size_t nUx = (size_t)(nX + 1) * nY, nUy = (size_t)nX * (nY + 1);
  const bool fused = mode != EP2D_MODE_PLAIN;
  const size_t vb = mode == EP2D_MODE_SPLIT ? sizeof(float) : sizeof(double);
  size_t t = a(nUx * 8) + a(nUy * 8);
  if (fused) t += a(nUx * 8) + a(nUy * 8);
  t += a(nUx * vb) + a(nUy * vb);
  if (fused) t += a(nUx * vb) + a(nUy * vb);
  t += a(nC) + a(mask_map_bytes(nX, nY)) + 5 * a(nC * 8) + 2 * a(nV * 8);
  t += a(4 * sizeof(unsigned long long)) + a((size_t)avg_partial_blocks(nX, nY) * NAVG * 8) +
       a(NAVG * 8);
  const size_t g = gather_entries(nX, nY, nPores);
  t += a(g * sizeof(int2)) + a(g * sizeof(double)) + a(8 * sizeof(double*));
  return t;
}

// Gather table of the sampled values the host post-processing needs:
// boundary rows/columns of P, tauXX, tauYY (EffPlast2D.cu:825-882, 925-942)
// and the pore-edge displacement windows (:549-594).
void Solver::buildGatherTable() {
  const long nX = nX_, nY = nY_;
  std::vector<int2> tab;
  auto cellLine = [&](Line& L, const std::vector<long>& flat) {
    L.offset = (int)tab.size();
    L.count = (int)flat.size();
    for (long q : flat)
      for (int f = 0; f < 3; f++) tab.push_back(make_int2(f, (int)q));  // P, tauXX, tauYY
  };
  std::vector<long> v;
  v.clear(); for (long i = 1; i < nX - 1; i++) v.push_back(0 * nX + i);            cellLine(hx0_, v);
  v.clear(); for (long i = 1; i < nX - 1; i++) v.push_back((nY - 1) * nX + i);     cellLine(hxN_, v);
  v.clear(); for (long j = 1; j < nY - 1; j++) v.push_back(j * nX + 0);            cellLine(hy0_, v);
  v.clear(); for (long j = 1; j < nY - 1; j++) v.push_back(j * nX + nY - 1);       cellLine(hyN_, v);
  if (nPores_ > 2) {
    const long nP = (long)nPores_;
    v.clear(); for (long i = nX / nP; i < nX * (nP - 1) / nP; i++) v.push_back(nY / nP * nX + i);            cellLine(px0_, v);
    v.clear(); for (long i = nX / nP; i < nX * (nP - 1) / nP; i++) v.push_back(nY * (nP - 1) / nP * nX + i); cellLine(pxN_, v);
    v.clear(); for (long j = nY / nP; j < nY * (nP - 1) / nPores_; j++) v.push_back(j * nX + nX / nP);       cellLine(py0_, v);
    v.clear(); for (long j = nY / nP; j < nY * (nP - 1) / nPores_; j++) v.push_back(j * nX + nX * (nP - 2) / nP); cellLine(pyN_, v);
  }
  // pore-edge windows, :552-594 (size_t truncation as in the reference)
  poreOffset_ = (int)tab.size();
  samplesValid_ = true;
  sampleInPore_.clear();
  const size_t nUx = (size_t)(nX + 1) * nY, nUy = (size_t)nX * (nY + 1);
  auto inPore = [&](bool isUy, long i, long j) { return uNodeInPore(isUy, i, j); };
  auto addSample = [&](bool isUy, size_t flat) {
    const size_t n = isUy ? nUy : nUx;
    const size_t m = isUy ? (size_t)nX : (size_t)(nX + 1);
    if (flat >= n) {
      samplesValid_ = false;
      flat = 0;
    }
    tab.push_back(make_int2(isUy ? 4 : 3, (int)flat));
    sampleInPore_.push_back(inPore(isUy, (long)(flat % m), (long)(flat / m)) ? 1 : 0);
  };
  for (int k = 0; k < nPores_; k++) {
    for (int l = 0; l < nPores_; l++) {
      const double cxdX = 0.5 * (nX - 1) * (1.0 - 1.0 / nPores_) - ((double)(nX - 1) / nPores_) * k;
      const double cydY = 0.5 * (nY - 1) * (1.0 - 1.0 / nPores_) - ((double)(nY - 1) / nPores_) * l;
      const size_t cyIdx = (size_t)(cydY + 0.5 * (nY - 1));
      size_t rxIdx = (size_t)(cxdX - rad_ / dX_ + 0.5 * nX);
      for (int i = 0; i < 5; i++) addSample(false, cyIdx * (nX + 1) + rxIdx - 1 + i);
      rxIdx = (size_t)(cxdX + rad_ / dX_ + 0.5 * nX);
      for (int i = 0; i < 5; i++) addSample(false, cyIdx * (nX + 1) + rxIdx - 2 + i);
      const size_t cxIdx = (size_t)(cxdX + 0.5 * (nX - 1));
      size_t ryIdx = (size_t)(cydY - rad_ / dY_ + 0.5 * nY);
      for (int j = 0; j < 5; j++) addSample(true, (ryIdx - 1 + j) * nX + cxIdx);
      ryIdx = (size_t)(cydY + rad_ / dY_ + 0.5 * nY);
      for (int j = 0; j < 5; j++) addSample(true, (ryIdx - 2 + j) * nX + cxIdx);
    }
  }
  nGather_ = (int)tab.size();
  dTable_ = (int2*)alloc(tab.size() * sizeof(int2));
  dGather_ = (double*)alloc(tab.size() * sizeof(double));
  dBases_ = (const double**)alloc(8 * sizeof(double*));
  hGather_ = (double*)hostAlloc(std::max<size_t>(tab.size(), 1) * sizeof(double));
  const double* bases[8] = {f_.P, f_.tauXX, f_.tauYY, r_.Ux[0], r_.Uy[0], r_.Ux[1], r_.Uy[1], nullptr};
  h2d(dTable_, tab.data(), tab.size());
  h2d(dBases_, bases, 8);
}

// ---- iteration ------------------------------------------------------------
void Solver::launchIteration() {
  const bool plastic = opt_.nTasks > 1;   // EffPlast2D.cu:392
  const bool timed = sampleEvery_ > 0 && (launchCount_ % sampleEvery_) == 0;
  std::pair<hipEvent_t, hipEvent_t> ev{};
  if (timed) {
    if (freeEvents_.empty()) {
      hipEvent_t a, b;
      EP2D_HIP(hipEventCreate(&a));
      EP2D_HIP(hipEventCreate(&b));
      freeEvents_.push_back({a, b});
    }
    ev = freeEvents_.back();
    freeEvents_.pop_back();
  }
  // timed launches carry their own start/stop events (hipExtLaunchKernelGGL),
  // which record the dispatch's execution window
  if (mode_ == EP2D_MODE_PLAIN) {
    // EffPlast2D.cu:390-399 (the reference synchronises the device after each
    // launch; stream order gives the same dependencies).  A timed iteration
    // is bracketed by events on the stream around all four kernels.
    if (timed) EP2D_HIP(hipEventRecord(ev.first, stream_));
    launch_plain_stress(r_, r_.Ux[0], r_.Uy[0], f_, stream_);
    if (plastic) {
      launch_plain_j2(r_, f_, stream_);
      launch_plain_plasticity(r_, f_, stream_);
    }
    launch_plain_disp(r_, r_.Ux[0], r_.Uy[0], f_, stream_);
    if (timed) EP2D_HIP(hipEventRecord(ev.second, stream_));
  } else {
    if (stress64_)
      launch_fused(r_, cur_, kModeSplitStress64, plastic, fplan64_, stream_,
                   timed ? ev.first : nullptr, timed ? ev.second : nullptr);
    else if (marchRows_ > 0)
      launch_march(dRve_, dSegs_, nSegs_, mode_, cur_, vIdx_, stream_,
                   timed ? ev.first : nullptr, timed ? ev.second : nullptr);
    else
      launch_fused(r_, cur_, mode_, plastic, fplan_, stream_, timed ? ev.first : nullptr,
                   timed ? ev.second : nullptr);
    cur_ ^= 1;
    stressStale_ = true;
  }
  if (timed) {
    pending_.push_back(ev);
    pendingIters_.push_back(1);
    if (pending_.size() >= 256) resolveTiming(false);
  }
  launchCount_++;
}

std::pair<hipEvent_t, hipEvent_t> Solver::takeTimingEvents() {
  if (freeEvents_.empty()) {
    hipEvent_t a, b;
    EP2D_HIP(hipEventCreate(&a));
    EP2D_HIP(hipEventCreate(&b));
    freeEvents_.push_back({a, b});
  }
  auto ev = freeEvents_.back();
  freeEvents_.pop_back();
  return ev;
}

// ipl_ iterations in one launch (each segment on a pair of waves or a chain
// of ipl_ waves, one iteration each); V is read from buffer vIdx_ ^ (q & 1)
// and written to the other, so after an odd number of such launches the
// other buffer is current (flipV).
void Solver::launchMulti(int q) {
  const bool timed = sampleEvery_ > 0 && (launchCount_ % sampleEvery_) == 0;
  std::pair<hipEvent_t, hipEvent_t> ev{};
  if (timed) ev = takeTimingEvents();
  if (ipl_ == 2)
    launch_march2(dRve_, dSegs2_, nSegs2_, iterMode(), cur_, vIdx_ ^ (q & 1), stream_,
                  timed ? ev.first : nullptr, timed ? ev.second : nullptr, true);
  else if (wide_)
    launch_wide(ipl_, dRve_, dSegs2_, nSegs2_, cur_, vIdx_ ^ (q & 1), stream_,
                timed ? ev.first : nullptr, timed ? ev.second : nullptr);
  else
    launch_chain(ipl_, dRve_, dSegs2_, nSegs2_, iterMode(), cur_, vIdx_ ^ (q & 1), stream_,
                 timed ? ev.first : nullptr, timed ? ev.second : nullptr, stageWaves_);
  cur_ ^= 1;
  stressStale_ = true;
  if (timed) {
    pending_.push_back(ev);
    pendingIters_.push_back(ipl_);
    if (pending_.size() >= 256) resolveTiming(false);
  }
  launchCount_++;
}

void Solver::flipV() {
  vIdx_ ^= 1;
  std::swap(r_.Vx, r_.Vx2);
  std::swap(r_.Vy, r_.Vy2);
}

// The FP64 finish of a split load step runs on the handle's multi-iteration
// shape when that shape has an FP64-arithmetic variant (wave pairs, three
// stages on one wave, four stages on two waves), else on tiles of 8 rows;
// both give the same bits.
bool Solver::multi64() const {
  return multiIter_ && !wide_ &&
         (ipl_ == 2 || (ipl_ == 3 && stageWaves_ == 1) || (ipl_ == 4 && stageWaves_ == 2));
}

int Solver::iterMode() const { return stress64_ ? kModeSplitStress64 : mode_; }

void Solver::iterate(long n) {
  EP2D_HIP(hipSetDevice(device_));
  if (!multiIter_ || (stress64_ && !multi64())) {
    for (long k = 0; k < n; k++) launchIteration();
    return;
  }
  // multi-iteration launches first, then 1..ipl_ single iterations (in
  // place), so the last launch reads U^{n-1}, the state the stress fields are
  // recomputed from
  if (n <= 0) return;
  long singles = (n - 1) % ipl_ + 1;
  long d = (n - singles) / ipl_;
  for (long q = 0; q < d; q++) launchMulti((int)(q & 1));
  // four closing iterations run as a three-iteration launch and the single,
  // three as a two-iteration launch and the single (the same results)
  if (singles == 4 && nSegsC3_ > 0) {
    launchClose(3, (int)(d & 1));
    d++;
    singles = 1;
  } else if (singles == 3 && nSegsC2_ > 0) {
    launchClose(2, (int)(d & 1));
    d++;
    singles = 1;
  }
  if (d & 1) flipV();
  for (long q = 0; q < singles; q++) launchIteration();
}

// A three-iteration chain (three waves per segment) or a wave-pair launch
// closing a run of multi-iteration launches; V is read from buffer
// vIdx_ ^ q and written to the other, as in launchMulti.
void Solver::launchClose(int ns, int q) {
  if (ns == 3)
    launch_chain(3, dRve_, dSegsC3_, nSegsC3_, iterMode(), cur_, vIdx_ ^ q, stream_, nullptr,
                 nullptr, 1);
  else
    launch_march2(dRve_, dSegsC2_, nSegsC2_, iterMode(), cur_, vIdx_ ^ q, stream_, nullptr,
                  nullptr, true);
  cur_ ^= 1;
  stressStale_ = true;
}

void Solver::resolveTiming(bool all) {
  const size_t keep = all ? 0 : 64;
  while (pending_.size() > keep) {
    auto ev = pending_.front();
    const int its = pendingIters_.front();
    pending_.erase(pending_.begin());
    pendingIters_.erase(pendingIters_.begin());
    EP2D_HIP(hipEventSynchronize(ev.second));
    float ms = 0.f;
    EP2D_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
    timedMs_ += ms;          // averages are per iteration
    timedSamples_ += its;
    freeEvents_.push_back(ev);
  }
}

void Solver::setTiming(long sampleEvery) {
  resolveTiming(true);
  sampleEvery_ = sampleEvery < 0 ? 0 : sampleEvery;
  timedMs_ = 0.0;
  timedSamples_ = 0;
  launchCount_ = 0;
}

void Solver::getTiming(double* avgMs, long* samples, const char** name) {
  resolveTiming(true);
  if (avgMs) *avgMs = timedSamples_ ? timedMs_ / timedSamples_ : 0.0;
  if (samples) *samples = timedSamples_;
  if (name) {
    // held by the handle: the pointer stays valid until the next call or
    // the handle's destruction
    std::string& nm = timingName_;
    nm = fused_kernel_name(mode_);
    if (marchRows_ > 0 && multiIter_)
      nm = (ipl_ == 2 ? std::string(mode_ == EP2D_MODE_SPLIT ? "k_march2p<float, float>"
                                                              : "k_march2p<double, double>")
            : wide_ ? "k_wide<" + std::to_string(ipl_) + ">"
                    : std::string(mode_ == EP2D_MODE_SPLIT ? "k_chain<float, float, "
                                                           : "k_chain<double, double, ") +
                          std::to_string(ipl_) + ", " + std::to_string(stageWaves_) + ">") +
           " rows=" + std::to_string(marchRows_) + " segments=" + std::to_string(nSegs2_) +
           " (per iteration)";
    else if (marchRows_ > 0)
      nm = std::string(mode_ == EP2D_MODE_SPLIT ? "k_march<float, float>" : "k_march<double, double>") +
           " rows=" + std::to_string(marchRows_) + " segments=" + std::to_string(nSegs_);
    else if (mode_ != EP2D_MODE_PLAIN) nm += " TY=" + std::to_string(fplan_.ty) + " grid=" +
                                        std::to_string(fplan_.grid) + " tiles=" +
                                        std::to_string(fplan_.ntiles) + " waves=" +
                                        std::to_string(fplan_.waves) + " blocks/CU=" +
                                        std::to_string(fplan_.blocksPerCU);
    *name = nm.c_str();
  }
}

void Solver::sync() { EP2D_HIP(hipStreamSynchronize(stream_)); }

// Stress fields as the reference holds them after the loop: the trial stress
// of U^{n-1} after the return mapping (the fused schemes keep it on chip, so
// it is recomputed here in FP64 from the previous displacement buffer).
void Solver::ensureStress() {
  if (!stressStale_) return;
  const int prev = cur_ ^ 1;
  launch_plain_stress(r_, r_.Ux[prev], r_.Uy[prev], f_, stream_);
  if (opt_.nTasks > 1) {
    launch_plain_j2(r_, f_, stream_);
    launch_plain_plasticity(r_, f_, stream_);
  }
  stressStale_ = false;
}

void Solver::applyAffine(double dUxdx, double dUydy, double dUxdy, double dUydx) {
  launch_affine(r_.Ux[cur_], r_.Uy[cur_], nX_, nY_, dX_, dY_, dUxdx, dUydy, dUxdy, dUydx, stream_);
}

// EffPlast2D.cu:407-412 (step-level entry point).
double Solver::residual(const double e[3], long* nonfinite) {
  launch_maxabs(r_, mode_ == EP2D_MODE_SPLIT, dMax_, stream_);
  d2h(hMax_, dMax_, 3);
  double mx, my;
  std::memcpy(&mx, &hMax_[0], 8);
  std::memcpy(&my, &hMax_[1], 8);
  if (nonfinite) *nonfinite = (long)hMax_[2];
  return (mx / (dX_ * (nX_ - 1)) + my / (dY_ * (nY_ - 1))) * dT_ /
         (std_max(std::fabs(e[0]), std_max(e[1], e[2])));
}

// Volume averages, EffPlast2D.cu:453-503.
void Solver::averages(ep2d_record& r) {
  ensureStress();
  const int pox = nX_ / nPores_;   // :453-454 (double division, truncated)
  const int poy = nY_ / nPores_;
  launch_averages(r_, f_, r_.Ux[cur_], r_.Uy[cur_], nu0_, pox, poy, dAvgPartial_, dAvg_, stream_);
  d2h(hAvg_, dAvg_, NAVG);
  finish_averages(hAvg_, nX_, nY_, pox, poy, r);
}

// The reference's normalisations of the sums (EffPlast2D.cu:491-503): sigma
// over nX*nY (0, 1) and (nX-1)*(nY-1) (2), the periodic window over its area.
void finish_averages(const double* sums, long nX, long nY, int pox, int poy, ep2d_record& r) {
  r.sigma[0] = sums[A_S0] / (nX * nY);
  r.sigma[1] = sums[A_S1] / (nX * nY);
  r.sigma[2] = sums[A_S2] / ((nX - 1) * (nY - 1));
  r.sigma[3] = 0.0 / (nX * nY);
  const long perArea = (nX - 2 * pox) * (nY - 2 * poy);
  r.sigmaPer[0] = sums[A_SP0] / perArea;
  r.sigmaPer[1] = sums[A_SP1] / perArea;
  r.sigmaPer[2] = sums[A_SP2] / perArea;
  r.sigmaPer[3] = sums[A_SP3] / perArea;
  r.epsilonPer[0] = sums[A_E0] / perArea;
  r.epsilonPer[1] = sums[A_E1] / perArea;
  r.epsilonPer[2] = sums[A_E2] / perArea;
  r.sigmaXYclean = sums[A_SXY] / (double)((nX - 1) * (nY - 1));
}

// getDeltaP_honest (:825-853), getDeltaP_periodic (:854-882),
// getTauInfty_honest (:925-942): summed on the host in the reference order
// from gathered boundary values.
void Solver::boundary(double& deltaP, double& deltaPper, double& tauInfty) {
  ensureStress();
  launch_gather(dBases_, dTable_, nGather_, cur_, dGather_, stream_);
  d2h(hGather_, dGather_, (size_t)nGather_);
  const double* g = hGather_;
  auto P = [&](const Line& L, int q) { return g[(L.offset + 3 * q) + 0]; };
  auto XX = [&](const Line& L, int q) { return g[(L.offset + 3 * q) + 1]; };
  auto YY = [&](const Line& L, int q) { return g[(L.offset + 3 * q) + 2]; };
  {
    double dPx = 0.0, dPy = 0.0;
    for (int q = 0; q < hx0_.count; q++) {
      dPx += XX(hx0_, q) - P(hx0_, q);
      dPx += YY(hx0_, q) - P(hx0_, q);
      dPx += XX(hxN_, q) - P(hxN_, q);
      dPx += YY(hxN_, q) - P(hxN_, q);
    }
    dPx /= (nX_ - 2);
    for (int q = 0; q < hy0_.count; q++) {
      dPy += XX(hy0_, q) - P(hy0_, q);
      dPy += YY(hy0_, q) - P(hy0_, q);
      dPy += XX(hyN_, q) - P(hyN_, q);
      dPy += YY(hyN_, q) - P(hyN_, q);
    }
    dPy /= (nY_ - 2);
    deltaP = -0.125 * (dPx + dPy);
  }
  if (nPores_ <= 2) {
    deltaPper = 0.0;
  } else {
    const long nP = (long)nPores_;
    double dPx = 0.0, dPy = 0.0;
    for (int q = 0; q < px0_.count; q++) {
      dPx += XX(px0_, q) - P(px0_, q);
      dPx += YY(px0_, q) - P(px0_, q);
      dPx += XX(pxN_, q) - P(pxN_, q);
      dPx += YY(pxN_, q) - P(pxN_, q);
    }
    dPx /= (long)nX_ * (nP - 2) / nP;
    for (int q = 0; q < py0_.count; q++) {
      dPy += XX(py0_, q) - P(py0_, q);
      dPy += YY(py0_, q) - P(py0_, q);
      dPy += XX(pyN_, q) - P(pyN_, q);
      dPy += YY(pyN_, q) - P(pyN_, q);
    }
    dPy /= (long)nY_ * (nP - 2) / nP;
    deltaPper = -0.125 * (dPx + dPy);
  }
  {
    double tx = 0.0, ty = 0.0;
    for (int q = 0; q < hx0_.count; q++) {
      tx += XX(hx0_, q) - YY(hx0_, q);
      tx += XX(hxN_, q) - YY(hxN_, q);
    }
    tx /= (nX_ - 2);
    for (int q = 0; q < hy0_.count; q++) {
      ty += XX(hy0_, q) - YY(hy0_, q);
      ty += XX(hyN_, q) - YY(hyN_, q);
    }
    ty /= (nY_ - 2);
    tauInfty = -0.125 * (tx + ty);
  }
}

namespace {
double find_max_abs_signed(const double* v, int n) {   // EffPlast2D.cu:814-822
  double m = 0.0;
  for (int k = 0; k < n; k++)
    if (std::fabs(v[k]) > std::fabs(m)) m = v[k];
  return m;
}
}  // namespace

// Pore-edge hole area, EffPlast2D.cu:514-620.  Uses the values gathered by
// the last boundary() call (same state).
void Solver::holeArea(ep2d_record& r) {
  if (!samplesValid_)
    throw StatusError(EP2D_ERR_GEOMETRY,
                      "pore-edge sample window outside the grid (pores too large or too close "
                      "to the boundary)");
  launch_gather(dBases_, dTable_, nGather_, cur_, dGather_, stream_);
  d2h(hGather_, dGather_, (size_t)nGather_);
  double HoleAreaPi = 0.0, InternalHoleAreaPi = 0.0;
  int s = 0;
  auto take = [&](double* w) {
    for (int q = 0; q < 5; q++, s++)
      w[q] = sampleInPore_[s] ? 0.0 : hGather_[poreOffset_ + s];
  };
  const double nP = nPores_;
  for (int k = 0; k < nP; k++) {
    for (int l = 0; l < nP; l++) {
      double left[5], right[5], bottom[5], top[5];
      take(left);
      take(right);
      take(bottom);
      take(top);
      const double dRx = -0.5 * (find_max_abs_signed(left, 5) - find_max_abs_signed(right, 5));
      const double dRy = -0.5 * (find_max_abs_signed(bottom, 5) - find_max_abs_signed(top, 5));
      HoleAreaPi += (rad_ + dRx) * (rad_ + dRy);
      if (k > 0 && l > 0 && k < nP - 1 && l < nP - 1) InternalHoleAreaPi += (rad_ + dRx) * (rad_ + dRy);
    }
  }
  const long nX = nX_, nY = nY_;
  const double rad = rad_, nPores = nPores_, dX = dX_, dY = dY_;
  r.holeAreaPi = HoleAreaPi;
  r.internalHoleAreaPi = InternalHoleAreaPi;
  r.Phi0 = 3.1415926 * rad * rad * nPores * nPores / (dX * (nX - 1) * dY * (nY - 1));
  r.Phi = 3.1415926 * HoleAreaPi / (dX * (nX - 1) * dY * (nY - 1));
  r.PhiPer = nPores > 2 ?
    3.1415926 * InternalHoleAreaPi / (dX * (nX - 1) * dY * (nY - 1) * (nPores - 2) * (nPores - 2) / nPores / nPores) :
    0.0;
  r.dPhi = 3.1415926 * std::fabs(HoleAreaPi - rad * rad * nPores * nPores) / (dX * (nX - 1) * dY * (nY - 1));
  r.dPhiPer = std::fabs(r.PhiPer - r.Phi0);
}

// ---- ComputeEffModuli / ComputeEffDamping as a resumable job -------------------
// The reference runs a fixed sequence of load steps (EffPlast2D.cu:179-243),
// each a loop of time steps (ComputeEffParams, :311-698) with an iteration loop
// inside.  Here the sequence is a plan, and a time step is split into begin /
// iterate-and-check / finish so that the same code drives one RVE on its own
// (runJob) and many RVEs in one batched launch (the Batch engine).
void Solver::jobBegin(int kind, double initLoadValue, const double loadType[3], double loadValue) {
  const int NL = opt_.nTasks;
  if (kind == 0 && NL > 3) throw ArgError("ERROR:  Wrong number of loads in ComputeEffModuli()!");
  if (kind == 1 && NL < 4) throw ArgError("ERROR:  Wrong number of loads in ComputeEffModuli()!");
  EP2D_HIP(hipSetDevice(device_));
  jobT0_ = std::chrono::steady_clock::now();
  std::memset(&res_, 0, sizeof(res_));
  records_.clear();
  recordStep_.clear();
  totalIters_ = 0;
  allConverged_ = true;
  std::memcpy(loadType_, loadType, sizeof(loadType_));
  jobKind_ = kind;
  plan_.clear();
  const long T = opt_.nTimeSteps;
  plan_.push_back({initLoadValue, {loadType[0], loadType[1], loadType[2]}, T});
  if (kind == 0) {   // ComputeEffModuli, :186-205
    const double incPercent = 0.005;   // EffPlast2D.h:105
    const double sph[3] = {0.5 * (loadType[0] + loadType[1]), 0.5 * (loadType[0] + loadType[1]), 0.0};
    double dev[3] = {loadType[0] - sph[0], loadType[1] - sph[1], loadType[2]};
    if (opt_.useShearLoadType) std::memcpy(dev, opt_.shearLoadType, sizeof(dev));
    if (NL > 1) plan_.push_back({initLoadValue * incPercent, {sph[0], sph[1], sph[2]}, 1});
    if (NL == 3) plan_.push_back({initLoadValue * incPercent, {dev[0], dev[1], dev[2]}, 1});
  } else {           // ComputeEffDamping, :239-243
    for (long i = 1; i < NL; i++) {
      const double v = loadValue * (std::sin(i * 2.0 * 3.1415926 / (NL - 1)) -
                                    std::sin((i - 1) * 2.0 * 3.1415926 / (NL - 1)));
      plan_.push_back({v, {loadType[0], loadType[1], loadType[2]}, 1});
    }
  }
  task_ = 0;
  tstep_ = 0;
  printCalculationType();
}

// Start of a time step (EffPlast2D.cu:312-384): banner, macro strain, affine U increment.
void Solver::tsBegin() {
  const TaskPlan& tp = plan_[task_];
  const int step = (int)task_;
  if (tstep_ == 0) {
    printStepInfo(step);
    if (step == 0) {   // :329-333: the macro strain and U restart, V does not
      curEffStrain_[0] = curEffStrain_[1] = curEffStrain_[2] = 0.0;
      launch_fill(r_.Ux[cur_], (size_t)(nX_ + 1) * nY_, 0.0, stream_);
      launch_fill(r_.Uy[cur_], (size_t)nX_ * (nY_ + 1), 0.0, stream_);
    }
  }
  std::memset(&rec_, 0, sizeof(rec_));
  const long nT = tp.nT;
  say("Time step ", tstep_ + 1, " from ", nT, "\n");
  const double dUxdx = tp.load * tp.lt[0] / (double)nT;   // :349-352
  const double dUydy = tp.load * tp.lt[1] / (double)nT;
  const double dUxdy = tp.load * tp.lt[2] / (double)nT;
  const double dUydx = dUxdy;
  curEffStrain_[0] += dUxdx;
  curEffStrain_[1] += dUydy;
  curEffStrain_[2] += dUxdy;
  for (int k = 0; k < 3; k++) rec_.epsilon[k] = curEffStrain_[k];
  say("Macro strain: (", curEffStrain_[0], ", ", curEffStrain_[1], ", ", curEffStrain_[2], ")\n");
  applyAffine(dUxdx, dUydy, dUxdy, dUydx);
  tsDone_ = 0;
  tsError_ = 0.0;
  tsConv_ = false;
  nextCheck_ = opt_.checkInterval;
  lastCheckAt_ = -1;
  lastCheckErr_ = 0.0;
  refErr_ = 0.0;
  sinceHalving_ = 0;
  stress64_ = false;
}

long Solver::tsChunk() const {
  return std::min(nextCheck_ - tsDone_, opt_.nIter - tsDone_);
}

// Next residual check after the one at tsDone_ with residual `error`.  Fixed
// schedule: every checkInterval iterations (EffPlast2D.cu:407, output_step).
// Adaptive: the residual of the pseudo-transient iteration decays about
// geometrically, so the rate over the last two checks predicts the iteration
// where it reaches eIter; the next check goes there (at least K/64 and at
// most K iterations on, so it never checks later than the fixed schedule; K
// when the residual did not fall).  A missed prediction costs one more short
// gap, a check costs a few iterations' time.
void Solver::scheduleNextCheck(double error) {
  const long K = opt_.checkInterval;
  long gap = K;
  if (opt_.adaptiveCheck && lastCheckAt_ >= 0 && error > 0.0 && error < lastCheckErr_) {
    const double rate = (std::log(error) - std::log(lastCheckErr_)) /
                        (double)(tsDone_ - lastCheckAt_);
    const double remaining = std::log(opt_.eIter / error) / rate;
    const long lo = std::max<long>(1, K / 64), hi = K;
    gap = remaining < (double)lo ? lo : (remaining > (double)hi ? hi : (long)std::ceil(remaining));
  }
  lastCheckAt_ = tsDone_;
  lastCheckErr_ = error;
  nextCheck_ = tsDone_ + gap;
}

bool Solver::tsOver() const { return tsConv_ || tsDone_ >= opt_.nIter; }

// Split scheme: the FP32 stresses carry rounding noise of about 2^-24 of
// their size, which puts a floor under the velocity residual (the floor
// depends on the grid, the load and the residual's normaliser max(|e0|,
// max(e1, e2)), EffPlast2D.cu:412, which can be far below the strain of the
// load).  The pseudo-transient residual does not fall at every check (it
// rises and falls by factors of two near eIter in FP64 too), but it halves
// within a few checks; on the floor it stops halving.  When the residual is
// within 1000 eIter of the target and has not halved over the last eight
// checks, the rest of the time step runs the FP64 arithmetic on the same
// state: FP64 U, the FP32 velocity storage.
void Solver::noteStall(double error) {
  if (mode_ != EP2D_MODE_SPLIT || opt_.keepFp32Stress || stress64_ || !(error > 0.0)) return;
  if (!(refErr_ > 0.0) || error < 0.5 * refErr_) {
    refErr_ = error;
    sinceHalving_ = 0;
    return;
  }
  if (++sinceHalving_ >= 8 && error < 1e3 * opt_.eIter) {
    stress64_ = true;
    escalations_++;
    say("    Residual stalled under the FP32 stress: FP64 arithmetic from iteration ", tsDone_,
        "\n");
  }
}

// Residual of EffPlast2D.cu:407-412: launch the reduction and the copy of its
// three words; tsCheckFinish reads them after the stream is synchronised.
void Solver::tsCheckLaunch() {
  launch_maxabs(r_, mode_ == EP2D_MODE_SPLIT, dMax_, stream_);
  EP2D_HIP(hipMemcpyAsync(hMax_, dMax_, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          stream_));
}

// The velocity fields tsCheckLaunch reduces, for an engine that reduces many
// RVEs in one launch.
MaxabsJob Solver::residualJob() const {
  MaxabsJob j;
  j.vx = r_.Vx;
  j.vy = r_.Vy;
  j.nx = (unsigned long long)(nX_ + 1) * nY_;
  j.ny = (unsigned long long)nX_ * (nY_ + 1);
  return j;
}

bool Solver::tsCheckFinish(const unsigned long long* words) {
  if (words) std::memcpy(hMax_, words, 3 * sizeof(unsigned long long));
  double mx, my;
  std::memcpy(&mx, &hMax_[0], 8);
  std::memcpy(&my, &hMax_[1], 8);
  const double* e = curEffStrain_;
  const double error = (mx / (dX_ * (nX_ - 1)) + my / (dY_ * (nY_ - 1))) * dT_ /
                       (std_max(std::fabs(e[0]), std_max(e[1], e[2])));
  if (hMax_[2])
    throw StatusError(EP2D_ERR_NONFINITE, "non-finite velocity at iteration " +
                                              std::to_string(tsDone_) + " of load step " +
                                              std::to_string(task_));
  tsError_ = error;
  say("    Iteration ", tsDone_, ": Error is ", error, "\n");
  if (error < opt_.eIter) {
    say("Number of iterations is ", tsDone_, "\n\n");
    tsConv_ = true;
  } else if (tsDone_ >= opt_.nIter) {
    say("WARNING: Maximum number of iterations reached!\nError is ", error, "\n\n");
  } else {
    noteStall(error);
  }
  scheduleNextCheck(error);
  return tsOver();
}

// End of a time step (EffPlast2D.cu:430-696): averages, boundary means, hole
// area, per-step analytic moduli; then move to the next time step or task.
void Solver::tsFinish() {
  const int NL = opt_.nTasks;
  const int step = (int)task_;
  ep2d_record& r = rec_;
  ep2d_result* res = &res_;
  r.error = tsError_;
  r.iters = (double)tsDone_;
  totalIters_ += tsDone_;
  if (!tsConv_) allConverged_ = false;

  averages(r);
  boundary(r.deltaP, r.deltaPper, r.tauInfty);
  holeArea(r);
  logOnly("    dPhi = ", r.dPhi, "\n");
  if (nPores_ > 2) logOnly("    dPhiPer = ", r.dPhiPer, "\n");
  say("    P / Y = ", r.deltaP / Y_, "\n");
  if (nPores_ > 2) say("    Pper / Y = ", r.deltaPper / Y_, "\n");
  say("    tau / Y = ", r.tauInfty / Y_, "\n");

  // analytic moduli of this step (:643-696)
  const long nX = nX_, nY = nY_;
  const double Phi0 = r.Phi0, Phi = r.Phi;
  const double phi = 3.1415926 * rad_ * rad_ / (dX_ * (nX - 1) * dY_ * (nY - 1));
  const double Y = Y_, G0 = G0_, K0 = K0_;
  if (NL <= 1) {
    res->KphiExactElast = G0 / Phi0;
    logOnly("    KphiExactElast = ", res->KphiExactElast, "\n");
    res->an.Kphi = G0 / Phi0;
    res->KdExactElast = (1.0 - phi) / (1.0 / K0 + 1.0 / res->KphiExactElast);
    logOnly("    KdExactElast = ", res->KdExactElast, "\n");
    res->an.Kd = G0 / Phi0;
  } else if (step == 0) {
    res->KphiExactPlast = G0 / Phi / std::exp(std::fabs(r.deltaP) / Y - 1.0) /
                          (1.0 + r.tauInfty * r.tauInfty / Y / Y);
    logOnly("    KphiExactPlast = ", res->KphiExactPlast, "\n");
    res->an.Kphi = res->KphiExactPlast;
    res->KdExactPlast = (1.0 - phi) / (1.0 / K0 + 1.0 / res->KphiExactPlast);
    logOnly("    KdExactPlast = ", res->KdExactPlast, "\n");
    res->an.Kd = res->KdExactPlast;
    res->KphiExactPlastPer = G0 / Phi0 / std::exp(std::fabs(r.deltaPper) / Y - 1.0) /
                             (1.0 + r.tauInfty * r.tauInfty / Y / Y);
    logOnly("    KphiExactPlastPer = ", res->KphiExactPlastPer, "\n");
    res->an_per.Kphi = res->KphiExactPlastPer;
    res->KdExactPlastPer = (1.0 - phi) / (1.0 / K0 + 1.0 / res->KphiExactPlastPer);
    logOnly("    KdExactPlastPer = ", res->KdExactPlastPer, "\n");
    res->an_per.Kd = res->KdExactPlastPer;
  }
  if (NL == 3 && step == 0) {
    res->GexactElast = G0 / (1.0 + Phi0);
    logOnly("    GexactElast = ", res->GexactElast, "\n");
    res->GexactPlast = G0 / (1.0 + Phi * std::exp(std::fabs(r.deltaP) / Y - 1.0));
    logOnly("    GexactPlast = ", res->GexactPlast, "\n");
    res->an.G = res->GexactPlast;
    res->GexactPlastPer = G0 / (1.0 + Phi0 * std::exp(std::fabs(r.deltaPper) / Y - 1.0));
    logOnly("    GexactPlastPer = ", res->GexactPlastPer, "\n");
    res->an_per.G = res->GexactPlastPer;
  }
  records_.push_back(r);
  recordStep_.push_back(step);
  if (++tstep_ >= plan_[task_].nT) {
    tstep_ = 0;
    task_++;
  }
}

// Moduli getters (EffPlast2D.cu:1412-1539) or the tensor summary of
// ComputeEffDamping (:245-299), then the closing output.
void Solver::jobFinish(ep2d_result* out) {
  const int NL = opt_.nTasks;
  const long T = opt_.nTimeSteps;
  ep2d_result& res = res_;
  if (jobKind_ == 0) {
    const ep2d_record r0 = records_[T - 1];
    if (NL == 1) {   // calcBulkModuli_PureElast, :1412-1481
      res.num.Kphi = r0.deltaP / r0.dPhi;
      logOnly("    ==============\n", "    Kphi = ", res.num.Kphi, "\n");
      res.num_per.Kphi = r0.deltaPper / r0.dPhiPer;
      logOnly("    KphiPer = ", res.num_per.Kphi, "\n");
      const double pInc = -0.33333333 * (r0.sigma[0] + r0.sigma[1] + r0.sigma[2]);
      const double epsInc = r0.epsilon[0] + r0.epsilon[1];
      res.num.Kd = -pInc / epsInc;
      logOnly("    Kd = ", res.num.Kd, "\n");
      const double pIncPer = -0.33333333 * (r0.sigmaPer[0] + r0.sigmaPer[1] + r0.sigmaPer[2]);
      const double epsIncPer = r0.epsilonPer[0] + r0.epsilonPer[1];
      res.num_per.Kd = -pIncPer / epsIncPer;
      logOnly("    KdPer = ", res.num_per.Kd, "\n");
    } else {         // calcBulkModuli_ElastPlast, :1429-1514
      const ep2d_record r1 = records_[T];
      res.num.Kphi = (r1.deltaP - r0.deltaP) / (r1.dPhi - r0.dPhi);
      logOnly("==============\n", "    Kphi = ", res.num.Kphi, "\n");
      res.num_per.Kphi = (r1.deltaPper - r0.deltaPper) / (r1.dPhiPer - r0.dPhiPer);
      logOnly("    KphiPer = ", res.num_per.Kphi, "\n");
      const double pInc = -0.33333333 * (r1.sigma[0] + r1.sigma[1] + r1.sigma[2] - r0.sigma[0] - r0.sigma[1] - r0.sigma[2]);
      const double epsInc = r1.epsilon[0] + r1.epsilon[1] - r0.epsilon[0] - r0.epsilon[1];
      res.num.Kd = -pInc / epsInc;
      logOnly("    Kd = ", res.num.Kd, "\n");
      const double pIncPer = -0.33333333 * (r1.sigmaPer[0] + r1.sigmaPer[1] + r1.sigmaPer[2] - r0.sigmaPer[0] - r0.sigmaPer[1] - r0.sigmaPer[2]);
      const double epsIncPer = r1.epsilonPer[0] + r1.epsilonPer[1] - r0.epsilonPer[0] - r0.epsilonPer[1];
      res.num_per.Kd = -pIncPer / epsIncPer;
      logOnly("    KdPer = ", res.num_per.Kd, "\n");
    }
    if (NL == 3) {   // calcShearModulus, :1446-1539
      const ep2d_record r1 = records_[T];
      const ep2d_record r2 = records_[T + 1];
      res.num.G = 0.5 * (r2.sigma[0] - r1.sigma[0] - r2.sigma[1] + r1.sigma[1]) /
                  (r2.epsilon[0] - r1.epsilon[0] - r2.epsilon[1] + r1.epsilon[1]);
      logOnly("==============\n", "    G = ", res.num.G, "\n");
      res.num_per.G = 0.5 * (r2.sigmaPer[0] - r1.sigmaPer[0] - r2.sigmaPer[1] + r1.sigmaPer[1]) /
                      (r2.epsilonPer[0] - r1.epsilonPer[0] - r2.epsilonPer[1] + r1.epsilonPer[1]);
      logOnly("    Gper = ", res.num_per.G, "\n");
      res.Gxy = 0.5 * (r2.sigmaXYclean - r1.sigmaXYclean) / (r2.epsilon[2] - r1.epsilon[2]);
      res.GxyPer = 0.5 * (r2.sigmaPer[2] - r1.sigmaPer[2]) / (r2.epsilonPer[2] - r1.epsilonPer[2]);
    }
    printEffectiveModuli(res);
    printWarnings();
  } else {
    auto printTensors = [&](const char* title, auto get, int n) {
      say("\n", title, ":\n", "0) ");
      for (int j = 0; j < n; j++) say(get(records_[T - 1])[j], " ");
      say("\n");
      for (long i = 1; i < NL; i++) {
        say(i, ") ");
        for (int j = 0; j < n; j++) say(get(records_[T - 1 + i])[j], " ");
        say("\n");
      }
    };
    printTensors("Strain tensors", [](const ep2d_record& r) { return r.epsilon; }, 3);
    printTensors("Strain tensors periodic", [](const ep2d_record& r) { return r.epsilonPer; }, 3);
    printTensors("Stress tensors", [](const ep2d_record& r) { return r.sigma; }, 4);
    printTensors("Stress tensors periodic", [](const ep2d_record& r) { return r.sigmaPer; }, 4);
  }
  const auto t1 = std::chrono::steady_clock::now();
  res.seconds = std::chrono::duration<double>(t1 - jobT0_).count();
  res.nRecords = (int)records_.size();
  res.totalIters = totalIters_;
  res.converged = allConverged_ ? 1 : 0;
  printDuration((int)res.seconds);
  if (out) *out = res;
}

// One RVE on its own: the iteration loop of EffPlast2D.cu:389-428 with a
// residual check every checkInterval iterations (or on the adaptive schedule).
void Solver::runJob(ep2d_result* out) {
  while (!jobDone()) {
    tsBegin();
    while (!tsOver()) {
      const long c = tsChunk();
      iterate(c);
      tsDone_ += c;
      if (tsAtCheck()) {
        tsCheckLaunch();
        EP2D_HIP(hipStreamSynchronize(stream_));
        tsCheckFinish();
      }
    }
    tsFinish();
  }
  jobFinish(out);
}

void Solver::computeEffModuli(double initLoadValue, const double loadType[3], ep2d_result* out) {
  jobBegin(0, initLoadValue, loadType, 0.0);
  runJob(out);
}

void Solver::computeEffDamping(double initLoadValue, const double loadType[3], double loadValue,
                               ep2d_result* out) {
  jobBegin(1, initLoadValue, loadType, loadValue);
  runJob(out);
}

// Iterations executed by a batched launch on this RVE's buffers.
void Solver::noteBatchIterations(long n, long launches, long doubles) {
  if (mode_ != EP2D_MODE_PLAIN) {
    if (launches & 1) cur_ ^= 1;
    if (doubles & 1) flipV();
    if (n) stressStale_ = true;
  }
  tsDone_ += n;
}

// ---- fields -------------------------------------------------------------------
void Solver::fieldShape(int field, int& m, int& n) const {
  switch (field) {
    case EP2D_FIELD_P: case EP2D_FIELD_TAUXX: case EP2D_FIELD_TAUYY: case EP2D_FIELD_TAUXYAV:
    case EP2D_FIELD_J2: case EP2D_FIELD_K: case EP2D_FIELD_G:
      m = nX_; n = nY_; return;
    case EP2D_FIELD_TAUXY: case EP2D_FIELD_J2XY:
      m = nX_ - 1; n = nY_ - 1; return;
    case EP2D_FIELD_UX: case EP2D_FIELD_VX:
      m = nX_ + 1; n = nY_; return;
    case EP2D_FIELD_UY: case EP2D_FIELD_VY:
      m = nX_; n = nY_ + 1; return;
    default:
      throw ArgError("unknown field id " + std::to_string(field));
  }
}

bool Solver::uNodeInPore(bool isUy, long i, long j) const {
  const long nX = nX_, nY = nY_;
  const double Lx = dX_ * (nX - 1), Ly = dY_ * (nY - 1);
  const double x = isUy ? -0.5 * dX_ * (nX - 1) + dX_ * (int)i : -0.5 * dX_ * nX + dX_ * (int)i;
  const double y = isUy ? -0.5 * dY_ * nY + dY_ * (int)j : -0.5 * dY_ * (nY - 1) + dY_ * (int)j;
  bool in = false;
  for (int k = 0; k < nPores_; k++)
    for (int l = 0; l < nPores_; l++)
      if (std::sqrt((x - 0.5 * Lx * (1.0 - 1.0 / nPores_) + (Lx / nPores_) * k) * (x - 0.5 * Lx * (1.0 - 1.0 / nPores_) + (Lx / nPores_) * k) +
                    (y - 0.5 * Ly * (1.0 - 1.0 / nPores_) + (Ly / nPores_) * l) * (y - 0.5 * Ly * (1.0 - 1.0 / nPores_) + (Ly / nPores_) * l)) < rad_)
        in = true;
  return in;
}

// SaveAnStatic2D (EffPlast2D.cu:1060-1292) on the host copies of the current
// state: the stresses as ComputeEffParams leaves them and U with the pore
// nodes zeroed (:514-547).
void Solver::anStatic(double deltaP, double tauInfty, const double lt[3], ep2d_an_static* out) {
  const size_t nC = (size_t)nX_ * nY_, nV = (size_t)(nX_ - 1) * (nY_ - 1);
  const size_t nUx = (size_t)(nX_ + 1) * nY_, nUy = (size_t)nX_ * (nY_ + 1);
  std::vector<double> P(nC), txx(nC), tyy(nC), txy(nV), Ux(nUx), Uy(nUy);
  getField(EP2D_FIELD_P, P.data(), nC);
  getField(EP2D_FIELD_TAUXX, txx.data(), nC);
  getField(EP2D_FIELD_TAUYY, tyy.data(), nC);
  getField(EP2D_FIELD_TAUXY, txy.data(), nV);
  getField(EP2D_FIELD_UX, Ux.data(), nUx);
  getField(EP2D_FIELD_UY, Uy.data(), nUy);
  for (long j = 0; j < nY_; j++)
    for (long i = 0; i <= nX_; i++)
      if (uNodeInPore(false, i, j)) Ux[(size_t)j * (nX_ + 1) + i] = 0.0;
  for (long j = 0; j <= nY_; j++)
    for (long i = 0; i < nX_; i++)
      if (uNodeInPore(true, i, j)) Uy[(size_t)j * nX_ + i] = 0.0;
  AnGeometry g{nX_, nY_, dX_, dY_, rad_, coh_ / std::sqrt(2.0), G0_, K0_, (int)nPores_};
  an_static_fields(g, P.data(), txx.data(), tyy.data(), txy.data(), Ux.data(), Uy.data(), deltaP,
                   tauInfty, lt, out);
}

void Solver::saveAnStatic(double deltaP, double tauInfty, const double lt[3],
                          const std::string& dir, ep2d_an_static* stats) {
  const size_t nC = (size_t)nX_ * nY_, nV = (size_t)(nX_ - 1) * (nY_ - 1);
  const bool an = nPores_ == 1;
  std::vector<double> Unu(nC), J1nu(nV), J2nu(nV), Uan, eU, J1an, J2an, e1, e2;
  ep2d_an_static o;
  std::memset(&o, 0, sizeof(o));
  o.UnuAbs = Unu.data();
  o.J1nu = J1nu.data();
  o.J2nu = J2nu.data();
  if (an) {
    for (auto* v : {&Uan, &eU}) v->resize(nC);
    for (auto* v : {&J1an, &J2an, &e1, &e2}) v->resize(nV);
    o.UanAbs = Uan.data();
    o.errorUabs = eU.data();
    o.J1an = J1an.data();
    o.J2an = J2an.data();
    o.errorJ1 = e1.data();
    o.errorJ2 = e2.data();
  }
  anStatic(deltaP, tauInfty, lt, &o);
  auto write = [&](const char* name, const std::vector<double>& a) {   // SaveVector, :800-803
    const std::string path = dir + "/" + name + "_" + std::to_string(nX_) + "_.dat";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw StatusError(EP2D_ERR_IO, "cannot create " + path);
    const size_t w = std::fwrite(a.data(), sizeof(double), a.size(), f);
    std::fclose(f);
    if (w != a.size()) throw StatusError(EP2D_ERR_IO, "write failed " + path);
  };
  if (an) {
    say("\n", "Uabs max error  = ", o.errorUabsMax, ", avg = ", o.errorUabsAvg, "\n",
        "J1 max error = ", o.errorJ1Max, ", avg = ", o.errorJ1Avg, "\n", "J2 max error = ",
        o.errorJ2Max, ", avg = ", o.errorJ2Avg, "\n");
    write("UanAbs", Uan);
    write("errorUabs", eU);
    write("J1an", J1an);
    write("J2an", J2an);
    write("errorJ1", e1);
    write("errorJ2", e2);
  }
  write("UnuAbs", Unu);
  write("J1nu", J1nu);
  write("J2nu", J2nu);
  if (stats) {
    ep2d_an_static keep = *stats;
    *stats = o;
    // the caller's array pointers are not ours to replace
    stats->UnuAbs = keep.UnuAbs;
    stats->J1nu = keep.J1nu;
    stats->J2nu = keep.J2nu;
    stats->UanAbs = keep.UanAbs;
    stats->errorUabs = keep.errorUabs;
    stats->J1an = keep.J1an;
    stats->J2an = keep.J2an;
    stats->errorJ1 = keep.errorJ1;
    stats->errorJ2 = keep.errorJ2;
  }
}

void Solver::getField(int field, double* host, size_t count) {
  int m, n;
  fieldShape(field, m, n);
  const size_t N = (size_t)m * n;
  if (count < N) throw ArgError("host buffer too small for field");
  EP2D_HIP(hipSetDevice(device_));
  switch (field) {
    case EP2D_FIELD_P: ensureStress(); d2h(host, f_.P, N); return;
    case EP2D_FIELD_TAUXX: ensureStress(); d2h(host, f_.tauXX, N); return;
    case EP2D_FIELD_TAUYY: ensureStress(); d2h(host, f_.tauYY, N); return;
    case EP2D_FIELD_TAUXY: ensureStress(); d2h(host, f_.tauXY, N); return;
    case EP2D_FIELD_TAUXYAV: ensureStress(); d2h(host, f_.tauXYav, N); return;
    case EP2D_FIELD_J2: ensureStress(); d2h(host, f_.J2, N); return;
    case EP2D_FIELD_J2XY: ensureStress(); d2h(host, f_.J2XY, N); return;
    case EP2D_FIELD_UX: d2h(host, r_.Ux[cur_], N); return;
    case EP2D_FIELD_UY: d2h(host, r_.Uy[cur_], N); return;
    case EP2D_FIELD_VX: case EP2D_FIELD_VY: {
      void* src = field == EP2D_FIELD_VX ? r_.Vx : r_.Vy;
      if (mode_ == EP2D_MODE_SPLIT) {
        std::vector<float> tmp(N);
        d2h(tmp.data(), (const float*)src, N);
        for (size_t k = 0; k < N; k++) host[k] = tmp[k];
      } else {
        d2h(host, (const double*)src, N);
      }
      return;
    }
    case EP2D_FIELD_K: case EP2D_FIELD_G: {
      std::vector<unsigned char> msk(N);
      d2h(msk.data(), r_.mask, N);
      const double v0 = field == EP2D_FIELD_K ? K0_ : G0_;
      const double vp = field == EP2D_FIELD_K ? r_.Kp : r_.Gp;
      for (size_t k = 0; k < N; k++) host[k] = (msk[k] & M_CELL_MAT) ? vp : v0;
      return;
    }
  }
}

// Device buffer of a field in the current state (stress fields finalised
// first); the stream is synchronised, so the buffer is ready for any stream.
const void* Solver::fieldDevice(int field) {
  EP2D_HIP(hipSetDevice(device_));
  const void* p = nullptr;
  switch (field) {
    case EP2D_FIELD_P: ensureStress(); p = f_.P; break;
    case EP2D_FIELD_TAUXX: ensureStress(); p = f_.tauXX; break;
    case EP2D_FIELD_TAUYY: ensureStress(); p = f_.tauYY; break;
    case EP2D_FIELD_TAUXY: ensureStress(); p = f_.tauXY; break;
    case EP2D_FIELD_TAUXYAV: ensureStress(); p = f_.tauXYav; break;
    case EP2D_FIELD_J2: ensureStress(); p = f_.J2; break;
    case EP2D_FIELD_J2XY: ensureStress(); p = f_.J2XY; break;
    case EP2D_FIELD_UX: p = r_.Ux[cur_]; break;
    case EP2D_FIELD_UY: p = r_.Uy[cur_]; break;
    case EP2D_FIELD_VX: p = r_.Vx; break;
    case EP2D_FIELD_VY: p = r_.Vy; break;
    default: throw ArgError("K and G are not stored on the device (the mask selects them)");
  }
  EP2D_HIP(hipStreamSynchronize(stream_));
  return p;
}

void Solver::setField(int field, const double* host, size_t count) {
  int m, n;
  fieldShape(field, m, n);
  const size_t N = (size_t)m * n;
  if (count < N) throw ArgError("host buffer too small for field");
  EP2D_HIP(hipSetDevice(device_));
  switch (field) {
    case EP2D_FIELD_UX: h2d(r_.Ux[cur_], host, N); return;
    case EP2D_FIELD_UY: h2d(r_.Uy[cur_], host, N); return;
    case EP2D_FIELD_VX: case EP2D_FIELD_VY: {
      // both buffers, so the velocities multi-iteration launches never write
      // (the boundary) agree
      void* dst[2] = {field == EP2D_FIELD_VX ? r_.Vx : r_.Vy,
                      field == EP2D_FIELD_VX ? r_.Vx2 : r_.Vy2};
      const int nb = dst[0] == dst[1] ? 1 : 2;
      if (mode_ == EP2D_MODE_SPLIT) {
        std::vector<float> tmp(host, host + N);
        for (int k = 0; k < nb; k++) h2d((float*)dst[k], tmp.data(), N);
      } else {
        for (int k = 0; k < nb; k++) h2d((double*)dst[k], host, N);
      }
      return;
    }
    default:
      throw ArgError("only Ux, Uy, Vx and Vy can be set");
  }
}

void Solver::getMask(unsigned char* host, size_t count) {
  const size_t N = (size_t)nX_ * nY_;
  if (count < N) throw ArgError("host buffer too small for mask");
  d2h(host, r_.mask, N);
}

// ---- console / log output (EffPlast2D.cu:1295-1375) -----------------------------
template <typename... A>
void Solver::say(const A&... a) {
  if (!opt_.verbose && !log_.is_open()) return;
  std::ostringstream os;
  (os << ... << a);
  if (opt_.verbose) std::cout << os.str();
  if (log_.is_open()) log_ << os.str();
}

template <typename... A>
void Solver::logOnly(const A&... a) {
  if (!log_.is_open()) return;
  std::ostringstream os;
  (os << ... << a);
  log_ << os.str();
}

// printDuration, EffPlast2D.cu:1371-1409: console lines end in one newline,
// log lines in three.
void Solver::printDuration(int sec) {
  std::ostringstream os;
  os << "\nCalculation time is ";
  int mn = sec / 60, hr = mn / 60, dy = hr / 24;
  if (sec < 60) {
    os << sec << " sec";
  } else if (mn < 60) {
    os << mn << " min " << sec % 60 << " sec";
  } else if (hr < 24) {
    os << hr << " hours " << mn % 60 << " min " << sec % 60 << " sec";
  } else if (dy < 7) {
    os << dy << " days " << hr % 24 << " hours " << mn % 60 << " min " << sec % 60 << " sec";
  } else {
    os << dy / 7 << " weeks " << dy % 7 << " days " << hr % 24 << " hours " << mn % 60 << " min "
       << sec % 60 << " sec";
  }
  if (opt_.verbose) std::cout << os.str() << "\n";
  if (log_.is_open()) log_ << os.str() << "\n\n\n";
}

void Solver::printStepInfo(int step) {
  const int NL = opt_.nTasks;
  say("\nLOAD STEP ", step + 1, " FROM ", NL, ": ");
  if (NL <= 3) {
    switch (step) {
      case 0: say("PRELOADING\n"); break;
      case 1: say("SMALL HYDROSTATIC INCREMENT\n"); break;
      case 2: say("SMALL DEVIATORIC INCREMENT\n"); break;
      default: throw ArgError("ERROR:  Wrong step index!");
    }
  }
  say("Porosity is ", porosity_ * 100, "%\n");
  say("Grid resolution is ", nX_, "x", nY_, "\n\n");
}

void Solver::printCalculationType() {
  switch (opt_.nTasks) {
    case 1: say("\nPURE ELASTIC CALCULATION\nESTIMATION OF THE EFFECTIVE BULK MODULI\n"); break;
    case 2: say("\nELASTOPLASTIC CALCULATION\nESTIMATION OF THE EFFECTIVE BULK MODULI\n"); break;
    case 3: say("\nELASTOPLASTIC CALCULATION\nESTIMATION OF THE EFFECTIVE BULK MODULI AND THE EFFECTIVE SHEAR MODULUS\n"); break;
    default: say("\nELASTOPLASTIC CALCULATION\nESTIMATION OF THE EFFECTIVE DAMPING\n");
  }
}

void Solver::printEffectiveModuli(const ep2d_result& r) {
  auto out = [&](const ep2d_moduli& m) {
    say("        Kphi = ", m.Kphi, "\n");
    say("        Kd = ", m.Kd, "\n");
    if (opt_.nTasks > 2) say("        G = ", m.G, "\n");
  };
  say("\nEFFECTIVE MODULI ESTIMATED\nNON-PERIODIC\n    Analytical:\n");
  out(r.an);
  say("    Numerical:\n");
  out(r.num);
  if (nPores_ > 2) {
    say("PERIODIC\n    Analytical:\n");
    out(r.an_per);
    say("    Numerical:\n");
    out(r.num_per);
  }
}

void Solver::printWarnings() {
  if (opt_.nTasks <= 1 || records_.empty()) return;
  const ep2d_record& r0 = records_[opt_.nTimeSteps - 1];
  if (std::fabs(r0.deltaP) / Y_ < 1.0 || std::fabs(r0.deltaP) / Y_ > 5.0 ||
      std::fabs(r0.tauInfty) / Y_ > 0.42) {
    say("\nWARNING:  Analytical solution is not applicable for this load!\n");
    say("TODO: Set the load so that 1 < |P|/Y < 5 and 0 < |tau|/Y < 0.42\n");
  }
  if (std::fabs(loadType_[0] - loadType_[1]) > 0.05 * std::max(std::fabs(loadType_[0]), std::fabs(loadType_[1])) ||
      std::fabs(loadType_[2]) > 0.01 * std::max(std::fabs(loadType_[0]), std::fabs(loadType_[1]))) {
    say("\nWARNING:  The numerical effective Kphi is not calculated quite accurately for non-hydrostatic load!\n");
  }
}

}  // namespace ep2d
