// ep2d_solver.h -- host driver of one RVE (the body of an ep2d_handle).
#pragma once

#include <chrono>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/ep2d.h"
#include "ep2d_internal.h"

namespace ep2d {

struct ArgError : std::invalid_argument {
  explicit ArgError(const std::string& s) : std::invalid_argument(s) {}
};
struct StatusError : std::runtime_error {
  int status;
  StatusError(int st, const std::string& s) : std::runtime_error(s), status(st) {}
};

// Derivation of pa[] from the physical inputs (get_sigma_2D.m:3-32).
void derive_params(const ep2d_physics& ph, double pa[EP2D_NPARS], int* nX);
// Volume averages of a record from the NAVG device sums (EffPlast2D.cu:491-503);
// pox, poy: the periodic window offsets nX / nPores, nY / nPores.
void finish_averages(const double* sums, long nX, long nY, int pox, int poy, ep2d_record& r);

// Device and pinned host blocks kept for reuse, by exact size.  A batch
// engine's solvers take their buffers from it and give them back when they
// retire, so admitting the next RVE of a size seen before allocates nothing
// and retiring one frees nothing (hipMalloc / hipFree cost tens to hundreds of
// microseconds each and hipFree waits for the device).  The owner trims it.
class MemPool {
 public:
  MemPool() = default;
  MemPool(const MemPool&) = delete;
  MemPool& operator=(const MemPool&) = delete;
  ~MemPool() { trim(); }
  void* take(size_t bytes, bool host);           // nullptr when none is held
  void give(void* p, size_t bytes, bool host);
  void trim();                                    // free every held block
  size_t heldBytes() const { return held_; }

 private:
  std::unordered_multimap<size_t, void*> dev_, host_;
  size_t held_ = 0;
};

class Solver {
 public:
  // stream: run on this stream (owned by the caller) instead of a new one;
  // pool: take buffers from it and give them back on destruction
  Solver(const double pa[EP2D_NPARS], int nX, int nY, const ep2d_options& opt,
         hipStream_t stream = nullptr, MemPool* pool = nullptr);
  ~Solver();
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // Device bytes a Solver of this grid, scheme and pore count allocates when
  // it runs on one-iteration shapes (the batched engine's solvers).
  static size_t plannedDeviceBytes(int nX, int nY, int mode, double nPores);

  // EffPlast2D::ComputeEffModuli / ComputeEffDamping (EffPlast2D.cu:179-310)
  void computeEffModuli(double initLoadValue, const double loadType[3], ep2d_result* out);
  void computeEffDamping(double initLoadValue, const double loadType[3], double loadValue,
                         ep2d_result* out);
  const std::vector<ep2d_record>& records() const { return records_; }

  // The same sequence as a resumable job (kind 0: moduli, 1: damping), for
  // drivers that run many RVEs in one launch.
  void jobBegin(int kind, double initLoadValue, const double loadType[3], double loadValue);
  bool jobDone() const { return task_ >= plan_.size(); }
  void tsBegin();                 // start a time step (affine increment)
  void scheduleNextCheck(double error);
  void noteStall(double error);   // split: switch to the FP64 arithmetic on a stalled residual
  long tsChunk() const;           // iterations to the next check or the cap
  bool tsOver() const;            // converged or iteration cap reached
  long tsDoneIters() const { return tsDone_; }
  bool tsAtCheck() const { return tsDone_ > 0 && tsDone_ == nextCheck_; }
  void tsCheckLaunch();           // residual reduction + async copy of its result
  // after the stream is synchronised; words: the three result words of a
  // reduction the caller launched over residualJob() instead
  bool tsCheckFinish(const unsigned long long* words = nullptr);
  MaxabsJob residualJob() const;  // the current velocity fields the check reduces
  bool splitVelocity() const { return mode_ == EP2D_MODE_SPLIT; }
  void tsFinish();                // averages, boundary, hole area, analytic moduli
  void jobFinish(ep2d_result* out);
  void runJob(ep2d_result* out);
  // after the batched engine ran n iterations of this RVE in `launches`
  // launches, `doubles` of them two-iteration launches
  void noteBatchIterations(long n, long launches, long doubles);
  const RveDev& dev() const { return r_; }
  const FusedPlan& fusedPlan() const { return fplan_; }
  int cur() const { return cur_; }
  // split scheme: this time step finishes in the FP64 arithmetic (the tile
  // kernel; the batched engine runs such RVEs in a list of their own)
  bool escalated() const { return stress64_; }
  long escalations() const { return escalations_; }
  int vBuffer() const { return vIdx_; }
  int mode() const { return mode_; }
  bool plastic() const { return opt_.nTasks > 1; }
  int nX() const { return nX_; }
  int nY() const { return nY_; }
  hipStream_t stream() const { return stream_; }

  // step-level pieces
  void iterate(long n);
  void applyAffine(double dUxdx, double dUydy, double dUxdy, double dUydx);
  double residual(const double cur[3], long* nonfinite);
  void averages(ep2d_record& r);
  void boundary(double& deltaP, double& deltaPper, double& tauInfty);
  void holeArea(ep2d_record& r);

  // fields
  void fieldShape(int field, int& m, int& n) const;
  void getField(int field, double* host, size_t count);
  // analytic static fields beside the current numerical ones (SaveAnStatic2D)
  void anStatic(double deltaP, double tauInfty, const double loadType[3], ep2d_an_static* out);
  void saveAnStatic(double deltaP, double tauInfty, const double loadType[3],
                    const std::string& dir, ep2d_an_static* stats);
  // host U zeroing test of the reference (EffPlast2D.cu:517-547)
  bool uNodeInPore(bool isUy, long i, long j) const;
  void setField(int field, const double* host, size_t count);
  const void* fieldDevice(int field);
  void getMask(unsigned char* host, size_t count);

  void setTiming(long sampleEvery);
  void getTiming(double* avgMs, long* samples, const char** name);
  void sync();
  size_t deviceBytes() const { return devBytes_; }

 private:
  void init(const double pa[EP2D_NPARS], hipStream_t stream);
  void release();
  void ensureStress();
  void launchIteration();
  void resolveTiming(bool all);
  void* alloc(size_t bytes);
  void* hostAlloc(size_t bytes);
  void buildGatherTable();
  template <typename T> void h2d(T* dst, const T* src, size_t n);
  template <typename T> void d2h(T* dst, const T* src, size_t n);

  // console/log output of the reference (EffPlast2D.cu:1295-1409)
  template <typename... A> void say(const A&... a);
  void printDuration(int sec);
  template <typename... A> void logOnly(const A&... a);
  void printStepInfo(int step);
  void printCalculationType();
  void printEffectiveModuli(const ep2d_result& r);
  void printWarnings();

  // configuration
  double pa_[EP2D_NPARS];
  int nX_, nY_;
  ep2d_options opt_;
  std::string logPath_;
  int mode_;
  double dX_, dY_, dT_, K0_, G0_, rho_, dampX_, dampY_, coh_, rad_, nPores_;
  double Y_, nu0_, E0_, porosity_;
  double loadType_[3] = {0, 0, 0};

  // device state
  hipStream_t stream_ = nullptr;
  bool ownStream_ = true;
  int device_ = 0;
  RveDev r_{};
  StressDev f_{};
  FusedPlan fplan_{};
  int marchRows_ = 0, nSegs_ = 0;   // row-marching kernel when marchRows_ > 0
  RveDev* dRve_ = nullptr;          // device copy of r_ with Vx/Vy = vbuf 0, Vx2/Vy2 = vbuf 1
  int4* dSegs_ = nullptr;
  bool multiIter_ = false;          // more than one iteration per launch
  int ipl_ = 1;                     // iterations per launch: 1, 2 (pairs), 3 or 4 (chains)
  bool wide_ = false;               // chains on two columns per lane (split scheme)
  int stageWaves_ = 1;              // chains: waves side by side per stage
  int nSegs2_ = 0;
  int4* dSegs2_ = nullptr;
  // segments of the three- and two-iteration launches that close a run of a
  // four- or three-iteration shape (chains of three waves, 42 rows; wave
  // pairs, 26 rows)
  int nSegsC3_ = 0, nSegsC2_ = 0;
  int4* dSegsC3_ = nullptr;
  int4* dSegsC2_ = nullptr;
  int vIdx_ = 0;                    // velocity buffer holding the current V
  void flipV();                     // the other velocity buffer becomes current
  std::pair<hipEvent_t, hipEvent_t> takeTimingEvents();
  void launchMulti(int q);         // ipl_ iterations, q-th such launch of a run
  void launchClose(int ns, int q);  // ns = 3 or 2 iterations closing a run
  bool multi64() const;            // the FP64 finish runs on the multi-iteration shape
  int iterMode() const;            // kernel arithmetic of the next iterations
  int cur_ = 0;
  bool stressStale_ = false;
  unsigned long long* dMax_ = nullptr;
  unsigned long long* hMax_ = nullptr;
  double* dAvgPartial_ = nullptr;
  double* dAvg_ = nullptr;
  double* hAvg_ = nullptr;
  const double** dBases_ = nullptr;
  int2* dTable_ = nullptr;
  double* dGather_ = nullptr;
  double* hGather_ = nullptr;
  int nGather_ = 0;
  std::vector<std::pair<void*, size_t>> allocs_;   // device blocks and their sizes
  std::vector<std::pair<void*, size_t>> hostAllocs_;  // pinned host blocks
  MemPool* pool_ = nullptr;
  size_t devBytes_ = 0;

  // gather table layout (host side)
  struct Line { int offset, count; };
  Line hx0_{}, hxN_{}, hy0_{}, hyN_{}, px0_{}, pxN_{}, py0_{}, pyN_{};
  int poreOffset_ = 0;            // first pore-sample entry
  bool samplesValid_ = true;
  std::vector<unsigned char> sampleInPore_;  // per pore sample: node inside a pore

  // timing
  long sampleEvery_ = 0;
  long launchCount_ = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending_, freeEvents_;
  std::vector<int> pendingIters_;   // iterations in each pending timed launch
  double timedMs_ = 0.0;
  long timedSamples_ = 0;
  std::string timingName_;         // kernel name getTiming() hands out

  // job state
  struct TaskPlan {
    double load;
    double lt[3];
    long nT;
  };
  std::vector<TaskPlan> plan_;
  size_t task_ = 0;
  long tstep_ = 0;
  long tsDone_ = 0;
  long nextCheck_ = 0;       // iteration of the next residual check in this time step
  long lastCheckAt_ = -1;    // the previous check and its residual (adaptive schedule)
  double lastCheckErr_ = 0.0;
  double refErr_ = 0.0;      // the residual at the check where it last halved
  int sinceHalving_ = 0;     // checks since then
  long escalations_ = 0;     // time steps finished in the FP64 arithmetic
  // split scheme: the rest of this time step runs the FP64 arithmetic (FP64
  // U, FP32 V kept) because the FP32 stress stopped resolving the residual
  bool stress64_ = false;
  FusedPlan fplan64_;        // tile plan of that arithmetic
  double tsError_ = 0.0;
  bool tsConv_ = false;
  int jobKind_ = 0;
  ep2d_record rec_{};
  ep2d_result res_{};
  std::chrono::steady_clock::time_point jobT0_;

  // results
  double curEffStrain_[3] = {0, 0, 0};
  std::vector<ep2d_record> records_;
  std::vector<int> recordStep_;
  long totalIters_ = 0;
  bool allConverged_ = true;
  std::ofstream log_;
};

}  // namespace ep2d
