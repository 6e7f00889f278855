// ep2d_solver.h -- host driver of one RVE (the body of an ep2d_handle).
#pragma once

#include <fstream>
#include <string>
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

class Solver {
 public:
  Solver(const double pa[EP2D_NPARS], int nX, int nY, const ep2d_options& opt);
  ~Solver();
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // EffPlast2D::ComputeEffModuli / ComputeEffDamping (EffPlast2D.cu:179-310)
  void computeEffModuli(double initLoadValue, const double loadType[3], ep2d_result* out);
  void computeEffDamping(double initLoadValue, const double loadType[3], double loadValue,
                         ep2d_result* out);
  const std::vector<ep2d_record>& records() const { return records_; }

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
  void setField(int field, const double* host, size_t count);
  void getMask(unsigned char* host, size_t count);

  void setTiming(long sampleEvery);
  void getTiming(double* avgMs, long* samples, const char** name);
  void sync();
  size_t deviceBytes() const { return devBytes_; }

 private:
  // ComputeEffParams (EffPlast2D.cu:311-698)
  void computeEffParams(int step, double loadStepValue, const double loadType[3], long nT,
                        ep2d_result* res);
  void ensureStress();
  void launchIteration();
  void resolveTiming(bool all);
  void* alloc(size_t bytes);
  void buildGatherTable();
  template <typename T> void h2d(T* dst, const T* src, size_t n);
  template <typename T> void d2h(T* dst, const T* src, size_t n);

  // console/log output of the reference (EffPlast2D.cu:1295-1409)
  template <typename... A> void say(const A&... a);
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
  int device_ = 0;
  RveDev r_{};
  StressDev f_{};
  FusedPlan plan_{};
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
  std::vector<void*> allocs_;
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
  double timedMs_ = 0.0;
  long timedSamples_ = 0;

  // results
  double curEffStrain_[3] = {0, 0, 0};
  std::vector<ep2d_record> records_;
  std::vector<int> recordStep_;
  long totalIters_ = 0;
  bool allConverged_ = true;
  std::ofstream log_;
};

}  // namespace ep2d
