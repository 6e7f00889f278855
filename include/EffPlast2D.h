/*
 * EffPlast2D.h -- drop-in C++ class for the reference's EffPlast2D
 * (EffPlast2D.h:53-198), header-only on top of the C ABI in ep2d.h
 * (link with -lep2d).
 *
 * A program written against the reference keeps compiling unchanged:
 *   EffPlast2D eff_plast(time_steps);                       (main.cu:30)
 *   eff_plast.ComputeEffModuli(init_load_value, load_type, load_value);
 * The configuration macros of the reference (EffPlast2D.h:17-39) keep their
 * names and defaults and become the runtime options of the solver; the
 * parameters are read from data/pa.dat as in ReadParams
 * (EffPlast2D.cu:700-707); the fields Pc and tauXXc are written to
 * data/<name>_<32*NGRID>_.dat after ComputeEffModuli (EffPlast2D.cu:214-215)
 * and, as SaveAnStatic2D does (EffPlast2D.cu:210-212, 301-303), the
 * analytic static fields UnuAbs, J1nu, J2nu (and with one pore UanAbs,
 * errorUabs, J1an, J2an, errorJ1, errorJ2).
 *
 * Differences: device errors throw std::runtime_error instead of exit(-1)
 * (EffPlast2D.h:42-49); the object is not copyable; the results are also
 * available through result() and records().
 */
#ifndef EFFPLAST2D_DROPIN_H
#define EFFPLAST2D_DROPIN_H

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "ep2d.h"

#if !defined(NGRID)
#define NGRID 32
#endif
#if !defined(NPARS)
#define NPARS 11
#endif
#if !defined(NL)
#define NL 3
#endif
#if !defined(NITER)
#define NITER 100'000
#endif
#if !defined(EITER)
#define EITER 1.0e-10
#endif
#if !defined(DEVICE_IDX)
#define DEVICE_IDX 0
#endif
/* iteration scheme: EP2D_MODE_PLAIN, EP2D_MODE_FP64 or EP2D_MODE_SPLIT */
#if !defined(EP2D_MODE)
#define EP2D_MODE EP2D_MODE_FP64
#endif
/* residual check interval (output_step, EffPlast2D.cu:1608) */
#if !defined(EP2D_CHECK_INTERVAL)
#define EP2D_CHECK_INTERVAL 10'000
#endif

class EffPlast2D {
 public:
  explicit EffPlast2D(const unsigned int nTimeSteps = 1)
      : h_(nullptr), nTasks_(NL), nTimeSteps_(nTimeSteps) {
    static_assert(NPARS == EP2D_NPARS, "pa.dat holds 11 parameters");
    double pa[EP2D_NPARS];
    if (ep2d_read_pa("data/pa.dat", pa) != EP2D_OK)
      throw std::runtime_error("ERROR:  Cannot open file data/pa.dat!\n");
    ep2d_options o;
    ep2d_default_options(&o);
    o.nTasks = NL;
    o.nIter = NITER;
    o.eIter = EITER;
    o.device = DEVICE_IDX;
    o.nTimeSteps = (int)nTimeSteps;
    o.mode = EP2D_MODE;
    o.checkInterval = EP2D_CHECK_INTERVAL;
    o.verbose = 1;
    o.logPath = "EffPlast2D.log";
    check(ep2d_create(pa, 32 * NGRID, 32 * NGRID, &o, &h_));
  }
  /* Extension: configure from the physical inputs at run time. */
  EffPlast2D(const ep2d_physics& phys, const ep2d_options& opt)
      : h_(nullptr), nTasks_(opt.nTasks), nTimeSteps_((unsigned)opt.nTimeSteps) {
    check(ep2d_create_from_physics(&phys, &opt, &h_));
  }
  ~EffPlast2D() { ep2d_destroy(h_); }
  EffPlast2D(const EffPlast2D&) = delete;
  EffPlast2D& operator=(const EffPlast2D&) = delete;

  double ComputeEffModuli(double initLoadValue, const std::array<double, 3>& loadType,
                          double loadValue = 0.0) {
    check(ep2d_compute_eff_moduli(h_, initLoadValue, loadType.data(), loadValue, &res_));
    if (nTasks_ > 1 && nTimeSteps_) saveAnStatic(loadType);
    save(EP2D_FIELD_P, "data/Pc");
    save(EP2D_FIELD_TAUXX, "data/tauXXc");
    return 0.0;
  }
  double ComputeEffDamping(double initLoadValue, const std::array<double, 3>& loadType,
                           double loadValue = 0.0) {
    check(ep2d_compute_eff_damping(h_, initLoadValue, loadType.data(), loadValue, &res_));
    if (nTimeSteps_) saveAnStatic(loadType);
    return 0.0;
  }

  const ep2d_result& result() const { return res_; }
  std::vector<ep2d_record> records() const {
    int n = 0;
    check(ep2d_get_records(h_, nullptr, 0, &n));
    std::vector<ep2d_record> v((size_t)n);
    check(ep2d_get_records(h_, v.data(), n, &n));
    return v;
  }
  ep2d_handle* handle() { return h_; }

 private:
  static void check(int rc) {
    if (rc != EP2D_OK) throw std::runtime_error(std::string("ep2d: ") + ep2d_last_error());
  }
  void save(int field, const std::string& stem) {
    int m = 0, n = 0;
    check(ep2d_field_shape(h_, EP2D_FIELD_P, &m, &n));
    check(ep2d_save_field(h_, field, (stem + "_" + std::to_string(m) + "_.dat").c_str()));
  }
  // SaveAnStatic2D with deltaP and tauInfty of load step 0, last time step
  void saveAnStatic(const std::array<double, 3>& loadType) {
    const std::vector<ep2d_record> r = records();
    const ep2d_record& s0 = r.at(nTimeSteps_ - 1);
    check(ep2d_save_an_static(h_, s0.deltaP, s0.tauInfty, loadType.data(), "data", nullptr));
  }
  ep2d_handle* h_;
  int nTasks_;
  unsigned nTimeSteps_;
  ep2d_result res_{};
};

#endif
