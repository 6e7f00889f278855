// ep2d_capi.cpp -- extern "C" entry points of libep2d.so (include/ep2d.h).
// Every call converts C++ exceptions into ep2d_status codes and records the
// message for ep2d_last_error(); nothing terminates the process.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ep2d_batch.h"

struct ep2d_handle {
  ep2d::Solver* solver;
};

struct ep2d_batch {
  ep2d::Batch* engine;
};

namespace {
// Message of the calling thread's last failure.  A plain character array:
// nothing is registered to run at thread or process exit.
thread_local char g_last_error[512];

void set_error(const char* msg) {
  std::snprintf(g_last_error, sizeof(g_last_error), "%s", msg ? msg : "");
}

// Set by the library's unload hook.  Destroy calls made after it (for
// instance from a language runtime's own teardown, after the HIP runtime or
// a profiler attached to it has shut down) release host memory only and
// make no HIP call; the process is ending and the driver reclaims the device
// memory.
bool g_unloading = false;
__attribute__((destructor)) void on_unload() { g_unloading = true; }

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return EP2D_OK;
  } catch (const ep2d::StatusError& e) {
    set_error(e.what());
    return e.status;
  } catch (const ep2d::ArgError& e) {
    set_error(e.what());
    return EP2D_ERR_ARG;
  } catch (const ep2d::HipError& e) {
    set_error(e.what());
    return EP2D_ERR_HIP;
  } catch (const std::invalid_argument& e) {
    set_error(e.what());
    return EP2D_ERR_ARG;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return EP2D_ERR_MEMORY;
  } catch (const std::exception& e) {
    set_error(e.what());
    return EP2D_ERR_STATE;
  }
}

}  // namespace

namespace ep2d {
// The message ep2d_last_error() returns, for entry points outside this file.
void record_error(const char* msg) { set_error(msg); }
}  // namespace ep2d

namespace {
ep2d::Solver* S(ep2d_handle* h) {
  if (!h || !h->solver) throw ep2d::ArgError("null handle");
  return h->solver;
}

ep2d::Batch* B(ep2d_batch* b) {
  if (!b || !b->engine) throw ep2d::ArgError("null batch");
  return b->engine;
}
}  // namespace

extern "C" {

const char* ep2d_last_error(void) { return g_last_error; }

const char* ep2d_version(void) { return "ep2d 0.1 gfx950"; }

int ep2d_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int ep2d_derive_params(const ep2d_physics* phys, double pa[EP2D_NPARS], int* nX) {
  return guarded([&] {
    if (!phys || !pa) throw ep2d::ArgError("null argument");
    ep2d::derive_params(*phys, pa, nX);
  });
}

int ep2d_read_pa(const char* path, double pa[EP2D_NPARS]) {
  return guarded([&] {
    if (!path || !pa) throw ep2d::ArgError("null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) throw ep2d::StatusError(EP2D_ERR_IO, std::string("ERROR:  Cannot open file ") + path + "!");
    const size_t n = std::fread(pa, sizeof(double), EP2D_NPARS, f);
    std::fclose(f);
    if (n != EP2D_NPARS)
      throw ep2d::StatusError(EP2D_ERR_IO, std::string("short parameter file ") + path);
  });
}

int ep2d_write_pa(const char* path, const double pa[EP2D_NPARS]) {
  return guarded([&] {
    if (!path || !pa) throw ep2d::ArgError("null argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) throw ep2d::StatusError(EP2D_ERR_IO, std::string("cannot create ") + path);
    const size_t n = std::fwrite(pa, sizeof(double), EP2D_NPARS, f);
    std::fclose(f);
    if (n != EP2D_NPARS) throw ep2d::StatusError(EP2D_ERR_IO, std::string("write failed ") + path);
  });
}

size_t ep2d_struct_size(const char* name) {
  if (!name) return 0;
  const std::string n(name);
  if (n == "ep2d_physics") return sizeof(ep2d_physics);
  if (n == "ep2d_options") return sizeof(ep2d_options);
  if (n == "ep2d_moduli") return sizeof(ep2d_moduli);
  if (n == "ep2d_record") return sizeof(ep2d_record);
  if (n == "ep2d_result") return sizeof(ep2d_result);
  if (n == "ep2d_an_static") return sizeof(ep2d_an_static);
  if (n == "ep2d_batch_options") return sizeof(ep2d_batch_options);
  if (n == "ep2d_request") return sizeof(ep2d_request);
  if (n == "ep2d_batch_result") return sizeof(ep2d_batch_result);
  if (n == "ep2d_batch_stats") return sizeof(ep2d_batch_stats);
  return 0;
}

void ep2d_default_options(ep2d_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->nTasks = 3;            // NL, EffPlast2D.h:25-27
  o->nIter = 100000;        // NITER, EffPlast2D.h:29-31
  o->eIter = 1.0e-10;       // EITER, EffPlast2D.h:33-35
  o->checkInterval = 10000; // output_step, EffPlast2D.cu:1608
  o->mode = EP2D_MODE_FP64;
  o->device = 0;            // DEVICE_IDX, EffPlast2D.h:37-39
  o->nTimeSteps = 1;        // EffPlast2D(const unsigned int nTimeSteps = 1)
  o->itersPerLaunch = 0;
  o->verbose = 0;
  o->logPath = nullptr;
}

int ep2d_create(const double pa[EP2D_NPARS], int nX, int nY, const ep2d_options* opt,
                ep2d_handle** out) {
  return ep2d_create_on_stream(pa, nX, nY, opt, nullptr, out);
}

int ep2d_create_on_stream(const double pa[EP2D_NPARS], int nX, int nY, const ep2d_options* opt,
                          void* stream, ep2d_handle** out) {
  return guarded([&] {
    if (!pa || !out) throw ep2d::ArgError("null argument");
    *out = nullptr;
    ep2d_options o;
    if (opt) o = *opt; else ep2d_default_options(&o);
    auto* h = new ep2d_handle{nullptr};
    try {
      h->solver = new ep2d::Solver(pa, nX, nY, o, static_cast<hipStream_t>(stream));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int ep2d_create_from_physics(const ep2d_physics* phys, const ep2d_options* opt, ep2d_handle** out) {
  double pa[EP2D_NPARS];
  int nX = 0;
  const int rc = ep2d_derive_params(phys, pa, &nX);
  if (rc) return rc;
  return ep2d_create(pa, nX, nX, opt, out);
}

void ep2d_destroy(ep2d_handle* h) {
  if (!h) return;
  if (g_unloading) return;   // process teardown: no HIP call (see on_unload)
  try {
    delete h->solver;
  } catch (...) {
  }
  delete h;
}

int ep2d_compute_eff_moduli(ep2d_handle* h, double initLoadValue, const double loadType[3],
                            double loadValue, ep2d_result* out) {
  (void)loadValue;  // deprecated in the reference (EffPlast2D.h:57)
  return guarded([&] {
    if (!loadType) throw ep2d::ArgError("null loadType");
    S(h)->computeEffModuli(initLoadValue, loadType, out);
  });
}

int ep2d_compute_eff_damping(ep2d_handle* h, double initLoadValue, const double loadType[3],
                             double loadValue, ep2d_result* out) {
  return guarded([&] {
    if (!loadType) throw ep2d::ArgError("null loadType");
    S(h)->computeEffDamping(initLoadValue, loadType, loadValue, out);
  });
}

int ep2d_get_records(ep2d_handle* h, ep2d_record* out, int capacity, int* count) {
  return guarded([&] {
    const auto& rec = S(h)->records();
    if (count) *count = (int)rec.size();
    if (out) {
      const int n = std::min<int>(capacity, (int)rec.size());
      for (int k = 0; k < n; k++) out[k] = rec[k];
    }
  });
}

int ep2d_field_shape(ep2d_handle* h, int field, int* m, int* n) {
  return guarded([&] {
    int a = 0, b = 0;
    S(h)->fieldShape(field, a, b);
    if (m) *m = a;
    if (n) *n = b;
  });
}

int ep2d_get_field(ep2d_handle* h, int field, double* host, size_t count) {
  return guarded([&] {
    if (!host) throw ep2d::ArgError("null buffer");
    S(h)->getField(field, host, count);
  });
}

int ep2d_set_field(ep2d_handle* h, int field, const double* host, size_t count) {
  return guarded([&] {
    if (!host) throw ep2d::ArgError("null buffer");
    S(h)->setField(field, host, count);
  });
}

int ep2d_an_static_fields(ep2d_handle* h, double deltaP, double tauInfty,
                          const double loadType[3], ep2d_an_static* out) {
  return guarded([&] {
    if (!out || !loadType) throw ep2d::ArgError("null argument");
    S(h)->anStatic(deltaP, tauInfty, loadType, out);
  });
}

int ep2d_save_an_static(ep2d_handle* h, double deltaP, double tauInfty,
                        const double loadType[3], const char* dir, ep2d_an_static* stats) {
  return guarded([&] {
    if (!loadType || !dir) throw ep2d::ArgError("null argument");
    S(h)->saveAnStatic(deltaP, tauInfty, loadType, dir, stats);
  });
}

int ep2d_save_field(ep2d_handle* h, int field, const char* path) {
  return guarded([&] {
    if (!path) throw ep2d::ArgError("null path");
    int m, n;
    S(h)->fieldShape(field, m, n);
    std::vector<double> buf((size_t)m * n);
    S(h)->getField(field, buf.data(), buf.size());
    FILE* f = std::fopen(path, "wb");
    if (!f) throw ep2d::StatusError(EP2D_ERR_IO, std::string("cannot create ") + path);
    const size_t w = std::fwrite(buf.data(), sizeof(double), buf.size(), f);
    std::fclose(f);
    if (w != buf.size()) throw ep2d::StatusError(EP2D_ERR_IO, std::string("write failed ") + path);
  });
}

int ep2d_get_mask(ep2d_handle* h, unsigned char* host, size_t count) {
  return guarded([&] {
    if (!host) throw ep2d::ArgError("null buffer");
    S(h)->getMask(host, count);
  });
}

int ep2d_iterate(ep2d_handle* h, long n) {
  return guarded([&] {
    if (n < 0) throw ep2d::ArgError("negative iteration count");
    S(h)->iterate(n);
  });
}

int ep2d_apply_affine(ep2d_handle* h, double dUxdx, double dUydy, double dUxdy, double dUydx) {
  return guarded([&] { S(h)->applyAffine(dUxdx, dUydy, dUxdy, dUydx); });
}

int ep2d_residual(ep2d_handle* h, const double curEffStrain[3], double* error, long* nonfinite) {
  return guarded([&] {
    if (!curEffStrain || !error) throw ep2d::ArgError("null argument");
    *error = S(h)->residual(curEffStrain, nonfinite);
  });
}

int ep2d_field_device(ep2d_handle* h, int field, const void** ptr) {
  return guarded([&] {
    if (!ptr) throw ep2d::ArgError("null argument");
    *ptr = nullptr;
    if (field < 0 || field >= EP2D_FIELD_COUNT) throw ep2d::ArgError("unknown field");
    *ptr = S(h)->fieldDevice(field);
  });
}

int ep2d_averages(ep2d_handle* h, ep2d_record* rec) {
  return guarded([&] {
    if (!rec) throw ep2d::ArgError("null record");
    S(h)->averages(*rec);
  });
}

int ep2d_boundary(ep2d_handle* h, double* deltaP, double* deltaPper, double* tauInfty) {
  return guarded([&] {
    double a, b, c;
    S(h)->boundary(a, b, c);
    if (deltaP) *deltaP = a;
    if (deltaPper) *deltaPper = b;
    if (tauInfty) *tauInfty = c;
  });
}

int ep2d_hole_area(ep2d_handle* h, ep2d_record* rec) {
  return guarded([&] {
    if (!rec) throw ep2d::ArgError("null record");
    S(h)->holeArea(*rec);
  });
}

int ep2d_set_timing(ep2d_handle* h, long sampleEvery) {
  return guarded([&] { S(h)->setTiming(sampleEvery); });
}

int ep2d_get_timing(ep2d_handle* h, double* avgKernelMs, long* samples, const char** kernelName) {
  return guarded([&] { S(h)->getTiming(avgKernelMs, samples, kernelName); });
}

int ep2d_diag_chain_order(const int* segs, int n, int stageWaves, const int* dims, int nRves,
                          int resident, int* out, int outCapacity, int* nOut) {
  return guarded([&] {
    if (!segs || !dims || !nOut || n < 0 || nRves < 1 || (stageWaves != 1 && stageWaves != 2))
      throw ep2d::ArgError("ep2d_diag_chain_order: bad argument");
    std::vector<int4> in((size_t)n);
    for (int k = 0; k < n; k++) {
      in[(size_t)k] = make_int4(segs[4 * k], segs[4 * k + 1], segs[4 * k + 2], segs[4 * k + 3]);
      if (in[(size_t)k].x < 0 || in[(size_t)k].x >= nRves)
        throw ep2d::ArgError("ep2d_diag_chain_order: segment names no RVE");
    }
    std::vector<int2> d((size_t)nRves);
    for (int k = 0; k < nRves; k++) d[(size_t)k] = make_int2(dims[2 * k], dims[2 * k + 1]);
    const int m = ep2d::chain_work_order(in.data(), n, stageWaves, d.data(), resident, nullptr);
    *nOut = m;
    if (!out) return;
    if (outCapacity < m) throw ep2d::ArgError("ep2d_diag_chain_order: out holds fewer entries");
    std::vector<int4> o((size_t)m);
    ep2d::chain_work_order(in.data(), n, stageWaves, d.data(), resident, o.data());
    for (int k = 0; k < m; k++) {
      out[4 * k] = o[(size_t)k].x;
      out[4 * k + 1] = o[(size_t)k].y;
      out[4 * k + 2] = o[(size_t)k].z;
      out[4 * k + 3] = o[(size_t)k].w;
    }
  });
}

int ep2d_diag_copy(int device, size_t bytes, int reps, double* avgMs) {
  return guarded([&] {
    if (!avgMs || reps < 1 || bytes < 16 || bytes > ((size_t)64 << 30))
      throw ep2d::ArgError("bad argument");
    EP2D_HIP(hipSetDevice(device));
    const size_t n = bytes / 16 * 2;
    double *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&] {
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (s) (void)hipStreamDestroy(s);
    };
    try {
      EP2D_HIP(hipMalloc(&a, n * 8));
      EP2D_HIP(hipMalloc(&b, n * 8));
      EP2D_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      EP2D_HIP(hipEventCreate(&e0));
      EP2D_HIP(hipEventCreate(&e1));
      ep2d::launch_fill(a, n, 1.0, s);
      ep2d::launch_copy(a, b, n, s);   // warm-up
      EP2D_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++) ep2d::launch_copy(a, b, n, s);
      EP2D_HIP(hipEventRecord(e1, s));
      EP2D_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      EP2D_HIP(hipEventElapsedTime(&ms, e0, e1));
      *avgMs = ms / reps;
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int ep2d_sync(ep2d_handle* h) {
  return guarded([&] { S(h)->sync(); });
}

size_t ep2d_device_bytes(ep2d_handle* h) {
  if (!h || !h->solver) return 0;
  return h->solver->deviceBytes();
}

// ---- batched engine -------------------------------------------------------------
void ep2d_batch_default_options(ep2d_batch_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->mode = EP2D_MODE_SPLIT;
  o->checkInterval = 500;
  o->tileRows = 0;
  o->device = 0;
  o->memoryBudget = 0;
  o->maxActive = 0;
  o->useGraphs = 1;
  o->timingEvery = 0;
}

int ep2d_batch_request_bytes(const ep2d_batch_options* opt, const ep2d_request* rq,
                             size_t* bytes) {
  return guarded([&] {
    if (!opt || !rq || !bytes) throw ep2d::ArgError("null argument");
    if (opt->mode != EP2D_MODE_FP64 && opt->mode != EP2D_MODE_SPLIT)
      throw ep2d::ArgError("batch mode must be EP2D_MODE_FP64 or EP2D_MODE_SPLIT");
    if (rq->nX < 8) throw ep2d::ArgError("request: nX must be >= 8");
    *bytes = ep2d::Solver::plannedDeviceBytes(rq->nX, rq->nX, opt->mode,
                                              rq->pa[EP2D_PA_NPORES]);
  });
}

int ep2d_batch_create(const ep2d_batch_options* opt, ep2d_batch** out) {
  return guarded([&] {
    if (!out) throw ep2d::ArgError("null argument");
    *out = nullptr;
    ep2d_batch_options o;
    if (opt) o = *opt; else ep2d_batch_default_options(&o);
    auto* b = new ep2d_batch{nullptr};
    try {
      b->engine = new ep2d::Batch(o);
    } catch (...) {
      delete b;
      throw;
    }
    *out = b;
  });
}

void ep2d_batch_destroy(ep2d_batch* b) {
  if (!b) return;
  if (g_unloading) return;   // process teardown: no HIP call (see on_unload)
  try {
    delete b->engine;
  } catch (...) {
  }
  delete b;
}

int ep2d_batch_submit(ep2d_batch* b, const ep2d_request* rq, long long* id) {
  return guarded([&] {
    if (!rq) throw ep2d::ArgError("null request");
    const long long v = B(b)->submit(*rq);
    if (id) *id = v;
  });
}

int ep2d_batch_step(ep2d_batch* b, int* nActive, int* nQueued) {
  return guarded([&] { B(b)->step(nActive, nQueued); });
}

int ep2d_batch_run(ep2d_batch* b) {
  return guarded([&] { B(b)->run(); });
}

int ep2d_batch_poll(ep2d_batch* b, ep2d_batch_result* out, int capacity, int* count) {
  return guarded([&] {
    if (!out && capacity > 0) throw ep2d::ArgError("null buffer");
    const int n = B(b)->poll(out, capacity);
    if (count) *count = n;
  });
}

double ep2d_batch_clock(ep2d_batch* b) {
  if (!b || !b->engine) return 0.0;
  return b->engine->clock();
}

int ep2d_batch_stats_get(ep2d_batch* b, ep2d_batch_stats* out) {
  return guarded([&] {
    if (!out) throw ep2d::ArgError("null argument");
    B(b)->stats(out);
  });
}

}  // extern "C"
