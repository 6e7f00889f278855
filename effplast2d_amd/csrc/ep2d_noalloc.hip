// ep2d_noalloc.hip -- ep2d_residual_noalloc: the residual of ep2d_residual
// (EffPlast2D.cu:407-412, 805-813) on velocity fields the caller owns, with
// no device allocation, for callers that cannot allocate (code under HIP
// graph capture, serving loops).
//
// The reduction's scratch -- the bit patterns of max|Vx| and max|Vy| and the
// count of non-finite values, k_maxabs's three result words -- lives in a
// file-scope device array.  Each GPU holds its own copy until the process
// ends; the host looks its address up once per device (hipGetSymbolAddress,
// with that device current) and passes that address to the kernels.  Every
// call on a device shares the array: calls that overlap on one device race.
#include <cmath>
#include <mutex>

#include "ep2d.h"
#include "ep2d_internal.h"

namespace ep2d {

void record_error(const char* msg);  // ep2d_capi.cpp: ep2d_last_error()'s message

namespace {

constexpr int kScratchWords = 4;  // max|Vx| bits, max|Vy| bits, non-finite count, unused
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned long long* g_scratch[kMaxDevices];  // device address of `scratch`, per device

// The current device's copy of `scratch`, looked up on the first call on
// that device.
hipError_t scratch_on(int dev, unsigned long long** out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

// error = (max|Vx| / dXn + max|Vy| / dYn) * dT / norm, the host formula of
// ep2d_residual in the same order (dXn = dX (nX-1), dYn = dY (nY-1) and
// norm = max(|e0|, max(e1, e2)) formed on the host as there).
__global__ void k_residual_finish(const unsigned long long* __restrict__ s, double dXn,
                                  double dYn, double dT, double norm, double* __restrict__ error,
                                  unsigned long long* __restrict__ nonfinite) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double mx = __longlong_as_double((long long)s[0]);
  const double my = __longlong_as_double((long long)s[1]);
  *error = (mx / dXn + my / dYn) * dT / norm;
  if (nonfinite) *nonfinite = s[2];
}

double max_ref(double a, double b) { return (a < b) ? b : a; }  // std::max semantics

}  // namespace
}  // namespace ep2d

extern "C" int ep2d_residual_noalloc(const void* Vx, const void* Vy, int nX, int nY,
                                     int velocityF32, double dX, double dY, double dT,
                                     const double curEffStrain[3], double* error,
                                     unsigned long long* nonfinite, void* stream) {
  using namespace ep2d;
  if (!Vx || !Vy || !curEffStrain || !error) {
    record_error("ep2d_residual_noalloc: null argument");
    return EP2D_ERR_ARG;
  }
  if (nX < 3 || nY < 3 || nX > EP2D_NOALLOC_MAX_SIDE || nY > EP2D_NOALLOC_MAX_SIDE) {
    record_error("ep2d_residual_noalloc: nX and nY must lie in [3, EP2D_NOALLOC_MAX_SIDE]");
    return EP2D_ERR_ARG;
  }
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    record_error("ep2d_residual_noalloc: no current HIP device");
    return EP2D_ERR_HIP;
  }
  unsigned long long* sc = nullptr;
  e = scratch_on(dev, &sc);
  if (e != hipSuccess) {
    record_error("ep2d_residual_noalloc: hipGetSymbolAddress of the scratch array failed");
    return EP2D_ERR_HIP;
  }
  RveDev r{};
  r.Vx = const_cast<void*>(Vx);
  r.Vy = const_cast<void*>(Vy);
  r.nX = nX;
  r.nY = nY;
  hipStream_t s = static_cast<hipStream_t>(stream);
  try {
    // zeroes the three words with a kernel, then reduces into them
    launch_maxabs(r, velocityF32 != 0, sc, s);
  } catch (const HipError& ex) {
    record_error(ex.what());
    return EP2D_ERR_HIP;
  }
  const double* ce = curEffStrain;
  const double norm = max_ref(std::fabs(ce[0]), max_ref(ce[1], ce[2]));
  hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(64), 0, s, sc, dX * (nX - 1),
                     dY * (nY - 1), dT, norm, error, nonfinite);
  if (hipGetLastError() != hipSuccess) {
    record_error("ep2d_residual_noalloc: kernel launch failed");
    return EP2D_ERR_HIP;
  }
  return EP2D_OK;
}
