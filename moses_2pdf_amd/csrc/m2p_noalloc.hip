// m2p_noalloc.hip -- entry points that allocate no device memory.
//
// m2p_mass_noalloc is m2p_mass for a state the caller holds in device memory of its own,
// for callers that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid device allocation).  It runs the same two reductions as
// m2p_mass, the block partials going to a fixed device array of this file instead of a
// buffer the library allocates, and the final double-double sum on the device instead of
// the host, so its totals equal m2p_mass's bit for bit.

#include <hip/hip_runtime.h>

#include <climits>
#include <mutex>
#include <string>

#include "../../include/moses2pdf.h"
#include "m2p_kernels.h"

// The per-block partial sums of one call: 4 doubles (solid hi, lo, fluid hi, lo) for
// each of the 512 blocks of the mass reduction, 16 KiB, whatever the grid size (the
// blocks stride over the cells).  Each GPU the library runs on holds its own copy until
// the process ends.  Every call writes every slot before the final sum reads it, so
// nothing needs clearing; calls that overlap on one GPU share it and race.
constexpr int kMassScratch = 4 * 512;
static __device__ __align__(256) double scratch[kMassScratch];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};   // device address of `scratch` on each device
thread_local std::string g_err;    // message of this thread's last failed context-free call

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

// The address of the current device's copy of `scratch`, looked up once per device with
// that device current.  In host code the array's name is only a stand-in for the symbol;
// its device address must come from the runtime.
hipError_t scratch_ptr(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = (double*)p;
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" {

int m2p_mass_noalloc(const void* u, int precision, const double* detmin, int nx, int ny,
                     double* out, void* stream) {
  if (!u || !detmin || !out) return fail(M2P_EINVAL, "m2p_mass_noalloc: null pointer");
  if (precision != M2P_FP64 && precision != M2P_FP32)
    return fail(M2P_EINVAL, "m2p_mass_noalloc: precision must be M2P_FP64 or M2P_FP32");
  if (nx < 1 || ny < 1 || (long long)nx * ny > INT_MAX)
    return fail(M2P_EINVAL, "m2p_mass_noalloc: grid must have 1 <= nx, ny and nx*ny <= INT_MAX");
  if (m2p::mass_blocks() * 4 > kMassScratch)
    return fail(M2P_EINVAL, "m2p_mass_noalloc: reduction larger than its scratch array");
  double* part = nullptr;
  hipError_t e = scratch_ptr(&part);
  if (e != hipSuccess)
    return fail(M2P_EHIP, std::string("m2p_mass_noalloc: scratch lookup: ") + hipGetErrorString(e));
  m2p::Grid g{};
  g.nx = nx;
  g.ny = ny;
  g.nxd = nx + 2 * m2p::MD;
  g.nyd = ny + 2 * m2p::MD;
  g.P = (long long)g.nxd * g.nyd;
  hipStream_t s = (hipStream_t)stream;
  if (precision == M2P_FP32)
    m2p::launch_mass(g, (const float*)u, detmin, part, s);
  else
    m2p::launch_mass(g, (const double*)u, detmin, part, s);
  m2p::launch_mass_sum(part, m2p::mass_blocks(), out, s);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(M2P_EHIP, std::string("m2p_mass_noalloc: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"

namespace m2p {
const char* context_free_error() { return g_err.c_str(); }
}  // namespace m2p
