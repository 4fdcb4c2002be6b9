// amg_noalloc.hip -- the divergence report for callers that cannot allocate.
//
// amg_divergence_device (dns/Routines.F90:54-183) allocates its block
// partials with hipMalloc on its first call.  amg_divergence_device_noalloc
// runs the same kernels with the partials and the results in `scratch`, a
// file-scope device array of kNoAllocDoubles doubles (256 KiB), so that it
// can be called under graph capture, from serving loops or where the
// harness forbids allocation.  Each GPU the library runs on holds its own
// copy of the array until the process ends.  On the host the array's name
// is only a stand-in: the device address is resolved once per device with
// hipGetSymbolAddress (at the entry point's first call on that device, with
// that device current) and only that address is used afterwards.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it; calls that overlap on one GPU (two streams, two threads)
// race.  Nothing in the array needs to start at zero: every block writes
// its four partials and the fold writes the four results before they are
// read.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "amg.h"
#include "amg_handle.h"
#include "amg_internal.h"
#include "amg_solver.h"

namespace {
constexpr int kNoAllocMaxBlocks = 8192;   // grids of at most 33 554 432 cells
constexpr size_t kNoAllocDoubles = 4 * (size_t)kNoAllocMaxBlocks + 4;
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

// This device's address of `scratch`, resolved on first use with the
// device current (the caller's current device is restored); nullptr and an
// error message on failure.
double* scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    amg::set_error("amg_divergence_device_noalloc: device id out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ptr[dev]) return g_ptr[dev];
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess) cur = -1;
  if (cur != dev && hipSetDevice(dev) != hipSuccess) {
    amg::set_error("amg_divergence_device_noalloc: hipSetDevice failed");
    return nullptr;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p) {
    amg::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return nullptr;
  }
  g_ptr[dev] = static_cast<double*>(p);
  return g_ptr[dev];
}

}  // namespace

extern "C" int amg_divergence_device_noalloc(amg_handle h, const double* U, const double* V,
                                             const double* W, double* div, double* avg,
                                             double* maxabs, int loc[3]) {
  if (!h || !U || !V || !W) {
    amg::set_error("amg_divergence_device_noalloc: NULL argument");
    return AMG_EINVAL;
  }
  const int nb = amg::divergence_blocks(h->I->nx, h->I->ny, h->I->nz);
  if (nb > kNoAllocMaxBlocks) {
    amg::set_error("amg_divergence_device_noalloc: the grid has more than " +
                   std::to_string((long long)kNoAllocMaxBlocks * amg::kBlock * amg::kDivPer) +
                   " cells, the limit of its static scratch; use amg_divergence_device");
    return AMG_EBADSIZE;
  }
  double* ptr = scratch_on(h->ctx.device);
  if (!ptr) return AMG_EHIP;
  return amg::divergence_report(h, U, V, W, div, ptr, avg, maxabs, loc);
}
