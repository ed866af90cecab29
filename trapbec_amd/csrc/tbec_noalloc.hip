// Allocation-free diagnostics: tbec_diagnose with its reduction partials in a file-scope
// device array instead of a buffer the library allocates on demand, for callers that
// must not allocate device memory (code under graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// Every call on one GPU shares that GPU's array.  Each call writes every partial it later
// reads (one 4-double slot per workgroup of the rhs and Psi_info passes), so nothing has
// to start at zero; calls that overlap on one GPU (two handles on two streams, or two
// threads) race on the array and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "tbec_internal.hpp"

using namespace tbec;

// 4 doubles per workgroup: 32768 slots, 1 MiB held on each GPU the library runs on until
// the process ends.
static constexpr long long kScratchSlots = 32768;
static __device__ __align__(256) double scratch[4 * kScratchSlots];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on device `dev`, looked up once per device with that
// device current.  In host code the array's name is a host-side stand-in, not a device
// address: only the looked-up address is ever handed to kernels.
static int scratch_on(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices)
    return set_error(TBEC_ERR_ARG, "device id " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess && prev != dev) e = hipSetDevice(dev);
    void *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != dev) (void)hipSetDevice(prev);
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return set_error(TBEC_ERR_HIP, std::string("scratch symbol address: ") +
                                         hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" int tbec_diagnose_noalloc(tbec_handle *h, tbec_diag *out) {
  if (!h || !out) return set_error(TBEC_ERR_ARG, "null argument");
  const long long need = diagnose_partial_slots(h);
  if (need > kScratchSlots)
    return set_error(TBEC_ERR_ARG,
                     "tbec_diagnose_noalloc: the grid, tile and segment need " +
                         std::to_string(need) + " reduction slots; the static buffer holds " +
                         std::to_string(kScratchSlots) + " (use tbec_diagnose, or a longer "
                         "segment via tbec_set_tuning)");
  if (physics_would_allocate(h, h->t_now))
    return set_error(TBEC_ERR_STATE,
                     "tbec_diagnose_noalloc: the tabulated potential of the current time is "
                     "not built yet (a step or tbec_diagnose builds it)");
  double *part = nullptr;
  int rc = scratch_on(h->device, &part);
  if (rc) return rc;
  return diagnose_with(h, out, part, kScratchSlots);
}
