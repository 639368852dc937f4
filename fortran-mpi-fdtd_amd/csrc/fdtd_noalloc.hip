// fdtd_noalloc.hip -- entry points that allocate no device memory.
//
// fdtd_get_flux_noalloc is fdtd_get_flux for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses
// that forbid it): the per-frequency flux sums go to `scratch`, a fixed
// file-scope device array (one copy per GPU, held until the process ends),
// instead of a buffer the handle grows with stream-ordered allocation.
//
// `scratch` as written in host code is a host-side stand-in, not a device
// address.  Its device address on each GPU is taken once, with that GPU
// current, by hipGetSymbolAddress and kept in g_scratch[device]; host code
// passes only that pointer to kernels and copies.
//
// Every call on one GPU shares that GPU's array, which keeps the last
// call's contents: the flux kernel writes each of its nfreq entries whole,
// so nothing needs clearing between calls, but calls that overlap on one
// GPU (from several host threads) race and must be serialized by the
// caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "fdtd_internal.h"

static __device__ __align__(256) double scratch[FDTD_FLUX_NOALLOC_MAX_FREQ];

namespace fdtd {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` per GPU

// Device address of `scratch` on `device`, which must be current: looked up
// on the first call for that device, then reused.
int scratch_on(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return set_error(FDTD_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = (double*)p;
  }
  *out = g_scratch[device];
  return 0;
}

}  // namespace
}  // namespace fdtd

using namespace fdtd;

extern "C" int fdtd_get_flux_noalloc(fdtd_sim* s, int id, double* S) {
  if (!s) return set_error(FDTD_ERR_ARG, "null handle");
  if (!S || id < 0 || id >= (int)s->dfts.size()) return set_error(FDTD_ERR_ARG, "bad argument");
  // committing allocates: the handle must have been committed already
  if (!s->committed) return set_error(FDTD_ERR_STATE, "fdtd_get_flux_noalloc needs a committed handle");
  const int nfreq = s->dfts[id].nfreq;
  if (nfreq > FDTD_FLUX_NOALLOC_MAX_FREQ)
    return set_error(FDTD_ERR_ARG, "fdtd_get_flux_noalloc: plane has " + std::to_string(nfreq) +
                                       " frequencies, more than the " +
                                       std::to_string(FDTD_FLUX_NOALLOC_MAX_FREQ) + " it reduces without allocating");
  int rc = flux_check(s, id);
  if (rc) return rc;
  FDTD_HIP(hipSetDevice(s->device));
  double* dS = nullptr;
  rc = scratch_on(s->device, &dS);
  if (rc) return rc;
  return launch_flux_into(s, id, dS, S);
}
