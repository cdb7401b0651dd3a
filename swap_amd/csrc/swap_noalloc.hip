// swap_noalloc.hip -- swap_get_particle_array_noalloc: the export of
// swap_get_particle_array for callers that cannot allocate device memory.
//
// swap_get_particle_array(SWAP_P_XP) turns the split positions (cell +
// fraction) into absolute FP64 positions in a device staging buffer that it
// allocates and frees on every call.  The variant here stages through a
// fixed buffer of this file, `scratch`: room for SWAP_NOALLOC_MAX particles
// (3 x 2^20 doubles, 24 MiB), held on every GPU the library runs on until
// the process ends.  Its device address is taken once per device, at
// swap_create on that device; the host-side name `scratch` is never used as
// a device pointer.  Every call on one GPU shares the buffer (last call's
// contents stay); the export writes all 3 n values it copies back, so
// nothing needs clearing.  Calls that overlap on one GPU race.
#include <mutex>
#include "swap_internal.h"

namespace {

__device__ __align__(256) double scratch[3 * (size_t)SWAP_NOALLOC_MAX];

constexpr int MAX_DEVICES = 64;
void* g_scratch[MAX_DEVICES] = {};  // device address of `scratch`, per device
std::mutex g_mu;

}  // namespace

// once per device, with that device current (swap_create calls it)
int na_setup(swap_ctx* c) {
  if (c->device < 0 || c->device >= MAX_DEVICES)
    return set_err(c, SWAP_EINVAL, "no-alloc staging: device %d out of range", c->device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[c->device]) return 0;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_err(c, SWAP_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch[c->device] = p;
  return 0;
}

extern "C" int swap_get_particle_array_noalloc(swap_ctx* c, int which, void* host) {
  if (!c || !host) return SWAP_EINVAL;
  // every other array is a plain device-to-host copy with no staging
  if (which != SWAP_P_XP) return swap_get_particle_array(c, which, host);
  if (int e = compact(c)) return e;  // an open x: drop the ions that left (a sort, no allocation)
  if (c->n > SWAP_NOALLOC_MAX)
    return set_err(c, SWAP_EINVAL,
                   "get_particle_array_noalloc: %lld particles, the fixed staging buffer holds "
                   "at most %d (use swap_get_particle_array)",
                   (long long)c->n, SWAP_NOALLOC_MAX);
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (c->device >= 0 && c->device < MAX_DEVICES) p = g_scratch[c->device];
  }
  if (!p) return set_err(c, SWAP_ESTATE, "no-alloc staging not set up on device %d", c->device);
  if (c->n == 0) return 0;
  if (int e = pl_export_abs(c, (double*)p)) return e;
  HIPCHK(c, hipMemcpyAsync(host, p, 24 * (size_t)c->n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return 0;
}
