// Allocation-free total energy of device-resident arrays
// (wdy_energy_device_noalloc): the computation of wdy_energy_arrays, with
// every scratch buffer taken from one file-scope device array instead of
// hipMalloc, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops).
//
// Each GPU holds its own copy of `scratch` for the life of the process.  Its
// address on the current device is looked up once per device with
// hipGetSymbolAddress (the array's name in host code is only a stand-in) and
// cached; every kernel receives that address.  Layout: the radix sort's
// look-back words for the largest n first -- always at the same place, so
// the epoch tags of earlier calls keep them unambiguous -- then the rest of
// the workspace, re-initialised by the kernels of each call.  The sort epoch
// continues from call to call on each device.
#include <mutex>

#include "wdy_internal.h"

namespace wdy {
namespace {

constexpr u32 MAX_N = WDY_NOALLOC_MAX_N;
constexpr size_t STATUS_BYTES =  // 2 KiB per radix tile
    ((size_t)(MAX_N + RADIX_TILE_KEYS - 1) / RADIX_TILE_KEYS) * 256 * sizeof(u64);
constexpr size_t WORK_BYTES = (size_t)49 << 20;
constexpr size_t SCRATCH_BYTES = STATUS_BYTES + WORK_BYTES;

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

struct DeviceScratch {
  unsigned char *ptr = nullptr;  // address of `scratch` on the device
  u32 epoch = 1;                 // next radix-sort epoch
};
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
DeviceScratch g_dev[MAX_DEVICES];

}  // namespace
}  // namespace wdy

using namespace wdy;

extern "C" int wdy_energy_device_noalloc(int64_t n, const double *x,
                                         const double *v, const double *m,
                                         double twopiG, double omega2,
                                         double *e_out, void *stream) {
  if (n < 1 || n > (int64_t)MAX_N)
    return fail(WDY_EINVAL,
                "wdy_energy_device_noalloc: n=%lld outside [1, %u] (the size "
                "of its static scratch)",
                (long long)n, MAX_N);
  if (!x || !v || !m || !e_out)
    return fail(WDY_EINVAL, "wdy_energy_device_noalloc: NULL pointer");
  if (!(twopiG != 0.0) || !std::isfinite(twopiG) || !std::isfinite(omega2))
    return fail(WDY_EINVAL,
                "wdy_energy_device_noalloc: twopiG must be finite and nonzero");
  const u32 nn = (u32)n;
  if (radix_status_bytes(nn) > STATUS_BYTES ||
      energy_scratch_bytes(nn, true) > WORK_BYTES)
    return fail(WDY_ERUNTIME, "wdy_energy_device_noalloc: scratch layout "
                "exceeds the static array");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return fail(WDY_ENODEV, "wdy_energy_device_noalloc: no GPU visible");
  if (dev < 0 || dev >= MAX_DEVICES)
    return fail(WDY_EUNSUPPORTED, "wdy_energy_device_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(g_mu);
  DeviceScratch &ds = g_dev[dev];
  if (!ds.ptr) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(WDY_ERUNTIME,
                  "wdy_energy_device_noalloc: hipGetSymbolAddress failed: %s",
                  hipGetErrorString(e));
    ds.ptr = (unsigned char *)p;
  }
  return energy_enqueue((hipStream_t)stream, nn, x, v, m, 0.0, 1.0, twopiG,
                        omega2, e_out, (u64 *)ds.ptr, ds.ptr + STATUS_BYTES,
                        &ds.epoch, nullptr);
}
