// Static device workspace of the allocation-free reduce (ftx_reduce_noalloc_dev).
//
// The block sums of a reduce (at most REDUCE_BLOCKS_MAX points, whatever the
// input length) live in a file-scope device array instead of a caller
// workspace or a library allocation, for callers that cannot allocate.
// Every GPU the library runs on holds its own copy of the array, at its own
// address, until the process ends.  In host code the array's name is only a
// host-side stand-in: its device address is taken once per device with
// hipGetSymbolAddress (at ft_crypto_init for the library's device, else on
// the first call on another device) and only that address is passed on.
// Every call on one GPU shares that GPU's array and leaves the last call's
// block sums in it; calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.h"

namespace ft {

// four coordinate planes of NOALLOC_POINTS points each (512 KiB)
static __device__ __align__(256) uint32_t scratch[4 * NOALLOC_POINTS * 8];

namespace {
constexpr int NOALLOC_DEVICES = 64;
std::mutex noalloc_mu;
void *noalloc_ptr[NOALLOC_DEVICES];
}  // namespace

hipError_t noalloc_resolve(int device) {
  if (device < 0 || device >= NOALLOC_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (noalloc_ptr[device]) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur != device) {
    const hipError_t e2 = hipSetDevice(cur);
    if (e == hipSuccess) e = e2;
  }
  if (e != hipSuccess) return e;
  noalloc_ptr[device] = p;
  return hipSuccess;
}

hipError_t noalloc_scratch(PtView *v) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if ((e = noalloc_resolve(dev)) != hipSuccess) return e;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  *v = PtView{static_cast<uint32_t *>(noalloc_ptr[dev]), NOALLOC_POINTS};
  return hipSuccess;
}

}  // namespace ft
