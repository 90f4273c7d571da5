// Variable-base scalar multiplication kernel.  It is compiled in a file of
// its own because it runs at one wave per SIMD (its 8-entry point table lives
// in registers) and gains from the instruction scheduler's max-ILP strategy,
// which the other kernels do not (build.py passes the option for this file).
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "elem.h"

namespace ft {

// ------------------------------------------------------ variable base

// pts[i] = k_i * pts[i]; scalar stride 0 applies one scalar to every element
// (ft_scale, ftcrypto.cu:966-983), stride 8 words is ft_mul (:986-1011).
__global__ void __launch_bounds__(TPB) k_varbase(PtView pts, const uint32_t *sc, size_t sc_stride,
                                                 size_t n) {
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (i >= n) return;
  uint32_t k[8], e[16];
  ld_words(k, sc + i * sc_stride);
  recode16(e, k);
  ge_p3 p, h;
  load_p3(p, pts, i);
  varbase_windows(h, p, e, 0);
  store_p3(pts, i, h);
}

hipError_t launch_varbase(PtView pts, const uint32_t *sc, size_t sc_stride_words, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_varbase, dim3((unsigned)((n + TPB - 1) / TPB)), dim3(TPB), 0, s, pts, sc, sc_stride_words, n);
  return hipGetLastError();
}

}  // namespace ft
