// Variable-base scalar multiplication kernels, radix 16 and radix 8.  They
// are compiled in a file of their own because the radix-16 kernel runs at one
// wave per SIMD (its 8-entry point table lives in registers) and gains from
// the instruction scheduler's max-ILP strategy, which the other kernels do not
// (build.py passes the option for this file).
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

// the same in radix 8: a 4-entry table (two entries in per-lane LDS slots,
// 56 KiB per block), two waves per SIMD, the reference's result for every
// scalar (varbase_mul8)
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_varbase_r8(PtView pts, const uint32_t *sc, size_t sc_stride, size_t n) {
  static_assert(TPB == PARK_STRIDE, "one LDS slot per lane");
  __shared__ uint4 park[VB8_PARK_WORDS / 4 * PARK_STRIDE];
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (i >= n) return;
  uint32_t k[8];
  ld_words(k, sc + i * sc_stride);
  ge_p3 p, h;
  load_p3(p, pts, i);
  varbase_mul8(h, p, k, park + threadIdx.x);
  store_p3(pts, i, h);
}

hipError_t launch_varbase_w(PtView pts, const uint32_t *sc, size_t sc_stride_words, size_t n, int w,
                            hipStream_t s) {
  if (!n) return hipSuccess;
  const dim3 grid((unsigned)((n + TPB - 1) / TPB));
  if (w == 3)
    hipLaunchKernelGGL(k_varbase_r8, grid, dim3(TPB), 0, s, pts, sc, sc_stride_words, n);
  else if (w == 4)
    hipLaunchKernelGGL(k_varbase, grid, dim3(TPB), 0, s, pts, sc, sc_stride_words, n);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_varbase(PtView pts, const uint32_t *sc, size_t sc_stride_words, size_t n, hipStream_t s) {
  return launch_varbase_w(pts, sc, sc_stride_words, n, tune_for(n).vbw, s);
}

}  // namespace ft
