// Data-movement, predicate, validation and affine/extended I/O kernels
// (the reference's ftcrypto.cu:390-626, 659-892, 1014-1086, 1262-1325).
#include "misc.h"

namespace ft {

constexpr int MTPB = 256;
static inline unsigned mgrid(size_t n) { return (unsigned)((n + MTPB - 1) / MTPB); }

// dst[i] = src[idx ? idx[i] : start + i * step]
__global__ void __launch_bounds__(MTPB) k_gather(PtView dst, PtView src, const uint64_t *idx, size_t start,
                                                 size_t step, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  const size_t j = idx ? (size_t)idx[i] : start + i * step;
  copy_point(dst, i, src, j);
}

// dst[idx ? idx[i] : start + i * step] = src[i] (identity when src.w is null)
__global__ void __launch_bounds__(MTPB) k_scatter(PtView dst, PtView src, const uint64_t *idx, size_t start,
                                                  size_t step, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  const size_t j = idx ? (size_t)idx[i] : start + i * step;
  if (src.w) copy_point(dst, j, src, i);
  else store_identity(dst, j);
}

__global__ void __launch_bounds__(MTPB) k_mux(PtView t, const int64_t *cond, PtView f, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  if (cond[i] == 0) copy_point(t, i, f, i);
}

// res[i] = (v[i] != identity): X == 0 and Y == Z (ftcrypto.cu:170-179)
__global__ void __launch_bounds__(MTPB) k_index(int64_t *res, PtView v, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  ge_p3 p;
  load_p3(p, v, i);
  fe d;
  fe_sub(d, p.Y, p.Z);
  res[i] = (fe_iszero(p.X) & fe_iszero(d)) ? 0 : 1;
}

// res[i] = (a[i] == b[i]) xor neq, by X1 Z2 = X2 Z1 and Y1 Z2 = Y2 Z1
__global__ void __launch_bounds__(MTPB) k_equal(int64_t *res, PtView a, PtView b, size_t n, int neq) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  ge_p3 p, q;
  load_p3(p, a, i);
  load_p3(q, b, i);
  fe l, r, d;
  fe_mul(l, p.X, q.Z);
  fe_mul(r, q.X, p.Z);
  fe_sub(d, l, r);
  uint32_t eq = fe_iszero(d);
  fe_mul(l, p.Y, q.Z);
  fe_mul(r, q.Y, p.Z);
  fe_sub(d, l, r);
  eq &= fe_iszero(d);
  res[i] = (int64_t)(eq ^ (uint32_t)neq);
}

// on the curve ((Y^2 - X^2) Z^2 = Z^4 + d X^2 Y^2) and X(l P) = 0
// (the reference's is_on_curve / is_on_main_subgroup, sodium_ge.cu:849-882)
__global__ void __launch_bounds__(MTPB) k_validate(unsigned long long *bad, PtView v, size_t n,
                                                   const uint32_t *lwords) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  long long fail = -1;
  if (i < n) {
    ge_p3 p;
    load_p3(p, v, i);
    fe x2, y2, z2, z4, t0, t1;
    fe_sq(x2, p.X);
    fe_sq(y2, p.Y);
    fe_sq(z2, p.Z);
    fe_sub(t0, y2, x2);
    fe_mul(t0, t0, z2);
    fe_mul(t1, x2, y2);
    fe_mul(t1, t1, FE_D);
    fe_sq(z4, z2);
    fe_add(t1, t1, z4);
    fe_carry(t1);
    fe_sub(t0, t0, t1);
    const uint32_t on_curve = fe_iszero(t0);
    uint32_t k[8], e[16];
#pragma unroll
    for (int q = 0; q < 8; ++q) k[q] = lwords[q];
    recode16_words(e, k);
    ge_p3 h;
    varbase_mul(h, p, e);
    const uint32_t in_sub = fe_iszero(h.X);
    fail = (on_curve & in_sub) ? -1 : (long long)i;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const long long o = __shfl_xor(fail, off);
    fail = o > fail ? o : fail;
  }
  if ((threadIdx.x & 63) == 0 && fail >= 0) atomicMax(bad, (unsigned long long)fail + 1ull);
}

// affine (x, y), 64 bytes each, canonical (ftcrypto.cu:659-677)
__global__ void __launch_bounds__(MTPB) k_to_affine(uint32_t *out, PtView v, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  ge_p3 p;
  load_p3(p, v, i);
  fe zi, x, y;
  fe_invert(zi, p.Z);
  fe_mul(x, p.X, zi);
  fe_mul(y, p.Y, zi);
  uint32_t w[8];
  fe_canonical_words(w, x);
  st_words(out + i * 16, w);
  fe_canonical_words(w, y);
  st_words(out + i * 16 + 8, w);
}

// Z = 1, T = x y, bit 255 of each coordinate ignored (ftcrypto.cu:793-805)
__global__ void __launch_bounds__(MTPB) k_from_affine(PtView v, const uint32_t *in, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  uint32_t w[8];
  ge_p3 p;
  ld_words(w, in + i * 16);
  w[7] &= 0x7fffffffu;
  fe_from_words(p.X, w);
  ld_words(w, in + i * 16 + 8);
  w[7] &= 0x7fffffffu;
  fe_from_words(p.Y, w);
  fe_1(p.Z);
  fe_mul(p.T, p.X, p.Y);
  store_p3(v, i, p);
}

// raw 160-byte records: the four resident 32-byte words X|Y|Z|T and 32 zero bytes
__global__ void __launch_bounds__(MTPB) k_to_raw(uint32_t *out, PtView v, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  uint32_t w[8];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    ld_words(w, v.w + (c * v.cap + i) * 8);
    st_words(out + i * 40 + c * 8, w);
  }
#pragma unroll
  for (int q = 0; q < 8; ++q) w[q] = 0;
  st_words(out + i * 40 + 32, w);
}

__global__ void __launch_bounds__(MTPB) k_from_raw(PtView v, const uint32_t *in, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  uint32_t w[8];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    ld_words(w, in + i * 40 + c * 8);
    st_words(v.w + (c * v.cap + i) * 8, w);
  }
}

// ---------------------------------------------------------------- launchers

#define MCHK() return hipGetLastError()

hipError_t launch_gather(PtView dst, PtView src, const uint64_t *idx, size_t start, size_t step, size_t n,
                         hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3(mgrid(n)), dim3(MTPB), 0, s, dst, src, idx, start, step, n);
  MCHK();
}
hipError_t launch_scatter(PtView dst, PtView src, const uint64_t *idx, size_t start, size_t step, size_t n,
                          hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_scatter, dim3(mgrid(n)), dim3(MTPB), 0, s, dst, src, idx, start, step, n);
  MCHK();
}
hipError_t launch_mux(PtView t, const int64_t *cond, PtView f, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_mux, dim3(mgrid(n)), dim3(MTPB), 0, s, t, cond, f, n);
  MCHK();
}
hipError_t launch_index(int64_t *res, PtView v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_index, dim3(mgrid(n)), dim3(MTPB), 0, s, res, v, n);
  MCHK();
}
hipError_t launch_equal(int64_t *res, PtView a, PtView b, size_t n, bool neq, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_equal, dim3(mgrid(n)), dim3(MTPB), 0, s, res, a, b, n, (int)neq);
  MCHK();
}
hipError_t launch_validate(unsigned long long *bad, PtView v, size_t n, const uint32_t *lwords, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_validate, dim3(mgrid(n)), dim3(MTPB), 0, s, bad, v, n, lwords);
  MCHK();
}
hipError_t launch_to_affine(uint32_t *out, PtView v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_to_affine, dim3(mgrid(n)), dim3(MTPB), 0, s, out, v, n);
  MCHK();
}
hipError_t launch_from_affine(PtView v, const uint32_t *in, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_from_affine, dim3(mgrid(n)), dim3(MTPB), 0, s, v, in, n);
  MCHK();
}
hipError_t launch_to_raw(uint32_t *out, PtView v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_to_raw, dim3(mgrid(n)), dim3(MTPB), 0, s, out, v, n);
  MCHK();
}
hipError_t launch_from_raw(PtView v, const uint32_t *in, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_from_raw, dim3(mgrid(n)), dim3(MTPB), 0, s, v, in, n);
  MCHK();
}

}  // namespace ft
