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
// (the reference's is_on_curve / is_on_main_subgroup, sodium_ge.cu:849-882).
// l is public and the same for every lane, so l P is a double-and-add over
// its width-3 NAF (digits read as wave-uniform scalars: the branches are
// scalar, nothing depends on secret data): 252 doublings, about 30 additions
// of P or 3P, instead of a constant-time windowed multiply.
__global__ void __launch_bounds__(MTPB) k_validate(unsigned long long *bad, PtView v, size_t n,
                                                   const int8_t *naf, int len) {
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
    ge_cached c1, c3;
    ge_p1p1 r;
    ge_p3 q;
    ge_p3_to_cached(c1, p);
    ge_p3_dbl(r, p);
    ge_p1p1_to_p3(q, r);                        // 2P
    ge_add(r, q, c1);
    ge_p1p1_to_p3(q, r);                        // 3P
    ge_p3_to_cached(c3, q);
    ge_p2 s;                                    // the top digit (positive) starts the sum
    const int top = __builtin_amdgcn_readfirstlane((int)naf[len - 1]);
    if (top == 3) { s.X = q.X; s.Y = q.Y; s.Z = q.Z; }
    else { s.X = p.X; s.Y = p.Y; s.Z = p.Z; }
#pragma unroll 1
    for (int b = len - 2; b >= 0; --b)
      ge_dbl_add_public(s, __builtin_amdgcn_readfirstlane((int)naf[b]), c1, c3);
    const uint32_t in_sub = fe_iszero(s.X);
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

// raw 160-byte records: X|Y|Z as their resident 32-byte words, T (the
// resident fourth plane divided by 2d) as 32 bytes, then 32 zero bytes
__global__ void __launch_bounds__(MTPB) k_to_raw(uint32_t *out, PtView v, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  uint32_t w[8];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ld_words(w, v.w + (c * v.cap + i) * 8);
    st_words(out + i * 40 + c * 8, w);
  }
  fe t2d, t;
  load_fe(t2d, v.w + (3 * v.cap + i) * 8);
  fe_mul(t, t2d, FE_INV_D2);
  fe_store_words(w, t);
  st_words(out + i * 40 + 24, w);
#pragma unroll
  for (int q = 0; q < 8; ++q) w[q] = 0;
  st_words(out + i * 40 + 32, w);
}

__global__ void __launch_bounds__(MTPB) k_from_raw(PtView v, const uint32_t *in, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  uint32_t w[8];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ld_words(w, in + i * 40 + c * 8);
    st_words(v.w + (c * v.cap + i) * 8, w);
  }
  ld_words(w, in + i * 40 + 24);
  fe t, t2d;
  fe_from_words(t, w);
  fe_mul(t2d, t, FE_D2);
  store_fe(v.w + (3 * v.cap + i) * 8, t2d);
}

// ------------------------------------------------------- set membership
//
// Canonical encodings are unique per point, so "array1[i] is in array2" is
// membership of 32-byte encodings.  The second array's encodings go into an
// open-addressing table of mask + 1 uint32 slots (a power of two, at least
// twice their number) holding index + 1, 0 = empty, with linear probing.  An
// insert that meets an equal encoding stops, so duplicates share one slot and
// probe chains stay short whatever the input.  The table always keeps empty
// slots, so every probe sequence ends within mask + 1 steps.  The reference
// compares every pair (ftcrypto.cu:1046-1086); this is O(n1 + n2).

__device__ __forceinline__ void ld_enc(uint4 &a, uint4 &b, const uint4 *enc, size_t i) {
  a = enc[2 * i];
  b = enc[2 * i + 1];
}

__device__ __forceinline__ bool enc_eq(const uint4 &a, const uint4 &b, const uint4 &c, const uint4 &d) {
  return ((a.x ^ c.x) | (a.y ^ c.y) | (a.z ^ c.z) | (a.w ^ c.w) | (b.x ^ d.x) | (b.y ^ d.y) | (b.z ^ d.z) |
          (b.w ^ d.w)) == 0;
}

__device__ __forceinline__ size_t enc_hash(const uint4 &a, const uint4 &b) {
  const uint64_t w[4] = {(uint64_t)a.y << 32 | a.x, (uint64_t)a.w << 32 | a.z, (uint64_t)b.y << 32 | b.x,
                         (uint64_t)b.w << 32 | b.z};
  uint64_t h = 0x243F6A8885A308D3ull;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    h = (h ^ w[q]) * 0x9E3779B97F4A7C15ull;
    h ^= h >> 31;
  }
  return (size_t)h;
}

__global__ void __launch_bounds__(MTPB) k_fill_u32(uint32_t *p, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void __launch_bounds__(MTPB) k_set_insert(uint32_t *slots, size_t mask, const uint4 *enc, size_t n) {
  const size_t j = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (j >= n) return;
  uint4 a, b;
  ld_enc(a, b, enc, j);
  size_t h = enc_hash(a, b) & mask;
  for (size_t probe = 0; probe <= mask; ++probe) {
    const uint32_t prev = atomicCAS(&slots[h], 0u, (uint32_t)(j + 1));
    if (prev == 0) return;                       // inserted
    uint4 c, d;
    ld_enc(c, d, enc, prev - 1);
    if (enc_eq(a, b, c, d)) return;              // already present
    h = (h + 1) & mask;
  }
}

__global__ void __launch_bounds__(MTPB) k_set_lookup(int64_t *res, const uint32_t *slots, size_t mask,
                                                     const uint4 *set_enc, const uint4 *enc, size_t n) {
  const size_t i = blockIdx.x * (size_t)MTPB + threadIdx.x;
  if (i >= n) return;
  uint4 a, b;
  ld_enc(a, b, enc, i);
  size_t h = enc_hash(a, b) & mask;
  int64_t found = 0;
  for (size_t probe = 0; probe <= mask; ++probe) {
    const uint32_t cur = slots[h];
    if (cur == 0) break;
    uint4 c, d;
    ld_enc(c, d, set_enc, cur - 1);
    if (enc_eq(a, b, c, d)) {
      found = 1;
      break;
    }
    h = (h + 1) & mask;
  }
  res[i] = found;
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
hipError_t launch_validate(unsigned long long *bad, PtView v, size_t n, const int8_t *naf, int len, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_validate, dim3(mgrid(n)), dim3(MTPB), 0, s, bad, v, n, naf, len);
  MCHK();
}
size_t set_slots(size_t n_set) {
  size_t t = 64;
  while (t < 2 * n_set) t <<= 1;
  return t;
}
hipError_t launch_set_contains(int64_t *res, const uint32_t *enc, size_t n, const uint32_t *set_enc, size_t n_set,
                               uint32_t *slots, hipStream_t s) {
  if (!n) return hipSuccess;
  const size_t t = set_slots(n_set);
  if (n_set >= 0xffffffffu) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill_u32, dim3(mgrid(t)), dim3(MTPB), 0, s, slots, t);
  if (n_set)
    hipLaunchKernelGGL(k_set_insert, dim3(mgrid(n_set)), dim3(MTPB), 0, s, slots, t - 1,
                       reinterpret_cast<const uint4 *>(set_enc), n_set);
  hipLaunchKernelGGL(k_set_lookup, dim3(mgrid(n)), dim3(MTPB), 0, s, res, slots, t - 1,
                     reinterpret_cast<const uint4 *>(set_enc), reinterpret_cast<const uint4 *>(enc), n);
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
