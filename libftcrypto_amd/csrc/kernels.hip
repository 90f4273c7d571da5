// Elementwise, scalar-multiplication, encoding and reduction kernels for
// gfx950, plus their host-side launchers (namespace ft, kernels.h).
//
// One lane owns one element (point or ciphertext) for the whole kernel; the
// hot loops are pure VALU integer work in registers.  Secret scalar digits
// never drive a branch or an address: they are consumed from register shift
// streams in window order, and every table entry (registers or LDS) is read
// and merged with v_cndmask.
#include "kernels.h"
#include "elem.h"
#include "quad.h"

namespace ft {

static inline unsigned grid_for(size_t n, int tpb = TPB) { return (unsigned)((n + tpb - 1) / tpb); }

// ------------------------------------------------------------ elementwise

__global__ void __launch_bounds__(TPB) k_set_identity(PtView v, size_t lo, size_t hi) {
  const size_t i = lo + blockIdx.x * (size_t)TPB + threadIdx.x;
  if (i < hi) store_identity(v, i);
}

__global__ void __launch_bounds__(TPB) k_broadcast(PtView d, size_t n, PtView s, size_t idx) {
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (i < n) copy_point(d, i, s, idx);
}

__global__ void __launch_bounds__(TPB) k_copy(PtView d, PtView s, size_t n) {
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (i < n) copy_point(d, i, s, i);
}

// res[i] = res[i] +/- b[i]  (ftcrypto.cu:1089-1110, 1220-1241)
template <bool SUB>
__global__ void __launch_bounds__(TPB) k_add(PtView res, PtView b, size_t n) {
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (i >= n) return;
  add_one<SUB>(res, i, b, i);
}

__global__ void __launch_bounds__(TPB) k_neg(PtView v, size_t n) {
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (i >= n) return;
  ge_p3 p;
  load_p3(p, v, i);
  ge_p3_neg(p, p);
  store_p3(v, i, p);
}

// ------------------------------------------------------ fixed-base comb

// out[i] = k_i * B (the reference routes this through its variable-base
// routine on the GPU, ftcrypto.cu:895-913, and keeps a radix-16 comb for the
// CPU, sodium_ge.cu:743-792; the point is the same).  B's radix-32 comb has
// 51 rows of 16 entries, (j+1) 32^r B (89 KiB in global memory, read by
// scalar loads: no LDS): one mixed add per row and no doubling, 50 adds per
// 256-bit scalar where a radix-16 comb takes 64 adds and four doublings.
// The sum is the reference's multiple: k is rewritten to the multiple its
// radix-16 digit rule gives (scalar_effective; the same as k below 2^255),
// reduced mod l and negated mod l when negative, which leaves k B unchanged
// (l B = 0).  SRC: 0 = 32-byte scalars, 1 = uint64 scalars over 13 rows
// (ftcrypto.cu:937-963 packing).
template <int SRC>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(4, 4))) k_fixed_base32(PtView out, const void *sc, size_t n,
                                                    const uint32_t *__restrict__ tab) {
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (i >= n) return;
  constexpr int ND = SRC == 0 ? COMB32_ROWS : 13;
  uint32_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (SRC == 0) {
    uint32_t k[8];
    load_scalar(k, static_cast<const uint32_t *>(sc), i);
    const uint32_t neg = scalar_effective(m, k);
    scalar_mod_l(m);
    scalar_cneg_l(m, neg);
  } else {
    const uint64_t x = static_cast<const uint64_t *>(sc)[i];
    m[0] = (uint32_t)x;
    m[1] = (uint32_t)(x >> 32);
  }
  uint32_t e[(ND + 3) / 4];
  recode32<ND>(e, m);
  ge_p3 h;
  comb32_run<(ND + 3) / 4>(h, tab, e, ND);
  store_p3(out, i, h);
}

// a[i] = k_i * a[i] (and b[i] = k_i * b[i] when two): homomorphic
// multiplication of a ciphertext by a 32-bit constant.  Radix 16 keeps its
// 8-entry table in registers at one wave per SIMD; radix 8's 4-entry table
// leaves room for two waves per SIMD (a few spilled words), which measured
// 23 % faster than radix 16 at 2^22 ciphertexts.
// lanes [0, n) multiply a, lanes [n, 2n) multiply b (when two)
__global__ void __launch_bounds__(TPB) k_mul_u32_r16(PtView a, PtView b, int two, const uint32_t *k,
                                                     size_t n) {
  const size_t t = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (t >= (two ? 2 * n : n)) return;
  const size_t i = t < n ? t : t - n;
  mul_u32_point<4>(t < n ? a : b, i, k[i]);
}

__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_mul_u32_r8(PtView a, PtView b, int two, const uint32_t *k, size_t n) {
  const size_t t = blockIdx.x * (size_t)TPB + threadIdx.x;
  if (t >= (two ? 2 * n : n)) return;
  const size_t i = t < n ? t : t - n;
  mul_u32_point<3>(t < n ? a : b, i, k[i]);
}

// ------------------------------------------------------ comb table build

// table[r][j] = (j+1) * 256^r * P in affine (y+x, y-x, 2dxy) form
// (sodium_ge.cu:480 layout).  One block of 256 lanes, one entry per lane;
// P is public, so the generic constant-time routine is simply reused.
__global__ void __launch_bounds__(TPB) k_build_comb(uint32_t *table, PtView src, size_t idx,
                                                    int use_base) {
  const int t = threadIdx.x, r = t >> 3, j = t & 7;
  uint32_t k[8] = {0, 0, 0, 0, 0, 0, 0, 0}, e[16];
  const int bit = 8 * r;                       // (j+1) << (8 r)
  k[bit >> 5] = (uint32_t)(j + 1) << (bit & 31);
  recode16(e, k);
  ge_p3 p, h;
  if (use_base) {
    // base point B = (x, 4/5) from its encoding 58 66 .. 66 (ftcrypto.cu:120-125)
    uint32_t w[8];
    w[0] = 0x66666658u;
#pragma unroll
    for (int q = 1; q < 8; ++q) w[q] = 0x66666666u;
    ge_decode(p, w);
  } else {
    load_p3(p, src, idx);
  }
  varbase_windows(h, p, e, 0);
  ge_precomp en;
  ge_p3_to_precomp(en, h);
  comb_entry_store(table + (r * 8 + j) * COMB_ENTRY_WORDS, en);
}

// table[r][j] = (j+1) * 32^r * P in the same form, r < 51, j < 16, for P =
// B or src[idx]: one entry per lane
__global__ void __launch_bounds__(TPB) k_build_comb32(uint32_t *table, PtView src, size_t idx, int use_base) {
  const int t = blockIdx.x * TPB + threadIdx.x;
  if (t >= COMB32_ROWS * 16) return;
  const int r = t >> 4, j = t & 15;
  uint32_t k[8] = {0, 0, 0, 0, 0, 0, 0, 0}, e[16];
  const int bit = 5 * r, w = bit >> 5, sh = bit & 31;  // (j+1) << (5 r) < 2^255
  k[w] = (uint32_t)(j + 1) << sh;
  if (sh > 27) k[w + 1] = (uint32_t)(j + 1) >> (32 - sh);
  recode16(e, k);
  ge_p3 p, h;
  if (use_base) {
    uint32_t bw[8];
    bw[0] = 0x66666658u;
#pragma unroll
    for (int q = 1; q < 8; ++q) bw[q] = 0x66666666u;
    ge_decode(p, bw);
  } else {
    load_p3(p, src, idx);
  }
  varbase_windows(h, p, e, 0);
  ge_precomp en;
  ge_p3_to_precomp(en, h);
  comb_entry_store(table + r * COMB32_ROW_WORDS + j * COMB_ENTRY_WORDS, en);
}

// ------------------------------------------------------ ElGamal encryption

// A launch gated by a device word (the message-bound check of the bounded
// encryptions): when it is nonzero the whole launch writes nothing.  The word
// is the same for every lane, so whole blocks return together, before any
// barrier.
__device__ __forceinline__ bool gated(const unsigned long long *gate) { return gate && *gate != 0ull; }

// Message bound of an encryption (a public bit count 0 < bits < 64): fail
// becomes 1 + the largest index whose message has a bit at or above `bits`
// (fail must hold 0 before; a wave-wide max, then one atomic per wave).
__global__ void __launch_bounds__(TPB) k_msg_bound(const uint64_t *m, size_t n, int bits, unsigned long long *fail) {
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  long long f = -1;
  if (i < n && (m[i] >> bits) != 0ull) f = (long long)i;
  for (int off = 32; off > 0; off >>= 1) {
    const long long o = __shfl_xor(f, off);
    f = o > f ? o : f;
  }
  if ((threadIdx.x & 63) == 0 && f >= 0) atomicMax(fail, (unsigned long long)f + 1ull);
}

// c1[i] = r_i B, c2[i] = m_i B + r_i H (or, with ACC, c1 += r_i B, c2 += r_i H:
// rerandomisation).  The comb runs for both tables at once so the four
// middle doublings of r H are shared with m B; the message uses ND digits
// (ND = 0 for rerandomisation), a public bound on its bit length.
template <bool ACC, int BS>
__device__ __forceinline__ void encrypt_block(PtView c1, PtView c2, const uint32_t *btab, const uint32_t *htab,
                                              const uint64_t *m, int nd, const uint32_t *r, size_t n,
                                              const unsigned long long *gate) {
  if (gated(gate)) return;
  __shared__ __align__(16) uint32_t tb[COMB_TABLE_WORDS];
  __shared__ __align__(16) uint32_t th[COMB_TABLE_WORDS];
  lds_load_table(tb, btab);
  lds_load_table(th, htab);
  __syncthreads();
  const size_t i = blockIdx.x * (size_t)BS + threadIdx.x;
  if (i >= n) return;
  encrypt_one<ACC>(c1, c2, i, tb, th, ACC ? 0ull : m[i], nd, r + i * 8);
}

template <bool ACC>
__global__ void __launch_bounds__(TPB) k_encrypt(PtView c1, PtView c2, const uint32_t *btab,
                                                 const uint32_t *htab, const uint64_t *m, int nd,
                                                 const uint32_t *r, size_t n, const unsigned long long *gate) {
  encrypt_block<ACC, TPB>(c1, c2, btab, htab, m, nd, r, n, gate);
}

// Two waves per 64 ciphertexts over radix-32 tables (encrypt_half32): wave
// pair k holds ciphertexts [64 k, 64 k + 64), its even wave C1, its odd wave
// C2, so that each wave reads one table, at wave-uniform addresses, through
// scalar loads.  No LDS: BLK-lane blocks at four waves per SIMD.
template <bool ACC, int BLK>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_encrypt_w32(PtView c1, PtView c2, const uint32_t *__restrict__ btab32, const uint32_t *__restrict__ htab32,
              const uint64_t *m, int nd32, const uint32_t *r, size_t n, const unsigned long long *gate) {
  static_assert(BLK % 128 == 0, "whole wave pairs per block");
  if (gated(gate)) return;
  const size_t wv = (size_t)blockIdx.x * (BLK / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t half = (uint32_t)(wv & 1);
  const size_t i = ((wv >> 1) << 6) + (threadIdx.x & 63);
  if (i >= n) return;
  encrypt_half32<ACC>(c1, c2, i, half, btab32, htab32, ACC ? 0ull : m[i], nd32, r + i * 8);
}

// Four waves per 64 ciphertexts over radix-32 tables, for arrays that two
// waves per 64 leave at two waves per SIMD or fewer: wave quad k holds
// ciphertexts [64 k, 64 k + 64), its waves 0 and 1 C1 and waves 2 and 3 C2,
// the even wave of each pair the comb's first rows and the odd wave the rest
// (encrypt_part32).  The odd wave passes its sum to the even one through LDS
// (a limb plane per word, lane-contiguous), which adds the two and stores.
// Every wave reaches the barrier: lanes past n compute element n - 1 again.
// A quad wholly past n stores nothing (a wave-uniform branch); in the quad
// holding element n - 1, the lanes past it store that element's value, the
// same bits to the same place as its own lane, in the same store instruction
// of the same wave (with ACC all of them load it before any of them stores),
// so that no branch depends on the lane.
template <bool ACC>
__global__ void __launch_bounds__(ETPB) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_encrypt_w32q(PtView c1, PtView c2, const uint32_t *__restrict__ btab32, const uint32_t *__restrict__ htab32,
               const uint64_t *m, int nd32, const uint32_t *r, size_t n, const unsigned long long *gate) {
  static_assert(ETPB % 256 == 0, "whole wave quads per block");
  __shared__ uint32_t xs[ETPB / 128][36][64];
  if (gated(gate)) return;                        // the same for every wave of the grid
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  const size_t wv = (size_t)blockIdx.x * (ETPB / 64) + w;
  const uint32_t half = (uint32_t)((wv >> 1) & 1), part = (uint32_t)(wv & 1);
  const size_t i0 = ((wv >> 2) << 6) + lane;
  const size_t i = i0 < n ? i0 : n - 1;
  // C2's split point takes half of the message rows to the even wave's side
  const int split = half ? (COMB32_ROWS + nd32 + 1) / 2 : (COMB32_ROWS + 1) / 2;
  ge_p3 a;
  encrypt_part32<ACC>(a, half, part, split, btab32, htab32, ACC ? 0ull : m[i], nd32, r + i * 8);
  uint32_t (*x)[64] = xs[w >> 1];
  if (part) {
#pragma unroll
    for (int l = 0; l < 9; ++l) {
      x[l][lane] = a.X.v[l];
      x[9 + l][lane] = a.Y.v[l];
      x[18 + l][lane] = a.Z.v[l];
      x[27 + l][lane] = a.T.v[l];
    }
  }
  __syncthreads();
  if (part || ((wv >> 2) << 6) >= n) return;
  ge_p3 b;
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    b.X.v[l] = x[l][lane];
    b.Y.v[l] = x[9 + l][lane];
    b.Z.v[l] = x[18 + l][lane];
    b.T.v[l] = x[27 + l][lane];
  }
  ge_cached bc;
  ge_p3_to_cached(bc, b);
  ge_p1p1 t;
  ge_add(t, a, bc);
  ge_p1p1_to_p3(a, t);
  const PtView dst = half ? c2 : c1;
  if (ACC) add_stored(a, dst, i);
  store_p3(dst, i, a);
}

// Several ciphertext arrays encrypted in one launch over radix-32 tables,
// wave pairs as in k_encrypt_w32: the jobs' ciphertexts are laid end to end
// from multiples of 64, so that a wave pair's job is one, which each wave
// finds in the job table (a kernel argument) as a scalar.  The tables'
// kernel arguments are __restrict__: without it the compiler cannot show
// that no store of the kernel writes them and reads the rows through vector
// loads instead of scalar ones.
template <int BS>
__device__ __forceinline__ void encrypt_batch_body32(const EncBatch &B, const uint32_t *__restrict__ btab32,
                                                     const uint32_t *__restrict__ htab32, int nd32, uint64_t total,
                                                     const unsigned long long *gate) {
  static_assert(BS % 128 == 0, "whole wave pairs per block");
  if (gated(gate)) return;
  const uint64_t wv = (uint64_t)blockIdx.x * (BS / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t half = (uint32_t)(wv & 1);
  const uint64_t t0 = (wv >> 1) << 6;             // the wave pair's first ciphertext: wave-uniform,
  if (t0 >= total) return;                        // so the job and its fields are scalars
  int j = 0;
  for (int q = 1; q < B.J; ++q)
    if (B.job[q].base <= t0) j = q;
  const EncJob &J = B.job[j];
  const uint64_t i = t0 - J.base + (threadIdx.x & 63);
  if (i >= J.n) return;                           // padding up to the next job's start
  encrypt_half32<false>(J.c1, J.c2, i, half, btab32, htab32, J.m[i], nd32, J.r + i * 8);
}

template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_encrypt_batch32(const EncBatch B, const uint32_t *__restrict__ btab32, const uint32_t *__restrict__ htab32,
                  int nd32, uint64_t total, const unsigned long long *gate) {
  encrypt_batch_body32<BS>(B, btab32, htab32, nd32, total, gate);
}

// message bound of a batch: bit j of fail is set when a message of job j has
// a bit at or above `bits`
__global__ void __launch_bounds__(TPB) k_msg_bound_batch(const EncBatch B, int bits, uint64_t total,
                                                         unsigned long long *fail) {
  const uint64_t t = blockIdx.x * (uint64_t)TPB + threadIdx.x;
  if (t >= total) return;
  int j = 0;
  for (int q = 1; q < B.J; ++q)
    if (B.job[q].base <= t) j = q;
  const uint64_t i = t - B.job[j].base;
  if (i < B.job[j].n && (B.job[j].m[i] >> bits) != 0ull) atomicOr(fail, 1ull << j);
}

template <bool ACC>
__global__ void __launch_bounds__(ETPB) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_encrypt_wide(PtView c1, PtView c2, const uint32_t *btab, const uint32_t *htab, const uint64_t *m, int nd,
               const uint32_t *r, size_t n, const unsigned long long *gate) {
  encrypt_block<ACC, ETPB>(c1, c2, btab, htab, m, nd, r, n, gate);
}

// ------------------------------------------------------ encode / decode

// Compressed encodings (sodium_ge.cu:382-394) of n points; every lane batch-
// inverts BATCH Z coordinates with one field inversion (Montgomery's trick).
// A zero Z (never produced by the group law) encodes as 32 zero bytes, as the
// reference's invert(0) = 0 does, without disturbing the other lanes' batch.
constexpr int ENC_BATCH = 8;

__global__ void __launch_bounds__(TPB) k_encode(uint32_t *out, PtView v, size_t n) {
  const size_t lanes = (size_t)gridDim.x * TPB;
  const size_t i0 = blockIdx.x * (size_t)TPB + threadIdx.x;
  fe z[ENC_BATCH], acc[ENC_BATCH];
  uint32_t zero[ENC_BATCH];
  fe one;
  fe_1(one);
  fe run = one;
#pragma unroll
  for (int b = 0; b < ENC_BATCH; ++b) {
    const size_t i = i0 + b * lanes;
    if (i < n) load_fe(z[b], v.w + (2 * v.cap + i) * 8);
    else z[b] = one;
    zero[b] = fe_iszero(z[b]);
    fe_cmov(z[b], one, zero[b]);
    acc[b] = run;                  // product of z[0..b-1]
    fe_mul(run, run, z[b]);
  }
  fe inv;
  fe_invert(inv, run);
#pragma unroll
  for (int b = ENC_BATCH - 1; b >= 0; --b) {
    fe zi;
    fe_mul(zi, inv, acc[b]);       // 1 / z[b]
    fe_mul(inv, inv, z[b]);
    const size_t i = i0 + b * lanes;
    if (i < n) {
      fe X, Y, x, y;
      load_fe(X, v.w + (0 * v.cap + i) * 8);
      load_fe(Y, v.w + (1 * v.cap + i) * 8);
      fe_mul(x, X, zi);
      fe_mul(y, Y, zi);
      uint32_t w[8];
      ge_encode_xy(w, x, y);
#pragma unroll
      for (int q = 0; q < 8; ++q) w[q] = zero[b] ? 0u : w[q];
      st_words(out + i * 8, w);
    }
  }
}

// Decode n encodings; bad[0] receives 1 + the largest failing index (0: none).
// Two waves per SIMD (a few spilled words outside the square-root chain)
// hide the latency of that chain better than one wave in 334 registers.
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(2, 2))) k_decode(PtView v, const uint32_t *in, size_t n,
                                                unsigned long long *bad) {
  const size_t i = blockIdx.x * (size_t)TPB + threadIdx.x;
  long long fail = -1;
  if (i < n) {
    uint32_t w[8];
    ld_words(w, in + i * 8);
    ge_p3 p;
    const uint32_t ok = ge_decode(p, w);
    store_p3(v, i, p);
    fail = ok ? -1 : (long long)i;
  }
  // wave-wide max, then one atomic per wave; bad holds (largest index + 1)
  for (int off = 32; off > 0; off >>= 1) {
    const long long o = __shfl_xor(fail, off);
    fail = o > fail ? o : fail;
  }
  if ((threadIdx.x & 63) == 0 && fail >= 0) atomicMax(bad, (unsigned long long)fail + 1ull);
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v) { *p = v; }

// ------------------------------------------------------ reduction

// q = the point p of lane (lane + d) of the wave; lanes past the wave's end
// get their own p back (their sums are never used)
__device__ __forceinline__ void shfl_down_p3(ge_p3 &q, const ge_p3 &p, int d) {
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    q.X.v[l] = __shfl_down(p.X.v[l], d);
    q.Y.v[l] = __shfl_down(p.Y.v[l], d);
    q.Z.v[l] = __shfl_down(p.Z.v[l], d);
    q.T.v[l] = __shfl_down(p.T.v[l], d);
  }
}

// lane 0 of the wave receives the sum of the points of lanes 0 .. 2 top - 1:
// lane i adds lane i + d for d = top .. 1 (the unified addition is exact for
// every pair, so the lanes whose partner is past the end compute harmless
// extra sums)
__device__ __forceinline__ void wave_sum_p3(ge_p3 &acc, int top) {
#pragma unroll 1
  for (int d = top; d > 0; d >>= 1) {
    ge_p3 q;
    shfl_down_p3(q, acc, d);
    ge_p3_add(acc, acc, q);
  }
}

__device__ __forceinline__ void p3_to_lds(uint32_t *d, const ge_p3 &p) {
#pragma unroll
  for (int l = 0; l < 9; ++l) { d[l] = p.X.v[l]; d[9 + l] = p.Y.v[l]; d[18 + l] = p.Z.v[l]; d[27 + l] = p.T.v[l]; }
}

__device__ __forceinline__ void p3_from_lds(ge_p3 &p, const uint32_t *s) {
#pragma unroll
  for (int l = 0; l < 9; ++l) { p.X.v[l] = s[l]; p.Y.v[l] = s[9 + l]; p.Z.v[l] = s[18 + l]; p.T.v[l] = s[27 + l]; }
}

// Each block sums a grid-strided share of its points; then waves 2-3 hand
// their lanes' sums to waves 0-1 and wave 1 to wave 0 through LDS (raw
// limbs of class N, no packing), and wave 0 sums its lanes by shuffles (six
// levels, no barrier): nine wave-wide additions per block, as few as a tree
// over 256 lanes needs.  When the grid has one point per lane or fewer, the
// levels that would only add identities are skipped (the decisions depend
// on n and the block only, so every thread of a block takes the same
// branch): a block of one point stores it, one of 64 runs six levels.
// part[blockIdx.x] receives the block sum.
__global__ void __launch_bounds__(TPB) k_reduce_blocks(PtView part, PtView v, size_t n) {
  static_assert(TPB == 256, "four waves per block");
  __shared__ uint32_t sh[128 * 36];
  const size_t lanes = (size_t)gridDim.x * TPB;
  const size_t b0 = blockIdx.x * (size_t)TPB;
  size_t i = b0 + threadIdx.x;
  ge_p3 acc, q;
  ge_cached qc;
  PW nw;
  ge_p3_0(acc);
  if (i < n) { load_p3(acc, v, i); i += lanes; }
  if (i < n) load_cached(qc, v, i);
#pragma unroll 1
  for (; i < n; i += lanes) {
    // the next point's words are loaded before the add and unpacked after
    // it: the barrier keeps the loads first, pinning qc keeps the add after
    // them, pinning acc keeps the wait for the words after the add
    ld_pw(nw, v, i + lanes < n ? i + lanes : i);
    asm volatile("" ::: "memory");
    pin_cached(qc);
    ge_p1p1 t;
    ge_add(t, acc, qc);
    ge_p1p1_to_p3(acc, t);
    pin_p3(acc);
    pin_pw(nw);
    un_pw_cached(qc, nw);
  }
  // lanes of this block holding a point (all when the grid strides)
  const size_t act = n >= lanes ? TPB : (n > b0 ? (n - b0 < TPB ? n - b0 : TPB) : 0);
  const int t = threadIdx.x;
  if (act > 128) {
    if (t >= 128) p3_to_lds(sh + (t - 128) * 36, acc);
    __syncthreads();
    if (t < 128) { p3_from_lds(q, sh + t * 36); ge_p3_add(acc, acc, q); }
    __syncthreads();
  }
  if (act > 64) {
    if (t >= 64 && t < 128) p3_to_lds(sh + (t - 64) * 36, acc);
    __syncthreads();
    if (t < 64) { p3_from_lds(q, sh + t * 36); ge_p3_add(acc, acc, q); }
  }
  if (t >= 64) return;
  int top = 32;
  while (top > 1 && (size_t)top >= act) top >>= 1;      // 2 top >= act points
  if (act > 1) wave_sum_p3(acc, top);
  if (t == 0) store_p3(part, blockIdx.x, acc);
}

// ------------------------------------------------------ quad-lane reduction

constexpr int QTPB = 1024;          // 256 quads per block
constexpr int QPB = QTPB / 4;

// part[blockIdx.x] = sum of the block's grid-strided share of v[0, n): each
// quad sums its points in sequence, then the block's 256 quads are summed
// by a tree, across waves through LDS down to 16 quads, then inside wave 0
// by shuffles; levels that would only add identities are skipped (block-
// uniform decisions, as in k_reduce_blocks)
__global__ void __launch_bounds__(QTPB) k_reduce_quads(PtView part, PtView v, size_t n) {
  __shared__ uint32_t sh[(QPB / 2) * 4 * 9];
  const int c = threadIdx.x & 3;
  const int gq = threadIdx.x >> 2;
  const size_t nq = (size_t)gridDim.x * QPB;
  const size_t q0 = blockIdx.x * (size_t)QPB;
  size_t i = q0 + gq;
  fe acc;
  quad_identity(acc, c);
  if (i < n) {
    load_fe(acc, v.w + (c * v.cap + i) * 8);
    quad_from_mem(acc, c);
    i += nq;
  }
#pragma unroll 1
  for (; i < n; i += nq) {
    fe o;
    load_fe(o, v.w + (c * v.cap + i) * 8);
    quad_add<true>(acc, o, c);
  }
  const size_t act = n >= nq ? QPB : (n > q0 ? (n - q0 < QPB ? n - q0 : QPB) : 0);
#pragma unroll 1
  for (int s = QPB / 2; s >= 16; s >>= 1) {
    if (act > (size_t)s) {
      if (gq >= s && gq < 2 * s) {
        uint32_t *d = sh + ((gq - s) * 4 + c) * 9;
#pragma unroll
        for (int l = 0; l < 9; ++l) d[l] = acc.v[l];
      }
      __syncthreads();
      if (gq < s) {
        fe o;
        const uint32_t *src = sh + (gq * 4 + c) * 9;
#pragma unroll
        for (int l = 0; l < 9; ++l) o.v[l] = src[l];
        quad_add(acc, o, c);
      }
      __syncthreads();
    }
  }
  if (threadIdx.x >= 64) return;
  int top = 8;
  while (top > 1 && (size_t)top >= act) top >>= 1;     // 2 top >= act quads
  if (act > 1) {
#pragma unroll 1
    for (int d = top; d > 0; d >>= 1) {
      fe o;
      fe_shfl_down(o, acc, 4 * d);
      quad_add(acc, o, c);
    }
  }
  if (gq == 0) {
    quad_to_mem(acc, c);
    store_fe(part.w + (c * part.cap + blockIdx.x) * 8, acc);
  }
}

// ------------------------------------------------------ host launchers

#define FT_LAUNCH_CHECK() return hipGetLastError()

hipError_t launch_set_identity(PtView v, size_t lo, size_t hi, hipStream_t s) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_set_identity, dim3(grid_for(hi - lo)), dim3(TPB), 0, s, v, lo, hi);
  FT_LAUNCH_CHECK();
}

hipError_t launch_broadcast(PtView d, size_t n, PtView src, size_t idx, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_broadcast, dim3(grid_for(n)), dim3(TPB), 0, s, d, n, src, idx);
  FT_LAUNCH_CHECK();
}

hipError_t launch_copy(PtView d, PtView src, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3(grid_for(n)), dim3(TPB), 0, s, d, src, n);
  FT_LAUNCH_CHECK();
}

hipError_t launch_add(PtView res, PtView b, size_t n, bool sub, hipStream_t s) {
  if (!n) return hipSuccess;
  if (sub) hipLaunchKernelGGL(k_add<true>, dim3(grid_for(n)), dim3(TPB), 0, s, res, b, n);
  else hipLaunchKernelGGL(k_add<false>, dim3(grid_for(n)), dim3(TPB), 0, s, res, b, n);
  FT_LAUNCH_CHECK();
}

hipError_t launch_neg(PtView v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_neg, dim3(grid_for(n)), dim3(TPB), 0, s, v, n);
  FT_LAUNCH_CHECK();
}

hipError_t launch_fixed_base(PtView out, const void *sc, bool u64, size_t n, const uint32_t *table32,
                             hipStream_t s) {
  if (!n) return hipSuccess;
  if (u64) hipLaunchKernelGGL(k_fixed_base32<1>, dim3(grid_for(n)), dim3(TPB), 0, s, out, sc, n, table32);
  else hipLaunchKernelGGL(k_fixed_base32<0>, dim3(grid_for(n)), dim3(TPB), 0, s, out, sc, n, table32);
  FT_LAUNCH_CHECK();
}

hipError_t launch_mul_u32_w(PtView a, PtView b, bool two, const uint32_t *k, size_t n, int w, hipStream_t s) {
  if (!n) return hipSuccess;
  const dim3 grid(grid_for(two ? 2 * n : n));
  if (w == 3)
    hipLaunchKernelGGL(k_mul_u32_r8, grid, dim3(TPB), 0, s, a, b, (int)two, k, n);
  else if (w == 4)
    hipLaunchKernelGGL(k_mul_u32_r16, grid, dim3(TPB), 0, s, a, b, (int)two, k, n);
  else
    return hipErrorInvalidValue;
  FT_LAUNCH_CHECK();
}

hipError_t launch_mul_u32(PtView a, PtView b, bool two, const uint32_t *k, size_t n, hipStream_t s) {
  return launch_mul_u32_w(a, b, two, k, n, tune_for(n).u32w, s);
}

hipError_t launch_build_comb(uint32_t *table, PtView src, size_t idx, bool use_base, hipStream_t s) {
  hipLaunchKernelGGL(k_build_comb, dim3(1), dim3(TPB), 0, s, table, src, idx, (int)use_base);
  FT_LAUNCH_CHECK();
}

hipError_t launch_build_comb32(uint32_t *table, PtView src, size_t idx, bool use_base, hipStream_t s) {
  hipLaunchKernelGGL(k_build_comb32, dim3((COMB32_ROWS * 16 + TPB - 1) / TPB), dim3(TPB), 0, s, table, src, idx,
                     (int)use_base);
  FT_LAUNCH_CHECK();
}

// lanes: 1 (one lane per ciphertext), 2 (C1 and C2 on separate waves) or
// 4 (each comb's rows split over two waves as well)
hipError_t launch_encrypt_l(PtView c1, PtView c2, const uint32_t *btab, const uint32_t *htab,
                            const uint64_t *m, int nd, const uint32_t *r, size_t n, bool accumulate, int lanes,
                            hipStream_t s, const unsigned long long *gate, const uint32_t *btab32,
                            const uint32_t *htab32) {
  if (!n) return hipSuccess;
  if (lanes == 2) {
    if (!btab32 || !htab32) return hipErrorInvalidValue;
    const size_t waves = 2 * ((n + 63) / 64);
    const int nd32 = accumulate ? 0 : msg_digits32(nd);
    if (2 * n >= ENC_SPLIT_WIDE_MIN) {
      const dim3 grid((unsigned)((waves * 64 + ETPB - 1) / ETPB));
      if (accumulate)
        hipLaunchKernelGGL((k_encrypt_w32<true, ETPB>), grid, dim3(ETPB), 0, s, c1, c2, btab32, htab32, m, 0, r, n,
                           gate);
      else
        hipLaunchKernelGGL((k_encrypt_w32<false, ETPB>), grid, dim3(ETPB), 0, s, c1, c2, btab32, htab32, m, nd32, r,
                           n, gate);
    } else {
      const dim3 grid((unsigned)((waves * 64 + TPB - 1) / TPB));
      if (accumulate)
        hipLaunchKernelGGL((k_encrypt_w32<true, TPB>), grid, dim3(TPB), 0, s, c1, c2, btab32, htab32, m, 0, r, n,
                           gate);
      else
        hipLaunchKernelGGL((k_encrypt_w32<false, TPB>), grid, dim3(TPB), 0, s, c1, c2, btab32, htab32, m, nd32, r, n,
                           gate);
    }
    FT_LAUNCH_CHECK();
  }
  if (lanes == 1 && n >= ENC_WIDE_MIN) {
    const dim3 grid(grid_for(n, ETPB));
    if (accumulate)
      hipLaunchKernelGGL(k_encrypt_wide<true>, grid, dim3(ETPB), 0, s, c1, c2, btab, htab, m, 0, r, n, gate);
    else
      hipLaunchKernelGGL(k_encrypt_wide<false>, grid, dim3(ETPB), 0, s, c1, c2, btab, htab, m, nd, r, n, gate);
  } else if (lanes == 1) {
    const dim3 grid(grid_for(n));
    if (accumulate)
      hipLaunchKernelGGL(k_encrypt<true>, grid, dim3(TPB), 0, s, c1, c2, btab, htab, m, 0, r, n, gate);
    else
      hipLaunchKernelGGL(k_encrypt<false>, grid, dim3(TPB), 0, s, c1, c2, btab, htab, m, nd, r, n, gate);
  } else if (lanes == 4) {
    if (!btab32 || !htab32) return hipErrorInvalidValue;
    const size_t waves = 4 * ((n + 63) / 64);    // a wave quad per 64 ciphertexts
    const dim3 grid((unsigned)((waves * 64 + ETPB - 1) / ETPB));
    if (accumulate)
      hipLaunchKernelGGL(k_encrypt_w32q<true>, grid, dim3(ETPB), 0, s, c1, c2, btab32, htab32, m, 0, r, n, gate);
    else
      hipLaunchKernelGGL(k_encrypt_w32q<false>, grid, dim3(ETPB), 0, s, c1, c2, btab32, htab32, m,
                         msg_digits32(nd), r, n, gate);
  } else {
    return hipErrorInvalidValue;
  }
  FT_LAUNCH_CHECK();
}

// bases of the batch's jobs (multiples of 64 ciphertexts: a wave pair's) and
// their total
static uint64_t batch_bases(EncBatch &B) {
  uint64_t total = 0;
  for (int q = 0; q < B.J; ++q) {
    total = (total + 63) & ~(uint64_t)63;
    B.job[q].base = total;
    total += B.job[q].n;
  }
  return total;
}

hipError_t launch_encrypt_batch(EncBatch &B, const uint32_t *btab32, const uint32_t *htab32, int nd, hipStream_t s,
                                const unsigned long long *gate) {
  if (B.J < 0 || B.J > ENC_BATCH_MAX || !btab32 || !htab32) return hipErrorInvalidValue;
  const uint64_t total = batch_bases(B);            // a wave pair holds 64 ciphertexts
  if (!total) return hipSuccess;
  const uint64_t lanes = (total + 63) / 64 * 128;   // a wave pair per 64 ciphertexts
  const int nd32 = msg_digits32(nd);
  if (lanes >= ENC_SPLIT_WIDE_MIN)
    hipLaunchKernelGGL(k_encrypt_batch32<ETPB>, dim3((unsigned)((lanes + ETPB - 1) / ETPB)), dim3(ETPB), 0, s, B,
                       btab32, htab32, nd32, total, gate);
  else
    hipLaunchKernelGGL(k_encrypt_batch32<TPB>, dim3((unsigned)((lanes + TPB - 1) / TPB)), dim3(TPB), 0, s, B,
                       btab32, htab32, nd32, total, gate);
  FT_LAUNCH_CHECK();
}

hipError_t launch_msg_bound(const uint64_t *m, size_t n, int bits, unsigned long long *fail, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, s, fail, 0ull);
  if (n) hipLaunchKernelGGL(k_msg_bound, dim3(grid_for(n)), dim3(TPB), 0, s, m, n, bits, fail);
  FT_LAUNCH_CHECK();
}

hipError_t launch_msg_bound_batch(EncBatch &B, int bits, unsigned long long *fail, hipStream_t s) {
  if (B.J < 0 || B.J > ENC_BATCH_MAX) return hipErrorInvalidValue;
  const uint64_t total = batch_bases(B);
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, s, fail, 0ull);
  if (total) hipLaunchKernelGGL(k_msg_bound_batch, dim3(grid_for(total)), dim3(TPB), 0, s, B, bits, total, fail);
  FT_LAUNCH_CHECK();
}

hipError_t launch_encrypt(PtView c1, PtView c2, const uint32_t *btab, const uint32_t *htab,
                          const uint64_t *m, int nd, const uint32_t *r, size_t n, bool accumulate,
                          hipStream_t s, const unsigned long long *gate, const uint32_t *btab32,
                          const uint32_t *htab32) {
  return launch_encrypt_l(c1, c2, btab, htab, m, nd, r, n, accumulate, tune_for(n).encl, s, gate, btab32, htab32);
}

hipError_t launch_encode(uint32_t *out, PtView v, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  const size_t per = (size_t)TPB * ENC_BATCH;
  const unsigned grid = (unsigned)((n + per - 1) / per);
  hipLaunchKernelGGL(k_encode, dim3(grid), dim3(TPB), 0, s, out, v, n);
  FT_LAUNCH_CHECK();
}

hipError_t launch_decode(PtView v, const uint32_t *in, size_t n, unsigned long long *bad, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, s, bad, 0ull);
  if (!n) return hipGetLastError();
  hipLaunchKernelGGL(k_decode, dim3(grid_for(n)), dim3(TPB), 0, s, v, in, n, bad);
  FT_LAUNCH_CHECK();
}

unsigned reduce_grid(size_t n, unsigned max_blocks) {
  const size_t g = (n + TPB - 1) / TPB;
  return (unsigned)(g < max_blocks ? (g ? g : 1) : max_blocks);
}

// Latency-bound passes (one block over the partials, or at most
// REDUCE_QUAD_MAX points) take the quad-lane kernel: 256 quads per block,
// the same number of partials as k_reduce_blocks's 256 lanes per block.
// Longer first passes, where each lane sums many points and throughput
// counts, keep the one-lane-per-point kernel (fewer instructions per add).
hipError_t launch_reduce_blocks(PtView part, PtView v, size_t n, unsigned grid, hipStream_t s) {
  static_assert(QPB == TPB, "one partial per block in both kernels");
  if (grid == 1 || n <= REDUCE_QUAD_MAX)
    hipLaunchKernelGGL(k_reduce_quads, dim3(grid), dim3(QTPB), 0, s, part, v, n);
  else
    hipLaunchKernelGGL(k_reduce_blocks, dim3(grid), dim3(TPB), 0, s, part, v, n);
  FT_LAUNCH_CHECK();
}

}  // namespace ft
