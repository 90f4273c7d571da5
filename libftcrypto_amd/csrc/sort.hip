// Device-side key sort for ft_reduce_isum (res[idx[i]] += summand[i]).
//
// The reference sorts (key, index) pairs on the host with std::sort and runs
// one thread per run of equal keys (ftcrypto.cu:1150-1217).  Here the element
// indices are ordered by key on the device with a stable least-significant-
// digit radix sort (8-bit digits, as many passes as the key range needs) and
// each bucket's CSR offset is a binary search in the sorted keys, so the CSR
// order -- and with it every sum's exact representation -- is the same on
// every run: indices of equal keys stay in increasing order.  A key >= m
// (no such element of res) goes to an extra bucket m after every valid key,
// whose sum the caller discards.
//
// Radix pass over blocks of RS_ITEMS = 4096 elements (256 lanes x 16 rounds,
// element r * 256 + lane of the block in round r):
//   1. k_rs_hist: per-block digit histogram into hist[digit * nblocks + b];
//   2. exclusive scan of hist (digit-major), so a digit's elements of block b
//      start after every smaller digit and after the same digit's elements
//      of earlier blocks;
//   3. k_rs_scatter: each element's rank among equal digits earlier in its
//      block (wave ballots give the rank inside a wave, per-wave counts in
//      LDS the rank across waves, a running count the rank across rounds)
//      places it in LDS in digit order; the block then writes its elements
//      out in that order, so elements of one digit leave as one contiguous
//      run (about 16 per digit and block) instead of 4-byte scattered writes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "kernels.h"

namespace ft {

namespace {

constexpr int RS_TPB = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_ITEMS = RS_TPB * RS_ROUNDS;
constexpr int SCAN_TPB = 1024;

// ---------------------------------------------------------------- keys

// key32[i] = min(keys[i], m), idx[i] = i
__global__ void __launch_bounds__(256) k_keys_init(uint32_t *key32, uint32_t *idx, const uint64_t *keys, size_t n,
                                                   uint64_t m) {
  const size_t i = blockIdx.x * (size_t)256 + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = keys[i];
  key32[i] = (uint32_t)(k < m ? k : m);
  idx[i] = (uint32_t)i;
}

// *bad = 1 when some key is >= m (one atomic per wave that sees one)
__global__ void __launch_bounds__(256) k_keys_bound(unsigned long long *bad, const uint64_t *keys, size_t n,
                                                    uint64_t m) {
  const size_t i = blockIdx.x * (size_t)256 + threadIdx.x;
  const bool out = i < n && keys[i] >= m;
  if (__ballot(out) && (threadIdx.x & 63) == 0) atomicOr(bad, 1ull);
}

__global__ void k_bound_clear(unsigned long long *bad) { *bad = 0; }

// off[b] = first position of a key >= b in the sorted keys, b in [0, nb]
__global__ void __launch_bounds__(256) k_offsets(int64_t *off, const uint32_t *sorted, size_t n, size_t nb) {
  const size_t b = blockIdx.x * (size_t)256 + threadIdx.x;
  if (b > nb) return;
  size_t lo = 0, hi = n;
  while (lo < hi) {
    const size_t mid = lo + (hi - lo) / 2;
    if (sorted[mid] < b) lo = mid + 1;
    else hi = mid;
  }
  off[b] = (int64_t)lo;
}

// ------------------------------------------------- exclusive scan (i32 -> i64)

// block b: its SCAN_TPB values' sum into bsum[b]
__global__ void __launch_bounds__(SCAN_TPB) k_scan_bsum(int64_t *bsum, const int32_t *v, size_t n) {
  __shared__ int64_t sh[SCAN_TPB];
  const size_t i = blockIdx.x * (size_t)SCAN_TPB + threadIdx.x;
  sh[threadIdx.x] = i < n ? v[i] : 0;
  __syncthreads();
  for (int s = SCAN_TPB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = sh[0];
}

// one block: exclusive scan of the nb block sums in place (sequential chunks)
__global__ void __launch_bounds__(SCAN_TPB) k_scan_bsum_scan(int64_t *bsum, size_t nb) {
  __shared__ int64_t sh[SCAN_TPB];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (size_t base = 0; base < nb; base += SCAN_TPB) {
    const size_t i = base + threadIdx.x;
    const int64_t x = i < nb ? bsum[i] : 0;
    sh[threadIdx.x] = x;
    __syncthreads();
    for (int s = 1; s < SCAN_TPB; s <<= 1) {                  // inclusive Hillis-Steele
      const int64_t y = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
      __syncthreads();
      sh[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < nb) bsum[i] = carry + sh[threadIdx.x] - x;
    __syncthreads();
    if (threadIdx.x == SCAN_TPB - 1) carry += sh[SCAN_TPB - 1];
    __syncthreads();
  }
}

// out[i] = bsum[block] + exclusive prefix of v inside the block; out[n] = total
__global__ void __launch_bounds__(SCAN_TPB) k_scan_final(int64_t *out, const int32_t *v, size_t n,
                                                        const int64_t *bsum, size_t nb) {
  __shared__ int64_t sh[SCAN_TPB];
  const size_t i = blockIdx.x * (size_t)SCAN_TPB + threadIdx.x;
  const int64_t x = i < n ? v[i] : 0;
  sh[threadIdx.x] = x;
  __syncthreads();
  for (int s = 1; s < SCAN_TPB; s <<= 1) {
    const int64_t y = (int)threadIdx.x >= s ? sh[threadIdx.x - s] : 0;
    __syncthreads();
    sh[threadIdx.x] += y;
    __syncthreads();
  }
  const int64_t pre = bsum[blockIdx.x] + sh[threadIdx.x] - x;
  if (i < n) out[i] = pre;
  if (i + 1 == n) out[n] = pre + x;
  (void)nb;
}

// --------------------------------------------------------------- radix pass

__device__ __forceinline__ uint32_t digit_of(uint32_t k, int shift) { return (k >> shift) & 255u; }

__global__ void __launch_bounds__(RS_TPB) k_rs_hist(int32_t *hist, const uint32_t *keys, size_t n, int shift,
                                                    size_t nblocks) {
  __shared__ int32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = blockIdx.x * (size_t)RS_ITEMS;
#pragma unroll
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const size_t i = base + (size_t)r * RS_TPB + threadIdx.x;
    if (i < n) atomicAdd(&h[digit_of(keys[i], shift)], 1);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_TPB) k_rs_scatter(uint32_t *okeys, uint32_t *ovals, const uint32_t *keys,
                                                       const uint32_t *vals, size_t n, int shift,
                                                       const int32_t *hist, const int64_t *start, size_t nblocks) {
  __shared__ uint32_t sk[RS_ITEMS], sv[RS_ITEMS];
  __shared__ int32_t dstart[256];           // first LDS slot of each digit in this block
  __shared__ int32_t run[256];              // elements of each digit placed in earlier rounds
  __shared__ int32_t wc[RS_TPB / 64][256];  // this round: elements of each digit per wave
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  // exclusive scan of this block's digit counts (Hillis-Steele over 256)
  const int32_t cnt = hist[(size_t)t * nblocks + blockIdx.x];
  dstart[t] = cnt;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    const int32_t y = t >= d ? dstart[t - d] : 0;
    __syncthreads();
    dstart[t] += y;
    __syncthreads();
  }
  const int32_t mystart = dstart[t] - cnt;
  __syncthreads();
  dstart[t] = mystart;
  run[t] = 0;
#pragma unroll
  for (int q = 0; q < RS_TPB / 64; ++q) wc[q][t] = 0;
  __syncthreads();
  const size_t base = blockIdx.x * (size_t)RS_ITEMS;
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int r = 0; r < RS_ROUNDS; ++r) {
    const size_t i = base + (size_t)r * RS_TPB + t;
    const bool valid = i < n;
    const uint32_t k = valid ? keys[i] : 0u;
    const uint32_t d = digit_of(k, shift);
    // lanes of this wave holding the same digit
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t ones = __ballot(valid && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? ones : ~ones;
    }
    const int lrank = __popcll(same & lt);
    if (valid && lrank == 0) wc[w][d] = __popcll(same);
    __syncthreads();
    if (valid) {
      int pos = dstart[d] + run[d] + lrank;
      for (int q = 0; q < w; ++q) pos += wc[q][d];
      sk[pos] = k;
      sv[pos] = vals[i];
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < RS_TPB / 64; ++q) {
      add += wc[q][t];
      wc[q][t] = 0;
    }
    run[t] += add;
    __syncthreads();
  }
  // LDS slot p holds the element of local rank p - dstart[d] among digit d
  const size_t m = n - base < (size_t)RS_ITEMS ? n - base : (size_t)RS_ITEMS;
  for (size_t p = t; p < m; p += RS_TPB) {
    const uint32_t k = sk[p];
    const uint32_t d = digit_of(k, shift);
    const size_t pos = (size_t)start[(size_t)d * nblocks + blockIdx.x] + (p - (size_t)dstart[d]);
    okeys[pos] = k;
    ovals[pos] = sv[p];
  }
}

hipError_t scan_counts(int64_t *out, const int32_t *v, size_t n, int64_t *bsum, hipStream_t s) {
  const size_t nb = (n + SCAN_TPB - 1) / SCAN_TPB;
  hipLaunchKernelGGL(k_scan_bsum, dim3((unsigned)nb), dim3(SCAN_TPB), 0, s, bsum, v, n);
  hipLaunchKernelGGL(k_scan_bsum_scan, dim3(1), dim3(SCAN_TPB), 0, s, bsum, nb);
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(SCAN_TPB), 0, s, out, v, n, bsum, nb);
  return hipGetLastError();
}

}  // namespace

// *bad (a device word) = 1 when some of the n keys is >= m, else 0
hipError_t keys_out_of_range(unsigned long long *bad, const uint64_t *keys, size_t n, size_t m, hipStream_t s) {
  hipLaunchKernelGGL(k_bound_clear, dim3(1), dim3(1), 0, s, bad);
  if (n) hipLaunchKernelGGL(k_keys_bound, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, bad, keys, n, (uint64_t)m);
  return hipGetLastError();
}

// Device scratch of csr_sort_keys, in bytes, for n keys.
size_t csr_sort_bytes(size_t n) {
  const size_t nblocks = (n + RS_ITEMS - 1) / RS_ITEMS;
  const size_t hist_n = 256 * nblocks;
  const size_t nb = (hist_n + SCAN_TPB - 1) / SCAN_TPB;
  return 16 * n              // keys and indices, two buffers each
         + 4 * hist_n        // digit histograms
         + 8 * (hist_n + 1)  // scanned histograms
         + 8 * (nb + 1);     // scan block sums
}

// off[0..m+1] = CSR offsets (int64, device) of the m + 1 buckets: keys
// 0..m-1, then every key >= m; *order = element indices stably sorted by
// bucket (n int32 inside scratch).  scratch: csr_sort_bytes(n) bytes,
// 8-byte aligned.  Needs m < 2^31 - 1 and n < 2^31.  Stream-ordered, no host
// synchronisation, no memset (capturable).
hipError_t csr_sort_keys(int64_t *off, const int32_t **order, const uint64_t *keys, size_t n, size_t m,
                         void *scratch, hipStream_t s) {
  const size_t nblocks = (n + RS_ITEMS - 1) / RS_ITEMS;
  const size_t hist_n = 256 * nblocks;
  uint32_t *ka = static_cast<uint32_t *>(scratch);
  uint32_t *kb = ka + n;
  uint32_t *va = kb + n;
  uint32_t *vb = va + n;
  int32_t *hist = reinterpret_cast<int32_t *>(vb + n);
  int64_t *hstart = reinterpret_cast<int64_t *>(hist + hist_n);   // hist_n is a multiple of 256
  int64_t *bsum = hstart + hist_n + 1;
  auto grid = [](size_t x) { return dim3((unsigned)((x + 255) / 256)); };
  if (n) hipLaunchKernelGGL(k_keys_init, grid(n), dim3(256), 0, s, ka, va, keys, n, (uint64_t)m);
  int bits = 0;                      // bucket numbers lie in [0, m]
  while (bits < 32 && ((uint64_t)1 << bits) <= (uint64_t)m) ++bits;
  const int passes = n ? (bits + 7) / 8 : 0;
  for (int q = 0; q < passes; ++q) {
    const int shift = 8 * q;
    hipLaunchKernelGGL(k_rs_hist, dim3((unsigned)nblocks), dim3(RS_TPB), 0, s, hist, ka, n, shift, nblocks);
    hipError_t e = scan_counts(hstart, hist, hist_n, bsum, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rs_scatter, dim3((unsigned)nblocks), dim3(RS_TPB), 0, s, kb, vb, ka, va, n, shift, hist,
                       hstart, nblocks);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  hipLaunchKernelGGL(k_offsets, grid(m + 2), dim3(256), 0, s, off, ka, n, m + 1);
  *order = reinterpret_cast<const int32_t *>(va);
  return hipGetLastError();
}

}  // namespace ft
