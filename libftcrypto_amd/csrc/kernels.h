// Host-side launchers of the gfx950 kernels (kernels.hip, aggregate.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "layout.h"

namespace ft {

hipError_t launch_set_identity(PtView v, size_t lo, size_t hi, hipStream_t s);
hipError_t launch_broadcast(PtView d, size_t n, PtView src, size_t idx, hipStream_t s);
hipError_t launch_copy(PtView d, PtView src, size_t n, hipStream_t s);
hipError_t launch_add(PtView res, PtView b, size_t n, bool sub, hipStream_t s);
hipError_t launch_neg(PtView v, size_t n, hipStream_t s);
// out[i] = k_i B over B's radix-32 comb table (launch_build_comb32)
hipError_t launch_fixed_base(PtView out, const void *sc, bool u64, size_t n, const uint32_t *table32,
                             hipStream_t s);
hipError_t launch_varbase(PtView pts, const uint32_t *sc, size_t sc_stride_words, size_t n, hipStream_t s);
hipError_t launch_varbase_w(PtView pts, const uint32_t *sc, size_t sc_stride_words, size_t n, int w,
                            hipStream_t s);
hipError_t launch_mul_u32(PtView a, PtView b, bool two, const uint32_t *k, size_t n, hipStream_t s);
hipError_t launch_mul_u32_w(PtView a, PtView b, bool two, const uint32_t *k, size_t n, int w, hipStream_t s);
hipError_t launch_build_comb(uint32_t *table, PtView src, size_t idx, bool use_base, hipStream_t s);
hipError_t launch_build_comb32(uint32_t *table, PtView src, size_t idx, bool use_base, hipStream_t s);
// gate: a device word; when it is nonzero at launch time the launch writes
// nothing (the message-bound check of bounded encryptions), null: ungated
// btab / htab: radix-16 comb tables of B and H (staged in LDS by the one-
// and four-lane schedules); btab32 / htab32: their radix-32 tables, which
// the two-lane schedule reads through scalar loads (required for it)
hipError_t launch_encrypt(PtView c1, PtView c2, const uint32_t *btab, const uint32_t *htab,
                          const uint64_t *m, int nd, const uint32_t *r, size_t n, bool accumulate,
                          hipStream_t s, const unsigned long long *gate = nullptr,
                          const uint32_t *btab32 = nullptr, const uint32_t *htab32 = nullptr);
hipError_t launch_encrypt_l(PtView c1, PtView c2, const uint32_t *btab, const uint32_t *htab,
                            const uint64_t *m, int nd, const uint32_t *r, size_t n, bool accumulate, int lanes,
                            hipStream_t s, const unsigned long long *gate = nullptr,
                            const uint32_t *btab32 = nullptr, const uint32_t *htab32 = nullptr);
// fail = 1 + the largest index i with m[i] >= 2^bits, 0 when none (0 < bits < 64)
hipError_t launch_msg_bound(const uint64_t *m, size_t n, int bits, unsigned long long *fail, hipStream_t s);
// several ciphertext arrays encrypted in one launch (two lanes per ciphertext);
// the table passes by value as a kernel argument; bases (multiples of 64
// ciphertexts) are filled in
constexpr int ENC_BATCH_MAX = 16;
struct EncJob {
  PtView c1, c2;
  const uint64_t *m;
  const uint32_t *r;
  uint64_t base, n;
};
struct EncBatch {
  int J;
  EncJob job[ENC_BATCH_MAX];
};
// over the radix-32 tables of B and H (two waves per 64 ciphertexts)
hipError_t launch_encrypt_batch(EncBatch &B, const uint32_t *btab32, const uint32_t *htab32, int nd, hipStream_t s,
                                const unsigned long long *gate = nullptr);
// radix-32 message digits for a radix-16 digit count nd (bits / 4 rounded
// up, plus one; 17 for any uint64): (bits + 5) / 5 for the largest bits the
// count allows, the top digit keeping the last carry (at most 16)
inline int msg_digits32(int nd) {
  if (nd <= 0) return 0;
  const int bits = nd >= 17 ? 64 : 4 * (nd - 1);
  return (bits + 5) / 5;
}
// bit j of fail set when a message of job j is >= 2^bits, 0 when none
hipError_t launch_msg_bound_batch(EncBatch &B, int bits, unsigned long long *fail, hipStream_t s);
hipError_t launch_encode(uint32_t *out, PtView v, size_t n, hipStream_t s);
hipError_t launch_decode(PtView v, const uint32_t *in, size_t n, unsigned long long *bad, hipStream_t s);
unsigned reduce_grid(size_t n, unsigned max_blocks);
hipError_t launch_reduce_blocks(PtView part, PtView v, size_t n, unsigned grid, hipStream_t s);

// tune.hip: launch schedules of the aggregation kernels and the window widths
// of the variable-base kernels, per size class of n
// (n < 2^14, < 2^18, < 2^22, larger), picked by measurement (ftx_autotune)
// or left at the defaults.
constexpr int TUNE_CLASSES = 4;
struct AggTune {
  int seg1;          // log2 elements per lane, segmented-sum level 1
  int segu;          // log2 units per quad and pass of the segmented-sum block walk (1..4)
  int scan;          // log2 elements per lane, every scan level
  int reduce;        // most blocks of 256 lanes in the first reduce pass
  int u32w;          // window width of the 32-bit-multiplier kernels (3 or 4)
  int vbw;           // window width of the 256-bit variable-base kernel (3 or 4)
  int encl;          // lanes per ciphertext of an encryption (1, or 2: C1 and C2 apart)
  int scanlb;        // scan: 0 chunk levels (two walks, upper levels between), 1 single-pass look-back
};
constexpr int CH_MIN = 4, CH_MAX = 7;             // allowed chunk log2 range
constexpr int TREE_LR_MIN = 1, TREE_LR_MAX = 4;   // allowed log2 units per block-walk pass
constexpr unsigned REDUCE_BLOCKS_MAX = 4096;
// reduces of at most this many points run their first pass on quad lanes
constexpr size_t REDUCE_QUAD_MAX = (size_t)1 << 18;
int tune_class(size_t n);
size_t tune_class_size(int c);                   // the length each class is tuned at
AggTune tune_for(size_t n);
void tune_init_from_env();

// aggregate.hip
size_t scan_scratch_points(size_t n);             // enough for every schedule
size_t lookback_scratch_points(size_t n);         // the single-pass look-back scan's
hipError_t scan_lookback_launch(PtView v, size_t n, bool exclusive, PtView scratch, hipStream_t s);
hipError_t scan_launch(PtView v, size_t n, bool exclusive, PtView scratch, hipStream_t s);
hipError_t scan_launch_ch(PtView v, size_t n, bool exclusive, PtView scratch, int ch, hipStream_t s,
                          bool level1_done = false);
hipError_t launch_seg_empty(PtView out, const int64_t *off, size_t nseg, hipStream_t s);
size_t segsum_scratch_points(size_t n_elems);     // enough for every schedule
// gather form with n_src source points: a scratch of at least
// segsum_gather_scratch_points(n_elems, n_src) points also holds a
// point-contiguous copy of the source, from which every gathered point is
// one 128-byte line (a smaller scratch gathers from the four planes)
size_t segsum_gather_scratch_points(size_t n_elems, size_t n_src);
hipError_t segsum_launch(PtView out, const int64_t *d_off, size_t nseg, size_t n_elems, const int32_t *gather,
                         PtView src, size_t n_src, PtView scratch, hipStream_t s);
hipError_t segsum_launch_ch(PtView out, const int64_t *d_off, size_t nseg, size_t n_elems,
                            const int32_t *gather, PtView src, size_t n_src, PtView scratch, int ch1, int lr,
                            hipStream_t s);

// Batched aggregation: up to AGG_BATCH_MAX jobs whose element walks share
// one persistent, load-balanced launch (aggregate.hip).  Jobs run
// concurrently: no job may write an array another job reads or writes.
enum AggKind { AGG_SEGSUM = 0, AGG_REDUCE = 1, AGG_SCAN = 2, AGG_XSCAN = 3 };
constexpr int AGG_BATCH_MAX = 32;
struct AggJob {
  int kind;
  PtView out;          // segmented sum: nseg sums; reduce: one sum; scan: the array, in place
  PtView src;          // segmented sum, reduce: the summed points
  const int64_t *off;  // segmented sum: nseg + 1 device CSR offsets, or null: segments of len
  size_t len;          // uniform segment length (nseg = ceil(n / len))
  size_t nseg;
  size_t n;            // elements walked
};
size_t agg_batch_scratch_points(const AggJob *jobs, int J);   // enough for every schedule
hipError_t agg_batch_launch(const AggJob *jobs, int J, PtView scratch, hipStream_t s);

// sort.hip: stable device sort of n keys into the CSR order of m + 1
// buckets (keys 0..m-1, then every key >= m)
size_t csr_sort_bytes(size_t n);
hipError_t keys_out_of_range(unsigned long long *bad, const uint64_t *keys, size_t n, size_t m, hipStream_t s);
hipError_t csr_sort_keys(int64_t *off, const int32_t **order, const uint64_t *keys, size_t n, size_t m,
                         void *scratch, hipStream_t s);

// noalloc.hip: the static block-sum workspace of the allocation-free reduce,
// REDUCE_BLOCKS_MAX points in a file-scope device array (one copy per GPU).
// noalloc_resolve takes its address on a device (once; ft_crypto_init does it
// for the library's device), noalloc_scratch gives the current device's copy.
constexpr size_t NOALLOC_POINTS = REDUCE_BLOCKS_MAX;
hipError_t noalloc_resolve(int device);
hipError_t noalloc_scratch(PtView *v);

// api.hip: device tables owned by the library
const uint32_t *base_table();
const uint32_t *base_table32();
const uint32_t *key_table(const struct ftx_key_s *key);
const uint32_t *key_table32(const struct ftx_key_s *key);

}  // namespace ft
