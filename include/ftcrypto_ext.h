/*
 * ftcrypto_ext.h -- extensions of the MI355X library beyond the reference's
 * 38 functions.  The reference has no ciphertext type and no encryption
 * routine: its caller composes Enc(m; r) = (r B, m B + r H) from
 * ft_array_from_scalars, ft_array_init_point, ft_mul and ft_add
 * (ftcrypto.h:86-97, 275-277, 335-337, 383-385).  The ftx_ functions below
 * fuse those compositions, add the CSR segmented sum and sparse gather-sum
 * of the CPU path (ftcrypto.c:435-474), and give device-resident, stream-
 * ordered variants for callers that keep their data in HBM.
 *
 * Streams are passed as `void *` (a hipStream_t; NULL = the library stream).
 * Functions with a _dev suffix take device pointers and return as soon as
 * the work is enqueued; all others consume host buffers before returning.
 */
#ifndef FT_CRYPTO_EXT_H
#define FT_CRYPTO_EXT_H

#include "ftcrypto.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque comb tables of a public key H: 32 x 8 affine multiples (28 KiB, the
   radix-16 comb) and 51 x 16 (89 KiB, the radix-32 comb of the two-lane
   encryption schedule), built once by ftx_key_create. */
struct ftx_key_s;
typedef struct ftx_key_s *ftx_key;

/* ---------------------------------------------------------- arrays */

/* New array of n elements whose contents are unspecified until written. */
ft_error ftx_array_alloc(ft_ge25519_array *array, size_t n);
/* Device address of coordinate plane 0 and the plane stride (elements). */
ft_error ftx_array_device_view(const_ft_ge25519_array array, void **planes, size_t *capacity);
/* Library stream (hipStream_t) used by the reference-compatible functions. */
void *ftx_default_stream(void);
ft_error ftx_synchronize(void *stream);

/* ---------------------------------------------------------- ElGamal */

/* Comb table of H = pk[pk_idx]; reused by every encryption under that key. */
ft_error ftx_key_create(ftx_key *key, const_ft_ge25519_array pk, size_t pk_idx);
void ftx_key_free(ftx_key key);

/* c1[i] = r_i B, c2[i] = m_i B + r_i H for host m (uint64) and r (32-byte
 * scalars).  msg_bits is a public bound on the bit length of every m_i: it
 * fixes the number of message windows, so the time taken depends on it and
 * not on the messages.  0 or 64: any uint64 (17 windows).  With a bound
 * 0 < msg_bits < 64, a message >= 2^msg_bits gives FT_ERR_OUT_OF_BOUNDS and
 * nothing is allocated. */
ft_error ftx_encrypt(ft_ge25519_array *c1, ft_ge25519_array *c2, ftx_key key,
                     const uint64_t *m, const uint8_t *r, size_t n, int msg_bits);
/* c1[i] += r_i B, c2[i] += r_i H */
ft_error ftx_rerandomize(ft_ge25519_array c1, ft_ge25519_array c2, ftx_key key,
                         const uint8_t *r);
/* c1[i] = k_i c1[i] and c2[i] = k_i c2[i] (c2 may be NULL) for uint32 k.
 * In this and every other call taking a ciphertext's two halves, c1 and c2
 * must be different arrays (FT_ERR_BAD_SIZE otherwise). */
ft_error ftx_mul_u32(ft_ge25519_array c1, ft_ge25519_array c2, const uint32_t *k, size_t n);

/* ---------------------------------------------------------- aggregation */

/* res[s] = sum of src[offsets[s] .. offsets[s+1]-1]; empty segment -> zero.
 * offsets: n_seg + 1 non-decreasing int64 values, offsets[n_seg] <= len(src). */
ft_error ftx_segmented_sum(ft_ge25519_array *res, const int64_t *offsets, size_t n_seg,
                           const_ft_ge25519_array src);
/* res[v] = sum of src[cols[j]] for j in [row_off[v], row_off[v+1]) -- the CSR
 * form of the CPU path's ft_sparse_matmul (ftcrypto.c:435-474). */
ft_error ftx_gather_sum(ft_ge25519_array *res, const int64_t *row_off, const int32_t *cols,
                        size_t n_rows, const_ft_ge25519_array src);

/* ---------------------------------------------------------- device-resident */

ft_error ftx_from_scalars_dev(ft_ge25519_array out, const void *d_scalars, void *stream);
ft_error ftx_mul_dev(ft_ge25519_array array, const void *d_scalars, void *stream);
ft_error ftx_mul_u32_dev(ft_ge25519_array c1, ft_ge25519_array c2, const void *d_k, void *stream);
ft_error ftx_add_dev(ft_ge25519_array res, const_ft_ge25519_array summand, void *stream);
/* Device form of ftx_encrypt (d_m: uint64 messages, d_r: 32-byte scalars).
 * A bound 0 < msg_bits < 64 is checked on the device: the device uint64 at
 * d_fail becomes 1 + the largest index whose message is >= 2^msg_bits (0 when
 * none), and when it is nonzero the encryption writes nothing (c1 and c2 keep
 * their contents).  d_fail is required with such a bound (FT_ERR_BAD_SIZE
 * otherwise); with msg_bits 0 or 64 it may be NULL, and is set to 0 if not. */
ft_error ftx_encrypt_dev(ft_ge25519_array c1, ft_ge25519_array c2, ftx_key key,
                         const void *d_m, const void *d_r, int msg_bits, void *d_fail, void *stream);
ft_error ftx_rerandomize_dev(ft_ge25519_array c1, ft_ge25519_array c2, ftx_key key,
                             const void *d_r, void *stream);
/* Several encryptions under one key in ONE launch (no counterpart in the
 * reference, whose caller composes each encryption from separate calls):
 * for every job, c1[i] = r_i B and c2[i] = m_i B + r_i H for i < len(c1),
 * with d_m (uint64) and d_r (32-byte scalars) in device memory as for
 * ftx_encrypt_dev; msg_bits bounds every job's messages, checked on the
 * device as there: bit j of the device uint64 at d_fail is set when a message
 * of job j is >= 2^msg_bits, and then no job is written.  Two lanes per
 * ciphertext, the jobs' ciphertexts laid end to end, so small arrays fill
 * the GPU together.  At most FTX_ENC_MAX_JOBS jobs; the jobs' output arrays
 * must be pairwise distinct (FT_ERR_BAD_SIZE otherwise). */
#define FTX_ENC_MAX_JOBS 16
typedef struct {
  ft_ge25519_array c1, c2;
  const void *d_m, *d_r;
} ftx_enc_job;
ft_error ftx_encrypt_batch_dev(const ftx_enc_job *jobs, size_t n_jobs, ftx_key key, int msg_bits,
                               void *d_fail, void *stream);
/* out must have n_seg elements and be another array than src
 * (FT_ERR_BAD_SIZE); d_gather may be NULL (plain CSR), otherwise
 * n_elems int32 source indices, each below len(src) (device data: the caller
 * guarantees this; ftx_gather_sum checks it on the host); d_scratch must hold
 * ftx_segsum_scratch_points(n_elems) points (NULL: library-owned).
 * With d_gather, a scratch of ftx_gather_scratch_points(n_elems, len(src))
 * points also takes a point-contiguous copy of src, from which each gathered
 * point is one 128-byte line instead of four (faster random gathers). */
size_t ftx_segsum_scratch_points(size_t n_elems);
size_t ftx_gather_scratch_points(size_t n_elems, size_t n_src);
ft_error ftx_segmented_sum_dev(ft_ge25519_array out, const void *d_offsets, size_t n_seg,
                               size_t n_elems, const void *d_gather,
                               const_ft_ge25519_array src, ft_ge25519_array scratch, void *stream);
/* Several aggregation jobs whose element walks share ONE persistent,
 * load-balanced launch (no counterpart in the reference, which launches
 * each reduce, scan and per-segment sum separately: ftcrypto.cu:1113-1217,
 * 1342-1586).  Every wave takes the next tile of any job from a device
 * counter, so uneven jobs and segment lengths balance over the GPU; run-tree
 * passes, the sums of the reduces' partials and the upper scan levels follow
 * on the same stream.  Jobs run concurrently: no job's out may be another
 * job's out or src.  Kinds:
 *   FTX_AGG_SEGSUM  out[s] = sum of src[off[s] .. off[s+1]-1], s < n_seg;
 *                   d_offsets: n_seg + 1 device int64 values (caller-checked
 *                   as for ftx_segmented_sum_dev), n_elems = off[n_seg]; or
 *                   d_offsets NULL: segment s = [s seg_len, min((s+1)
 *                   seg_len, n_elems)) with n_seg = ceil(n_elems / seg_len)
 *   FTX_AGG_REDUCE  out[0] = sum of src (all len(src) elements)
 *   FTX_AGG_SCAN    inclusive scan of out in place (src unused)
 *   FTX_AGG_XSCAN   exclusive scan of out in place
 * scratch: ftx_aggregate_batch_scratch_points(jobs, n_jobs) points (NULL: the
 * call allocates it and synchronises).  At most FTX_AGG_MAX_JOBS jobs. */
#define FTX_AGG_SEGSUM 0
#define FTX_AGG_REDUCE 1
#define FTX_AGG_SCAN 2
#define FTX_AGG_XSCAN 3
#define FTX_AGG_MAX_JOBS 32
typedef struct {
  int kind;
  ft_ge25519_array out;
  const_ft_ge25519_array src;
  const void *d_offsets;
  size_t n_seg;
  size_t n_elems;
  size_t seg_len;
} ftx_agg_job;
size_t ftx_aggregate_batch_scratch_points(const ftx_agg_job *jobs, size_t n_jobs);
ft_error ftx_aggregate_batch_dev(const ftx_agg_job *jobs, size_t n_jobs, ft_ge25519_array scratch,
                                 void *stream);
/* Device-key form of ft_reduce_isum (ftcrypto.h:338-348, ftcrypto.cu:1150-1217):
 * res[keys[i]] += summand[i] for the len(summand) uint64 keys at d_keys.  A key
 * >= len(res) adds nothing (ft_reduce_isum rejects such keys on the host).
 * Keys are sorted on the device (stable radix sort), so the summands of one
 * key are added in increasing i and results are the same on every run.
 * scratch: ftx_reduce_isum_scratch_points(n_keys, len(res)) points; d_work:
 * ftx_reduce_isum_work_bytes(n_keys, len(res)) bytes of device memory, 8-byte
 * aligned.  Either may be NULL: the call then allocates it and synchronises. */
size_t ftx_reduce_isum_scratch_points(size_t n_keys, size_t n_res);
size_t ftx_reduce_isum_work_bytes(size_t n_keys, size_t n_res);
ft_error ftx_reduce_isum_dev(ft_ge25519_array res, const void *d_keys, const_ft_ge25519_array summand,
                             ft_ge25519_array scratch, void *d_work, void *stream);
/* out[0] = sum of src (out needs 1 element).  scratch: an array of at least
 * ftx_reduce_scratch_points(len(src)) points, or NULL (then the call allocates
 * and synchronises). */
size_t ftx_reduce_scratch_points(size_t n);
ft_error ftx_reduce_dev(ft_ge25519_array out, const_ft_ge25519_array src, ft_ge25519_array scratch,
                        void *stream);
/* ftx_reduce_dev that allocates nothing (no counterpart in the reference):
 * for callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  out[0] =
 * sum of src (out needs 1 element), enqueued on `stream`; the block sums go
 * to a static device array of FTX_REDUCE_NOALLOC_POINTS points (512 KiB, one
 * per GPU, held until the process ends) instead of a scratch array.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (on
 * different streams, or from several threads) race, so serialise them, e.g.
 * on one stream.  Any len(src) fits (the first pass uses at most 4096 blocks);
 * a schedule needing more than the array holds is refused (FT_ERR_BAD_SIZE). */
#define FTX_REDUCE_NOALLOC_POINTS 4096
ft_error ftx_reduce_noalloc_dev(ft_ge25519_array out, const_ft_ge25519_array src, void *stream);
/* in-place inclusive (exclusive = 0) or exclusive scan; scratch as above with
 * ftx_scan_scratch_points(len(array)) points */
size_t ftx_scan_scratch_points(size_t n);
ft_error ftx_scan_dev(ft_ge25519_array array, int exclusive, ft_ge25519_array scratch, void *stream);
/* 32-byte canonical compressed encodings into device memory d_out */
ft_error ftx_encode_dev(void *d_out, const_ft_ge25519_array array, void *stream);
/* Device form of ft_array_from_bytes_folded (ftcrypto.h:193-212): decodes
 * len(out) 32-byte encodings at d_in into out; the device uint64 at d_fail
 * becomes 1 + the largest index that failed to decode, 0 when all decoded
 * (the host function returns that index, or -1). */
ft_error ftx_decode_dev(ft_ge25519_array out, const void *d_in, void *d_fail, void *stream);

/* ---------------------------------------------------------- streams, graphs */

/* Opaque instantiated HIP graph of captured library calls. */
struct ftx_graph_s;
typedef struct ftx_graph_s *ftx_graph;

/* A non-blocking stream at default priority. */
ft_error ftx_stream_create(void **stream);
void ftx_stream_destroy(void *stream);
/* Capture the _dev calls issued on `stream` (not the library stream) between
 * begin and end into a graph.  Capturable: ftx_from_scalars_dev, ftx_mul_dev,
 * ftx_mul_u32_dev, ftx_add_dev, ftx_encrypt_dev, ftx_encrypt_batch_dev, ftx_rerandomize_dev,
 * ftx_encode_dev, and ftx_segmented_sum_dev / ftx_reduce_dev / ftx_scan_dev /
 * ftx_reduce_isum_dev / ftx_aggregate_batch_dev with caller-provided scratch
 * (and work memory). */
ft_error ftx_capture_begin(void *stream);
ft_error ftx_capture_end(void *stream, ftx_graph *graph);
ft_error ftx_graph_launch(ftx_graph graph, void *stream);
void ftx_graph_free(ftx_graph graph);

/* ---------------------------------------------------------- serving */

/* A request loop over ciphertext vectors resident in HBM (one public key).
 * Vectors are named by uint64 handles.  Requests are asynchronous: each
 * returns a request id; ftx_server_wait blocks until it completed and gives
 * its latency (submission to observed completion).  Queued requests of one
 * kind are batched into one kernel launch, batches run on `n_streams`
 * concurrent streams, and requests that create vectors are admitted only when
 * the `hbm_budget_bytes` arena can hold them.  Requests on one vector execute
 * in submission order.  A batch holds at most `max_batch_elems` ciphertexts
 * (0: FTX_SERVER_DEFAULT_BATCH); an encryption or multiply longer than that
 * runs as consecutive slices of that many ciphertexts, at most two in flight
 * and slices together on at most half of the streams, so that shorter requests
 * behind it are not held up for its whole length. */
#define FTX_SERVER_DEFAULT_BATCH ((size_t)1 << 18)
struct ftx_server_s;
typedef struct ftx_server_s *ftx_server;

/* flags: FTX_SERVER_BORROW_INPUTS -- host inputs passed to the submit calls
 * are read when their batch launches (not copied at submission) and must stay
 * valid and unchanged until the request completes; pinned (hipHostMalloc)
 * inputs are then transferred by DMA without a host copy. */
#define FTX_SERVER_BORROW_INPUTS 1
ft_error ftx_server_create(ftx_server *server, ftx_key key, size_t hbm_budget_bytes, int n_streams,
                           size_t max_batch_elems, int flags);
/* waits for every submitted request, then frees the server */
void ftx_server_destroy(ftx_server server);
ft_error ftx_server_encrypt(ftx_server server, const uint64_t *m, const uint8_t *r, size_t n, int msg_bits,
                            uint64_t *req, uint64_t *handle);
ft_error ftx_server_add(ftx_server server, uint64_t dst, uint64_t src, uint64_t *req);
ft_error ftx_server_mul_u32(ftx_server server, uint64_t handle, const uint32_t *k, size_t n, uint64_t *req);
ft_error ftx_server_segsum(ftx_server server, uint64_t handle, const int64_t *offsets, size_t n_seg,
                           uint64_t *req, uint64_t *out_handle);
ft_error ftx_server_wait(ftx_server server, uint64_t req, double *latency_us);
/* as ftx_server_wait, and the part of the latency the request spent queued
 * before its first batch launched (queued_us may be NULL) */
ft_error ftx_server_wait_ex(ftx_server server, uint64_t req, double *latency_us, double *queued_us);
ft_error ftx_server_release(ftx_server server, uint64_t handle);
/* folded encodings of both halves of a vector (waits for its pending work) */
ft_error ftx_server_read(ftx_server server, uint64_t handle, uint8_t *c1_enc, uint8_t *c2_enc);
ft_error ftx_server_stats(ftx_server server, uint64_t *bytes_in_use, uint64_t *batches, uint64_t *completed);

/* ------------------------------------------------ aggregation schedules */

/* Launch schedules of the reduce, scan and segmented sum and the window widths
 * of the variable-base mults, one per size class of the input length (class
 * 0: n < 2^14, 1: < 2^18, 2: < 2^22, 3: larger).  seg1 and scan are log2
 * elements per lane (4..7) at segmented-sum level 1 and at the scan's first
 * level (the levels above it take 16 per lane); segu
 * (1..4) is log2 of the units a quad walks per pass when the partial sums of
 * segments spanning several level-1 chunks are added; reduce_blocks (1..4096) bounds
 * the first reduce pass; u32_window
 * (ciphertext x u32) and varbase_window (ft_mul, ft_scale) are 4 (radix-16
 * digits, 8-entry table) or 3 (radix 8, 4 entries); lanes per ciphertext of
 * an encryption (ftx_tune_get_ct_lanes) is 1, 2 or 4.  Results are the same
 * under every schedule.
 * ftx_autotune times every candidate for classes 0..max_class on the device
 * (class 3 at n = 2^24, about 3.8 GB while it runs), keeps the fastest and
 * writes the table to path unless it is NULL; ft_crypto_init loads the file
 * named by the FTX_TUNE_FILE environment variable when set. */
ft_error ftx_autotune(int max_class, const char *path);
ft_error ftx_tune_save(const char *path);
ft_error ftx_tune_load(const char *path);
ft_error ftx_tune_get(int size_class, int *seg1, int *segu, int *scan, int *reduce_blocks, int *u32_window,
                      int *varbase_window);
ft_error ftx_tune_set(int size_class, int seg1, int segu, int scan, int reduce_blocks, int u32_window,
                      int varbase_window);
/* Lanes per ciphertext of an encryption or rerandomisation in a size class:
 * 1, or 2 (C1 and C2 on separate lanes: twice the waves of a small batch).
 * A field element's limbs always stay in one lane.  Tuned by ftx_autotune,
 * saved and loaded with the table (eighth column). */
ft_error ftx_tune_get_ct_lanes(int size_class, int *lanes);
ft_error ftx_tune_set_ct_lanes(int size_class, int lanes);
/* Scan form of a size class: 0, chunk levels (a walk summing chunks, the
 * upper levels, a second walk from each chunk's prefix); 1, one pass whose
 * 8,192-point tiles find their prefixes by a decoupled look-back over the
 * tiles before them (no counterpart in the reference, which scans
 * recursively: ftcrypto.cu:1474-1562).  Tuned by ftx_autotune, saved and
 * loaded with the table (ninth column).  Batched scans
 * (ftx_aggregate_batch_dev) always use the chunk levels. */
ft_error ftx_tune_get_scan_lookback(int size_class, int *lookback);
ft_error ftx_tune_set_scan_lookback(int size_class, int lookback);
int ftx_tune_size_class(size_t n);

/* ---------------------------------------------------------- inputs */

/* n scalars uniform mod l: 64 bytes of SplitMix64(seed, stream) output per
 * scalar, reduced modulo l = 2^252 + 27742317777372353535851937790883648493. */
void ftx_random_scalars(uint8_t *out, size_t n, uint64_t seed, uint64_t stream);
/* 64-byte little-endian integers -> 32-byte residues mod l */
void ftx_scalar_reduce64(uint8_t *out, const uint8_t *in, size_t n);

#ifdef __cplusplus
}
#endif

#endif
