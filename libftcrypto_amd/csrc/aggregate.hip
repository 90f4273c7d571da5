// Scan and CSR segmented-sum kernels for gfx950 (combine = point addition).
//
// Both split the input into fixed chunks of 2^CH consecutive elements, one
// chunk per lane, so the work per lane is the same whatever the segment
// lengths are (the reference runs one thread per segment, ftcrypto.cu:1116-1127,
// and a recursive Blelloch scan, :1474-1562).  The chunk size of each kernel
// is a template parameter in [2^4, 2^7]; the schedule used for an input size
// comes from the tuning table (tune.hip).
//
// Scan: chunk sums are taken level by level until a level has at most 2^16
// points; that level is scanned by blocks (wave shuffles, LDS), then every
// lane of each level below walks its chunk once more from its chunk's
// prefix.
//
// Segmented sum: level 1 walks the elements; a segment that lies inside one
// chunk is written to the output directly, one that crosses a chunk edge
// leaves a partial sum and its segment id in the chunk's head slot (segment
// began before the chunk) or tail slot (segment continues after it).  A
// segment spanning chunks c0..c1 thus has one partial per chunk: c0's tail
// slot, then the head slots of c0+1..c1: in record order a segment's
// records are consecutive.  They are summed by a block walk on quads
// (k_seg_up, and k_agg_up over every job of a batched launch): each pass
// walks the records of 2^LR units (chunks, then blocks of them; LR 1..4,
// tuned) per quad, with a point spread over the quad's four lanes (three
// dependent multiplies per add), writes the segments that end in the block
// and leaves one head and one tail record per block for the next pass, whose
// work is 2^LR times smaller.  (A radix tree over each run with one lane per
// chunk, which this replaced, spent its first passes on waves where a few
// lanes added and the rest waited.)
#include <algorithm>
#include <vector>
#include <atomic>
#include <cstring>
#include <mutex>

#include "kernels.h"
#include "quad.h"

namespace ft {

constexpr int ATPB = 256;

static size_t tiles(size_t n, int ch) { return (n + ((size_t)1 << ch) - 1) >> ch; }

// ------------------------------------------------------------------ scan

// Lane-transposed load of one coordinate of a group for a wave of whole
// chunks: eight coalesced loads (eight lanes per 128-byte line, eight whole
// lines per load) land in r[0..7]; through LDS every lane then gets its own
// line (its four points' 32-byte pieces) back in r.
template <int CH>
__device__ __forceinline__ void ld_line_issue(uint4 r[8], const uint8_t *src, int lane) {
  const size_t off = ((size_t)(lane >> 3) << (CH + 5)) + (size_t)(lane & 7) * 16;
#pragma unroll
  for (int k = 0; k < 8; ++k) r[k] = *reinterpret_cast<const uint4 *>(src + off + ((size_t)(8 * k) << (CH + 5)));
}
__device__ __forceinline__ void ld_line_tr(uint4 r[8], uint4 *tr, int lane) {
  const int pc = lane & 7;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int L = 8 * k + (lane >> 3);
    tr[L * 8 + (pc ^ ((L >> 1) & 7))] = r[k];
  }
  __builtin_amdgcn_wave_barrier();
  const int sw = (lane >> 1) & 7;
#pragma unroll
  for (int p = 0; p < 8; ++p) r[p] = tr[lane * 8 + (p ^ sw)];
  __builtin_amdgcn_wave_barrier();
}
__device__ __forceinline__ void line_to_pw(PW &a, PW &b, PW &c, PW &d, int co, const uint4 r[8]) {
  uint32_t *pa = co == 0 ? a.x : co == 1 ? a.y : co == 2 ? a.z : a.t;
  uint32_t *pb = co == 0 ? b.x : co == 1 ? b.y : co == 2 ? b.z : b.t;
  uint32_t *pc = co == 0 ? c.x : co == 1 ? c.y : co == 2 ? c.z : c.t;
  uint32_t *pd = co == 0 ? d.x : co == 1 ? d.y : co == 2 ? d.z : d.t;
  pa[0] = r[0].x; pa[1] = r[0].y; pa[2] = r[0].z; pa[3] = r[0].w;
  pa[4] = r[1].x; pa[5] = r[1].y; pa[6] = r[1].z; pa[7] = r[1].w;
  pb[0] = r[2].x; pb[1] = r[2].y; pb[2] = r[2].z; pb[3] = r[2].w;
  pb[4] = r[3].x; pb[5] = r[3].y; pb[6] = r[3].z; pb[7] = r[3].w;
  pc[0] = r[4].x; pc[1] = r[4].y; pc[2] = r[4].z; pc[3] = r[4].w;
  pc[4] = r[5].x; pc[5] = r[5].y; pc[6] = r[5].z; pc[7] = r[5].w;
  pd[0] = r[6].x; pd[1] = r[6].y; pd[2] = r[6].z; pd[3] = r[6].w;
  pd[4] = r[7].x; pd[5] = r[7].y; pd[6] = r[7].z; pd[7] = r[7].w;
}

template <int CH, bool TR>
__device__ __forceinline__ void chunk_sum_walk(ge_p3 &acc, PtView v, size_t lo, size_t hi, uint4 *tr, size_t wave0,
                                               int lane) {
#pragma unroll 1
  for (size_t gi = lo; gi < hi; gi += 4) {
    size_t g = gi;
    asm volatile("" : "+v"(g));
    PW q0, q1, q2, q3;
    if (TR) {
      const uint8_t *base = reinterpret_cast<const uint8_t *>(v.w) + ((wave0 << CH) + (g - lo)) * 32;
      const size_t plane = v.cap * 32;
      uint4 r[4][8];
#pragma unroll
      for (int c = 0; c < 4; ++c) ld_line_issue<CH>(r[c], base + c * plane, lane);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        ld_line_tr(r[c], tr, lane);
        line_to_pw(q0, q1, q2, q3, c, r[c]);
      }
    } else {
      FT_LOAD4(q0, q1, q2, q3, v, g, hi);
    }
    if (g == lo) un_pw(acc, q0);
    else add_pw(acc, q0);
    if (TR || g + 1 < hi) add_pw(acc, q1);
    if (TR || g + 2 < hi) add_pw(acc, q2);
    if (TR || g + 3 < hi) add_pw(acc, q3);
  }
}

// sums[t] = sum of elements [2^CH t, 2^CH (t + 1)), read a cache line at a time
template <int CH>
__global__ void __launch_bounds__(ATPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_chunk_sums(PtView sums, PtView v, size_t n) {
  __shared__ uint4 trs[ATPB / 64][64 * 8];
  const size_t t = blockIdx.x * (size_t)ATPB + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const size_t wave0 = t - (size_t)lane;
  const size_t lo = t << CH;
  if (lo >= n) return;
  const size_t hi = (lo + (1u << CH) < n) ? lo + (1u << CH) : n;
  ge_p3 acc;
  ge_p3_0(acc);
  if (((wave0 + 64) << CH) <= n)   // wave-uniform: all 64 chunks whole
    chunk_sum_walk<CH, true>(acc, v, lo, hi, trs[threadIdx.x >> 6], wave0, lane);
  else
    chunk_sum_walk<CH, false>(acc, v, lo, hi, nullptr, wave0, lane);
  store_p3(sums, t, acc);
}

// one scan element: q holds the element on entry and its output on return
template <bool EX>
__device__ __forceinline__ void scan_step(ge_p3 &run, PW &q) {
  ge_cached x;
  ge_p1p1 t;
  un_pw_cached(x, q);
  if (EX) pk_pw(q, run);
  ge_add(t, run, x);
  ge_p1p1_to_p3(run, t);
  if (!EX) pk_pw(q, run);
}

// Lane-transposed store of one coordinate of a group: every lane of the wave
// holds one whole 128-byte line (its four points' 32-byte pieces); the lines
// go through LDS so that each store instruction writes eight whole lines
// (eight lanes per line) instead of a 16-byte piece of 64 different lines.
// Slot (line L, piece p) sits at L * 8 + (p ^ ((L >> 1) & 7)): conflict-free
// for the line writes and the piece reads.
template <int CH>
__device__ __forceinline__ void st_line_tr(uint4 *tr, uint8_t *dst, int lane, const uint32_t *a, const uint32_t *b,
                                           const uint32_t *c, const uint32_t *d) {
  const int sw = (lane >> 1) & 7;
  uint4 *row = tr + lane * 8;
  row[0 ^ sw] = make_uint4(a[0], a[1], a[2], a[3]);
  row[1 ^ sw] = make_uint4(a[4], a[5], a[6], a[7]);
  row[2 ^ sw] = make_uint4(b[0], b[1], b[2], b[3]);
  row[3 ^ sw] = make_uint4(b[4], b[5], b[6], b[7]);
  row[4 ^ sw] = make_uint4(c[0], c[1], c[2], c[3]);
  row[5 ^ sw] = make_uint4(c[4], c[5], c[6], c[7]);
  row[6 ^ sw] = make_uint4(d[0], d[1], d[2], d[3]);
  row[7 ^ sw] = make_uint4(d[4], d[5], d[6], d[7]);
  __builtin_amdgcn_wave_barrier();
  const int pc = lane & 7;
  const size_t off = ((size_t)(lane >> 3) << (CH + 5)) + (size_t)pc * 16;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int L = 8 * k + (lane >> 3);
    const uint4 x = tr[L * 8 + (pc ^ ((L >> 1) & 7))];
    *reinterpret_cast<uint4 *>(dst + off + ((size_t)(8 * k) << (CH + 5))) = x;
  }
  __builtin_amdgcn_wave_barrier();
}

template <int CH, bool TR, bool EX>
__device__ __forceinline__ void chunk_walk(PtView v, size_t lo, size_t hi, ge_p3 &run, uint4 *tr,
                                           size_t wave0, int lane) {
#pragma unroll 1
  for (size_t gi = lo; gi < hi; gi += 4) {
    // an opaque copy of the element index: the plane addresses are then
    // formed each group from it and the uniform plane bases
    size_t g = gi;
    asm volatile("" : "+v"(g));
    PW q0, q1, q2, q3;
    const uint8_t *base = reinterpret_cast<const uint8_t *>(v.w) + ((wave0 << CH) + (g - lo)) * 32;
    const size_t plane = v.cap * 32;
    if (TR) {
      uint4 r[4][8];
#pragma unroll
      for (int c = 0; c < 4; ++c) ld_line_issue<CH>(r[c], base + c * plane, lane);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        ld_line_tr(r[c], tr, lane);
        line_to_pw(q0, q1, q2, q3, c, r[c]);
      }
    } else {
      FT_LOAD4(q0, q1, q2, q3, v, g, hi);
    }
    scan_step<EX>(run, q0);
    if (TR || g + 1 < hi) scan_step<EX>(run, q1);
    if (TR || g + 2 < hi) scan_step<EX>(run, q2);
    if (TR || g + 3 < hi) scan_step<EX>(run, q3);
    if (TR) {
      uint8_t *out = const_cast<uint8_t *>(base);
      st_line_tr<CH>(tr, out, lane, q0.x, q1.x, q2.x, q3.x);
      st_line_tr<CH>(tr, out + plane, lane, q0.y, q1.y, q2.y, q3.y);
      st_line_tr<CH>(tr, out + 2 * plane, lane, q0.z, q1.z, q2.z, q3.z);
      st_line_tr<CH>(tr, out + 3 * plane, lane, q0.t, q1.t, q2.t, q3.t);
    } else {
      st_pw4(v, g, (int)(hi - g < 4 ? hi - g : 4), q0, q1, q2, q3);
    }
  }
}

// walk chunk t from its exclusive prefix (identity when pre.w is null); the
// four outputs of a group are written together.  A wave whose 64 chunks are
// all whole writes them through LDS, a whole 128-byte line per eight lanes
// and store; the array's last wave stores each coordinate's line by
// consecutive stores of its lane.
template <int CH, bool EX>
__global__ void __launch_bounds__(ATPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_chunk_scan(PtView v, size_t n, PtView pre) {
  __shared__ uint4 trs[ATPB / 64][64 * 8];
  const size_t t = blockIdx.x * (size_t)ATPB + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const size_t wave0 = t - (size_t)lane;
  const size_t lo = t << CH;
  if (lo >= n) return;
  const size_t hi = (lo + (1u << CH) < n) ? lo + (1u << CH) : n;
  ge_p3 run;
  if (pre.w) load_p3(run, pre, t);
  else ge_p3_0(run);
  if (((wave0 + 64) << CH) <= n)   // wave-uniform: all 64 chunks whole
    chunk_walk<CH, true, EX>(v, lo, hi, run, trs[threadIdx.x >> 6], wave0, lane);
  else
    chunk_walk<CH, false, EX>(v, lo, hi, run, nullptr, wave0, lane);
}

static hipError_t launch_chunk_sums(PtView sums, PtView v, size_t n, int ch, hipStream_t s) {
  const size_t T = tiles(n, ch);
  if (!T) return hipSuccess;
  const dim3 grid((unsigned)((T + ATPB - 1) / ATPB));
  switch (ch) {
    case 4: hipLaunchKernelGGL(k_chunk_sums<4>, grid, dim3(ATPB), 0, s, sums, v, n); break;
    case 5: hipLaunchKernelGGL(k_chunk_sums<5>, grid, dim3(ATPB), 0, s, sums, v, n); break;
    case 6: hipLaunchKernelGGL(k_chunk_sums<6>, grid, dim3(ATPB), 0, s, sums, v, n); break;
    case 7: hipLaunchKernelGGL(k_chunk_sums<7>, grid, dim3(ATPB), 0, s, sums, v, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static hipError_t launch_chunk_scan(PtView v, size_t n, PtView pre, bool exclusive, int ch, hipStream_t s) {
  const size_t T = tiles(n, ch);
  if (!T) return hipSuccess;
  const dim3 grid((unsigned)((T + ATPB - 1) / ATPB));
#define FT_CSCAN(C)                                                                                     \
  if (exclusive) hipLaunchKernelGGL((k_chunk_scan<C, true>), grid, dim3(ATPB), 0, s, v, n, pre);        \
  else hipLaunchKernelGGL((k_chunk_scan<C, false>), grid, dim3(ATPB), 0, s, v, n, pre)
  switch (ch) {
    case 4: FT_CSCAN(4); break;
    case 5: FT_CSCAN(5); break;
    case 6: FT_CSCAN(6); break;
    case 7: FT_CSCAN(7); break;
    default: return hipErrorInvalidValue;
  }
#undef FT_CSCAN
  return hipGetLastError();
}

// ---------------------------------------------- block scan (short levels)
//
// A level of at most BSCAN_MAX points is scanned by blocks of 256 points
// (k_block_scan_q below): shuffle-and-add steps scan each wave, the wave
// totals are scanned the same way, and block totals are scanned the same
// way (exclusive) and added back.  That is about ten adds per point, against
// two for a chunk level, but a chain about ten adds deep instead of two
// chains of 2^ch: the short top levels of a scan are latency-bound, the
// long bottom ones not.
constexpr size_t BSCAN_MAX = (size_t)1 << 16;

__device__ __forceinline__ void shfl_up_p3(ge_p3 &r, const ge_p3 &a, int d) {
  fe_shfl_up(r.X, a.X, d);
  fe_shfl_up(r.Y, a.Y, d);
  fe_shfl_up(r.Z, a.Z, d);
  fe_shfl_up(r.T, a.T, d);
}

// points of scratch the block scan of m points needs (its block totals)
static size_t block_scan_points(size_t m) {
  size_t tot = 0;
  for (size_t k = tiles(m, 8); k > 1; k = tiles(k, 8)) tot += k;
  return tot;
}

// ------------------------------------------- quad-lane scan (short levels)
//
// Levels with few chunks are latency-bound: each lane's chain of 2^ch adds
// (and stores) is the critical path and the GPU holds far more waves than
// the level has.  There a chunk walk goes to the four lanes of a quad (one
// coordinate each, quad.h), whose point add takes three dependent
// multiplies instead of nine, and the block scan of the top level scans 256
// points per block of 256 quads.  Results are the same points as the
// one-lane kernels'.  Four lanes per point also issue about 1.7 times the
// instructions of one lane per point, so the quad form pays only where a
// level is short: at 2^16 chunks of 16 the walk that stores every output is
// 26 % faster on quads (96 against 130 us), the one that only sums them is
// not (83 against 82 us) and stays on one lane per chunk; the block scan of
// 2^16 points takes 36 + 33 us against 49 + 44.
constexpr size_t QUAD_CHUNKS_MAX = (size_t)1 << 16;
constexpr int QBTPB = 1024;                 // block-scan blocks: 256 quads

__device__ __forceinline__ const uint32_t *coord(PtView v, int c, size_t i) { return v.w + (c * v.cap + i) * 8; }
__device__ __forceinline__ uint32_t *coord_w(PtView v, int c, size_t i) { return v.w + (c * v.cap + i) * 8; }

// walk chunk t on quad t from its exclusive prefix (identity when pre.w is
// null), storing each output
template <int CH>
__global__ void __launch_bounds__(ATPB) k_chunk_scan_q(PtView v, size_t n, PtView pre, int exclusive) {
  const int c = threadIdx.x & 3;
  const size_t t = (blockIdx.x * (size_t)ATPB + threadIdx.x) >> 2;
  const size_t lo = t << CH;
  if (lo >= n) return;
  const size_t hi = (lo + (1u << CH) < n) ? lo + (1u << CH) : n;
  fe run, x;
  if (pre.w) {
    load_fe(run, coord(pre, c, t));
    quad_from_mem(run, c);
  } else {
    quad_identity(run, c);
  }
  load_fe(x, coord(v, c, lo));
#pragma unroll 1
  for (size_t g = lo; g < hi; ++g) {
    fe nx, m;
    load_fe(nx, coord(v, c, g + 1 < hi ? g + 1 : g));
    if (exclusive) {
      m = run;
      quad_to_mem(m, c);
      store_fe(coord_w(v, c, g), m);
    }
    quad_add<true>(run, x, c);
    if (!exclusive) {
      m = run;
      quad_to_mem(m, c);
      store_fe(coord_w(v, c, g), m);
    }
    x = nx;
  }
}

static hipError_t launch_chunk_scan_q(PtView v, size_t n, PtView pre, bool exclusive, int ch, hipStream_t s) {
  const size_t T = tiles(n, ch);
  if (!T) return hipSuccess;
  const dim3 grid((unsigned)((4 * T + ATPB - 1) / ATPB));
  switch (ch) {
    case 4: hipLaunchKernelGGL(k_chunk_scan_q<4>, grid, dim3(ATPB), 0, s, v, n, pre, (int)exclusive); break;
    case 5: hipLaunchKernelGGL(k_chunk_scan_q<5>, grid, dim3(ATPB), 0, s, v, n, pre, (int)exclusive); break;
    case 6: hipLaunchKernelGGL(k_chunk_scan_q<6>, grid, dim3(ATPB), 0, s, v, n, pre, (int)exclusive); break;
    case 7: hipLaunchKernelGGL(k_chunk_scan_q<7>, grid, dim3(ATPB), 0, s, v, n, pre, (int)exclusive); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// in-block inclusive (exclusive) scan of v[0, m) in place, 256 points per
// block on its 256 quads: each wave scans its 16 quads (shuffles up by 4, 8,
// 16, 32 lanes), wave 0 scans the 16 wave totals the same way, and each wave
// adds the total of the waves before it: nine dependent adds per point.
// Block totals to tot when tot.w is set.
__global__ void __launch_bounds__(QBTPB) k_block_scan_q(PtView v, size_t m, PtView tot, int exclusive) {
  constexpr int NW = QBTPB / 64;
  __shared__ uint32_t wt[NW * 4 * 9];
  const int c = threadIdx.x & 3, q = threadIdx.x >> 2;
  const int wq = q & 15, w = q >> 4;
  const size_t i = blockIdx.x * (size_t)(QBTPB / 4) + q;
  fe x;
  if (i < m) {
    load_fe(x, coord(v, c, i));
    quad_from_mem(x, c);
  } else {
    quad_identity(x, c);
  }
#pragma unroll 1
  for (int d = 1; d < 16; d <<= 1) {
    fe y, z;
    fe_shfl_up(y, x, 4 * d);
    z = x;
    quad_add(z, y, c);
#pragma unroll
    for (int l = 0; l < 9; ++l) x.v[l] = wq >= d ? z.v[l] : x.v[l];
  }
  if (wq == 15) {
#pragma unroll
    for (int l = 0; l < 9; ++l) wt[(w * 4 + c) * 9 + l] = x.v[l];
  }
  __syncthreads();
  if (w == 0) {
    fe t;
#pragma unroll
    for (int l = 0; l < 9; ++l) t.v[l] = wt[(wq * 4 + c) * 9 + l];
#pragma unroll 1
    for (int d = 1; d < 16; d <<= 1) {
      fe y, z;
      fe_shfl_up(y, t, 4 * d);
      z = t;
      quad_add(z, y, c);
#pragma unroll
      for (int l = 0; l < 9; ++l) t.v[l] = wq >= d ? z.v[l] : t.v[l];
    }
#pragma unroll
    for (int l = 0; l < 9; ++l) wt[(wq * 4 + c) * 9 + l] = t.v[l];    // waves 0 .. wq
  }
  __syncthreads();
  fe pre;                                        // total of the earlier waves
  if (w > 0) {
#pragma unroll
    for (int l = 0; l < 9; ++l) pre.v[l] = wt[((w - 1) * 4 + c) * 9 + l];
    quad_add(x, pre, c);
  } else {
    quad_identity(pre, c);
  }
  if (tot.w && q == QBTPB / 4 - 1) {
    fe t = x;
    quad_to_mem(t, c);
    store_fe(coord_w(tot, c, blockIdx.x), t);
  }
  if (exclusive) {
    fe y;
    fe_shfl_up(y, x, 4);
#pragma unroll
    for (int l = 0; l < 9; ++l) x.v[l] = wq == 0 ? pre.v[l] : y.v[l];
  }
  quad_to_mem(x, c);
  if (i < m) store_fe(coord_w(v, c, i), x);
}

// v[i] = pre[b] + v[i] for the points of blocks b >= 1 (256 per block)
__global__ void __launch_bounds__(QBTPB) k_block_add_q(PtView v, size_t m, PtView pre) {
  const int c = threadIdx.x & 3;
  const size_t b = blockIdx.x + 1;
  const size_t i = b * (QBTPB / 4) + (threadIdx.x >> 2);
  if (i >= m) return;
  fe x, p;
  load_fe(x, coord(v, c, i));
  load_fe(p, coord(pre, c, b));
  quad_from_mem(x, c);
  quad_add<true>(x, p, c);
  quad_to_mem(x, c);
  store_fe(coord_w(v, c, i), x);
}

static hipError_t block_scan_q(PtView v, size_t m, bool exclusive, PtView work, hipStream_t s) {
  const size_t nb = tiles(m, 8);
  if (nb <= 1) {
    hipLaunchKernelGGL(k_block_scan_q, dim3(1), dim3(QBTPB), 0, s, v, m, PtView{nullptr, 0}, (int)exclusive);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_block_scan_q, dim3((unsigned)nb), dim3(QBTPB), 0, s, v, m, work, (int)exclusive);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = block_scan_q(work, nb, true, PtView{work.w + nb * 8, work.cap}, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_block_add_q, dim3((unsigned)(nb - 1)), dim3(QBTPB), 0, s, v, m, work);
  return hipGetLastError();
}

// scratch of a scan of n points with chunks of 2^ch: the chunk sums of every
// level longer than BSCAN_MAX, then the block totals of the top level
// Chunk size of level l of a scan whose first level has 2^ch points per
// lane: the levels above the first are short and latency-bound (a lane's
// chain of 2^ch dependent adds sets their time), so they take the smallest
// chunks (2^18 points at 64 per lane: 250 us; at 16 per lane: about 60).
static int scan_level_ch(int l, int ch) { return l == 0 ? ch : CH_MIN; }

static size_t scan_plan_points(size_t n, int ch) {
  size_t used = 0, m = n;
  for (int l = 0; m > BSCAN_MAX; ++l) {
    m = tiles(m, scan_level_ch(l, ch));
    used += m;
  }
  return used + block_scan_points(m);
}

size_t scan_scratch_points(size_t n) {
  size_t r = lookback_scratch_points(n);
  for (int ch = CH_MIN; ch <= CH_MAX; ++ch) r = std::max(r, scan_plan_points(n, ch));
  return r;
}

// inclusive / exclusive scan of n points in place: chunk levels (2^ch points
// per lane) until a level has at most BSCAN_MAX points, a block scan of that
// level, then the chunk walks back down from each chunk's prefix; sums and
// totals live in scratch, no allocation, no sync.  level1_done: the first
// level's chunk sums are already in scratch slots [0, ceil(n / 2^ch)) (the
// batched launch computed them).
hipError_t scan_launch_ch(PtView v, size_t n, bool exclusive, PtView scratch, int ch, hipStream_t s,
                          bool level1_done) {
  if (n == 0) return hipSuccess;
  if (ch < CH_MIN || ch > CH_MAX) return hipErrorInvalidValue;
  PtView lv[48];
  size_t ln[48];
  int L = 0;
  lv[0] = v;
  ln[0] = n;
  size_t used = 0;
  while (ln[L] > BSCAN_MAX) {
    const size_t m = tiles(ln[L], scan_level_ch(L, ch));
    lv[L + 1] = PtView{scratch.w + used * 8, scratch.cap};
    ln[L + 1] = m;
    used += m;
    if (L > 0 || !level1_done) {
      hipError_t e = launch_chunk_sums(lv[L + 1], lv[L], ln[L], scan_level_ch(L, ch), s);
      if (e != hipSuccess) return e;
    }
    ++L;
  }
  hipError_t e = block_scan_q(lv[L], ln[L], L > 0 ? true : exclusive, PtView{scratch.w + used * 8, scratch.cap}, s);
  for (int l = L - 1; l >= 0 && e == hipSuccess; --l) {
    const bool ex = l > 0 ? true : exclusive;
    const int c = scan_level_ch(l, ch);
    e = ln[l + 1] <= QUAD_CHUNKS_MAX ? launch_chunk_scan_q(lv[l], ln[l], lv[l + 1], ex, c, s)
                                     : launch_chunk_scan(lv[l], ln[l], lv[l + 1], ex, c, s);
  }
  return e;
}

hipError_t scan_launch(PtView v, size_t n, bool exclusive, PtView scratch, hipStream_t s) {
  const AggTune t = tune_for(n);
  if (t.scanlb) return scan_lookback_launch(v, n, exclusive, scratch, s);
  return scan_launch_ch(v, n, exclusive, scratch, t.scan, s);
}

// ------------------------------------------- single-pass scan (look-back)
//
// Decoupled look-back with a point addition as the combine (the reference
// scans recursively, ftcrypto.cu:1474-1562).  One block of LB_TPB lanes per
// tile of LB_TILE points; blocks take tile numbers from a device counter in
// launch order, so every lower tile belongs to a block that is already
// running and can finish its first walk without waiting for anything.  A tile:
//   walk 1   each lane sums its 2^LB_CH points (whole-line groups);
//   in-tile  waves scan their 64 lane sums with shuffles (six adds), wave
//            totals meet in LDS, and wave 3 publishes the tile's aggregate A
//            (the point, then flag 1);
//   look-back  wave 0 reads the flags of the 64 preceding tiles (one per
//            lane), waits until each has at least its aggregate, and adds
//            the aggregates back to the nearest tile that published its
//            inclusive prefix (flag 2) or to the array's start -- a wave
//            reduction of six adds per window of 64 tiles -- then publishes
//            the tile's inclusive prefix (flag 2);
//   walk 2   every lane walks its points again from its exclusive prefix and
//            writes the outputs.
// A flag wait is bounded: a tile whose aggregate has not appeared after
// LB_POLLS polls is summed from its points by the waiting lane (it checks
// the flag again afterwards, see below), so every block finishes whatever
// the order the hardware runs them in.  The raw
// coordinates of the outputs depend on how far back each look-back went
// (point addition is associative and commutative, so the points do not).
// 32 points per lane (8,192 per tile) measured best on MI355X for 2^24
// points: 3.5 ms against 4.7 ms with 16 per lane (a chain of 2,048 tiles),
// 3.8 with 64 and 5.8 with 128 (fewer waves for the two walks)
constexpr int LB_CH = 5, LB_TPB = 256;
constexpr size_t LB_TILE = (size_t)LB_TPB << LB_CH;
constexpr int LB_POLLS = 1 << 22;

static size_t lb_tiles(size_t n) { return (n + LB_TILE - 1) / LB_TILE; }

// scratch (points): aggregates [0, T), inclusive prefixes [T, 2T), then the
// T flags and the tile counter as 32-bit words in plane 0
size_t lookback_scratch_points(size_t n) {
  const size_t T = lb_tiles(n);
  return 2 * T + (T + 1 + 7) / 8 + 1;
}

__device__ __forceinline__ void shfl_xor_p3(ge_p3 &r, const ge_p3 &a, int d) {
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    r.X.v[k] = (uint32_t)__shfl_xor((int)a.X.v[k], d, 64);
    r.Y.v[k] = (uint32_t)__shfl_xor((int)a.Y.v[k], d, 64);
    r.Z.v[k] = (uint32_t)__shfl_xor((int)a.Z.v[k], d, 64);
    r.T.v[k] = (uint32_t)__shfl_xor((int)a.T.v[k], d, 64);
  }
}

// acc += points [lo, hi) of v, read a line per coordinate at a time
__device__ __forceinline__ void chunk_sum(ge_p3 &acc, const PtView v, size_t lo, size_t hi) {
#pragma unroll 1
  for (size_t g = lo; g < hi; g += 4) {
    PW q0, q1, q2, q3;
    FT_LOAD4(q0, q1, q2, q3, v, g, hi);
    add_pw(acc, q0);
    if (g + 1 < hi) add_pw(acc, q1);
    if (g + 2 < hi) add_pw(acc, q2);
    if (g + 3 < hi) add_pw(acc, q3);
  }
}

__global__ void k_lb_init(unsigned *flag, size_t T) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i <= T) flag[i] = 0u;                      // flags, then the tile counter
}

template <bool EX>
__global__ void __launch_bounds__(LB_TPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_scan_lookback(PtView v, size_t n, PtView agg, PtView inc, unsigned *flag, unsigned *ctr) {
  // each lane's exclusive prefix within its wave, then the wave totals, the
  // tile's aggregate and its exclusive prefix (packed points), so that no
  // point stays in registers across the look-back
  __shared__ __align__(16) PW s_lane[LB_TPB];
  __shared__ __align__(16) PW s_pt[LB_TPB / 64 + 2];
  __shared__ unsigned s_tile;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  constexpr int NW = LB_TPB / 64;
  if (threadIdx.x == 0) s_tile = atomicAdd(ctr, 1u);
  __syncthreads();
  const size_t tile = s_tile;
  const size_t lo = tile * LB_TILE + ((size_t)threadIdx.x << LB_CH);
  const size_t hi = lo + ((size_t)1 << LB_CH) < n ? lo + ((size_t)1 << LB_CH) : n;
  {
    // walk 1 and the wave's inclusive scan of its lane sums
    ge_p3 I;
    ge_p3_0(I);
    if (lo < n) chunk_sum(I, v, lo, hi);
#pragma unroll 1
    for (int d = 1; d < 64; d <<= 1) {
      ge_p3 y;
      shfl_up_p3(y, I, d);
      if (lane >= d) ge_p3_add(I, I, y);
    }
    ge_p3 Ix;
    shfl_up_p3(Ix, I, 1);
    if (lane == 0) ge_p3_0(Ix);
    pk_pw(s_lane[threadIdx.x], Ix);
    if (lane == 63) pk_pw(s_pt[w], I);
  }
  __syncthreads();
  if (w == NW - 1) {
    // the tile's aggregate, published by one lane: point, then flag 1
    ge_p3 A;
    un_pw(A, s_pt[0]);
#pragma unroll 1
    for (int q = 1; q < NW; ++q) add_pw(A, s_pt[q]);
    if (lane == 63) {
      pk_pw(s_pt[NW], A);
      store_p3(agg, tile, A);
      __hip_atomic_store(flag + tile, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      // the flag is visible before this tile's second walk overwrites its
      // points (a waiting helper that saw flag 0 after reading them read inputs)
      __threadfence();
    }
  }
  ge_p3 E;
  ge_p3_0(E);
  if (w == 0 && tile > 0) {
    // the tile's exclusive prefix from the tiles before it, 64 at a time
    bool first = true;
    int64_t base = (int64_t)tile - 1;
#pragma unroll 1
    for (;;) {
      const int64_t pred = base - lane;
      unsigned f = 2;                            // before the array's start: nothing to add
      if (pred >= 0) {
        f = __hip_atomic_load(flag + pred, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll 1
        for (int it = 0; f == 0 && it < LB_POLLS; ++it)
          f = __hip_atomic_load(flag + pred, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      }
      // nearest tile (lowest lane) with an inclusive prefix, or the start
      const unsigned long long m = __ballot(f == 2);
      const int k0 = m ? __builtin_ctzll(m) : 64;
      ge_p3 x;
      ge_p3_0(x);
      if (pred >= 0 && lane <= k0) {
        if (lane == k0) load_p3(x, inc, (size_t)pred);
        else if (f == 1) load_p3(x, agg, (size_t)pred);
        else {
          // not published within the polls: sum the tile's points here; if
          // its aggregate appeared meanwhile, that one is used (the points
          // may already be outputs), else the points read were its inputs
          const size_t plo = (size_t)pred * LB_TILE;
          const size_t phi = plo + LB_TILE < n ? plo + LB_TILE : n;
          chunk_sum(x, v, plo, phi);
          if (__hip_atomic_load(flag + pred, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != 0)
            load_p3(x, agg, (size_t)pred);
        }
      }
#pragma unroll 1
      for (int d = 1; d < 64; d <<= 1) {
        ge_p3 y;
        shfl_xor_p3(y, x, d);
        ge_p3_add(x, x, y);
      }
      if (first) E = x;
      else ge_p3_add(E, E, x);
      first = false;
      if (k0 < 64) break;
      base -= 64;
    }
  }
  __syncthreads();
  if (w == 0) {
    // the tile's inclusive prefix E + A, published by one lane: point, then flag 2
    ge_p3 P;
    if (tile > 0) {
      P = E;
      add_pw(P, s_pt[NW]);
    } else {
      un_pw(P, s_pt[NW]);
    }
    if (lane == 0) {
      pk_pw(s_pt[NW + 1], E);
      store_p3(inc, tile, P);
      __hip_atomic_store(flag + tile, 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  if (lo >= n) return;
  // walk 2 from E + (the earlier waves' totals) + (the wave's lane sums before this lane)
  ge_p3 run;
  un_pw(run, s_pt[NW + 1]);
#pragma unroll 1
  for (int q = 0; q < w; ++q) add_pw(run, s_pt[q]);
  if (lane > 0) add_pw(run, s_lane[threadIdx.x]);
#pragma unroll 1
  for (size_t gi = lo; gi < hi; gi += 4) {
    // an opaque copy of the element index: the plane addresses are then
    // formed each group from it and the uniform plane bases, instead of
    // four per-lane plane pointers kept across the loop (which spill, and
    // whose reloads wait on the previous group's stores)
    size_t g = gi;
    asm volatile("" : "+v"(g));
    PW q0, q1, q2, q3;
    FT_LOAD4(q0, q1, q2, q3, v, g, hi);
    scan_step<EX>(run, q0);
    if (g + 1 < hi) scan_step<EX>(run, q1);
    if (g + 2 < hi) scan_step<EX>(run, q2);
    if (g + 3 < hi) scan_step<EX>(run, q3);
    st_pw4(v, g, (int)(hi - g < 4 ? hi - g : 4), q0, q1, q2, q3);
  }
}

// single-pass scan of n points in place; scratch of lookback_scratch_points(n)
hipError_t scan_lookback_launch(PtView v, size_t n, bool exclusive, PtView scratch, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t T = lb_tiles(n);
  if (T >= (1ull << 32) || scratch.cap < lookback_scratch_points(n)) return hipErrorInvalidValue;
  const PtView agg{scratch.w, scratch.cap}, inc{scratch.w + T * 8, scratch.cap};
  unsigned *flag = reinterpret_cast<unsigned *>(scratch.w + 2 * T * 8);
  hipLaunchKernelGGL(k_lb_init, dim3((unsigned)((T + 1 + 255) / 256)), dim3(256), 0, s, flag, T);
  if (exclusive)
    hipLaunchKernelGGL(k_scan_lookback<true>, dim3((unsigned)T), dim3(LB_TPB), 0, s, v, n, agg, inc, flag, flag + T);
  else
    hipLaunchKernelGGL(k_scan_lookback<false>, dim3((unsigned)T), dim3(LB_TPB), 0, s, v, n, agg, inc, flag, flag + T);
  return hipGetLastError();
}

// --------------------------------------------------------- segmented sum

// Segment boundaries: CSR offsets in device memory, or uniform segments
// s = [s len, min((s + 1) len, n)) computed on the fly.
struct OffArr {
  const int64_t *p;
  __device__ __forceinline__ int64_t operator[](size_t s) const { return p[s]; }
};
struct OffUni {
  int64_t len, n;
  __device__ __forceinline__ int64_t operator[](size_t s) const {
    const int64_t v = (int64_t)s * len;
    return v < n ? v : n;
  }
};

// first segment in [a, nseg) whose level-1 range ends after position p
template <class OffT>
__device__ __forceinline__ size_t seg_find1(const OffT off, size_t a, size_t nseg, int64_t p) {
  size_t b = nseg;
  while (a < b) {
    const size_t mid = (a + b) >> 1;
    if (off[mid + 1] <= p) a = mid + 1;
    else b = mid;
  }
  return a;
}

// Segment s's sum over this chunk goes to out[s] when the segment is complete
// in the chunk, else to the chunk's head record (it began before the chunk)
// or tail record (it continues after it).  One pack and one store sequence
// with a per-lane destination: lanes of a wave that flush to different
// places do not run three copies of the store.
__device__ __forceinline__ void seg_flush(PtView out, PtView rec, int64_t *keys, size_t s, int64_t lo,
                                          int64_t hi, int64_t clo, int64_t chi, size_t t, const ge_p3 &acc) {
  const bool inside = lo >= clo && hi <= chi;
  const size_t slot = 2 * t + (lo < clo ? 0 : 1);
  const PtView v{inside ? out.w : rec.w, inside ? out.cap : rec.cap};
  store_p3(v, inside ? s : slot, acc);
  if (!inside) keys[slot] = (int64_t)s;
}

// Flush queue of a wave walking whole chunks (LDS): a lane whose segment
// ends writes its unpacked sum (36 limbs) and destination to the next free
// entry; the wave packs and stores the queued sums together, one entry per
// lane, when half the queue is used (at a group's start) and at the end of
// the walk.  Lanes' segment ends fall at different elements, so a wave that
// packed and stored at each end would run the pack at most crossings of any
// lane; from the queue it runs once per drain.  A full queue sends further
// sums straight to memory.  dst: the output index, or the record slot with
// bit 63 set (key = segment id).
constexpr int SEGQ_N = 32;
struct SegQ {
  uint4 p[9][SEGQ_N];
  uint64_t dst[SEGQ_N];
  int64_t key[SEGQ_N];
};

__device__ __forceinline__ void segq_put(SegQ *q, int pos, const ge_p3 &a, uint64_t dst, int64_t key) {
  const uint32_t *w[4] = {a.X.v, a.Y.v, a.Z.v, a.T.v};
  uint32_t f[36];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int l = 0; l < 9; ++l) f[9 * c + l] = w[c][l];
#pragma unroll
  for (int j = 0; j < 9; ++j) q->p[j][pos] = make_uint4(f[4 * j], f[4 * j + 1], f[4 * j + 2], f[4 * j + 3]);
  q->dst[pos] = dst;
  q->key[pos] = key;
}

// entries [0, qn) to memory, entry l by lane l
__device__ __forceinline__ void segq_drain(SegQ *q, unsigned qn, PtView out, PtView rec, int64_t *keys, int lane) {
  __builtin_amdgcn_wave_barrier();
  if ((unsigned)lane < qn) {
    uint32_t f[36];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      const uint4 u = q->p[j][lane];
      f[4 * j] = u.x; f[4 * j + 1] = u.y; f[4 * j + 2] = u.z; f[4 * j + 3] = u.w;
    }
    ge_p3 a;
    uint32_t *w[4] = {a.X.v, a.Y.v, a.Z.v, a.T.v};
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int l = 0; l < 9; ++l) w[c][l] = f[9 * c + l];
    const uint64_t d = q->dst[lane];
    const bool r = d >> 63;
    const size_t i = (size_t)(d & ~(1ull << 63));
    const PtView v{r ? rec.w : out.w, r ? rec.cap : out.cap};
    store_p3(v, i, a);
    if (r) keys[i] = q->key[lane];
  }
  __builtin_amdgcn_wave_barrier();
}

// Level 1: lane t walks elements [2^CH t, 2^CH (t + 1)) of the input (through
// the gather indices when given), segment boundaries from the offsets.
// MODE 0: elements in order; 1: gathered from the planes of src; 2:
// gathered from the point-contiguous copy `aos`.  Two waves per SIMD (the
// gather forms would otherwise take a few registers over 256).
// TR (MODE 0, a wave of 64 whole chunks starting at chunk wave0): each
// group's points arrive by lane-transposed loads through the wave's LDS area tr.
template <int CH, int MODE, class OffT, bool TR = false>
__device__ __forceinline__ void seg_walk(PtView out, PtView rec, int64_t *keys, PtView src, size_t n,
                                         const OffT off, size_t nseg, const int32_t *gather, PtView aos,
                                         size_t aos_base, size_t t, uint4 *tr = nullptr, size_t wave0 = 0,
                                         SegQ *sq = nullptr) {
  const int64_t clo = (int64_t)(t << CH);
  if (clo >= (int64_t)n) return;
  const int64_t chi = clo + (1 << CH);
  const int64_t cend = chi < (int64_t)n ? chi : (int64_t)n;
  if (rec.w) {                                   // slots left unused keep key -1
    keys[2 * t] = -1;
    keys[2 * t + 1] = -1;
    // a queued record's key may be written by another lane of the wave
    if (TR) __threadfence_block();
  }
  const int lane = (int)(t - wave0);
  unsigned qn = 0;                               // queued entries (wave-uniform)
  size_t s = seg_find1(off, 0, nseg, clo);
  // segment s = [lo, bnd[0]); bnd[k] = end of segment s + k (INT64_MAX past
  // the last segment).  The ends are reloaded at the start of every group of
  // four elements, together with the group's points, so that no crossing
  // waits on a load: on this GPU stores count in the same in-order memory
  // counter as loads, and a boundary loaded at one crossing and read at the
  // next would wait for every flush store issued in between.  Four crossings
  // at most fit a group (one per non-empty segment), so five ends suffice;
  // runs of empty segments take the search path.
  constexpr int NB = 5;
  int64_t lo = s < nseg ? off[s] : 0;
  int64_t bnd[NB];
  auto refill = [&]() {
#pragma unroll
    for (int k = 0; k < NB; ++k) {
      const size_t j = s + 1 + (size_t)k;
      const int64_t e = off[j <= nseg ? j : nseg];
      bnd[k] = j <= nseg ? e : INT64_MAX;
    }
  };
  ge_p3 acc;
  ge_p3_0(acc);
  bool touched = false;
  // advance the segment cursor to element i; true when i lies in segment s
  // queue (TR) or store the sum of segment s, which ends before element i
  auto flush = [&](bool need) {
    if (!TR) {
      if (need) seg_flush(out, rec, keys, s, lo, bnd[0], clo, chi, t, acc);
      return;
    }
    const uint64_t m = __ballot(need);
    if (!m) return;
    if (need) {
      const unsigned pos =
          qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
      if (pos < (unsigned)SEGQ_N) {
        const bool inside = lo >= clo && bnd[0] <= chi;
        const size_t slot = 2 * t + (lo < clo ? 0 : 1);
        segq_put(sq, (int)pos, acc, inside ? (uint64_t)s : (slot | (1ull << 63)), (int64_t)s);
      } else {
        seg_flush(out, rec, keys, s, lo, bnd[0], clo, chi, t, acc);
      }
    }
    qn = min(qn + (unsigned)__popcll(m), (unsigned)SEGQ_N);
  };
  auto take = [&](int64_t i) -> bool {
    const bool cross = s < nseg && bnd[0] <= i;
    if (TR) flush(cross && touched);
    if (cross) {
      if (!TR) flush(touched);
      touched = false;
      ge_p3_0(acc);
      ++s;
      lo = bnd[0];
#pragma unroll
      for (int k = 0; k + 1 < NB; ++k) bnd[k] = bnd[k + 1];
      if (s < nseg && bnd[0] <= i) {             // empty segments: search
        s = seg_find1(off, s, nseg, i);
        if (s < nseg) lo = off[s];
        refill();
      }
    }
    return s < nseg && lo <= i;
  };
  // every element is added to the running sum, which a segment starts at
  // the identity: the same code on every lane whether or not its segment
  // just began
  auto accum = [&](const PW &q) {
    add_pw(acc, q);
    touched = true;
  };
  if (MODE != 0) {
    // four gathered points in flight per group: their indices (in one 16-byte
    // load when the group is whole and the index array 16-byte aligned; the
    // chunk start is a multiple of 16), then all their lines; from the
    // point-contiguous copy when there is one (one line per point)
    const bool vec = ((uintptr_t)gather & 15) == 0;
#pragma unroll 1
    for (int64_t g = clo; g < cend; g += 4) {
      refill();
      int4 ix;
      if (vec && g + 4 <= cend) {
        ix = *reinterpret_cast<const int4 *>(gather + g);
      } else {
        ix.x = gather[g];
        ix.y = g + 1 < cend ? gather[g + 1] : ix.x;
        ix.z = g + 2 < cend ? gather[g + 2] : ix.x;
        ix.w = g + 3 < cend ? gather[g + 3] : ix.x;
      }
      const size_t j0 = (size_t)ix.x;
      const size_t j1 = (size_t)(g + 1 < cend ? ix.y : ix.x);
      const size_t j2 = (size_t)(g + 2 < cend ? ix.z : ix.x);
      const size_t j3 = (size_t)(g + 3 < cend ? ix.w : ix.x);
      PW q0, q1, q2, q3;
      if (MODE == 2) {
        ld_pw_aos(q0, aos, aos_base, j0);
        ld_pw_aos(q1, aos, aos_base, j1);
        ld_pw_aos(q2, aos, aos_base, j2);
        ld_pw_aos(q3, aos, aos_base, j3);
      } else {
        ld_pw(q0, src, j0);
        ld_pw(q1, src, j1);
        ld_pw(q2, src, j2);
        ld_pw(q3, src, j3);
      }
      pin_pw(q0); pin_pw(q1); pin_pw(q2); pin_pw(q3);
      if (take(g)) accum(q0);
      if (g + 1 < cend && take(g + 1)) accum(q1);
      if (g + 2 < cend && take(g + 2)) accum(q2);
      if (g + 3 < cend && take(g + 3)) accum(q3);
    }
  } else {
#pragma unroll 1
    for (int64_t g = clo; g < cend; g += 4) {
      if (TR && qn >= (unsigned)SEGQ_N / 2) {
        segq_drain(sq, qn, out, rec, keys, lane);
        qn = 0;
      }
      refill();
      PW q0, q1, q2, q3;
      if (TR) {
        const uint8_t *base = reinterpret_cast<const uint8_t *>(src.w) + ((wave0 << CH) + (size_t)(g - clo)) * 32;
        const size_t plane = src.cap * 32;
        uint4 r[4][8];
#pragma unroll
        for (int c = 0; c < 4; ++c) ld_line_issue<CH>(r[c], base + c * plane, lane);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          ld_line_tr(r[c], tr, lane);
          line_to_pw(q0, q1, q2, q3, c, r[c]);
        }
      } else {
        FT_LOAD4(q0, q1, q2, q3, src, (size_t)g, (size_t)cend);
      }
      if (take(g)) accum(q0);
      if (g + 1 < cend && take(g + 1)) accum(q1);
      if (g + 2 < cend && take(g + 2)) accum(q2);
      if (g + 3 < cend && take(g + 3)) accum(q3);
    }
  }
  if (TR) {
    flush(touched);
    if (qn) segq_drain(sq, qn, out, rec, keys, lane);
  } else if (touched) {
    seg_flush(out, rec, keys, s, lo, bnd[0], clo, chi, t, acc);
  }
}

template <int CH, int MODE>
__global__ void __launch_bounds__(ATPB) __attribute__((amdgpu_waves_per_eu(2, 2))) k_seg_level1(PtView out, PtView rec, int64_t *keys, PtView src,
                                                     size_t n, const int64_t *off, size_t nseg,
                                                     const int32_t *gather, PtView aos, size_t aos_base) {
  const size_t t = blockIdx.x * (size_t)ATPB + threadIdx.x;
  if constexpr (MODE == 0) {
    __shared__ uint4 trs[ATPB / 64][64 * 8];
    __shared__ SegQ sqs[ATPB / 64];
    const size_t wave0 = t - (threadIdx.x & 63);
    if (((wave0 + 64) << CH) <= n)   // wave-uniform: all 64 chunks whole
      seg_walk<CH, 0, OffArr, true>(out, rec, keys, src, n, OffArr{off}, nseg, gather, aos, aos_base, t,
                                    trs[threadIdx.x >> 6], wave0, &sqs[threadIdx.x >> 6]);
    else
      seg_walk<CH, 0>(out, rec, keys, src, n, OffArr{off}, nseg, gather, aos, aos_base, t);
  } else {
    seg_walk<CH, MODE>(out, rec, keys, src, n, OffArr{off}, nseg, gather, aos, aos_base, t);
  }
}

// Block walk over the chunk records, on quads (a latency-bound sum: a
// few adds per quad and pass).  The records of level p are two per unit (a unit is 2^(p LR)
// level-1 chunks): a head record (the partial of the segment entering the
// unit from before it) and a tail record (the partial of the segment that
// starts in the unit and leaves it), key -1 when absent.  In record order
// (head then tail, unit by unit) the records of one segment are consecutive:
// the tail record of its first unit, then the head records of the units it
// covers.  Quad b walks the records of units [2^LR b, 2^LR (b + 1)) in that
// order with one coordinate per lane (a point add is three dependent field
// multiplies on the quad), sums each segment's records, and writes the sum
// of a segment that started in the block and ends in it to the output; a
// segment entering the block from before leaves the block's head record, one
// leaving it the block's tail record (level p + 1).  The last level has one
// block, and every sum goes to the output.  Work per pass shrinks by 2^LR.
template <int LR, class OffT>
__device__ __forceinline__ void seg_up_quad(PtView out, PtView rin, const int64_t *kin, size_t nu, PtView rout,
                                            int64_t *kout, const OffT off, int ch1, int ulog, size_t b, int c) {
  const size_t u0 = b << LR;
  if (u0 >= nu) return;
  const size_t u1 = u0 + ((size_t)1 << LR) < nu ? u0 + ((size_t)1 << LR) : nu;
  int64_t key = -1, last = 0;           // open segment and its last level-1 chunk
  bool before = false;                  // it entered the block from before
  fe acc, x;
  int64_t hk = -1, tk = -1;             // keys of the block's records
#pragma unroll 1
  for (size_t u = u0; u < u1; ++u) {
    const int64_t uend = (int64_t)((u + 1) << ulog);       // first chunk after unit u
#pragma unroll
    for (int slot = 0; slot < 2; ++slot) {
      const int64_t k = kin[2 * u + slot];
      if (k < 0) continue;
      load_fe(x, coord(rin, c, 2 * u + slot));
      if (k == key) {
        quad_add<true>(acc, x, c);
      } else {                          // a segment's first record in the block
        acc = x;
        quad_from_mem(acc, c);
        key = k;
        before = slot == 0 && u == u0;
        last = (off[k + 1] - 1) >> ch1;
      }
      if (last < uend) {                // the segment ends in unit u
        fe m = acc;
        quad_to_mem(m, c);
        if (before) {
          store_fe(coord_w(rout, c, 2 * b), m);
          hk = key;
        } else {
          store_fe(coord_w(out, c, (size_t)key), m);
        }
        key = -1;
      }
    }
  }
  if (key >= 0) {                       // the open segment leaves the block
    quad_to_mem(acc, c);
    store_fe(coord_w(rout, c, 2 * b + (before ? 0 : 1)), acc);
    if (before) hk = key;
    else tk = key;
  }
  if (c == 0 && kout) {
    kout[2 * b] = hk;
    kout[2 * b + 1] = tk;
  }
}

template <int LR>
__global__ void __launch_bounds__(ATPB) k_seg_up(PtView out, PtView rin, const int64_t *kin, size_t nu,
                                                 PtView rout, int64_t *kout, const int64_t *off, int ch1, int ulog) {
  const size_t g = blockIdx.x * (size_t)ATPB + threadIdx.x;
  seg_up_quad<LR>(out, rin, kin, nu, rout, kout, OffArr{off}, ch1, ulog, g >> 2, (int)(g & 3));
}

// point-contiguous copy of src into the scratch planes' tails (aos_slot)
__global__ void __launch_bounds__(ATPB) k_to_aos(PtView aos, size_t aos_base, PtView src, size_t n) {
  const size_t j = blockIdx.x * (size_t)ATPB + threadIdx.x;
  if (j >= n) return;
  PW q;
  ld_pw(q, src, j);
  uint32_t *d = aos_point(aos, aos_base, j);
  st_words(d, q.x);
  st_words(d + 8, q.y);
  st_words(d + 16, q.z);
  st_words(d + 24, q.t);
}

__global__ void __launch_bounds__(ATPB) k_seg_empty(PtView out, const int64_t *off, size_t nseg) {
  const size_t s = blockIdx.x * (size_t)ATPB + threadIdx.x;
  if (s < nseg && off[s] == off[s + 1]) store_identity(out, s);
}

static hipError_t launch_seg_level1(PtView out, PtView rec, int64_t *keys, PtView elems, size_t n,
                                    const int64_t *off, size_t nseg, const int32_t *gather, int ch, PtView aos,
                                    size_t aos_base, hipStream_t s) {
  const size_t T = tiles(n, ch);
  if (!T || !nseg) return hipSuccess;
  const dim3 grid((unsigned)((T + ATPB - 1) / ATPB));
  const int mode = !gather ? 0 : (aos.w ? 2 : 1);
#define FT_SEG1(C, M)                                                                                    \
  hipLaunchKernelGGL((k_seg_level1<C, M>), grid, dim3(ATPB), 0, s, out, rec, keys, elems, n, off, nseg, gather, \
                     aos, aos_base)
  switch (ch * 4 + mode) {
    case 16: FT_SEG1(4, 0); break;
    case 17: FT_SEG1(4, 1); break;
    case 18: FT_SEG1(4, 2); break;
    case 20: FT_SEG1(5, 0); break;
    case 21: FT_SEG1(5, 1); break;
    case 22: FT_SEG1(5, 2); break;
    case 24: FT_SEG1(6, 0); break;
    case 25: FT_SEG1(6, 1); break;
    case 26: FT_SEG1(6, 2); break;
    case 28: FT_SEG1(7, 0); break;
    case 29: FT_SEG1(7, 1); break;
    case 30: FT_SEG1(7, 2); break;
    default: return hipErrorInvalidValue;
  }
#undef FT_SEG1
  return hipGetLastError();
}

// the block walk's passes, levels alternating between (r0, k0) and (r1, k1)
static hipError_t launch_seg_up(PtView out, PtView r0, int64_t *k0, PtView r1, int64_t *k1, size_t T,
                                const int64_t *off, int ch1, int lr, hipStream_t s) {
  size_t nu = T;
  int ulog = 0;
  for (int p = 0; nu > 1; ++p) {
    const size_t nb = (nu + ((size_t)1 << lr) - 1) >> lr;
    const dim3 grid((unsigned)((4 * nb + ATPB - 1) / ATPB));
    PtView ri = (p & 1) ? r1 : r0, ro = (p & 1) ? r0 : r1;
    int64_t *ki = (p & 1) ? k1 : k0, *ko = nb > 1 ? ((p & 1) ? k0 : k1) : nullptr;
    switch (lr) {
      case 1: hipLaunchKernelGGL(k_seg_up<1>, grid, dim3(ATPB), 0, s, out, ri, ki, nu, ro, ko, off, ch1, ulog); break;
      case 2: hipLaunchKernelGGL(k_seg_up<2>, grid, dim3(ATPB), 0, s, out, ri, ki, nu, ro, ko, off, ch1, ulog); break;
      case 3: hipLaunchKernelGGL(k_seg_up<3>, grid, dim3(ATPB), 0, s, out, ri, ki, nu, ro, ko, off, ch1, ulog); break;
      case 4: hipLaunchKernelGGL(k_seg_up<4>, grid, dim3(ATPB), 0, s, out, ri, ki, nu, ro, ko, off, ch1, ulog); break;
      default: return hipErrorInvalidValue;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    nu = nb;
    ulog += lr;
  }
  return hipSuccess;
}

// Scratch of the segmented sum, as points of one array: the level-1 records
// (two per chunk, for the smallest chunks so that any schedule fits) in plane
// slots [0, R), their int64 segment keys in plane-0 slots [R, R + R/4); the
// block walk's levels alternate between those and a second area of R/2 + 2
// records from slot B = R + R/4 + 8 (keys in plane-0 slots after them).
static size_t segsum_area2(size_t n) {
  const size_t R = 2 * tiles(n, CH_MIN);
  return R + R / 4 + 8;
}
size_t segsum_scratch_points(size_t n) {
  const size_t R = 2 * tiles(n, CH_MIN);
  const size_t R2 = R / 2 + 2;
  return segsum_area2(n) + R2 + R2 / 4 + 8;
}

// the point-contiguous source copy starts at plane slot segsum_scratch_points(n)
// (after the records and, in plane 0, their keys); plane q holds points
// q, q + 4, ... at four slots each: 4 ceil(n_src / 4) slots per plane
size_t segsum_gather_scratch_points(size_t n_elems, size_t n_src) {
  return segsum_scratch_points(n_elems) + (n_src + 3) / 4 * 4;
}

hipError_t segsum_launch_ch(PtView out, const int64_t *d_off, size_t nseg, size_t n_elems,
                            const int32_t *gather, PtView src, size_t n_src, PtView scratch, int ch1, int lr,
                            hipStream_t s) {
  if (nseg == 0) return hipSuccess;
  if (ch1 < CH_MIN || ch1 > CH_MAX || lr < TREE_LR_MIN || lr > TREE_LR_MAX) return hipErrorInvalidValue;
  hipError_t e = launch_seg_empty(out, d_off, nseg, s);
  if (e != hipSuccess) return e;
  PtView aos{nullptr, 0};
  const size_t aos_base = segsum_scratch_points(n_elems);
  if (gather && n_src && scratch.cap >= segsum_gather_scratch_points(n_elems, n_src)) {
    aos = scratch;
    hipLaunchKernelGGL(k_to_aos, dim3((unsigned)((n_src + ATPB - 1) / ATPB)), dim3(ATPB), 0, s, aos, aos_base,
                       src, n_src);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const size_t T = tiles(n_elems, ch1);
  const PtView rec = T > 1 ? scratch : PtView{nullptr, 0};
  int64_t *keys = reinterpret_cast<int64_t *>(scratch.w + 2 * tiles(n_elems, CH_MIN) * 8);
  e = launch_seg_level1(out, rec, keys, src, n_elems, d_off, nseg, gather, ch1, aos, aos_base, s);
  if (e != hipSuccess || T <= 1) return e;
  const size_t a2 = segsum_area2(n_elems), R2 = tiles(n_elems, CH_MIN) + 2;
  const PtView rec2{scratch.w + a2 * 8, scratch.cap};
  int64_t *keys2 = reinterpret_cast<int64_t *>(scratch.w + (a2 + R2) * 8);
  return launch_seg_up(out, rec, keys, rec2, keys2, T, d_off, ch1, lr, s);
}

// Level 1's chunk for n elements: the class's tuned chunk, made smaller
// while the level would have fewer lanes than it had at the length the class
// was tuned at (up to 2^17 lanes, two waves on each of 1024 SIMDs).  A class
// spans a factor of 16 in length; without this, the top class's chunk timed
// at 2^24 elements left half the GPU idle at 2^22.
static int seg1_for(size_t n) {
  const AggTune t = tune_for(n);
  const size_t lanes_t = tune_class_size(tune_class(n)) >> t.seg1;
  const size_t want = lanes_t < ((size_t)1 << 17) ? lanes_t : ((size_t)1 << 17);
  int ch = t.seg1;
  while (ch > CH_MIN && (n >> ch) < want) --ch;
  return ch;
}

hipError_t segsum_launch(PtView out, const int64_t *d_off, size_t nseg, size_t n_elems, const int32_t *gather,
                         PtView src, size_t n_src, PtView scratch, hipStream_t s) {
  const AggTune t = tune_for(n_elems);
  return segsum_launch_ch(out, d_off, nseg, n_elems, gather, src, n_src, scratch, seg1_for(n_elems), t.segu, s);
}

hipError_t launch_seg_empty(PtView out, const int64_t *off, size_t nseg, hipStream_t s) {
  if (!nseg) return hipSuccess;
  hipLaunchKernelGGL(k_seg_empty, dim3((unsigned)((nseg + ATPB - 1) / ATPB)), dim3(ATPB), 0, s, out, off,
                     nseg);
  return hipGetLastError();
}

// ------------------------------------------------ batched aggregation
//
// Several reduce / segmented-sum / scan jobs in ONE persistent launch.  The
// element work of every job is cut into wave tiles: 64 level-1 chunks of a
// segmented sum (a scan's first level is a segmented sum with uniform
// segments of 2^ch), or 64 lanes x 2^r elements of a reduce, which each
// lane walks with a stride of 64 (coalesced) and leaves as one partial (r
// from 0 to RCH_MAX, so that a reduce has about REDUCE_TILES tiles; a reduce
// of at most 64 REDUCE_TILES elements has no tiles and is summed straight
// from its source afterwards).  The
// grid holds as many waves as the GPU keeps resident; each wave takes the
// next tile from an atomic counter until none is left, so short and long
// jobs, and tiles that flush more or fewer segments, balance over the whole
// GPU.  No wave waits for another.  The dependent work follows in ordinary
// launches: one run-tree pass per radix step over all jobs' records, the
// sum of each reduce's partials, and the upper scan levels with the second
// walk.

constexpr int RCH_MAX = 8;                 // reduce: at most 2^8 elements per lane and tile
constexpr size_t REDUCE_TILES = 4096;      // twice the waves the GPU keeps resident

// reduces this short are summed straight from their source after the launch
constexpr size_t REDUCE_DIRECT_MAX = 64 * REDUCE_TILES;

// log2 elements per lane of a reduce tile for n elements
static int reduce_rch(size_t n) {
  int r = 0;
  while (r < RCH_MAX && (n >> (6 + r)) > REDUCE_TILES) ++r;
  return r;
}

// One job as the kernels see it: 88 bytes, so that a table of AGG_BATCH_MAX
// jobs passes by value as a kernel argument (well under the 4 KiB limit) and
// a captured graph holds its own copy.  Tile, chunk, segment and partial
// counts of one GPU's arrays fit 32 bits (2^32 points would be 512 GiB).
struct BJob {
  PtView out, src;
  const int64_t *off;                      // CSR offsets; null: uniform segments of len
  int64_t len;                             // reduce: log2 elements per lane of its tiles, -1: no tiles
  uint64_t nseg, n;
  uint32_t tb;                             // first global tile
  uint32_t cb, cn;                         // chunk-kind jobs: first global chunk (multiple of 64), chunks
  uint32_t sb;                             // CSR jobs: first global segment (multiple of 64)
  uint32_t pb;                             // reduce: first partial slot
  int kind;                                // 0 chunk walk (segmented sum, scan level 1), 1 reduce
};
static_assert(sizeof(BJob) == 88, "job table layout");

struct BBatch {
  int J;
  unsigned tiles;
  BJob job[AGG_BATCH_MAX];
};
static_assert(sizeof(BBatch) + 48 <= 3072, "job table must stay well inside the kernel-argument limit");

// reduce tile: lane walks elements base, base + 64, ... (2^rch of them, below
// n) with the next point's words in flight during each add
__device__ __forceinline__ void reduce_tile(PtView part, size_t slot, PtView v, size_t n, size_t base, int rch) {
  ge_p3 acc;
  ge_cached q;
  PW nw;
  ge_p3_0(acc);
  const size_t end0 = base + ((size_t)64 << rch);
  const size_t end = end0 < n ? end0 : n;
  size_t i = base;
  if (i < n) { load_p3(acc, v, i); i += 64; }
  if (i < n && i < end) load_cached(q, v, i);
#pragma unroll 1
  for (; i < n && i < end; i += 64) {
    const size_t nx = (i + 64 < n && i + 64 < end) ? i + 64 : i;
    ld_pw(nw, v, nx);
    asm volatile("" ::: "memory");
    pin_cached(q);
    ge_p1p1 t;
    ge_add(t, acc, q);
    ge_p1p1_to_p3(acc, t);
    pin_p3(acc);
    pin_pw(nw);
    un_pw_cached(q, nw);
  }
  store_p3(part, slot, acc);
}

template <int CH>
__global__ void __launch_bounds__(ATPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_agg_batch(const BBatch B, PtView rec, int64_t *keys, PtView part, unsigned *ctr) {
  __shared__ uint4 trs[ATPB / 64][64 * 8];
  __shared__ SegQ sqs[ATPB / 64];
  uint4 *tr = trs[threadIdx.x >> 6];
  SegQ *sq = &sqs[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
#pragma unroll 1
  for (;;) {
    unsigned w = 0;
    if (lane == 0) w = atomicAdd(ctr, 1u);
    const unsigned tile = __builtin_amdgcn_readfirstlane(__shfl(w, 0));
    if (tile >= B.tiles) break;
    int j = 0;
#pragma unroll 1
    for (int q = 1; q < B.J; ++q)
      if (B.job[q].tb <= tile) j = q;
    const size_t lt = (size_t)(tile - B.job[j].tb);
    if (B.job[j].kind == 1) {
      const int rch = (int)B.job[j].len;
      reduce_tile(part, (size_t)B.job[j].pb + lt * 64 + lane, B.job[j].src, B.job[j].n, (lt << (6 + rch)) + lane,
                  rch);
    } else {
      const PtView r{rec.w + (size_t)B.job[j].cb * 2 * 8, rec.cap};
      int64_t *k = keys + 2 * (size_t)B.job[j].cb;
      const size_t c = lt * 64 + lane;
      const bool whole = ((lt * 64 + 64) << CH) <= B.job[j].n;   // wave-uniform
      if (B.job[j].off) {
        const OffArr o{B.job[j].off};
        if (whole)
          seg_walk<CH, 0, OffArr, true>(B.job[j].out, r, k, B.job[j].src, B.job[j].n, o, B.job[j].nseg, nullptr,
                                        PtView{nullptr, 0}, 0, c, tr, lt * 64, sq);
        else
          seg_walk<CH, 0>(B.job[j].out, r, k, B.job[j].src, B.job[j].n, o, B.job[j].nseg, nullptr,
                          PtView{nullptr, 0}, 0, c);
      } else {
        const OffUni o{B.job[j].len, (int64_t)B.job[j].n};
        if (whole)
          seg_walk<CH, 0, OffUni, true>(B.job[j].out, r, k, B.job[j].src, B.job[j].n, o, B.job[j].nseg, nullptr,
                                        PtView{nullptr, 0}, 0, c, tr, lt * 64, sq);
        else
          seg_walk<CH, 0>(B.job[j].out, r, k, B.job[j].src, B.job[j].n, o, B.job[j].nseg, nullptr,
                          PtView{nullptr, 0}, 0, c);
      }
    }
  }
}

// segments of a job with CSR offsets (uniform segments are never empty)
__device__ __forceinline__ size_t csr_segs(const BJob &b) { return b.off ? (size_t)b.nseg : 0; }

// job owning global index g of the chunk (seg = false) or segment ranges;
// ranges start at multiples of 64, so the answer is the same for a wave
__device__ __forceinline__ int batch_job_of(const BBatch &B, size_t g, bool seg) {
  int j = -1;
#pragma unroll 1
  for (int q = 0; q < B.J; ++q) {
    const size_t b = seg ? B.job[q].sb : B.job[q].cb;
    const size_t m = seg ? csr_segs(B.job[q]) : B.job[q].cn;
    if (m && b <= g) j = q;
  }
  return j;
}

// units of job j at level q of the block walk (level 0: its chunks; a job
// already down to one unit has no further level)
__device__ __forceinline__ size_t agg_units(const BJob &b, int q, int lr) {
  size_t nu = b.cn;
  for (int i = 0; i < q && nu > 1; ++i) nu = (nu + ((size_t)1 << lr) - 1) >> lr;
  return nu;
}
__device__ __forceinline__ size_t agg_blocks(const BJob &b, int q, int lr) {
  const size_t nu = agg_units(b, q, lr);
  return nu > 1 ? (nu + ((size_t)1 << lr) - 1) >> lr : 0;
}

// Pass p (1-based) of the block walk over every chunk-kind job of a batch:
// the jobs' blocks are numbered one job after another (job j from the sum of
// the blocks of the jobs before it, which also places its records of the
// pass's output level); level 0 is the persistent walk's records at the
// job's chunk base, odd levels live in the second area (r1, k1), even ones
// in the first.
template <int LR>
__global__ void __launch_bounds__(ATPB) k_agg_up(const BBatch B, PtView r0, int64_t *k0, PtView r1, int64_t *k1,
                                                 int ch1, int p) {
  const size_t g = (blockIdx.x * (size_t)ATPB + threadIdx.x) >> 2;
  const int c = (int)(threadIdx.x & 3);
  size_t base_out = 0, base_in = 0;
  int j = -1;
#pragma unroll 1
  for (int q = 0; q < B.J; ++q) {
    if (B.job[q].kind != 0) continue;
    const size_t nb = agg_blocks(B.job[q], p - 1, LR);
    if (g < base_out + nb) { j = q; break; }
    base_out += nb;
    base_in += agg_blocks(B.job[q], p - 2, LR);
  }
  if (j < 0) return;
  const BJob &b = B.job[j];
  const size_t nu = agg_units(b, p - 1, LR);
  PtView rin, rout;
  const int64_t *kin;
  int64_t *kout;
  if (p == 1) {
    rin = PtView{r0.w + (size_t)b.cb * 2 * 8, r0.cap};
    kin = k0 + 2 * (size_t)b.cb;
  } else {
    const PtView ra = (p - 1) & 1 ? r1 : r0;
    int64_t *ka = (p - 1) & 1 ? k1 : k0;
    rin = PtView{ra.w + base_in * 2 * 8, ra.cap};
    kin = ka + 2 * base_in;
  }
  {
    const PtView ra = p & 1 ? r1 : r0;
    int64_t *ka = p & 1 ? k1 : k0;
    rout = PtView{ra.w + base_out * 2 * 8, ra.cap};
    kout = ka + 2 * base_out;
  }
  const int ulog = (p - 1) * LR;
  if (b.off)
    seg_up_quad<LR>(b.out, rin, kin, nu, rout, kout, OffArr{b.off}, ch1, ulog, g - base_out, c);
  else
    seg_up_quad<LR>(b.out, rin, kin, nu, rout, kout, OffUni{b.len, (int64_t)b.n}, ch1, ulog, g - base_out, c);
}

__global__ void __launch_bounds__(ATPB) k_agg_empty(const BBatch B) {
  const size_t g = blockIdx.x * (size_t)ATPB + threadIdx.x;
  const int j = batch_job_of(B, g & ~(size_t)63, true);
  if (j < 0) return;
  const size_t s = g - B.job[j].sb;
  if (s < csr_segs(B.job[j]) && B.job[j].off[s] == B.job[j].off[s + 1]) store_identity(B.job[j].out, s);
}

__global__ void k_agg_ctr0(unsigned *ctr) { *ctr = 0u; }

namespace {

struct BatchPlan {
  BBatch B;
  size_t C = 0, Sg = 0, P = 0;             // global chunks, CSR segments, reduce partials
  size_t maxT = 0;                         // most chunks of one job
  int nred = 0;                            // reduce jobs (one block-sum region each)
  size_t rec_pts = 0, part_base = 0, red_base = 0, scan_base = 0, total = 0;
  size_t rec2_base = 0, rec2_n = 0;        // the block walk's second record area
  size_t scan_at[AGG_BATCH_MAX];           // scan jobs (by caller's index): their scratch region
  int scan_ch[AGG_BATCH_MAX];
  int from[AGG_BATCH_MAX];                 // caller's index of each laid-out job
  bool fits = true;                        // every count fits the table's 32-bit fields
};

size_t up64(size_t x) { return (x + 63) & ~(size_t)63; }

bool scan_kind(int k) { return k == AGG_SCAN || k == AGG_XSCAN; }

// lay out the jobs: reduces first (their tiles are the longest), then the
// chunk walks; ch1 is the level-1 chunk log2 (CH_MIN sizes the scratch for
// any schedule)
BatchPlan plan_batch(const AggJob *jobs, int J, int ch1) {
  BatchPlan p;
  std::memset(&p.B, 0, sizeof p.B);
  size_t nt = 0;
  int q = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (int i = 0; i < J; ++i) {
      const AggJob &a = jobs[i];
      const bool red = a.kind == AGG_REDUCE;
      if ((pass == 0) != red) continue;
      p.from[q] = i;
      BJob &b = p.B.job[q++];
      b.tb = (uint32_t)nt;
      if (red) {
        const int rch = reduce_rch(a.n);
        const size_t t = a.n <= REDUCE_DIRECT_MAX ? 0 : (a.n + ((size_t)64 << rch) - 1) >> (6 + rch);
        b.kind = 1;
        b.len = t ? rch : -1;
        ++p.nred;
        b.src = a.src;
        b.n = a.n;
        b.pb = (uint32_t)p.P;
        b.cb = (uint32_t)p.C;
        p.P += 64 * t;
        nt += t;
        continue;
      }
      const int sch = scan_kind(a.kind) ? tune_for(a.n).scan : 0;
      b.kind = 0;
      b.src = a.src;
      b.n = a.n;
      if (scan_kind(a.kind)) {
        b.src = a.out;
        b.len = (int64_t)1 << sch;
        b.nseg = a.n > BSCAN_MAX ? tiles(a.n, sch) : 0;
        if (!b.nseg) b.n = 0;                 // short scans have no chunk level
      } else {
        b.out = a.out;
        b.off = a.off;
        b.len = (int64_t)a.len;
        b.nseg = a.nseg;
      }
      const size_t T = (b.n && b.nseg) ? tiles(b.n, ch1) : 0;
      b.cb = (uint32_t)p.C;
      b.cn = (uint32_t)T;
      p.C += up64(T);
      p.maxT = std::max(p.maxT, T);
      b.sb = (uint32_t)p.Sg;
      p.Sg += up64(b.off ? b.nseg : 0);
      nt += up64(T) / 64;
    }
  }
  p.B.J = q;
  p.B.tiles = (unsigned)nt;
  p.fits = nt < (1ull << 32) && p.C < (1ull << 32) && p.Sg < (1ull << 32) && p.P < (1ull << 32);
  // scratch: records [0, 2C), keys (2C int64) and the tile counter in plane 0
  // after them, the block walk's second record area (level-1 blocks: at most
  // C + J records) with its keys in plane 0, reduce partials, one reduce's
  // block sums, scan regions
  p.rec_pts = 2 * p.C + (2 * p.C * 8 + 31) / 32 + 1;
  p.rec2_base = p.rec_pts;
  p.rec2_n = p.C + 2 * AGG_BATCH_MAX + 8;
  p.part_base = p.rec2_base + p.rec2_n + (p.rec2_n * 8 + 31) / 32 + 1;
  p.red_base = p.part_base + p.P;
  p.scan_base = p.red_base + REDUCE_BLOCKS_MAX * (size_t)p.nred;
  size_t at = p.scan_base;
  for (int i = 0; i < J; ++i) {
    p.scan_at[i] = at;
    p.scan_ch[i] = scan_kind(jobs[i].kind) ? tune_for(jobs[i].n).scan : 0;
    if (scan_kind(jobs[i].kind)) at += scan_scratch_points(jobs[i].n);
  }
  p.total = at;
  return p;
}

}  // namespace

// Fork of the dependent tails (scans' upper levels and second walks, reduces'
// block sums) onto library-owned streams (default priority, created on first
// use) and its join back into the caller's stream.  Held under a mutex from fork to join, so calls from several host
// threads enqueue their record / wait pairs in order.  While s is being
// captured into a graph, or when the caller asks for no fork (want false),
// next() returns s itself.
constexpr int AUX_STREAMS = 4;

class TailFork {
 public:
  explicit TailFork(hipStream_t s, bool want = true) : s_(s), lk_(mu(), std::defer_lock) {
    if (!want) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
    lk_.lock();
    if (!ready() || hipEventRecord(ev_fork(), s) != hipSuccess) {
      lk_.unlock();
      return;
    }
    on_ = true;
  }
  ~TailFork() { (void)join(); }
  // stream for the next tail
  hipStream_t next() {
    if (!on_) return s_;
    if (used_ == AUX_STREAMS) return aux()[rr_++ % AUX_STREAMS];   // all forked: round robin
    if (hipStreamWaitEvent(aux()[used_], ev_fork(), 0) != hipSuccess) return s_;
    return aux()[used_++];
  }
  hipError_t join() {
    if (!on_) return hipSuccess;
    hipError_t e = hipSuccess;
    for (int k = 0; k < used_; ++k) {
      hipError_t r = hipEventRecord(ev_join()[k], aux()[k]);
      if (r == hipSuccess) r = hipStreamWaitEvent(s_, ev_join()[k], 0);
      if (r != hipSuccess && e == hipSuccess) e = r;
    }
    on_ = false;
    lk_.unlock();
    return e;
  }

 private:
  static std::mutex &mu() { static std::mutex m; return m; }
  static hipStream_t *aux() { static hipStream_t a[AUX_STREAMS]; return a; }
  static hipEvent_t &ev_fork() { static hipEvent_t e; return e; }
  static hipEvent_t *ev_join() { static hipEvent_t e[AUX_STREAMS]; return e; }
  static bool ready() {            // under mu()
    static bool done = false;
    if (done) return true;
    for (int k = 0; k < AUX_STREAMS; ++k)
      if (hipStreamCreateWithFlags(&aux()[k], hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&ev_join()[k], hipEventDisableTiming) != hipSuccess)
        return false;
    if (hipEventCreateWithFlags(&ev_fork(), hipEventDisableTiming) != hipSuccess) return false;
    done = true;
    return true;
  }
  hipStream_t s_;
  std::unique_lock<std::mutex> lk_;
  bool on_ = false;
  int used_ = 0, rr_ = 0;
};

// resident blocks of k_agg_batch<CH> on the whole GPU (queried once)
template <int CH>
unsigned agg_resident_blocks() {
  static std::atomic<unsigned> cached{0};
  unsigned v = cached.load(std::memory_order_relaxed);
  if (v) return v;
  int nb = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_agg_batch<CH>, ATPB, 0) != hipSuccess || nb < 1) nb = 1;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
    cus = 1;
  v = (unsigned)(nb * cus);
  cached.store(v, std::memory_order_relaxed);
  return v;
}

template <int CH>
void launch_agg_batch(const BBatch &B, PtView rec, int64_t *keys, PtView part, unsigned *ctr, hipStream_t s) {
  const unsigned want = (B.tiles + ATPB / 64 - 1) / (ATPB / 64);
  const unsigned res = agg_resident_blocks<CH>();
  hipLaunchKernelGGL(k_agg_batch<CH>, dim3(want < res ? want : res), dim3(ATPB), 0, s, B, rec, keys, part, ctr);
}

size_t agg_batch_scratch_points(const AggJob *jobs, int J) {
  if (J <= 0 || J > AGG_BATCH_MAX) return 1;
  return plan_batch(jobs, J, CH_MIN).total;
}

hipError_t agg_batch_launch(const AggJob *jobs, int J, PtView scratch, hipStream_t s) {
  if (J <= 0) return hipSuccess;
  if (J > AGG_BATCH_MAX) return hipErrorInvalidValue;
  size_t walked = 0;
  for (int i = 0; i < J; ++i)
    if (jobs[i].kind != AGG_REDUCE) walked += jobs[i].n;
  const AggTune tu = tune_for(walked);
  const int ch1 = tu.seg1, lr = tu.segu;
  if (ch1 < CH_MIN || ch1 > CH_MAX || lr < TREE_LR_MIN || lr > TREE_LR_MAX) return hipErrorInvalidValue;
  BatchPlan p = plan_batch(jobs, J, ch1);
  if (!p.fits || scratch.cap < p.total) return hipErrorInvalidValue;
  auto at = [&](size_t slot) { return PtView{scratch.w + slot * 8, scratch.cap}; };
  for (int q = 0; q < p.B.J; ++q) {
    const int i = p.from[q];
    if (scan_kind(jobs[i].kind)) p.B.job[q].out = at(p.scan_at[i]);
  }
  const PtView rec = at(0);
  int64_t *keys = reinterpret_cast<int64_t *>(scratch.w + 2 * p.C * 8);
  unsigned *ctr = reinterpret_cast<unsigned *>(scratch.w + (p.rec_pts - 1) * 8);
  if (p.B.tiles) hipLaunchKernelGGL(k_agg_ctr0, dim3(1), dim3(1), 0, s, ctr);
  if (p.Sg) hipLaunchKernelGGL(k_agg_empty, dim3((unsigned)((p.Sg + ATPB - 1) / ATPB)), dim3(ATPB), 0, s, p.B);
  if (p.B.tiles) {
    switch (ch1) {
      case 4: launch_agg_batch<4>(p.B, rec, keys, at(p.part_base), ctr, s); break;
      case 5: launch_agg_batch<5>(p.B, rec, keys, at(p.part_base), ctr, s); break;
      case 6: launch_agg_batch<6>(p.B, rec, keys, at(p.part_base), ctr, s); break;
      default: launch_agg_batch<7>(p.B, rec, keys, at(p.part_base), ctr, s); break;
    }
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // block walk over the jobs' chunk records, a pass per level until every
  // job is down to one unit
  const PtView rec2 = at(p.rec2_base);
  int64_t *keys2 = reinterpret_cast<int64_t *>(scratch.w + (p.rec2_base + p.rec2_n) * 8);
  std::vector<size_t> nu(p.B.J);
  for (int q = 0; q < p.B.J; ++q) nu[q] = p.B.job[q].kind == 0 ? p.B.job[q].cn : 0;
  for (int pass = 1;; ++pass) {
    size_t nb = 0;
    for (int q = 0; q < p.B.J; ++q)
      if (nu[q] > 1) nb += (nu[q] + ((size_t)1 << lr) - 1) >> lr;
    if (!nb) break;
    const dim3 ugrid((unsigned)((4 * nb + ATPB - 1) / ATPB));
    switch (lr) {
      case 1: hipLaunchKernelGGL(k_agg_up<1>, ugrid, dim3(ATPB), 0, s, p.B, rec, keys, rec2, keys2, ch1, pass); break;
      case 2: hipLaunchKernelGGL(k_agg_up<2>, ugrid, dim3(ATPB), 0, s, p.B, rec, keys, rec2, keys2, ch1, pass); break;
      case 3: hipLaunchKernelGGL(k_agg_up<3>, ugrid, dim3(ATPB), 0, s, p.B, rec, keys, rec2, keys2, ch1, pass); break;
      default: hipLaunchKernelGGL(k_agg_up<4>, ugrid, dim3(ATPB), 0, s, p.B, rec, keys, rec2, keys2, ch1, pass); break;
    }
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int q = 0; q < p.B.J; ++q)
      if (nu[q] > 1) nu[q] = (nu[q] + ((size_t)1 << lr) - 1) >> lr;
  }
  // the scans' upper levels and second walks and the reduces' block sums go
  // to auxiliary streams forked from s here (eagerly; inside a graph capture
  // they stay on s), so that these tails, latency-bound at their small top
  // levels, fill each other's gaps; each has its own scratch region.  Only
  // reduces short enough to be summed straight from their source (two
  // launches of a few microseconds each) stay on s: there the fork's event
  // records and waits would cost more than the overlap gains.
  bool fork = false;
  for (int q = 0; q < p.B.J; ++q) {
    const BJob &b = p.B.job[q];
    if (b.kind == 1 ? b.len >= 0 : scan_kind(jobs[p.from[q]].kind)) fork = true;
  }
  TailFork fk(s, fork);
  for (int i = 0; i < J && e == hipSuccess; ++i) {
    if (!scan_kind(jobs[i].kind)) continue;
    e = scan_launch_ch(jobs[i].out, jobs[i].n, jobs[i].kind == AGG_XSCAN, at(p.scan_at[i]), p.scan_ch[i],
                       fk.next(), jobs[i].n > BSCAN_MAX);
  }
  if (e != hipSuccess) return e;
  int kred = 0;
  for (int q = 0; q < p.B.J; ++q) {
    const AggJob &a = jobs[p.from[q]];
    if (a.kind != AGG_REDUCE) continue;
    const hipStream_t rs = fk.next();
    const PtView red = at(p.red_base + REDUCE_BLOCKS_MAX * (size_t)kred++);
    const BJob &b = p.B.job[q];
    if (!a.n) {
      e = launch_set_identity(a.out, 0, 1, rs);
    } else {
      // from the tiles' partials, or straight from the source (no tiles)
      const bool tiled = b.len >= 0;
      const int rch = tiled ? (int)b.len : 0;
      const size_t np = tiled ? 64 * ((a.n + ((size_t)64 << rch) - 1) >> (6 + rch)) : a.n;
      const PtView from = tiled ? at(p.part_base + b.pb) : a.src;
      const unsigned g1 = reduce_grid(np, REDUCE_BLOCKS_MAX);
      e = launch_reduce_blocks(red, from, np, g1, rs);
      if (e == hipSuccess) e = launch_reduce_blocks(a.out, red, g1, 1, rs);
    }
    if (e != hipSuccess) return e;
  }
  return fk.join();
}

}  // namespace ft
