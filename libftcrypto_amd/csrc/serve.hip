// Serving loop: continuous batching of encrypt / add / multiply-by-u32 /
// segmented-sum requests over ciphertext vectors kept resident in HBM.
//
// * Memory: one device arena of the configured HBM budget; a request that
//   creates a vector is admitted only when the arena can hold its output
//   (first-fit with coalescing).  Admission control is that allocation.
// * Scheduling: a scheduler thread takes the oldest queued request and every
//   later queued request of the same kind that does not touch a vector an
//   earlier skipped request touches (per-vector order is kept), up to a batch
//   element cap, and launches them as ONE kernel whose lanes find their
//   request in a descriptor table; segmented sums (up to 16 requests) are
//   the jobs of one persistent aggregation launch whose waves take tiles of
//   any request from a device counter.  Batches go to the first idle of several
//   default-priority streams, so batches of different kinds overlap on the
//   device; a batch touching a vector still in flight on another stream waits
//   for that stream's event.
// * Slicing: an encryption or multiply longer than the batch cap runs as
//   slices of at most that many elements, each a batch of its own, at most
//   two in flight (the next slice's inputs are copied while the previous one
//   computes; its kernel waits for that slice's event), and slices of all
//   long requests together hold at most half of the streams (one of one).
//   Short requests queued behind a long one then run on the other streams
//   instead of waiting for the whole request.
// * Latency: submit time to the scheduler's observation of the batch's
//   completion event (steady clock).
#include <hip/hip_runtime.h>
#include <sys/prctl.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/ftcrypto_ext.h"
#include "arena.h"
#include "elem.h"
#include "kernels.h"

namespace ft {

// ------------------------------------------------------------ batched kernels

struct BItem {
  PtView a1, a2;        // destination (encrypt output, add dst, mul operand)
  PtView b1, b2;        // add source
  const uint64_t *m;    // encrypt messages
  const uint32_t *r;    // encrypt scalars, 8 words each
  const uint32_t *k;    // multipliers
  uint64_t start, n;    // lanes [start, start + n) of the launch
};

__device__ __forceinline__ int find_item(const BItem *it, int cnt, uint64_t t) {
  int lo = 0, hi = cnt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (it[mid].start <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int BS>
__device__ __forceinline__ void kb_encrypt_body(const BItem *items, int cnt, uint64_t total, const uint32_t *btab,
                                                const uint32_t *htab, int nd) {
  __shared__ __align__(16) uint32_t tb[COMB_TABLE_WORDS];
  __shared__ __align__(16) uint32_t th[COMB_TABLE_WORDS];
  lds_load_table(tb, btab);
  lds_load_table(th, htab);
  __syncthreads();
  const uint64_t t = blockIdx.x * (uint64_t)BS + threadIdx.x;
  if (t >= total) return;
  const BItem &it = items[find_item(items, cnt, t)];
  const uint64_t i = t - it.start;
  encrypt_one<false>(it.a1, it.a2, i, tb, th, it.m[i], nd, it.r + i * 8);
}

__global__ void __launch_bounds__(TPB) kb_encrypt(const BItem *items, int cnt, uint64_t total,
                                                  const uint32_t *btab, const uint32_t *htab, int nd) {
  kb_encrypt_body<TPB>(items, cnt, total, btab, htab, nd);
}

// two waves per 64 ciphertexts over the radix-32 tables (as k_encrypt_w32):
// the wave's half, hence its table, is uniform; each lane finds its item
template <int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(4, 4)))
kb_encrypt_w32(const BItem *items, int cnt, uint64_t total, const uint32_t *__restrict__ btab32,
               const uint32_t *__restrict__ htab32, int nd32) {
  static_assert(BS % 128 == 0, "whole wave pairs per block");
  const uint64_t wv = (uint64_t)blockIdx.x * (BS / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint32_t half = (uint32_t)(wv & 1);
  const uint64_t t = ((wv >> 1) << 6) + (threadIdx.x & 63);
  if (t >= total) return;
  const BItem &it = items[find_item(items, cnt, t)];
  const uint64_t i = t - it.start;
  encrypt_half32<false>(it.a1, it.a2, i, half, btab32, htab32, it.m[i], nd32, it.r + i * 8);
}

// large batches: 512-lane blocks at four waves per SIMD (as k_encrypt_wide)
__global__ void __launch_bounds__(ETPB) __attribute__((amdgpu_waves_per_eu(4, 4)))
kb_encrypt_wide(const BItem *items, int cnt, uint64_t total, const uint32_t *btab, const uint32_t *htab, int nd) {
  kb_encrypt_body<ETPB>(items, cnt, total, btab, htab, nd);
}

__global__ void __launch_bounds__(TPB) kb_add(const BItem *items, int cnt, uint64_t total) {
  const uint64_t t = blockIdx.x * (uint64_t)TPB + threadIdx.x;
  if (t >= total) return;
  const BItem &it = items[find_item(items, cnt, t)];
  const uint64_t i = t - it.start;
  add_one<false>(it.a1, i, it.b1, i);
  add_one<false>(it.a2, i, it.b2, i);
}

// lane 2 t + h multiplies half h of ciphertext t of the batch
__global__ void __launch_bounds__(TPB) kb_mul_u32_r16(const BItem *items, int cnt, uint64_t total) {
  const uint64_t t2 = blockIdx.x * (uint64_t)TPB + threadIdx.x;
  const uint64_t t = t2 >> 1;
  if (t >= total) return;
  const BItem &it = items[find_item(items, cnt, t)];
  const uint64_t i = t - it.start;
  mul_u32_point<4>((t2 & 1) ? it.a2 : it.a1, i, it.k[i]);
}

// radix 8 at two waves per SIMD, as k_mul_u32_r8
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(2, 2)))
kb_mul_u32_r8(const BItem *items, int cnt, uint64_t total) {
  const uint64_t t2 = blockIdx.x * (uint64_t)TPB + threadIdx.x;
  const uint64_t t = t2 >> 1;
  if (t >= total) return;
  const BItem &it = items[find_item(items, cnt, t)];
  const uint64_t i = t - it.start;
  mul_u32_point<3>((t2 & 1) ? it.a2 : it.a1, i, it.k[i]);
}

}  // namespace ft

// ------------------------------------------------------------ server state

namespace {

using namespace ft;
using Clock = std::chrono::steady_clock;

enum Kind { K_ENCRYPT = 0, K_ADD = 1, K_MUL = 2, K_SEGSUM = 3 };
// segmented sums per batch: two aggregation jobs each (the C1 and C2 halves)
constexpr size_t SEGSUM_BATCH_MAX = AGG_BATCH_MAX / 2;
struct Handle {
  size_t n = 0;             // ciphertexts
  bool allocated = false;
  size_t off = 0, bytes = 0;
  PtView c1{nullptr, 0}, c2{nullptr, 0};
  int inflight = 0;         // launched batches not yet complete
  int queued = 0;           // queued requests referencing it
  int reading = 0;          // ftx_server_read calls encoding it (writers wait)
  bool released = false;
  bool failed = false;      // its creating request failed
  int last_slot = -1;       // stream of the last batch that touched it
  uint64_t last_seq = 0;    // that batch's sequence number
};

struct Request {
  uint64_t id = 0;
  Kind kind = K_ENCRYPT;
  uint64_t dst = 0, src = 0;   // vectors: dst is written (created for encrypt/segsum)
  size_t n = 0, nseg = 0;
  int msg_bits = 0;
  std::vector<uint8_t> in;     // host inputs owned by the request: m|r, k, or offsets
  const uint8_t *bp[2] = {nullptr, nullptr};   // borrowed inputs (m, r) / (k) / (offsets)
  size_t bn[2] = {0, 0};
  Clock::time_point t_submit, t_launch, t_done;   // t_launch: its first batch
  int state = 0;               // 0 queued, 1 launched (wholly or in part), 2 done
  ft_error err = FT_ERR_NO_ERROR;
  size_t next = 0;             // sliced requests: elements launched so far
  int slices_inflight = 0;
};

// slices of one long request in flight at once
constexpr int SLICES_INFLIGHT = 2;

// elements [s0, s0 + len) of a request, launched in one batch; `last` when the
// request is complete with it; `slice` when it is part of a sliced request
struct Piece {
  std::shared_ptr<Request> r;
  size_t s0 = 0, len = 0;
  bool last = true;
  bool slice = false;
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t ev = nullptr;
  bool busy = false;
  bool slice = false;          // the batch in flight is a slice of a long request
  size_t scr_off = 0, scr_bytes = 0;   // arena scratch of the batch in flight
  uint64_t seq = 0;            // sequence number of the batch in flight / last done
  uint64_t done_seq = 0;
  char *dev = nullptr;         // device staging (inputs, descriptors, scratch)
  size_t dev_bytes = 0;
  char *host = nullptr;        // pinned staging
  size_t host_bytes = 0;
  std::vector<Piece> pieces;
  std::vector<uint64_t> touched;
};

}  // namespace

struct ftx_server_s {
  ftx_key key = nullptr;
  const uint32_t *btab = nullptr;
  const uint32_t *btab32 = nullptr;
  bool borrow = false;
  size_t max_batch = 0;
  Arena arena;
  void *arena_mem = nullptr;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
  std::unordered_map<uint64_t, Handle> handles;
  uint64_t next_req = 1, next_handle = 1;
  uint64_t n_batches = 0, n_done = 0;
  bool stop = false;
  bool draining = false;    // destroy called: queued requests memory cannot admit fail
  std::thread th;
  ft_error fatal = FT_ERR_NO_ERROR;
};

namespace {

// staging buffer growth (slot idle, so nothing of its is in flight)
bool ensure_staging(Slot &s, size_t dev_need, size_t host_need) {
  if (dev_need > s.dev_bytes) {
    if (s.dev) (void)hipFree(s.dev);
    s.dev_bytes = std::max(dev_need, (size_t)64 << 20);
    if (hipMalloc(&s.dev, s.dev_bytes) != hipSuccess) { s.dev = nullptr; s.dev_bytes = 0; return false; }
  }
  if (host_need > s.host_bytes) {
    if (s.host) (void)hipHostFree(s.host);
    s.host_bytes = std::max(host_need, (size_t)64 << 20);
    if (hipHostMalloc(&s.host, s.host_bytes, 0) != hipSuccess) { s.host = nullptr; s.host_bytes = 0; return false; }
  }
  return true;
}

// A request's inputs are one or two parts (m and r of an encryption; the
// multipliers; the offsets).  On the device the second part starts at the
// next multiple of 256 bytes after the first, so that 32-byte scalars are
// read by aligned 16-byte loads whatever n is.
size_t up256(size_t x) { return (x + 255) / 256 * 256; }
// host address of input part q: borrowed client memory, or the request's copy
// (parts back to back)
const uint8_t *in_ptr(const Request &r, int q) { return r.bp[0] ? r.bp[q] : r.in.data() + (q ? r.bn[0] : 0); }
// a slice of an encryption or multiply reads the same element range of each
// input part (8 + 32 or 4 bytes per element); a whole request reads all of it
size_t piece_bytes(const Piece &p, int q) {
  const Request &r = *p.r;
  return p.len == r.n ? r.bn[q] : r.bn[q] / r.n * p.len;
}
const uint8_t *piece_ptr(const Piece &p, int q) {
  const Request &r = *p.r;
  return in_ptr(r, q) + (p.len == r.n ? 0 : r.bn[q] / r.n * p.s0);
}
size_t piece_off(const Piece &p, int q) { return q ? up256(piece_bytes(p, 0)) : 0; }
size_t piece_in_bytes(const Piece &p) {
  return piece_bytes(p, 1) ? up256(piece_bytes(p, 0)) + piece_bytes(p, 1) : piece_bytes(p, 0);
}
// elements [s0, ...) of a point array: the same planes, shifted
PtView shifted(PtView v, size_t s0) { return PtView{v.w + 8 * s0, v.cap}; }

void maybe_free(ftx_server_s *S, uint64_t h) {
  auto it = S->handles.find(h);
  if (it == S->handles.end()) return;
  Handle &H = it->second;
  if (H.released && H.inflight == 0 && H.queued == 0) {
    if (H.allocated) S->arena.release(H.off, H.bytes);
    S->handles.erase(it);
  }
}

// point views of a vector of n ciphertexts at arena offset off
void place(ftx_server_s *S, Handle &H, size_t off) {
  H.off = off;
  H.bytes = H.n * 256;
  uint32_t *w = reinterpret_cast<uint32_t *>(S->arena.base + off);
  const size_t cap = H.n ? H.n : 1;
  H.c1 = PtView{w, cap};
  H.c2 = PtView{w + 32 * cap, cap};
  H.allocated = true;
}

void complete(ftx_server_s *S, Slot &sl) {
  const auto now = Clock::now();
  for (auto &p : sl.pieces) {
    p.r->slices_inflight--;
    if (!p.last) continue;
    p.r->state = 2;
    p.r->t_done = now;
    ++S->n_done;
  }
  for (uint64_t h : sl.touched) {
    auto it = S->handles.find(h);
    if (it != S->handles.end()) {
      it->second.inflight--;
      maybe_free(S, h);
    }
  }
  sl.pieces.clear();
  sl.touched.clear();
  if (sl.scr_bytes) S->arena.release(sl.scr_off, sl.scr_bytes);
  sl.scr_bytes = 0;
  sl.busy = false;
  sl.slice = false;
  sl.done_seq = sl.seq;
}

void fail_request(ftx_server_s *S, const std::shared_ptr<Request> &r, ft_error e) {
  r->err = e;
  if (r->kind == K_ENCRYPT || r->kind == K_SEGSUM) S->handles[r->dst].failed = true;
  if (r->state == 0) r->t_launch = Clock::now();
  r->state = 2;
  r->t_done = Clock::now();
  ++S->n_done;
  for (uint64_t h : {r->dst, r->src}) {
    auto it = S->handles.find(h);
    if (h && it != S->handles.end()) { it->second.queued--; maybe_free(S, h); }
  }
}

// a batch that could not launch: a request with slices still to come is
// still queued and leaves the queue with it
void fail_piece(ftx_server_s *S, const Piece &p, ft_error e) {
  if (!p.last) {
    auto it = std::find(S->queue.begin(), S->queue.end(), p.r);
    if (it != S->queue.end()) S->queue.erase(it);
  }
  fail_request(S, p.r, e);
}

// vector `h` still being written by an in-flight batch on another stream?
void wait_for(ftx_server_s *S, Slot &sl, int slot_idx, const Handle &H) {
  if (H.inflight > 0 && H.last_slot >= 0 && H.last_slot != slot_idx) {
    Slot &o = S->slots[H.last_slot];
    if (o.busy && o.seq == H.last_seq) (void)hipStreamWaitEvent(sl.st, o.ev, 0);
  }
}

// 1: ready, 0: waits (for a vector still being created, or for a read of the
// vector it writes), -1: a vector it needs failed
int request_ready(ftx_server_s *S, const Request &r) {
  const bool dst_in = (r.kind == K_ADD || r.kind == K_MUL);     // dst must already exist
  const bool src_in = (r.kind == K_ADD || r.kind == K_SEGSUM);
  for (int q = 0; q < 2; ++q) {
    const bool need = q == 0 ? dst_in : src_in;
    if (!need) continue;
    auto it = S->handles.find(q == 0 ? r.dst : r.src);
    if (it == S->handles.end() || it->second.failed) return -1;
    if (!it->second.allocated) return 0;
    if (q == 0 && it->second.reading > 0) return 0;
  }
  return 1;
}

// the aggregation jobs of a batch of segmented sums (offsets at dins[j] on
// the device; dins null: for sizing the scratch only, with the host offsets
// standing in, since the plan counts only the segments of jobs with offsets)
std::vector<AggJob> segsum_jobs(ftx_server_s *S, const std::vector<Piece> &batch, char *const *dins) {
  std::vector<AggJob> jobs;
  for (size_t j = 0; j < batch.size(); ++j) {
    const Request &r = *batch[j].r;
    const Handle &Sx = S->handles[r.src];
    const Handle &D = S->handles[r.dst];
    const size_t n_el = (size_t)reinterpret_cast<const int64_t *>(in_ptr(r, 0))[r.nseg];
    for (int h = 0; h < 2; ++h) {
      AggJob a{};
      a.kind = AGG_SEGSUM;
      a.out = h ? D.c2 : D.c1;
      a.src = h ? Sx.c2 : Sx.c1;
      a.off = dins ? reinterpret_cast<const int64_t *>(dins[j]) : reinterpret_cast<const int64_t *>(in_ptr(r, 0));
      a.nseg = r.nseg;
      a.n = n_el;
      jobs.push_back(a);
    }
  }
  return jobs;
}

// launch one batch of same-kind requests (or one slice of a long request) on
// slot idx (lock held)
void launch(ftx_server_s *S, int idx, std::vector<Piece> &batch) {
  Slot &sl = S->slots[idx];
  const Kind kind = batch[0].r->kind;
  // staging sizes
  size_t in_total = 0;
  for (auto &p : batch) in_total += up256(piece_in_bytes(p));
  const size_t desc = (batch.size() * sizeof(BItem) + 255) / 256 * 256;
  // segmented sums: scratch from the arena while the batch runs (no device
  // allocation, which could wait for the work in flight), else from the
  // slot's staging after the descriptors
  size_t scr_pts = 0, scr_off = 0;
  bool scr_arena = false;
  if (kind == K_SEGSUM) {
    const std::vector<AggJob> sizing = segsum_jobs(S, batch, nullptr);
    scr_pts = agg_batch_scratch_points(sizing.data(), (int)sizing.size());
    scr_arena = S->arena.alloc(scr_pts * 128, &scr_off);
  }
  if (!ensure_staging(sl, in_total + desc + (scr_arena ? 0 : scr_pts * 128), in_total + desc)) {
    if (scr_arena) S->arena.release(scr_off, scr_pts * 128);
    for (auto &p : batch) fail_piece(S, p, FT_ERR_OUT_OF_MEMORY);
    return;
  }
  // host staging: inputs then descriptors
  std::vector<BItem> items(batch.size());
  std::vector<char *> dins(batch.size());
  size_t off = 0;
  uint64_t start = 0;
  int nd = 1;
  for (size_t j = 0; j < batch.size(); ++j) {
    const Piece &p = batch[j];
    Request &r = *p.r;
    Handle &D = S->handles[r.dst];
    BItem &it = items[j];
    std::memset(&it, 0, sizeof it);
    it.a1 = shifted(D.c1, p.s0);
    it.a2 = shifted(D.c2, p.s0);
    if (kind == K_ADD) {
      Handle &Sx = S->handles[r.src];
      it.b1 = shifted(Sx.c1, p.s0);
      it.b2 = shifted(Sx.c2, p.s0);
    }
    char *din = sl.dev + off;
    dins[j] = din;
    if (kind == K_ENCRYPT) {
      it.m = reinterpret_cast<const uint64_t *>(din);
      it.r = reinterpret_cast<const uint32_t *>(din + piece_off(p, 1));
      // radix-16 message digits of the public bound (0 or 64: any uint64)
      const int d = (r.msg_bits <= 0 || r.msg_bits >= 64) ? 17 : (r.msg_bits + 3) / 4 + 1;
      nd = std::max(nd, d);
    } else if (kind == K_MUL) {
      it.k = reinterpret_cast<const uint32_t *>(din);
    }
    off += up256(piece_in_bytes(p));
    it.start = start;
    it.n = p.len;
    start += p.len;
  }
  // inputs: one DMA per request (borrowed, ideally pinned, client memory) or
  // a copy through the slot's pinned staging and one DMA for the batch
  std::memcpy(sl.host + off, items.data(), items.size() * sizeof(BItem));
  const BItem *ditems = reinterpret_cast<const BItem *>(sl.dev + off);
  size_t o2 = 0;
  for (auto &p : batch) {
    for (int q = 0; q < 2; ++q) {
      const size_t nb = piece_bytes(p, q);
      if (!nb) continue;
      if (S->borrow)
        (void)hipMemcpyAsync(sl.dev + o2 + piece_off(p, q), piece_ptr(p, q), nb, hipMemcpyHostToDevice, sl.st);
      else
        std::memcpy(sl.host + o2 + piece_off(p, q), piece_ptr(p, q), nb);
    }
    o2 += up256(piece_in_bytes(p));
  }
  // ordering against other streams (and a slice after the previous slice of
  // its request: that slice is the last batch that touched its vectors)
  for (auto &p : batch) {
    const Request *r = p.r.get();
    for (uint64_t h : {r->dst, r->src}) {
      if (!h) continue;
      auto it = S->handles.find(h);
      if (it != S->handles.end()) wait_for(S, sl, idx, it->second);
    }
  }
  hipError_t e = S->borrow
      ? hipMemcpyAsync(sl.dev + off, sl.host + off, items.size() * sizeof(BItem), hipMemcpyHostToDevice, sl.st)
      : hipMemcpyAsync(sl.dev, sl.host, off + items.size() * sizeof(BItem), hipMemcpyHostToDevice, sl.st);
  const unsigned grid = (unsigned)((start + TPB - 1) / TPB);
  if (e == hipSuccess && start > 0) {
    if (kind == K_ENCRYPT) {
      // two lanes per ciphertext for a two- or four-lane schedule (the
      // batched kernels have no four-lane form)
      const bool split = tune_for(start).encl >= 2;
      if (!split && start >= ENC_WIDE_MIN)
        hipLaunchKernelGGL(kb_encrypt_wide, dim3((unsigned)((start + ETPB - 1) / ETPB)), dim3(ETPB), 0, sl.st,
                           ditems, (int)batch.size(), start, S->btab, key_table(S->key), nd);
      else if (!split)
        hipLaunchKernelGGL(kb_encrypt, dim3(grid), dim3(TPB), 0, sl.st, ditems, (int)batch.size(), start, S->btab,
                           key_table(S->key), nd);
      else if (2 * start >= ENC_SPLIT_WIDE_MIN)
        hipLaunchKernelGGL(kb_encrypt_w32<ETPB>, dim3((unsigned)(((start + 63) / 64 * 128 + ETPB - 1) / ETPB)),
                           dim3(ETPB), 0, sl.st, ditems, (int)batch.size(), start, S->btab32, key_table32(S->key),
                           msg_digits32(nd));
      else
        hipLaunchKernelGGL(kb_encrypt_w32<TPB>, dim3((unsigned)(((start + 63) / 64 * 128 + TPB - 1) / TPB)),
                           dim3(TPB), 0, sl.st, ditems, (int)batch.size(), start, S->btab32, key_table32(S->key),
                           msg_digits32(nd));
    } else if (kind == K_ADD) {
      hipLaunchKernelGGL(kb_add, dim3(grid), dim3(TPB), 0, sl.st, ditems, (int)batch.size(), start);
    } else if (kind == K_MUL) {
      const dim3 grid2((unsigned)((2 * start + TPB - 1) / TPB));
      if (tune_for(start).u32w == 3)
        hipLaunchKernelGGL(kb_mul_u32_r8, grid2, dim3(TPB), 0, sl.st, ditems, (int)batch.size(), start);
      else
        hipLaunchKernelGGL(kb_mul_u32_r16, grid2, dim3(TPB), 0, sl.st, ditems, (int)batch.size(), start);
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess && kind == K_SEGSUM) {
    // every request of the batch is two jobs (C1 and C2 halves) of one
    // persistent aggregation launch; the batch's requests touch distinct
    // vectors, so no job writes what another reads
    std::vector<AggJob> jobs = segsum_jobs(S, batch, dins.data());
    char *scr = scr_arena ? S->arena.base + scr_off : sl.dev + off + desc;
    e = agg_batch_launch(jobs.data(), (int)jobs.size(), PtView{reinterpret_cast<uint32_t *>(scr), scr_pts}, sl.st);
  }
  if (e == hipSuccess) e = hipEventRecord(sl.ev, sl.st);
  if (e != hipSuccess) {
    if (scr_arena) S->arena.release(scr_off, scr_pts * 128);
    S->fatal = FT_ERR_INTERNAL_LIBRARY_ERROR;
    for (auto &p : batch) fail_piece(S, p, FT_ERR_INTERNAL_LIBRARY_ERROR);
    return;
  }
  sl.busy = true;
  sl.slice = batch[0].slice;
  sl.scr_off = scr_off;
  sl.scr_bytes = scr_arena ? scr_pts * 128 : 0;
  sl.seq++;
  ++S->n_batches;
  const auto now = Clock::now();
  for (auto &p : batch) {
    Request &r = *p.r;
    if (r.state == 0) r.t_launch = now;
    r.state = 1;
    r.slices_inflight++;
    for (uint64_t h : {r.dst, r.src}) {
      if (!h) continue;
      Handle &H = S->handles[h];
      if (p.last) H.queued--;        // a request with slices to come stays queued
      H.inflight++;
      H.last_slot = idx;
      H.last_seq = sl.seq;
      sl.touched.push_back(h);
    }
  }
  sl.pieces = batch;
}

// completion polling interval of the scheduler while batches are in flight
constexpr std::chrono::microseconds POLL{20};

void scheduler(ftx_server_s *S) {
  std::unique_lock<std::mutex> lk(S->mu);
  // the waits below are short: let them end on time (1 us timer slack)
  (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
  uint64_t notified = ~0ull;
  while (!S->stop) {
    // completions
    bool any_busy = false, completed = false;
    for (auto &sl : S->slots) {
      if (!sl.busy) continue;
      if (hipEventQuery(sl.ev) == hipSuccess) {
        complete(S, sl);
        completed = true;
      } else {
        any_busy = true;
      }
    }
    if (completed || S->n_done != notified) {
      notified = S->n_done;
      S->cv_done.notify_all();
    }
    int idle = -1, slices_busy = 0;
    for (size_t i = 0; i < S->slots.size(); ++i) {
      if (!S->slots[i].busy && idle < 0) idle = (int)i;
      slices_busy += S->slots[i].busy && S->slots[i].slice;
    }
    // slices of long requests take at most half of the streams (one of one),
    // so that a long request at the head of the queue does not take every
    // stream that frees up while short requests wait behind it
    const int slice_slots = std::max(1, (int)S->slots.size() / 2);
    // nothing to launch: wait for a submission (which notifies) or, while
    // batches are in flight, poll their events again after POLL; the lock is
    // free meanwhile, so clients submitting are not held off by the polling
    if (S->queue.empty() || idle < 0) {
      S->cv_work.wait_for(lk, any_busy || idle < 0 ? POLL : std::chrono::microseconds(200));
      continue;
    }
    // form a batch: the oldest ready request and later ready ones of its kind;
    // a request touching a vector of a skipped request, or of one already in
    // the batch, stays queued (per-vector order, no intra-batch aliasing).  A
    // long encryption or multiply contributes its next slice, alone.
    std::vector<Piece> batch;
    std::unordered_set<uint64_t> blocked, in_batch;
    size_t elems = 0;
    Kind kind = K_ENCRYPT;
    for (auto it = S->queue.begin(); it != S->queue.end();) {
      Request &r = **it;
      const bool sliced = (r.kind == K_ENCRYPT || r.kind == K_MUL) && r.n > S->max_batch;
      const bool touches_blocked = (r.dst && (blocked.count(r.dst) || in_batch.count(r.dst))) ||
                                   (r.src && (blocked.count(r.src) || in_batch.count(r.src)));
      int ready = touches_blocked ? 0 : request_ready(S, r);
      if (ready < 0) {
        auto rr = *it;
        it = S->queue.erase(it);
        fail_request(S, rr, FT_ERR_OUT_OF_BOUNDS);
        continue;
      }
      if (sliced && (r.slices_inflight >= SLICES_INFLIGHT || slices_busy >= slice_slots)) ready = 0;
      bool take = false;
      if (ready > 0) {
        if (batch.empty()) {
          take = true;
        } else if (!sliced && r.kind == kind && elems + r.n <= S->max_batch &&
                   (r.kind != K_SEGSUM || batch.size() < SEGSUM_BATCH_MAX)) {
          take = true;
        }
      }
      if (take && (r.kind == K_ENCRYPT || r.kind == K_SEGSUM)) {
        Handle &D = S->handles[r.dst];
        if (!D.allocated) {
          size_t off;
          if (!S->arena.alloc(D.n * 256, &off)) {
            take = false;                                 // admission: wait for memory
            bool any_inflight = false;
            for (auto &sl2 : S->slots) any_inflight |= sl2.busy;
            // never fits; or the server is being destroyed and nothing in
            // flight can free memory (the client releases nothing any more)
            if (batch.empty() && !any_inflight && (D.n * 256 > S->arena.size || S->draining)) {
              auto rr = *it;
              it = S->queue.erase(it);
              fail_request(S, rr, FT_ERR_OUT_OF_MEMORY);
              continue;
            }
          } else {
            place(S, D, off);
          }
        }
      }
      if (take) {
        if (batch.empty()) kind = r.kind;
        Piece p{*it, 0, r.n, true, sliced};
        if (sliced) {
          p.s0 = r.next;
          p.len = std::min(S->max_batch, r.n - r.next);
          p.last = p.s0 + p.len == r.n;
          r.next += p.len;
        }
        elems += p.len;
        if (r.dst) in_batch.insert(r.dst);
        if (r.src) in_batch.insert(r.src);
        batch.push_back(p);
        if (p.last) it = S->queue.erase(it);
        if (sliced || (kind == K_SEGSUM && batch.size() == SEGSUM_BATCH_MAX)) break;
      } else {
        if (r.dst) blocked.insert(r.dst);
        if (r.src) blocked.insert(r.src);
        ++it;
      }
    }
    if (batch.empty()) {
      S->cv_work.wait_for(lk, POLL);
      continue;
    }
    launch(S, idle, batch);
  }
}

ft_error submit(ftx_server_s *S, std::shared_ptr<Request> r, uint64_t *req) {
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->fatal) return S->fatal;
  r->id = S->next_req++;
  r->t_submit = Clock::now();
  for (uint64_t h : {r->dst, r->src})
    if (h) S->handles[h].queued++;
  S->reqs[r->id] = r;
  S->queue.push_back(r);
  *req = r->id;
  S->cv_work.notify_one();
  return FT_ERR_NO_ERROR;
}

}  // namespace

extern "C" {

ft_error ftx_server_create(ftx_server *out, ftx_key key, size_t hbm_budget_bytes, int n_streams, size_t max_batch,
                           int flags) {
  if (!key || n_streams < 1 || n_streams > 16 || hbm_budget_bytes < 4096) return FT_ERR_BAD_SIZE;
  auto *S = new (std::nothrow) ftx_server_s;
  if (!S) return FT_ERR_OUT_OF_MEMORY;
  S->key = key;
  S->borrow = (flags & FTX_SERVER_BORROW_INPUTS) != 0;
  S->btab = base_table();
  S->btab32 = base_table32();
  S->max_batch = max_batch ? max_batch : FTX_SERVER_DEFAULT_BATCH;
  if (hipMalloc(&S->arena_mem, hbm_budget_bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete S;
    return FT_ERR_OUT_OF_MEMORY;
  }
  S->arena.init(static_cast<char *>(S->arena_mem), hbm_budget_bytes);
  S->slots.resize(n_streams);
  for (auto &sl : S->slots) {
    if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
      ftx_server_destroy(S);
      return FT_ERR_INTERNAL_LIBRARY_ERROR;
    }
    // staging of the default size now, not inside the first batches' launches
    // (a pinned allocation takes milliseconds and the scheduler would hold
    // every queued request meanwhile)
    if (!ensure_staging(sl, 1, 1)) {
      ftx_server_destroy(S);
      return FT_ERR_OUT_OF_MEMORY;
    }
  }
  S->th = std::thread(scheduler, S);
  *out = S;
  return FT_ERR_NO_ERROR;
}

void ftx_server_destroy(ftx_server S) {
  if (!S) return;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    // drain: wait for queued and in-flight work; queued requests that memory
    // can no longer admit fail with FT_ERR_OUT_OF_MEMORY instead of waiting
    S->draining = true;
    S->cv_work.notify_all();
    S->cv_done.wait(lk, [&] {
      bool busy = false;
      for (auto &sl : S->slots) busy |= sl.busy;
      return S->queue.empty() && !busy;
    });
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->th.joinable()) S->th.join();
  for (auto &sl : S->slots) {
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.st) (void)hipStreamDestroy(sl.st);
    if (sl.dev) (void)hipFree(sl.dev);
    if (sl.host) (void)hipHostFree(sl.host);
  }
  if (S->arena_mem) (void)hipFree(S->arena_mem);
  delete S;
}

ft_error ftx_server_encrypt(ftx_server S, const uint64_t *m, const uint8_t *r, size_t n, int msg_bits,
                            uint64_t *req, uint64_t *handle) {
  if (msg_bits < 0 || msg_bits > 64) return FT_ERR_BAD_SIZE;
  if (msg_bits > 0 && msg_bits < 64)          // a public bound, checked before queueing (as ftx_encrypt)
    for (size_t i = 0; i < n; ++i)
      if (m[i] >> msg_bits) return FT_ERR_OUT_OF_BOUNDS;
  auto q = std::make_shared<Request>();
  q->kind = K_ENCRYPT;
  q->n = n;
  q->msg_bits = msg_bits ? msg_bits : 64;
  q->bn[0] = n * 8;
  q->bn[1] = n * 32;
  if (S->borrow) {
    q->bp[0] = reinterpret_cast<const uint8_t *>(m);
    q->bp[1] = r;
  } else {
    q->in.resize(n * 8 + n * 32);
    std::memcpy(q->in.data(), m, n * 8);
    std::memcpy(q->in.data() + n * 8, r, n * 32);
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    q->dst = S->next_handle++;
    S->handles[q->dst].n = n;
  }
  *handle = q->dst;
  return submit(S, q, req);
}

ft_error ftx_server_add(ftx_server S, uint64_t dst, uint64_t src, uint64_t *req) {
  auto q = std::make_shared<Request>();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto a = S->handles.find(dst), b = S->handles.find(src);
    if (a == S->handles.end() || b == S->handles.end() || a->second.released || b->second.released)
      return FT_ERR_OUT_OF_BOUNDS;
    if (a->second.n != b->second.n) return FT_ERR_SIZE_MISMATCH;
    q->n = a->second.n;
  }
  q->kind = K_ADD;
  q->dst = dst;
  q->src = src;
  return submit(S, q, req);
}

ft_error ftx_server_mul_u32(ftx_server S, uint64_t h, const uint32_t *k, size_t n, uint64_t *req) {
  auto q = std::make_shared<Request>();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto a = S->handles.find(h);
    if (a == S->handles.end() || a->second.released) return FT_ERR_OUT_OF_BOUNDS;
    if (a->second.n != n) return FT_ERR_SIZE_MISMATCH;
  }
  q->kind = K_MUL;
  q->n = n;
  q->dst = h;
  q->bn[0] = n * 4;
  if (S->borrow) q->bp[0] = reinterpret_cast<const uint8_t *>(k);
  else q->in.assign(reinterpret_cast<const uint8_t *>(k), reinterpret_cast<const uint8_t *>(k) + n * 4);
  return submit(S, q, req);
}

ft_error ftx_server_segsum(ftx_server S, uint64_t h, const int64_t *offsets, size_t nseg, uint64_t *req,
                           uint64_t *out_handle) {
  auto q = std::make_shared<Request>();
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto a = S->handles.find(h);
    if (a == S->handles.end() || a->second.released) return FT_ERR_OUT_OF_BOUNDS;
    if (offsets[0] < 0) return FT_ERR_OUT_OF_BOUNDS;
    for (size_t s = 0; s < nseg; ++s)
      if (offsets[s + 1] < offsets[s]) return FT_ERR_BAD_SIZE;
    if ((size_t)offsets[nseg] > a->second.n) return FT_ERR_OUT_OF_BOUNDS;
    q->dst = S->next_handle++;
    S->handles[q->dst].n = nseg;
  }
  q->kind = K_SEGSUM;
  q->n = nseg;
  q->nseg = nseg;
  q->src = h;
  q->bn[0] = (nseg + 1) * 8;
  if (S->borrow) q->bp[0] = reinterpret_cast<const uint8_t *>(offsets);
  else q->in.assign(reinterpret_cast<const uint8_t *>(offsets), reinterpret_cast<const uint8_t *>(offsets + nseg + 1));
  *out_handle = q->dst;
  return submit(S, q, req);
}

ft_error ftx_server_wait_ex(ftx_server S, uint64_t req, double *latency_us, double *queued_us) {
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(req);
  if (it == S->reqs.end()) return FT_ERR_OUT_OF_BOUNDS;
  auto r = it->second;
  S->cv_done.wait(lk, [&] { return r->state == 2; });
  if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
  if (queued_us) *queued_us = std::chrono::duration<double, std::micro>(r->t_launch - r->t_submit).count();
  S->reqs.erase(req);
  return r->err;
}

ft_error ftx_server_wait(ftx_server S, uint64_t req, double *latency_us) {
  return ftx_server_wait_ex(S, req, latency_us, nullptr);
}

ft_error ftx_server_release(ftx_server S, uint64_t h) {
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->handles.find(h);
  if (it == S->handles.end() || it->second.released) return FT_ERR_OUT_OF_BOUNDS;
  it->second.released = true;
  maybe_free(S, h);
  return FT_ERR_NO_ERROR;
}

ft_error ftx_server_read(ftx_server S, uint64_t h, uint8_t *c1_enc, uint8_t *c2_enc) {
  PtView v1, v2;
  size_t n;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->handles.find(h);
    if (it == S->handles.end()) return FT_ERR_OUT_OF_BOUNDS;
    S->cv_done.wait(lk, [&] {
      auto jt = S->handles.find(h);
      return jt == S->handles.end() || (jt->second.queued == 0 && jt->second.inflight == 0);
    });
    it = S->handles.find(h);
    if (it == S->handles.end() || !it->second.allocated) return FT_ERR_OUT_OF_BOUNDS;
    it->second.queued++;                 // pin while reading
    it->second.reading++;                // writers of h stay queued until the read ends
    v1 = it->second.c1;
    v2 = it->second.c2;
    n = it->second.n;
  }
  ft_error err = FT_ERR_NO_ERROR;
  void *d = nullptr;
  hipStream_t st = (hipStream_t)ftx_default_stream();
  if (hipMalloc(&d, n * 32 + 32) != hipSuccess) err = FT_ERR_OUT_OF_MEMORY;
  if (!err && launch_encode((uint32_t *)d, v1, n, st) != hipSuccess) err = FT_ERR_INTERNAL_LIBRARY_ERROR;
  if (!err && hipMemcpyAsync(c1_enc, d, n * 32, hipMemcpyDeviceToHost, st) != hipSuccess) err = FT_ERR_INTERNAL_LIBRARY_ERROR;
  if (!err && hipStreamSynchronize(st) != hipSuccess) err = FT_ERR_INTERNAL_LIBRARY_ERROR;
  if (!err && launch_encode((uint32_t *)d, v2, n, st) != hipSuccess) err = FT_ERR_INTERNAL_LIBRARY_ERROR;
  if (!err && hipMemcpyAsync(c2_enc, d, n * 32, hipMemcpyDeviceToHost, st) != hipSuccess) err = FT_ERR_INTERNAL_LIBRARY_ERROR;
  if (!err && hipStreamSynchronize(st) != hipSuccess) err = FT_ERR_INTERNAL_LIBRARY_ERROR;
  if (d) (void)hipFree(d);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    Handle &H = S->handles[h];
    H.queued--;
    H.reading--;
    maybe_free(S, h);
  }
  S->cv_work.notify_one();
  return err;
}

ft_error ftx_server_stats(ftx_server S, uint64_t *bytes_in_use, uint64_t *batches, uint64_t *done) {
  std::lock_guard<std::mutex> lk(S->mu);
  if (bytes_in_use) *bytes_in_use = S->arena.used;
  if (batches) *batches = S->n_batches;
  if (done) *done = S->n_done;
  return FT_ERR_NO_ERROR;
}

}  // extern "C"
