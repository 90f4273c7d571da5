// Launch schedules of the aggregation kernels and the autotuner that picks
// them on the device at hand.
//
// A schedule is chosen per size class of the input length n (n < 2^14,
// < 2^18, < 2^22, larger):
//   seg1    log2 elements per lane at segmented-sum level 1 (4..7)
//   segu    log2 of the units each quad walks per pass of the block walk
//           that adds the chunk partials of segments spanning several
//           level-1 chunks (1..4: 2, 4, 8 or 16 units)
//   scan    log2 elements per lane at a scan's first level (4..7; the
//           levels above it take 16)
//   reduce  most blocks of 256 lanes in the first reduce pass (512..4096)
//   u32w    window width of ciphertext x 32-bit constant: 4 (radix 16,
//           8-entry table, 9 windows) or 3 (radix 8, 4-entry table, 11)
//   vbw     window width of the 256-bit variable-base mult: 4 (radix 16,
//           64 windows) or 3 (radix 8, 86 windows)
//   scanlb  scan form: 0 chunk levels (a walk summing chunks, upper levels,
//           a second walk from each chunk's prefix), 1 single pass with a
//           decoupled look-back over 8,192-point tiles
//   encl    lanes per ciphertext of an encryption or rerandomisation: 1,
//           2 (C1 on one lane, C2 on the next; twice the waves of a small
//           batch, each lane with one comb pass) or 4 (each radix-32
//           comb's rows split over two waves, joined by one add through
//           LDS: twice the waves again, 2 % more work).  The nine limbs of a field
//           element always stay in one lane: splitting them would exchange
//           partial column sums between lanes inside every multiply.
// Every schedule gives the same group element (point addition is
// associative and commutative), so the choice only moves time.  Scratch
// sizes are those of the schedule needing the most, so buffers sized by
// ftx_*_scratch_points stay valid whatever is picked.
//
// ftx_autotune times every candidate of every class on synthetic points
// (s_i B for small s_i, lognormal segment lengths) with HIP events, keeps
// the fastest and can write the table to a text file; ftx_tune_load reads
// such a file back, and ft_crypto_init loads the file named by the
// FTX_TUNE_FILE environment variable when it is set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/ftcrypto.h"
#include "../../include/ftcrypto_ext.h"
#include "kernels.h"

namespace ft {

namespace {

// Defaults per size class: what the tuner picks on MI355X (elements per lane
// 2^4 below 2^22 and 2^6 above, blocks of 2 then 4 units in the block walk, radix-16
// windows for the variable base below 2^18, two lanes per ciphertext of an
// encryption), so an untuned process is close to a tuned one.
constexpr AggTune DEFAULT_TUNE[TUNE_CLASSES] = {
    {4, 1, 4, 2048, 3, 4, 4, 0}, {4, 1, 4, 512, 3, 4, 4, 0}, {4, 1, 4, 512, 3, 3, 2, 0}, {6, 2, 4, 2048, 3, 3, 2, 0}};

std::atomic<int> T_seg1[TUNE_CLASSES], T_segu[TUNE_CLASSES], T_scan[TUNE_CLASSES], T_reduce[TUNE_CLASSES],
    T_u32w[TUNE_CLASSES], T_vbw[TUNE_CLASSES], T_encl[TUNE_CLASSES], T_scanlb[TUNE_CLASSES];
std::atomic<bool> T_ready{false};

void tune_defaults() {
  for (int c = 0; c < TUNE_CLASSES; ++c) {
    T_seg1[c] = DEFAULT_TUNE[c].seg1;
    T_segu[c] = DEFAULT_TUNE[c].segu;
    T_scan[c] = DEFAULT_TUNE[c].scan;
    T_reduce[c] = DEFAULT_TUNE[c].reduce;
    T_u32w[c] = DEFAULT_TUNE[c].u32w;
    T_vbw[c] = DEFAULT_TUNE[c].vbw;
    T_encl[c] = DEFAULT_TUNE[c].encl;
    T_scanlb[c] = DEFAULT_TUNE[c].scanlb;
  }
  T_ready = true;
}

bool valid(const AggTune &t) {
  auto ch_ok = [](int c) { return c >= CH_MIN && c <= CH_MAX; };
  return ch_ok(t.seg1) && t.segu >= TREE_LR_MIN && t.segu <= TREE_LR_MAX && ch_ok(t.scan) && t.reduce >= 1 &&
         (unsigned)t.reduce <= REDUCE_BLOCKS_MAX && (t.u32w == 3 || t.u32w == 4) &&
         (t.vbw == 3 || t.vbw == 4) && (t.encl == 1 || t.encl == 2 || t.encl == 4) && (t.scanlb == 0 || t.scanlb == 1);
}

void tune_store(int c, const AggTune &t) {
  T_seg1[c] = t.seg1;
  T_segu[c] = t.segu;
  T_scan[c] = t.scan;
  T_reduce[c] = t.reduce;
  T_u32w[c] = t.u32w;
  T_vbw[c] = t.vbw;
  T_encl[c] = t.encl;
  T_scanlb[c] = t.scanlb;
}

}  // namespace

// the length at which ftx_autotune times each class
static const int TUNE_LOGN[TUNE_CLASSES] = {12, 16, 20, 24};
size_t tune_class_size(int c) { return (size_t)1 << TUNE_LOGN[c]; }

int tune_class(size_t n) {
  if (n < ((size_t)1 << 14)) return 0;
  if (n < ((size_t)1 << 18)) return 1;
  if (n < ((size_t)1 << 22)) return 2;
  return 3;
}

AggTune tune_for(size_t n) {
  if (!T_ready) tune_defaults();
  const int c = tune_class(n);
  return AggTune{T_seg1[c], T_segu[c], T_scan[c], T_reduce[c], T_u32w[c], T_vbw[c], T_encl[c], T_scanlb[c]};
}

}  // namespace ft

using namespace ft;

namespace {

// encryptions are timed at the class's own size up to 2^22 (about 20 ms per
// launch there): the lane split that wins at 2^18 need not win at 2^22
constexpr size_t ENC_TUNE_MAX = (size_t)1 << 22;

#define TCK(x)                                              \
  do {                                                      \
    hipError_t e_ = (x);                                    \
    if (e_ != hipSuccess) {                                 \
      (void)hipGetLastError();                              \
      return e_ == hipErrorOutOfMemory ? FT_ERR_OUT_OF_MEMORY : FT_ERR_INTERNAL_LIBRARY_ERROR; \
    }                                                       \
  } while (0)

struct DevMem {
  void *p = nullptr;
  ~DevMem() { if (p) (void)hipFree(p); }
};

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

struct Stream {
  hipStream_t s = nullptr;
  ~Stream() { if (s) (void)hipStreamDestroy(s); }
};

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// average milliseconds of `reps` launches of fn after one warm-up launch
template <class F>
ft_error time_ms(float *ms, const Events &ev, hipStream_t s, int reps, F fn) {
  TCK(fn());
  TCK(hipEventRecord(ev.a, s));
  for (int r = 0; r < reps; ++r) TCK(fn());
  TCK(hipEventRecord(ev.b, s));
  TCK(hipEventSynchronize(ev.b));
  float t = 0;
  TCK(hipEventElapsedTime(&t, ev.a, ev.b));
  *ms = t / reps;
  return FT_ERR_NO_ERROR;
}

}  // namespace

extern "C" {

ft_error ftx_tune_get(int size_class, int *seg1, int *segu, int *scan, int *reduce_blocks, int *u32_window,
                      int *varbase_window) {
  if (size_class < 0 || size_class >= TUNE_CLASSES) return FT_ERR_OUT_OF_BOUNDS;
  if (!T_ready) tune_defaults();
  *seg1 = T_seg1[size_class];
  *segu = T_segu[size_class];
  *scan = T_scan[size_class];
  *reduce_blocks = T_reduce[size_class];
  *u32_window = T_u32w[size_class];
  *varbase_window = T_vbw[size_class];
  return FT_ERR_NO_ERROR;
}

ft_error ftx_tune_set(int size_class, int seg1, int segu, int scan, int reduce_blocks, int u32_window,
                      int varbase_window) {
  if (size_class < 0 || size_class >= TUNE_CLASSES) return FT_ERR_OUT_OF_BOUNDS;
  if (!T_ready) tune_defaults();
  const AggTune t{seg1, segu, scan, reduce_blocks, u32_window, varbase_window, T_encl[size_class],
                  T_scanlb[size_class]};
  if (!valid(t)) return FT_ERR_BAD_SIZE;
  tune_store(size_class, t);
  return FT_ERR_NO_ERROR;
}

ft_error ftx_tune_get_ct_lanes(int size_class, int *lanes) {
  if (size_class < 0 || size_class >= TUNE_CLASSES) return FT_ERR_OUT_OF_BOUNDS;
  if (!T_ready) tune_defaults();
  *lanes = T_encl[size_class];
  return FT_ERR_NO_ERROR;
}

ft_error ftx_tune_set_ct_lanes(int size_class, int lanes) {
  if (size_class < 0 || size_class >= TUNE_CLASSES) return FT_ERR_OUT_OF_BOUNDS;
  if (lanes != 1 && lanes != 2 && lanes != 4) return FT_ERR_BAD_SIZE;
  if (!T_ready) tune_defaults();
  T_encl[size_class] = lanes;
  return FT_ERR_NO_ERROR;
}

ft_error ftx_tune_get_scan_lookback(int size_class, int *lookback) {
  if (size_class < 0 || size_class >= TUNE_CLASSES) return FT_ERR_OUT_OF_BOUNDS;
  if (!T_ready) tune_defaults();
  *lookback = T_scanlb[size_class];
  return FT_ERR_NO_ERROR;
}

ft_error ftx_tune_set_scan_lookback(int size_class, int lookback) {
  if (size_class < 0 || size_class >= TUNE_CLASSES) return FT_ERR_OUT_OF_BOUNDS;
  if (lookback != 0 && lookback != 1) return FT_ERR_BAD_SIZE;
  if (!T_ready) tune_defaults();
  T_scanlb[size_class] = lookback;
  return FT_ERR_NO_ERROR;
}

int ftx_tune_size_class(size_t n) { return tune_class(n); }

// Text format: a comment line, then one line per class:
// "class seg1 segu scan reduce u32w vbw encl scanlb" (files without the last
// columns, written before they existed, load with the defaults of each class).
ft_error ftx_tune_save(const char *path) {
  FILE *f = std::fopen(path, "w");
  if (!f) return FT_ERR_INTERNAL_LIBRARY_ERROR;
  if (!T_ready) tune_defaults();
  std::fprintf(f, "# libftcrypto_amd launch schedules: class seg1 segu scan reduce u32w vbw encl scanlb\n");
  for (int c = 0; c < TUNE_CLASSES; ++c)
    std::fprintf(f, "%d %d %d %d %d %d %d %d %d\n", c, (int)T_seg1[c], (int)T_segu[c], (int)T_scan[c],
                 (int)T_reduce[c], (int)T_u32w[c], (int)T_vbw[c], (int)T_encl[c], (int)T_scanlb[c]);
  return std::fclose(f) == 0 ? FT_ERR_NO_ERROR : FT_ERR_INTERNAL_LIBRARY_ERROR;
}

ft_error ftx_tune_load(const char *path) {
  FILE *f = std::fopen(path, "r");
  if (!f) return FT_ERR_OUT_OF_BOUNDS;
  AggTune t[TUNE_CLASSES];
  bool seen[TUNE_CLASSES] = {};
  char line[256];
  ft_error err = FT_ERR_NO_ERROR;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    int c, a, b, d, r, w, v, l, k;
    const int got = std::sscanf(line, "%d %d %d %d %d %d %d %d %d", &c, &a, &b, &d, &r, &w, &v, &l, &k);
    if (got < 7 || c < 0 || c >= TUNE_CLASSES) {
      err = FT_ERR_BAD_SIZE;
      break;
    }
    t[c] = AggTune{a, b, d, r, w, v, got >= 8 ? l : DEFAULT_TUNE[c].encl, got >= 9 ? k : DEFAULT_TUNE[c].scanlb};
    if (!valid(t[c])) { err = FT_ERR_BAD_SIZE; break; }
    seen[c] = true;
  }
  std::fclose(f);
  if (err != FT_ERR_NO_ERROR) return err;
  for (int c = 0; c < TUNE_CLASSES; ++c)
    if (!seen[c]) return FT_ERR_BAD_SIZE;
  if (!T_ready) tune_defaults();
  for (int c = 0; c < TUNE_CLASSES; ++c) tune_store(c, t[c]);
  return FT_ERR_NO_ERROR;
}

// Measure every candidate schedule of every size class (classes up to
// max_class, 0..3) and keep the fastest; write the table to path when it is
// not NULL.  Needs ft_crypto_init; allocates up to ~3.8 GB of device memory
// while it runs (class 3 is timed at n = 2^24, its encryptions at 2^22).
ft_error ftx_autotune(int max_class, const char *path) {
  if (max_class < 0 || max_class >= TUNE_CLASSES) return FT_ERR_OUT_OF_BOUNDS;
  if (!base_table()) return FT_ERR_INSUFFICIENT_DEVICE_OR_DRIVER;
  if (!T_ready) tune_defaults();
  const size_t nmax = tune_class_size(max_class);
  Stream st;
  TCK(hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking));
  Events ev;
  TCK(hipEventCreate(&ev.a));
  TCK(hipEventCreate(&ev.b));
  // points s_i B, s_i = i + 1; scratch for the segmented sum, scan and reduce
  DevMem pts, cpy, out, scr, sc;
  const size_t seg_pts = segsum_scratch_points(nmax);
  const size_t scan_pts = scan_scratch_points(nmax);
  const size_t red_pts = REDUCE_BLOCKS_MAX;
  const size_t scr_pts = std::max(seg_pts, std::max(scan_pts, red_pts));
  TCK(hipMalloc(&pts.p, nmax * 128));
  TCK(hipMalloc(&cpy.p, nmax * 128));
  TCK(hipMalloc(&scr.p, scr_pts * 128));
  TCK(hipMalloc(&sc.p, nmax * 8));
  std::vector<uint64_t> small(nmax);
  for (size_t i = 0; i < nmax; ++i) small[i] = i + 1;
  TCK(hipMemcpyAsync(sc.p, small.data(), nmax * 8, hipMemcpyHostToDevice, st.s));
  const PtView P{(uint32_t *)pts.p, nmax};
  TCK(launch_fixed_base(P, sc.p, true, nmax, base_table32(), st.s));
  TCK(hipStreamSynchronize(st.s));
  // skewed segment lengths: lognormal(ln 16, 1.5) clipped to [1, 65536]
  // (median 16, mean about 49, a few segments spanning thousands of chunks),
  // normal variates by Box-Muller from two uniforms of the SplitMix64 stream
  std::vector<int64_t> off;
  off.reserve(nmax / 8);
  off.push_back(0);
  for (uint64_t k = 0; off.back() < (int64_t)nmax; ++k) {
    const double u1 = ((mix64(0xF7C0AB00ull + 2 * k) >> 11) + 1) * (1.0 / 9007199254740993.0);
    const double u2 = (mix64(0xF7C0AB00ull + 2 * k + 1) >> 11) * (1.0 / 9007199254740992.0);
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    const double x = std::exp(std::log(16.0) + 1.5 * z);
    const int64_t len = std::min<int64_t>(65536, std::max<int64_t>(1, (int64_t)std::llround(x)));
    off.push_back(std::min<int64_t>(off.back() + len, (int64_t)nmax));
  }
  TCK(hipMalloc(&out.p, off.size() * 128));          // at least one slot per segment
  TCK(hipMemcpyAsync(cpy.p, pts.p, nmax * 128, hipMemcpyDeviceToDevice, st.s));
  const PtView C{(uint32_t *)cpy.p, nmax};
  const PtView S{(uint32_t *)scr.p, scr_pts};
  // encryption outputs and 32-byte r scalars for at most ENC_TUNE_MAX ciphertexts
  const size_t nenc = std::min(nmax, ENC_TUNE_MAX);
  DevMem e1, e2, er;
  TCK(hipMalloc(&e1.p, nenc * 128));
  TCK(hipMalloc(&e2.p, nenc * 128));
  TCK(hipMalloc(&er.p, nenc * 32));
  TCK(hipMemcpyAsync(er.p, pts.p, nenc * 32, hipMemcpyDeviceToDevice, st.s));
  const PtView E1{(uint32_t *)e1.p, nenc}, E2{(uint32_t *)e2.p, nenc};
  for (int c = 0; c <= max_class; ++c) {
    const size_t n = tune_class_size(c);
    const int reps = c < 2 ? 10 : 3;
    // segments covering [0, n): the prefix of the offsets, last one clipped
    size_t nseg = 0;
    while (nseg + 1 < off.size() && off[nseg] < (int64_t)n) ++nseg;
    std::vector<int64_t> offc(off.begin(), off.begin() + nseg + 1);
    offc[nseg] = (int64_t)n;
    DevMem doffc;
    TCK(hipMalloc(&doffc.p, offc.size() * 8));
    TCK(hipMemcpyAsync(doffc.p, offc.data(), offc.size() * 8, hipMemcpyHostToDevice, st.s));
    const PtView O{(uint32_t *)out.p, off.size()};
    AggTune best = tune_for(n);
    float bt = INFINITY, t = 0;
    for (int a = CH_MIN; a <= CH_MAX; ++a)
      for (int b = TREE_LR_MIN; b <= TREE_LR_MAX; ++b) {
        ft_error e = time_ms(&t, ev, st.s, reps, [&] {
          return segsum_launch_ch(O, (const int64_t *)doffc.p, nseg, n, nullptr, P, 0, S, a, b, st.s);
        });
        if (e != FT_ERR_NO_ERROR) return e;
        if (t < bt) { bt = t; best.seg1 = a; best.segu = b; }
      }
    bt = INFINITY;
    for (int a = CH_MIN; a <= CH_MAX; ++a) {
      // in place on a copy of the points (any points time the same)
      ft_error e = time_ms(&t, ev, st.s, reps, [&] { return scan_launch_ch(C, n, false, S, a, st.s); });
      if (e != FT_ERR_NO_ERROR) return e;
      if (t < bt) { bt = t; best.scan = a; best.scanlb = 0; }
    }
    {
      // the single-pass look-back scan against the best chunk schedule
      ft_error e = time_ms(&t, ev, st.s, reps, [&] { return scan_lookback_launch(C, n, false, S, st.s); });
      if (e != FT_ERR_NO_ERROR) return e;
      if (t < bt) { bt = t; best.scanlb = 1; }
    }
    bt = INFINITY;
    for (unsigned blocks = 512; blocks <= REDUCE_BLOCKS_MAX; blocks *= 2) {
      ft_error e = time_ms(&t, ev, st.s, reps, [&] {
        const unsigned g1 = reduce_grid(n, blocks);
        hipError_t x = launch_reduce_blocks(S, P, n, g1, st.s);
        if (x != hipSuccess) return x;
        return launch_reduce_blocks(O, S, g1, 1, st.s);
      });
      if (e != FT_ERR_NO_ERROR) return e;
      if (t < bt) { bt = t; best.reduce = (int)blocks; }
    }
    // ciphertext x u32 over the points as both halves (in place; any
    // points time the same), window width 3 against 4
    bt = INFINITY;
    for (int w = 3; w <= 4; ++w) {
      const size_t nm = std::min(n, (size_t)1 << 20);       // mults are costly: time at most 2^20
      ft_error e = time_ms(&t, ev, st.s, c < 2 ? 3 : 1, [&] {
        return launch_mul_u32_w(C, C, false, (const uint32_t *)sc.p, nm, w, st.s);
      });
      if (e != FT_ERR_NO_ERROR) return e;
      if (t < bt) { bt = t; best.u32w = w; }
    }
    // variable-base mults in place on the copy, radix 8 against radix 16;
    // one 8-word scalar for every lane (stride 0: sc holds 8 bytes per
    // element, not 32), which times the same as distinct scalars
    bt = INFINITY;
    for (int w = 3; w <= 4; ++w) {
      const size_t nm = std::min(n, (size_t)1 << 19);
      ft_error e = time_ms(&t, ev, st.s, c < 2 ? 3 : 1, [&] {
        return launch_varbase_w(C, (const uint32_t *)sc.p, 0, nm, w, st.s);
      });
      if (e != FT_ERR_NO_ERROR) return e;
      if (t < bt) { bt = t; best.vbw = w; }
    }
    // encryption with one lane per ciphertext against two (C1 and C2 on
    // separate lanes) and four (each comb split in two); B's comb table for
    // both keys and arbitrary r (any scalars time the same: the combs are
    // constant-time)
    bt = INFINITY;
    for (int l = 1; l <= 4; l *= 2) {
      const size_t ne = std::min(n, ENC_TUNE_MAX);
      ft_error e = time_ms(&t, ev, st.s, c < 2 ? 3 : 1, [&] {
        return launch_encrypt_l(E1, E2, base_table(), base_table(), (const uint64_t *)sc.p, 1,
                                (const uint32_t *)er.p, ne, false, l, st.s, nullptr, base_table32(), base_table32());
      });
      if (e != FT_ERR_NO_ERROR) return e;
      if (t < bt) { bt = t; best.encl = l; }
    }
    tune_store(c, best);
  }
  TCK(hipStreamSynchronize(st.s));
  return path ? ftx_tune_save(path) : FT_ERR_NO_ERROR;
}

}  // extern "C"

namespace ft {
void tune_init_from_env() {
  const char *p = std::getenv("FTX_TUNE_FILE");
  if (p && *p) (void)ftx_tune_load(p);
}
}  // namespace ft
