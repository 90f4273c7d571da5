// Host-side launchers of the data-movement, predicate and I/O kernels (misc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "scalarmult.h"

namespace ft {

hipError_t launch_gather(PtView dst, PtView src, const uint64_t *idx, size_t start, size_t step, size_t n,
                         hipStream_t s);
hipError_t launch_scatter(PtView dst, PtView src, const uint64_t *idx, size_t start, size_t step, size_t n,
                          hipStream_t s);
hipError_t launch_mux(PtView t, const int64_t *cond, PtView f, size_t n, hipStream_t s);
hipError_t launch_index(int64_t *res, PtView v, size_t n, hipStream_t s);
hipError_t launch_equal(int64_t *res, PtView a, PtView b, size_t n, bool neq, hipStream_t s);
// bad = 1 + the largest index of a point off the curve or with X(l P) != 0;
// naf: the len width-3 NAF digits of l, lowest first, in device memory
hipError_t launch_validate(unsigned long long *bad, PtView v, size_t n, const int8_t *naf, int len, hipStream_t s);
// res[i] = 1 when the 32-byte encoding enc[i] equals one of set_enc[0..n_set),
// else 0; slots: set_slots(n_set) uint32 of device scratch
size_t set_slots(size_t n_set);
hipError_t launch_set_contains(int64_t *res, const uint32_t *enc, size_t n, const uint32_t *set_enc, size_t n_set,
                               uint32_t *slots, hipStream_t s);
hipError_t launch_to_affine(uint32_t *out, PtView v, size_t n, hipStream_t s);
hipError_t launch_from_affine(PtView v, const uint32_t *in, size_t n, hipStream_t s);
hipError_t launch_to_raw(uint32_t *out, PtView v, size_t n, hipStream_t s);
hipError_t launch_from_raw(PtView v, const uint32_t *in, size_t n, hipStream_t s);

}  // namespace ft
