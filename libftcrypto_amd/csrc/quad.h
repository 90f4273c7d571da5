// Quad-lane point arithmetic for latency-bound sums and scans.
//
// For latency-bound work (few points per lane: short reduces, the block
// partials of every reduce, the upper levels of a scan) a point is spread over the four lanes of a quad: lane c
// holds coordinate c (X, Y, Z, T).  A point add then takes three field
// multiplies in sequence instead of nine, one per lane in each step, with
// the operands moved inside the quad by DPP quad permutes:
//   1: lane 0 (Y1 - X1)(Y2 - X2) = A, 1 (Y1 + X1)(Y2 + X2) = B, 2 T1 T2, 3 Z1 Z2
//   2: lane 2 C = 2d T1 T2, lane 3 D = 2 Z1 Z2 (lanes 0, 1 multiply by 1;
//      with the second point read from memory, whose fourth plane is 2d T,
//      step 1 already gives C and step 2 is a doubling on lane 3);
//      then E = B - A, H = B + A, G = D + C, F = D - C, carried to class N
//   3: lane 0 X3 = E F, 1 Y3 = G H, 2 Z3 = G F, 3 T3 = E H
// the same formulas as ge_add of a cached point followed by the p1p1 to p3
// conversion, so the sums are the same points.
#pragma once
#include "layout.h"

namespace ft {

// lane c of each quad receives the value of lane PAT[c] of the same quad
// (PAT: a DPP quad_perm selector, two bits per lane)
template <int PAT>
__device__ __forceinline__ void fe_qperm(fe &r, const fe &a) {
#pragma unroll
  for (int l = 0; l < 9; ++l) r.v[l] = (uint32_t)__builtin_amdgcn_mov_dpp((int)a.v[l], PAT, 0xf, 0xf, false);
}

// quad_perm selectors: lane i takes lane p[i] of its quad
constexpr int QP_SWAP = 1 | (0 << 2) | (3 << 4) | (2 << 6);   // [1, 0, 3, 2]
constexpr int QP_OP1 = 0 | (2 << 2) | (2 << 4) | (0 << 6);    // [0, 2, 2, 0]: E G G E
constexpr int QP_OP2 = 3 | (1 << 2) | (3 << 4) | (1 << 6);    // [3, 1, 3, 1]: F H F H

// the first factor of step 1 in lane c from coordinate m of the lane and n
// of its pair lane (c ^ 1): Y - X (c = 0: n = Y, m = X), Y + X (c = 1), T
// (c = 2: n = T) or Z (c = 3: n = Z)
__device__ __forceinline__ void quad_factor(fe &u, const fe &m, const fe &n, int c) {
  const uint32_t neg = c == 0 ? ~0u : 0u, keep = c < 2 ? ~0u : 0u;
  fe t;
  fe_cneg_mask(t, m, neg);
#pragma unroll
  for (int l = 0; l < 9; ++l) u.v[l] = n.v[l] + (t.v[l] & keep);
}

// Points read from memory hold 2d T in their fourth plane (the resident
// form, layout.h); a quad's running point holds T.  quad_from_mem and
// quad_to_mem convert coordinate c between the two (lane 3 multiplies by
// 1/(2d) or 2d, the others by 1: one multiply step of the quad).
__device__ __forceinline__ void quad_scale_t(fe &x, int c, const fe &k3) {
  fe k;
#pragma unroll
  for (int l = 0; l < 9; ++l) k.v[l] = c == 3 ? k3.v[l] : 0u;
  k.v[0] = c == 3 ? k3.v[0] : 1u;
  fe_mul(x, x, k);
}
__device__ __forceinline__ void quad_from_mem(fe &x, int c) { quad_scale_t(x, c, FE_INV_D2); }
__device__ __forceinline__ void quad_to_mem(fe &x, int c) { quad_scale_t(x, c, FE_D2); }

// acc += o: the quad's point plus the point whose coordinate c is o
// (MEM: o in the resident form, whose 2d T makes lane 2's product C at once,
// so the second step only doubles lane 3's Z1 Z2: two multiply steps)
template <bool MEM = false>
__device__ __forceinline__ void quad_add(fe &acc, const fe &o, int c) {
  fe n1, n2, u1, u2, p, r;
  fe_qperm<QP_SWAP>(n1, acc);
  fe_qperm<QP_SWAP>(n2, o);
  quad_factor(u1, acc, n1, c);            // A2 (c = 0), A1 (c = 1), N
  quad_factor(u2, o, n2, c);
  fe_carry(u2);                           // N, so every product is exact
  fe_mul(p, u1, u2);                      // A, B, T1 T2 (C for MEM), Z1 Z2
  if (MEM) {
    // lane 3 doubles (<= 2N), the others keep their product
    const uint32_t sh = c == 3 ? 1u : 0u;
#pragma unroll
    for (int l = 0; l < 9; ++l) r.v[l] = p.v[l] << sh;
  } else {
    // lanes 2 and 3 scale by 2d and 2, lanes 0 and 1 by 1
    fe k;
#pragma unroll
    for (int l = 0; l < 9; ++l) k.v[l] = c == 2 ? FE_D2.v[l] : 0u;
    k.v[0] = c == 2 ? FE_D2.v[0] : (c == 3 ? 2u : 1u);
    fe_mul(r, p, k);                      // A, B, C, D
  }
  fe s, w;
  fe_qperm<QP_SWAP>(s, r);                // B, A, D, C
  // lane 0: E = B - A, 1: H = B + A, 2: G = D + C, 3: F = D - C
  const uint32_t first = c == 0 ? ~0u : 0u, neg = (c == 0 || c == 3) ? ~0u : 0u;
  fe a, b;
#pragma unroll
  for (int l = 0; l < 9; ++l) { a.v[l] = first ? s.v[l] : r.v[l]; b.v[l] = first ? r.v[l] : s.v[l]; }
  fe_cneg_mask(b, b, neg);
  fe_add(w, a, b);                        // <= N + 2^30
  fe_carry(w);                            // E, H, G, F in class N
  fe x, y;
  fe_qperm<QP_OP1>(x, w);
  fe_qperm<QP_OP2>(y, w);
  fe_mul(acc, x, y);                      // X3, Y3, Z3, T3
}

__device__ __forceinline__ void quad_identity(fe &r, int c) {
  fe_0(r);
  r.v[0] = (c == 1 || c == 2) ? 1u : 0u;
}

__device__ __forceinline__ void fe_shfl_down(fe &r, const fe &a, int d) {
#pragma unroll
  for (int l = 0; l < 9; ++l) r.v[l] = __shfl_down(a.v[l], d);
}

// r = the value of lane (lane - d) for lanes d .. 63 of the wave (lanes below
// d get their own value back)
__device__ __forceinline__ void fe_shfl_up(fe &r, const fe &a, int d) {
#pragma unroll
  for (int l = 0; l < 9; ++l) r.v[l] = (uint32_t)__shfl_up((int)a.v[l], (unsigned)d, 64);
}

}  // namespace ft
