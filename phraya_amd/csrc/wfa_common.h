// wfa_common.h -- device helpers shared by the WFA kernels: packed-sequence
// match extension, cell validity and the CIGAR run writer.
#ifndef PHX_WFA_COMMON_H
#define PHX_WFA_COMMON_H

#include <hip/hip_runtime.h>

#include "phx_device.h"

namespace phx {

constexpr int32_t UNSET = INT32_MIN;

// Element types qualified with their address space (1 = global, 3 = LDS), so
// every access compiles to a global_ / ds_ instruction instead of flat.
template <class T>
using gmem_t = __attribute__((address_space(1))) T;
template <class T>
using lmem_t = __attribute__((address_space(3))) T;

// --- 2-bit / 8-bit match extension -------------------------------------------

// The same packed words seen as 32-bit halves, in the same address space.
__device__ __forceinline__ const lmem_t<const uint32_t>* as_dwords(const lmem_t<const uint64_t>* w) {
  return (const lmem_t<const uint32_t>*)w;
}
__device__ __forceinline__ const gmem_t<const uint32_t>* as_dwords(const gmem_t<const uint64_t>* w) {
  return (const gmem_t<const uint32_t>*)w;
}

// 64 bits of a packed symbol stream starting at bit `bit` (symbol 0 in the
// lowest bits of dword 0): two funnel shifts over three dwords.  Reads dwords
// bit / 32 .. bit / 32 + 2, inside the stream's words (each packed sequence
// carries one word past its last symbol).
template <class DP>
__device__ __forceinline__ uint64_t bit_window(DP d, uint32_t bit) {
  const uint32_t i = bit >> 5;
  const uint32_t d0 = d[i], d1 = d[i + 1], d2 = d[i + 2];
  const uint32_t lo = __builtin_amdgcn_alignbit(d1, d0, bit);  // shift = bit & 31
  const uint32_t hi = __builtin_amdgcn_alignbit(d2, d1, bit);
  return ((uint64_t)hi << 32) | lo;
}

// Upstream `extend` closure (wfa_simd.rs:252-262, :608-616): unchanged outside
// the matrix, else position + longest common prefix of the two suffixes.
// 32 bases (or 8 raw bytes) are compared per 64-bit XOR; the first differing
// symbol is the lowest set bit.  The loop is uniform over the wave: a lane
// that has found its mismatch (or was never inside the matrix) keeps its
// position and count while the others go on, so no per-lane branches.
// qsh / tsh: index of symbol 0 inside word 0 of each sequence.
template <bool B8, class WP>
__device__ __forceinline__ int32_t extend_diag(WP qw, uint32_t qsh, WP tw, uint32_t tsh,
                                               int32_t qn, int32_t tn, int32_t p, int32_t k) {
  constexpr uint32_t BITS = B8 ? 8 : 2;
  constexpr uint32_t CTZ_SHIFT = B8 ? 3 : 1;
  const int32_t j = p - k;
  const bool inside = p >= 0 && p < qn && j >= 0 && j < tn;
  bool act = inside;
  const int32_t lim = inside ? min(qn - p, tn - j) : 0;
  const uint32_t qb = (qsh + (uint32_t)(inside ? p : 0)) * BITS;
  const uint32_t tb = (tsh + (uint32_t)(inside ? j : 0)) * BITS;
  const auto qd = as_dwords(qw);
  const auto td = as_dwords(tw);
  int32_t n = 0;
  while (__ballot(act)) {
    const uint64_t d = bit_window(qd, qb + (uint32_t)n * BITS) ^ bit_window(td, tb + (uint32_t)n * BITS);
    // Equal symbols before the first difference; a whole window when none.
    const int32_t z = (int32_t)((uint32_t)__builtin_ctzg(d, 64) >> CTZ_SHIFT);
    n = act ? n + z : n;
    act = act && d == 0 && n < lim;
  }
  return inside ? p + min(n, lim) : p;
}

__device__ __forceinline__ bool pos_valid(int32_t pos, int32_t k, int32_t qn, int32_t tn) {
  // affine_pos_valid (wfa_simd.rs:563-569)
  if (pos == UNSET || pos < 0 || pos > qn) return false;
  const int32_t tp = pos - k;
  return tp >= 0 && tp <= tn;
}

__device__ __forceinline__ bool in_range(int32_t k, int32_t lo, int32_t hi) { return k >= lo && k <= hi; }

// --- CIGAR run writer (thread 0) -----------------------------------------------

// Runs are produced last-to-first by the traceback and stored from the end of
// the pair's slot range backwards, so they land in forward order.
struct RunWriter {
  uint32_t* base;  // runs + run_off
  uint32_t cap;
  uint32_t n;
  uint32_t cur_op;
  uint32_t cur_len;
  int32_t non_m;  // edit distance accumulated from non-M runs
  bool overflow;
  bool store;     // whether this thread performs the stores (one writer per pair)

  __device__ void init(uint32_t* b, uint32_t c, bool writer = true) {
    store = writer;
    base = b;
    cap = c;
    n = 0;
    cur_op = 0xFFFFFFFFu;
    cur_len = 0;
    non_m = 0;
    overflow = false;
  }
  __device__ void flush() {
    if (cur_len == 0) return;
    if (n < cap) {
      if (store) base[cap - 1 - n] = (cur_len << 4) | cur_op;
      ++n;
    } else {
      overflow = true;
    }
    if (cur_op != OP_M) non_m += (int32_t)cur_len;
    cur_len = 0;
  }
  __device__ void emit(uint32_t op, int32_t len) {
    if (len <= 0) return;
    if (op != cur_op) {
      flush();
      cur_op = op;
    }
    cur_len += (uint32_t)len;
  }
};

}  // namespace phx

#endif
