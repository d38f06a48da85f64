// wfa_common.h -- device helpers shared by the WFA kernels: packed-sequence
// match extension, cell validity and the CIGAR run writer.
#ifndef PHX_WFA_COMMON_H
#define PHX_WFA_COMMON_H

#include <hip/hip_runtime.h>

#include "phx_device.h"

namespace phx {

constexpr int32_t UNSET = INT32_MIN;

// --- 2-bit / 8-bit match extension -------------------------------------------

// Symbols [pos, pos + per_word) of a packed sequence as one little-endian word.
// WP: a pointer to const u64 words in LDS or global memory (address space
// carried by the type, so accesses compile to ds_ / global_ instructions).
template <bool B8, class WP>
__device__ __forceinline__ uint64_t seq_window(WP w, uint32_t pos) {
  constexpr uint32_t LOG = B8 ? 3 : 5;
  constexpr uint32_t BITS = B8 ? 8 : 2;
  const uint32_t wi = pos >> LOG;
  const uint32_t sh = (pos & ((1u << LOG) - 1)) * BITS;
  const uint64_t lo = w[wi] >> sh;
  const uint64_t hi = (w[wi + 1] << (63 - sh)) << 1;  // zero when sh == 0
  return lo | hi;
}

// Upstream `extend` closure (wfa_simd.rs:252-262, :608-616): unchanged outside
// the matrix, else position + longest common prefix of the two suffixes.
// 32 bases (or 8 raw bytes) are compared per 64-bit XOR; the first differing
// symbol is the lowest set bit.
// qsh / tsh: index of symbol 0 inside word 0 of each sequence.
template <bool B8, class WP>
__device__ __forceinline__ int32_t extend_diag(WP qw, uint32_t qsh, WP tw, uint32_t tsh,
                                               int32_t qn, int32_t tn, int32_t p, int32_t k) {
  const int32_t j = p - k;
  if (p < 0 || p >= qn || j < 0 || j >= tn) return p;
  const int32_t lim = min(qn - p, tn - j);
  constexpr int32_t SPW = B8 ? 8 : 32;
  constexpr uint32_t CTZ_SHIFT = B8 ? 3 : 1;
  int32_t n = 0;
  for (;;) {
    const uint64_t d = seq_window<B8>(qw, qsh + (uint32_t)(p + n)) ^
                       seq_window<B8>(tw, tsh + (uint32_t)(j + n));
    if (d) {
      n += (int32_t)((uint32_t)__builtin_ctzll(d) >> CTZ_SHIFT);
      break;
    }
    n += SPW;
    if (n >= lim) break;
  }
  return p + min(n, lim);
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
