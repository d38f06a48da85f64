// wfa_common.h -- device helpers shared by the WFA kernels: packed-sequence
// match extension, cell validity and the CIGAR run writer.
#ifndef PHX_WFA_COMMON_H
#define PHX_WFA_COMMON_H

#include <hip/hip_runtime.h>

#include <type_traits>

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

// Whether a word pointer addresses LDS (address space 3).
template <class WP>
struct in_lds : std::false_type {};
template <>
struct in_lds<const lmem_t<const uint64_t>*> : std::true_type {};

// 64 bits of a packed symbol stream starting at bit `bit` (symbol 0 in the
// lowest bits of word 0), as two 32-bit funnel shifts over three halves.
// LDS: the three 32-bit halves at bit / 32 .. + 2 (two read instructions,
// no selects).  Global memory: the two aligned words bit / 64 and + 1 (one
// paired load each, fewer memory instructions) and a select of the three
// halves that matter.  Either way the reads stay inside the stream's words
// (each packed sequence carries one word past its last symbol).
template <class WP>
__device__ __forceinline__ uint64_t bit_window(WP w, uint32_t bit) {
  uint32_t x0, x1, x2;
  if constexpr (in_lds<WP>::value) {
    const lmem_t<const uint32_t>* d = (const lmem_t<const uint32_t>*)w;
    const uint32_t i = bit >> 5;
    x0 = d[i];
    x1 = d[i + 1];
    x2 = d[i + 2];
  } else {
    const uint32_t wi = bit >> 6;
    const uint64_t w0 = w[wi], w1 = w[wi + 1];
    const bool upper = (bit & 32u) != 0;
    x0 = upper ? (uint32_t)(w0 >> 32) : (uint32_t)w0;
    x1 = upper ? (uint32_t)w1 : (uint32_t)(w0 >> 32);
    x2 = upper ? (uint32_t)(w1 >> 32) : (uint32_t)w1;
  }
  const uint32_t lo = __builtin_amdgcn_alignbit(x1, x0, bit);  // shift = bit & 31
  const uint32_t hi = __builtin_amdgcn_alignbit(x2, x1, bit);
  return ((uint64_t)hi << 32) | lo;
}

// Upstream `extend` closure (wfa_simd.rs:252-262, :608-616): unchanged outside
// the matrix, else position + longest common prefix of the two suffixes.
// 32 bases (or 8 raw bytes) are compared per 64-bit XOR; the first differing
// symbol is the lowest set bit.  Each lane leaves at its own first
// difference, so finished lanes issue no further reads.  (A wave-uniform
// loop in which finished lanes keep re-reading a valid window measured
// slower on every tier: 4% on 150 bp pairs in LDS, 2-8% in global memory.)
// qsh / tsh: index of symbol 0 inside word 0 of each sequence.
template <bool B8, class WP>
__device__ __forceinline__ int32_t extend_diag(WP qw, uint32_t qsh, WP tw, uint32_t tsh,
                                               int32_t qn, int32_t tn, int32_t p, int32_t k) {
  constexpr uint32_t BITS = B8 ? 8 : 2;
  constexpr uint32_t CTZ_SHIFT = B8 ? 3 : 1;
  constexpr int32_t SPW = B8 ? 8 : 32;
  const int32_t j = p - k;
  if (p < 0 || p >= qn || j < 0 || j >= tn) return p;
  const int32_t lim = min(qn - p, tn - j);
  int32_t n = 0;
  for (;;) {
    const uint64_t d = bit_window(qw, (qsh + (uint32_t)(p + n)) * BITS) ^
                       bit_window(tw, (tsh + (uint32_t)(j + n)) * BITS);
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
