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

// Sequential 64-bit windows of a packed symbol stream (symbol 0 in the
// lowest bits of word 0), starting at bit `bit` and moving 64 bits per
// advance(): each window is two 32-bit funnel shifts over three 32-bit
// halves, and the shift stays the same from one window to the next.
// LDS: the three halves at dword bit / 32 .. + 2 (two read instructions, no
// selects).  Global memory: the two aligned words bit / 64 and + 1 (one
// paired load each, fewer memory instructions) and a select of the three
// halves that matter.  Either way the reads stay inside the stream's words
// (each packed sequence carries one word past its last symbol).
template <class WP, bool LDS = in_lds<WP>::value>
struct BitCursor;

template <class WP>
struct BitCursor<WP, true> {
  const lmem_t<const uint32_t>* d;
  uint32_t sh;
  __device__ __forceinline__ BitCursor(WP w, uint32_t bit)
      : d((const lmem_t<const uint32_t>*)w + (bit >> 5)), sh(bit) {}
  __device__ __forceinline__ uint64_t read() const {
    const uint32_t x0 = d[0], x1 = d[1], x2 = d[2];
    const uint32_t lo = __builtin_amdgcn_alignbit(x1, x0, sh);  // shift = sh & 31
    const uint32_t hi = __builtin_amdgcn_alignbit(x2, x1, sh);
    return ((uint64_t)hi << 32) | lo;
  }
  __device__ __forceinline__ void advance() { d += 2; }
};

template <class WP>
struct BitCursor<WP, false> {
  WP base;       // the stream's first word (wave-uniform in the HBM tiers)
  uint32_t off;  // byte offset of the current word: base + 32-bit offset
                 // addressing, no 64-bit pointer per lane
  uint32_t sh;
  __device__ __forceinline__ BitCursor(WP p, uint32_t bit) : base(p), off((bit >> 6) << 3), sh(bit) {}
  __device__ __forceinline__ uint64_t word(uint32_t o) const {
    return *reinterpret_cast<WP>(reinterpret_cast<const gmem_t<const uint8_t>*>(base) + o);
  }
  __device__ __forceinline__ uint64_t read() const {
    const uint64_t w0 = word(off), w1 = word(off + 8);
    const bool upper = (sh & 32u) != 0;
    const uint32_t x0 = upper ? (uint32_t)(w0 >> 32) : (uint32_t)w0;
    const uint32_t x1 = upper ? (uint32_t)w1 : (uint32_t)(w0 >> 32);
    const uint32_t x2 = upper ? (uint32_t)(w1 >> 32) : (uint32_t)w1;
    const uint32_t lo = __builtin_amdgcn_alignbit(x1, x0, sh);
    const uint32_t hi = __builtin_amdgcn_alignbit(x2, x1, sh);
    return ((uint64_t)hi << 32) | lo;
  }
  __device__ __forceinline__ void advance() { off += 8; }
};

// Index of the lowest set bit of a 64-bit word, 64 when it is zero.
__device__ __forceinline__ uint32_t lowest_bit64(uint64_t d) {
  return (uint32_t)__builtin_ctzg(d, 64);
}

// Upstream `extend` closure (wfa_simd.rs:252-262, :608-616): unchanged outside
// the matrix, else position + longest common prefix of the two suffixes.
// 32 bases (or 8 raw bytes) are compared per 64-bit XOR; the first differing
// symbol is the lowest set bit, and a window with no difference counts a
// full word.  Each lane leaves at its own first difference (or at the end
// of the shorter suffix), so finished lanes issue no further reads.  (A
// wave-uniform loop in which finished lanes keep re-reading a valid window
// measured slower on every tier: 4% on 150 bp pairs in LDS, 2-8% in global
// memory.)  qsh / tsh: index of symbol 0 inside word 0 of each sequence.
template <bool B8, class WP>
__device__ __forceinline__ int32_t extend_diag(WP qw, uint32_t qsh, WP tw, uint32_t tsh,
                                               int32_t qn, int32_t tn, int32_t p, int32_t k) {
  constexpr uint32_t BITS = B8 ? 8 : 2;
  constexpr uint32_t CTZ_SHIFT = B8 ? 3 : 1;
  const int32_t j = p - k;
  if ((uint32_t)p >= (uint32_t)qn || (uint32_t)j >= (uint32_t)tn) return p;
  const int32_t lim = min(qn - p, tn - j);
  BitCursor<WP> qc(qw, (qsh + (uint32_t)p) * BITS), tc(tw, (tsh + (uint32_t)j) * BITS);
  int32_t n = 0;
  for (;;) {
    const uint32_t z = lowest_bit64(qc.read() ^ tc.read());
    n += (int32_t)(z >> CTZ_SHIFT);  // no difference: a whole word, 32 bases / 8 bytes
    if (z < 64u || n >= lim) break;
    qc.advance();
    tc.advance();
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
