// wfa_wave.hip -- wave-per-pair WFA (gap-affine and gap-linear).
//
// Each wave64 of a 256-thread workgroup owns one pair at a time and pulls the
// next pair from the device work queue on its own; the four waves never
// synchronise with each other.  Two storage tiers share one implementation:
//
//  * LDS tier (short pairs, qn + tn <= 2048): the packed query and target, a
//    ring of the last R score frames (int16 offsets, 128 diagonals, two per
//    lane), per-step cell offsets and 16-bit history cells live in the wave's
//    LDS region.  A pair that outgrows it reports an overflow status.
//  * HBM tier (any length): ring (int32, width a power of two >= qn + tn + 1)
//    and history (32-bit cells) live in the wave's HBM slot; sequences are
//    read through the caches.  When the history slot is full the wave keeps
//    going without recording history, so the pair returns the exact number of
//    score steps it needs and the host re-runs it in a slot of exactly that
//    size.
//
// History cell: bits 31..4 (15..4 in LDS) = pre-extension offset + 1 (0 =
// unset), bits 3..0 = ops (affine: m_op | ins_op << 2 | del_op << 3; linear:
// 1 X, 2 query-only, 3 target-only).  Every M run of the upstream traceback is
// "current offset - pre-extension offset of the cell", so nothing else is kept.
//
// Diagonal ranges are arithmetic: every diagonal shift costs at least one
// gap-extend unit (one score step of g = gcd(x, o, e)), so after u steps all
// set cells lie in |k| <= u.  Step u computes and writes every k in
// [max(-u, -tn), min(u, qn)] (-1 = unset), so a ring read is valid exactly when
// the predecessor step's range holds the diagonal; cells outside the reachable
// set come out unset, as in upstream's full-width frames.
//
// Recurrences, candidate orders and tie rules follow
// phraya-align/src/wfa_simd.rs: linear :283-330 (last maximum wins, bounds
// check after the choice), fitting end :940-955, linear traceback :959-1130,
// affine :581-779 (strict-greater first candidate), affine traceback :785-886.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "phx_device.h"
#include "wfa_common.h"

namespace phx {

constexpr int WV_W = 128;    // LDS tier ring width (diagonals)
constexpr int WV_WAVES = 4;  // waves (pairs) per workgroup

// Orders one wave's memory traffic between steps when the wave owns its pair
// alone.  A wave's own LDS and global accesses are performed in program
// order, so a wavefront-scope fence (no waits; it only keeps the compiler
// from reordering) is all a lane needs to see what another lane of the same
// wave stored in an earlier step, in LDS or in its HBM slot alike.
template <bool LDS>
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup barrier that orders LDS only: global stores issued before it may
// still be in flight (they are not read by other waves until a full barrier).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Synchronisation between the threads of one pair per score step: a
// wave-level fence for a wave-per-pair launch; a workgroup barrier when NT
// threads share the pair, ordering LDS only when the ring lives in LDS (RL),
// else LDS and global memory.
template <bool LDS, int NT, bool RL = false>
__device__ __forceinline__ void pair_sync() {
  if (NT == 64) wave_sync<LDS>();
  else if (RL) lds_barrier();
  else __syncthreads();
}

__device__ __forceinline__ int32_t readlane(int32_t v, int32_t lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

__device__ __forceinline__ int32_t step_lo(int32_t u, int32_t tn) { return max(-u, -tn); }
__device__ __forceinline__ int32_t step_hi(int32_t u, int32_t qn) { return min(u, qn); }

// Diagonals whose cells of step s are computed and stored.  Without a bound
// (bound = 0): the step's whole range.  With an upper bound B on the
// optimum (score steps), a cell of step s on diagonal k lies on an
// alignment of score <= B g only if the gap bases still needed to reach an
// end diagonal fit: e |k - kend| <= (B - s) g, i.e. |k - kend| <= (B - s) /
// en (global; fitting ends lie on any diagonal >= kend, so only k < kend
// is limited).  A cell outside this band is on no optimal alignment, and
// neither is any cell whose value it would decide: a candidate that attains
// a cell's maximum, ties included, extends that cell's optimal alignments
// backwards.  So the cells of every optimal alignment, their values and
// their tie-breaks are unchanged, and so are the score, the end and the
// traceback.  If B is below the optimum the band empties before the end is
// reached (lo > hi), and the pair is re-run without a bound.
struct Band {
  int32_t bound;  // 0 = none
  int32_t kend;   // qn - tn
  int32_t en;     // gap extend in score steps
  bool fitting;
};

__device__ __forceinline__ void band_range(const Band& b, int32_t s, int32_t qn, int32_t tn, int32_t& lo,
                                           int32_t& hi) {
  lo = step_lo(s, tn);
  hi = step_hi(s, qn);
  if (b.bound > 0) {
    if (s > b.bound) {
      lo = 1;
      hi = 0;
      return;
    }
    const int32_t r = (b.bound - s) / b.en;
    lo = max(lo, b.kend - r);
    if (!b.fitting) hi = min(hi, b.kend + r);
  }
}

template <bool IN_LDS, class T>
using mem_t = typename std::conditional<IN_LDS, lmem_t<T>, gmem_t<T>>::type;

// Element `idx` of a wave-uniform array through a 32-bit byte offset, so a
// global access is "uniform base + 32-bit lane offset" (no per-lane 64-bit
// address arithmetic).  idx * sizeof(T) stays below 4 GiB for every ring
// frame and history step.
template <class T>
__device__ __forceinline__ gmem_t<T>& elem(gmem_t<T>* base, uint32_t idx) {
  return *(gmem_t<T>*)((gmem_t<char>*)base + idx * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ const gmem_t<T>& elem(const gmem_t<T>* base, uint32_t idx) {
  return *(const gmem_t<T>*)((const gmem_t<char>*)base + idx * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ lmem_t<T>& elem(lmem_t<T>* base, uint32_t idx) {
  return base[idx];
}
template <class T>
__device__ __forceinline__ const lmem_t<T>& elem(const lmem_t<T>* base, uint32_t idx) {
  return base[idx];
}

// A wave-uniform 64-bit pointer held in scalar registers.
template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const uint64_t v = (uint64_t)(uintptr_t)p;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(v >> 32));
  return (T*)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

// Storage of one pair: all in LDS (LDS tier), all in HBM, or (RL) ring and
// sequences in LDS with history in HBM.
template <bool LDS, bool RL = false>
struct WaveStore {
  using ring_v = typename std::conditional<LDS, int16_t, int32_t>::type;
  using cell_v = typename std::conditional<LDS, uint16_t, uint32_t>::type;
  using ring_t = mem_t<LDS || RL, ring_v>;
  using cell_t = mem_t<LDS, cell_v>;
  using seq_t = mem_t<LDS || RL, const uint64_t>;
  ring_t* ring;
  cell_t* cells;
  cell_t* toff_base;  // LDS: toff[step]; HBM: entries grow down from toff_base
  uint64_t cap;       // LDS: cells; HBM: 32-bit words of the slot
  uint32_t steps_cap; // LDS only
  uint32_t wmask;
  seq_t* qw;
  seq_t* tw;
  uint32_t qsh, tsh;  // index of symbol 0 in word 0

  __device__ __forceinline__ cell_t& toff(int32_t step) const {
    return LDS ? toff_base[step] : *(toff_base - 1 - step);
  }
  // Room for `width` more cells (and the step's offset entry)?
  __device__ __forceinline__ bool fits(int32_t step, uint64_t used, int32_t width) const {
    if (LDS) return (uint32_t)step < steps_cap && used + (uint64_t)width <= cap;
    return used + (uint64_t)width + (uint64_t)step + 1 <= cap;
  }
};

template <bool B8>
__device__ __forceinline__ uint32_t seq_words(int32_t n) {
  return B8 ? ((uint32_t)n + 7) / 8 + 1 : ((uint32_t)n + 31) / 32 + 1;
}

template <bool LDS, bool B8, int NT, bool RL = false>
__device__ __forceinline__ WaveStore<LDS, RL> make_store(const WfaArgs& a, const WaveGeom& G,
                                                     uint8_t* lds_base, const PairDesc& pd,
                                                     int lane) {
  using S_t = WaveStore<LDS, RL>;
  S_t S;
  if constexpr (LDS) {
    S.ring = (typename S_t::ring_t*)(lds_base + G.ring_off);
    // Both sequences are re-packed to start at bit 0 of their first word (a
    // funnel shift per word, the last word the zero pad), so the step loop's
    // extension needs no start offsets.  Reads stay inside the source
    // words: data word j takes source words j and j + 1 <= the last one.
    lmem_t<uint64_t>* seq = (lmem_t<uint64_t>*)(lds_base + G.seq_off);
    const uint32_t nq = seq_words<B8>(pd.qn), nt = seq_words<B8>(pd.tn);
    const uint32_t qb = pd.q_shift * (B8 ? 8u : 2u), tb = pd.t_shift * (B8 ? 8u : 2u);
    const gmem_t<const uint64_t>* gq = (const gmem_t<const uint64_t>*)pd.q;
    const gmem_t<const uint64_t>* gt = (const gmem_t<const uint64_t>*)pd.t;
    for (uint32_t i = (uint32_t)lane; i < nq + nt; i += 64) {
      const bool isq = i < nq;
      const uint32_t j = isq ? i : i - nq, n = isq ? nq : nt, b = isq ? qb : tb;
      const gmem_t<const uint64_t>* src = isq ? gq : gt;
      uint64_t v = 0;
      if (j + 1 < n) {
        const uint64_t lo = src[j];
        v = b == 0 ? lo : (lo >> b) | (src[j + 1] << (64 - b));
      }
      seq[i] = v;
    }
    S.qw = seq;
    S.tw = seq + nq;
    S.toff_base = (typename S_t::cell_t*)(lds_base + G.tab_off);
    S.cells = (typename S_t::cell_t*)(lds_base + G.cell_off);
    S.cap = G.cells_cap;
    S.steps_cap = G.steps_cap;
    S.wmask = WV_W - 1;
    S.qsh = 0;
    S.tsh = 0;
    wave_sync<true>();
    return S;
  } else {
    // Storage of the pair's own (request server), else one HBM slot per pair
    // in flight: per wave, or per workgroup when the whole workgroup shares a
    // pair.
    uint8_t* h;
    uint64_t hbytes;
    if (pd.hist != nullptr) {
      S.ring = (typename S_t::ring_t*)pd.ring;
      h = pd.hist;
      hbytes = pd.hist_bytes;
      S.wmask = (1u << pd.ring_log2) - 1;
    } else {
      // (wave kernels hold WV_WAVES pairs per workgroup; a workgroup kernel,
      // RL included the one-wave register tier, one)
      const uint32_t wv = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(threadIdx.x >> 6));
      const uint64_t gw = NT == 64 && !RL ? (uint64_t)blockIdx.x * WV_WAVES + wv : (uint64_t)blockIdx.x;
      S.ring = (typename S_t::ring_t*)(a.ring_g + gw * a.ring_stride);
      h = a.hist + gw * a.hist_stride;
      hbytes = a.hist_stride;
      S.wmask = (1u << a.wcap_log2) - 1;
    }
    S.cells = (typename S_t::cell_t*)h;
    S.toff_base = (typename S_t::cell_t*)(h + hbytes);
    S.cap = hbytes / 4;
    S.steps_cap = 0;
    if constexpr (RL) {
      // Ring (int32, 1 << wcap_log2 diagonals) and both sequences in the
      // workgroup's LDS; history stays in the HBM slot.
      S.ring = (typename S_t::ring_t*)(lds_base + G.ring_off);
      S.wmask = (1u << a.wcap_log2) - 1;
      lmem_t<uint64_t>* seq = (lmem_t<uint64_t>*)(lds_base + G.seq_off);
      const uint32_t nq = seq_words<B8>((int32_t)pd.q_shift + pd.qn);
      const uint32_t nt = seq_words<B8>((int32_t)pd.t_shift + pd.tn);
      const gmem_t<const uint64_t>* gq = (const gmem_t<const uint64_t>*)pd.q;
      const gmem_t<const uint64_t>* gt = (const gmem_t<const uint64_t>*)pd.t;
      for (uint32_t i = (uint32_t)lane; i < nq + nt; i += NT) seq[i] = i < nq ? gq[i] : gt[i - nq];
      S.qw = seq;
      S.tw = seq + nq;
      S.qsh = pd.q_shift;
      S.tsh = pd.t_shift;
      lds_barrier();
      return S;
    }
    S.qw = uniform_ptr((typename S_t::seq_t*)pd.q);
    S.tw = uniform_ptr((typename S_t::seq_t*)pd.t);
  }
  S.qsh = pd.q_shift;
  S.tsh = pd.t_shift;
  return S;
}

// Parallel search along a gap: lane j looks at the cell of step s - j*en on
// diagonal k + dir*j and reports whether the gap was opened there (op bit 0).
// Returns the gap bases up to and including the opening cell (64 and
// *opened = false when the opening lies beyond this window).
template <class S_t>
__device__ __forceinline__ int32_t gap_run(const S_t& S, const Band& band, int32_t s, int32_t k,
                                           int32_t dir, int32_t en, uint32_t bit, int32_t qn,
                                           int32_t tn, int lane, bool* opened) {
  const int32_t sj = s - lane * en;
  const int32_t kj = k + dir * lane;
  bool open = false;
  if (sj >= 1) {
    int32_t lo, hi;
    band_range(band, sj, qn, tn, lo, hi);
    if (kj >= lo && kj <= hi) {
      const uint32_t cell = S.cells[(uint64_t)S.toff(sj) + (uint64_t)(kj - lo)];
      open = ((cell >> bit) & 1u) == 0;
    } else {
      open = true;  // not on a valid path; the caller's checks report it
    }
  }
  const uint64_t mask = __ballot(open);
  *opened = mask != 0;
  return mask ? (int32_t)__builtin_ctzll(mask) + 1 : 64;
}

// Thread 0 stores the pair's result and, on an overflow, raises the launch's
// host-visible flag (a system-scope store to pinned host memory, so the host
// learns whether any pair must be re-run without a copy).
__device__ __forceinline__ void publish(const WfaArgs& a, PairOut* po, const PairOut& r, int lane) {
  if (lane == 0) {
    *po = r;
    if (r.status >= ST_OVERFLOW_HIST && a.overflow)
      __hip_atomic_store(a.overflow, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Endpoint reduction for multi-wave pairs: smallest ((k + tn) << 2 | layer),
// triple-buffered so a step can reset the next buffer before its barrier.
struct PairShared {
  int32_t key[3];
  uint32_t next[2];  // next work item, by parity of the item count
};

// --- register wavefront (LDS tier) -------------------------------------------
//
// The LDS tier keeps the wavefront in registers.  Phase 1: while the
// pair's diagonal range fits 64 diagonals, lane L owns diagonal kb + L for
// the whole pair and keeps the last score frames of its diagonal in
// registers; the k - 1 / k + 1 neighbours come from the adjacent lanes by a
// wave-wide DPP shift.  Phase 2: once the range outgrows that window, the
// frames move (one cross-lane permute per register) to a window of 128
// diagonals, two per lane (kb2 + L and kb2 + 64 + L); the neighbours across
// the two halves come from one lane read each.  A lane outside a step's
// range holds "no cell" for that step, which is exactly what the ring
// path's range checks produce, so no range checks are needed on reads.
// History cells go to the ring path's LDS layout (cells at toff[step] + k -
// lo), so the traceback is shared.  A pair whose range outgrows 128
// diagonals stops with ST_OVERFLOW_WIDTH (the ring path's limit as well) and
// is re-run on an HBM tier.  With register phases the LDS tier needs no LDS
// ring at all (the host leaves it out of the layout), so more waves fit a
// CU.

// Register frames hold P = offset + 1, with 0 for "no cell", so the lane
// shifts below can fill the wave's ends with 0 through the DPP bound
// control instead of first loading a -1 into the destination.

// Value of the lane below (k - 1); lane 0 reads 0.  Call with every lane
// active: a source lane outside EXEC also reads as 0.
__device__ __forceinline__ int32_t from_below(int32_t v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true);  // wave_shr:1
}
// Value of the lane above (k + 1); lane 63 reads 0.
__device__ __forceinline__ int32_t from_above(int32_t v) {
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true);  // wave_shl:1
}
// The same shifts with the end lane taking `edge` instead of 0 (the DPP
// move leaves a lane with no source lane unwritten, so it keeps `edge`).
__device__ __forceinline__ int32_t below_or(int32_t edge, int32_t v) {
  return __builtin_amdgcn_update_dpp(edge, v, 0x138, 0xF, 0xF, false);  // wave_shr:1
}
__device__ __forceinline__ int32_t above_or(int32_t edge, int32_t v) {
  return __builtin_amdgcn_update_dpp(edge, v, 0x130, 0xF, 0xF, false);  // wave_shl:1
}

// Neighbours over a window of CH chunks of 64 diagonals (chunk c, lane L =
// diagonal kb + 64 c + L): the lane below / above, across chunk edges too;
// 0 past either end of the window.
template <int CH>
__device__ __forceinline__ void neighbours(const int32_t (&v)[CH], int lane, int32_t (&below)[CH],
                                           int32_t (&above)[CH]) {
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    below[c] = from_below(v[c]);
    above[c] = from_above(v[c]);
  }
#pragma unroll
  for (int c = 0; c + 1 < CH; ++c) {
    const int32_t top = readlane(v[c], 63), bottom = readlane(v[c + 1], 0);
    above[c] = lane == 63 ? bottom : above[c];
    below[c + 1] = lane == 0 ? top : below[c + 1];
  }
}

// Loop state shared by the register phases and the ring loop.
struct StepState {
  int32_t step;
  uint64_t used;
  bool found;
  int32_t end_k, end_layer;
  int32_t status;
};

// First diagonal of a register window of n diagonals: centred on 0, moved
// down when the query is short (the range never passes qn) and up to -tn
// when the target is short (it never passes -tn).
__device__ __forceinline__ int32_t reg_window_base(int32_t n, int32_t qn, int32_t tn) {
  return max(-tn, min(-(n / 2), qn - (n - 1)));
}

// Moves a phase-1 register (lane L = diagonal kb + L) into the phase-2
// window (chunk c, lane L = diagonal kb2 + 64 c + L); d = kb - kb2 >= 0.
__device__ __forceinline__ void widen(int32_t v, int32_t d, int lane, int32_t (&out)[2]) {
  const int32_t s0 = lane - d, s1 = lane + 64 - d;
  const int32_t p0 = __builtin_amdgcn_ds_bpermute((s0 & 63) << 2, v);
  const int32_t p1 = __builtin_amdgcn_ds_bpermute((s1 & 63) << 2, v);
  out[0] = s0 >= 0 ? p0 : 0;
  out[1] = s1 <= 63 ? p1 : 0;
}

// Writes the toff entries of the completed steps 1 .. done (<= 127) held
// by the lanes (lane s: step s, and step 64 + s in the second register).
__device__ __forceinline__ void flush_toff(const WaveStore<true>& S, uint32_t t0, uint32_t t1,
                                           int32_t done, int lane) {
  if (lane >= 1 && lane <= done) S.toff(lane) = (uint16_t)t0;
  if (64 + lane <= done) S.toff(64 + lane) = (uint16_t)t1;
}

// Frames of the gap-affine register wavefront, per chunk.  The m layer keeps
// MD frames, each gap layer EN; I[c][EN - d] and D[c][EN - d] hold step
// s - d while step s is computed.  The m layer rotates when MD is small:
// step s writes its frame into slot R = (s - s0) mod MD, so step s - d sits
// in slot (R - d) mod MD and no frame moves between registers (the step
// loop is unrolled MD times, each copy with its R fixed).  Between phases,
// and always when MD is large, the layout is that of R = 0: M[c][MD - d]
// holds step s - d and the frames shift down one slot per step.
template <int CH, int MD, int EN>
struct AffineFrames {
  int32_t M[CH][MD];
  int32_t I[CH][EN];
  int32_t D[CH][EN];
};

constexpr bool rotate_frames(int md) { return md <= 5; }

template <int MD, int R>
__device__ __forceinline__ constexpr int mslot(int d) {
  return rotate_frames(MD) ? ((R - d) % MD + MD) % MD : MD - d;
}

// One gap-affine score step in registers over a window of CH chunks, with
// the m layer at rotation R.  XN, OEN, EN: mismatch, open + extend and
// extend costs in score steps.  Recurrence and tie rules as in wave_affine's
// ring loop.  Returns 0 to go on, 1 when the step's range leaves the window
// (step not taken; frames left in the R = 0 layout) and 2 when the pair is
// done (end reached, step limit or history full).
template <int R, int CH, int XN, int OEN, int EN, bool B8>
__device__ __forceinline__ int affine_step(const WaveStore<true>& S, int32_t qn, int32_t tn,
                                           int32_t kb, bool fitting, int32_t max_step, int lane,
                                           AffineFrames<CH, (XN > OEN ? XN : OEN), EN>& F,
                                           uint32_t& used, uint32_t& t0, uint32_t& t1,
                                           StepState& st) {
  constexpr int MD = XN > OEN ? XN : OEN;
  const int32_t s = st.step + 1;
  const int32_t lo = step_lo(s, tn), hi = step_hi(s, qn);
  if (lo < kb || hi > kb + 64 * CH - 1 || s > 64 * CH - 1) {
    if constexpr (rotate_frames(MD) && R != 0) {
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        int32_t t[MD];
#pragma unroll
        for (int d = 1; d <= MD; ++d) t[MD - d] = F.M[c][mslot<MD, R>(d)];
#pragma unroll
        for (int j = 0; j < MD; ++j) F.M[c][j] = t[j];
      }
    }
    return 1;
  }
  st.step = s;
  if (s > max_step) { st.status = ST_ALIGNMENT_FAILED; return 2; }
  const uint32_t width = (uint32_t)(hi - lo + 1);
  if ((uint32_t)s >= S.steps_cap || used + width > (uint32_t)S.cap) {
    st.status = ST_OVERFLOW_HIST;
    return 2;
  }
  if (s < 64) t0 = lane == s ? used : t0;
  else t1 = lane == s - 64 ? used : t1;
  int32_t om[CH], ie[CH], de[CH], vol[CH], vor[CH], vil[CH], vdr[CH], dummy[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    om[c] = F.M[c][mslot<MD, R>(OEN)];
    ie[c] = F.I[c][0];
    de[c] = F.D[c][0];
  }
  neighbours<CH>(om, lane, vol, vor);
  neighbours<CH>(ie, lane, vil, dummy);
  neighbours<CH>(de, lane, dummy, vdr);
  const int32_t kend_g = qn - tn;
  const int32_t ke_lo = max(kend_g, lo);  // fitting: any diagonal from qn - tn up
  // Global: the end cell is on diagonal qn - tn only; its chunk and lane.
  const int32_t ke_off = kend_g - kb;
  const bool ke_in = kend_g >= lo && kend_g <= hi;
  const uint32_t uq = (uint32_t)qn;
  uint64_t fmask = 0;
  int32_t fc = 0, flayer = -1;
#pragma unroll
  for (int c = 0; c < CH; ++c) {
    const int32_t k = kb + 64 * c + lane;
    const uint32_t kl = (uint32_t)(k - lo);
    const bool inr = kl <= width - 1;
    // Unsigned tests of P - 1 (the offset) against a bound also reject P = 0.
    // tn + k >= 0 on every diagonal of the range, which is all that is kept.
    const uint32_t utk = (uint32_t)(tn + k);
    const int32_t vx = F.M[c][mslot<MD, R>(XN)];
    int32_t ins = (uint32_t)(vol[c] - 1) < uq ? vol[c] + 1 : 0;
    uint32_t iop = 0;
    if ((uint32_t)(vil[c] - 1) < uq && vil[c] + 1 > ins) { ins = vil[c] + 1; iop = 1; }
    int32_t del = (uint32_t)(vor[c] - 1) <= utk ? vor[c] : 0;
    uint32_t dop = 0;
    if ((uint32_t)(vdr[c] - 1) <= utk && vdr[c] > del) { del = vdr[c]; dop = 1; }
    int32_t best = (uint32_t)(vx - 1) < min(uq, utk) ? vx + 1 : 0;
    uint32_t mop = 0;
    if (ins > best) { best = ins; mop = 1; }
    if (del > best) { best = del; mop = 2; }
    const int32_t mval = extend_diag<B8>(S.qw, S.qsh, S.tw, S.tsh, qn, tn, inr ? best - 1 : -1, k);
    if (inr)
      elem(S.cells, used + kl) = (uint16_t)(((uint32_t)best << 4) | mop | (iop << 2) | (dop << 3));
    if (fitting) {
      int32_t layer = -1;
      if (inr && k >= ke_lo) layer = mval == qn ? 0 : ins == qn + 1 ? 1 : del == qn + 1 ? 2 : -1;
      const uint64_t m = __ballot(layer >= 0);
      if (m && !fmask) {  // the lowest diagonal reaching the end
        fmask = m;
        fc = c;
        flayer = readlane(layer, (int32_t)__builtin_ctzll(m));
      }
    } else if (ke_in && (ke_off >> 6) == c && !fmask) {
      // Global: three lane reads of the end diagonal instead of a test on
      // every lane.
      const int32_t ln = ke_off & 63;
      const int32_t layer = readlane(mval, ln) == qn           ? 0
                            : readlane(ins, ln) == qn + 1 ? 1
                            : readlane(del, ln) == qn + 1 ? 2
                                                          : -1;
      if (layer >= 0) {
        fmask = 1ull << ln;
        fc = c;
        flayer = layer;
      }
    }
    if constexpr (rotate_frames(MD)) {
      F.M[c][R] = inr ? mval + 1 : 0;
    } else {
#pragma unroll
      for (int j = 0; j + 1 < MD; ++j) F.M[c][j] = F.M[c][j + 1];
      F.M[c][MD - 1] = inr ? mval + 1 : 0;
    }
#pragma unroll
    for (int j = 0; j + 1 < EN; ++j) { F.I[c][j] = F.I[c][j + 1]; F.D[c][j] = F.D[c][j + 1]; }
    F.I[c][EN - 1] = inr ? ins : 0;
    F.D[c][EN - 1] = inr ? del : 0;
  }
  used += width;
  if (fmask) {
    st.found = true;
    st.end_k = kb + 64 * fc + (int32_t)__builtin_ctzll(fmask);
    st.end_layer = flayer;
    return 2;
  }
  return 0;
}

// Steps R, R + 1, .. U - 1 of one pass of the unrolled step loop.
template <int R, int U, int CH, int XN, int OEN, int EN, bool B8>
__device__ __forceinline__ int affine_pass(const WaveStore<true>& S, int32_t qn, int32_t tn,
                                           int32_t kb, bool fitting, int32_t max_step, int lane,
                                           AffineFrames<CH, (XN > OEN ? XN : OEN), EN>& F,
                                           uint32_t& used, uint32_t& t0, uint32_t& t1,
                                           StepState& st) {
  const int r = affine_step<R, CH, XN, OEN, EN, B8>(S, qn, tn, kb, fitting, max_step, lane, F, used,
                                                    t0, t1, st);
  if constexpr (R + 1 < U) {
    if (r) return r;
    return affine_pass<R + 1, U, CH, XN, OEN, EN, B8>(S, qn, tn, kb, fitting, max_step, lane, F,
                                                      used, t0, t1, st);
  }
  return r;
}

// Gap-affine steps in registers over a window of CH chunks, until the pair
// is done (true) or the next step's range leaves the window (false; the
// state is then that of the last completed step, frames in the R = 0
// layout).
template <int CH, int XN, int OEN, int EN, bool B8>
__device__ bool affine_steps(const WaveStore<true>& S, int32_t qn, int32_t tn, int32_t kb,
                             bool fitting, int32_t max_step, int lane,
                             AffineFrames<CH, (XN > OEN ? XN : OEN), EN>& F, uint32_t& used,
                             uint32_t& t0, uint32_t& t1, StepState& st) {
  constexpr int MD = XN > OEN ? XN : OEN;
  constexpr int U = rotate_frames(MD) ? MD : 1;
  for (;;) {
    const int r = affine_pass<0, U, CH, XN, OEN, EN, B8>(S, qn, tn, kb, fitting, max_step, lane, F,
                                                         used, t0, t1, st);
    if (r) return r == 2;
  }
}

// The whole register wavefront of one gap-affine pair (phase 1, then phase
// 2 if needed).
template <int XN, int OEN, int EN, bool B8>
__device__ void affine_regs(const WaveStore<true>& S, int32_t qn, int32_t tn, int32_t m0,
                            bool fitting, int32_t max_step, int lane, StepState& st) {
  constexpr int MD = XN > OEN ? XN : OEN;
  st = StepState{0, 0, false, 0, 0, ST_OK};
  const int32_t kb = reg_window_base(64, qn, tn);
  AffineFrames<1, MD, EN> F;
#pragma unroll
  for (int j = 0; j < MD; ++j) F.M[0][j] = 0;
#pragma unroll
  for (int j = 0; j < EN; ++j) F.I[0][j] = F.D[0][j] = 0;
  F.M[0][MD - 1] = kb + lane == 0 ? m0 + 1 : 0;  // step 0, R = 0 layout
  uint32_t used = 0, t0 = 0, t1 = 0;
  bool done = affine_steps<1, XN, OEN, EN, B8>(S, qn, tn, kb, fitting, max_step, lane, F, used, t0,
                                               t1, st);
  if (!done) {
    const int32_t kb2 = reg_window_base(128, qn, tn);
    AffineFrames<2, MD, EN> F2;
    int32_t w[2];
#pragma unroll
    for (int j = 0; j < MD; ++j) { widen(F.M[0][j], kb - kb2, lane, w); F2.M[0][j] = w[0]; F2.M[1][j] = w[1]; }
#pragma unroll
    for (int j = 0; j < EN; ++j) {
      widen(F.I[0][j], kb - kb2, lane, w); F2.I[0][j] = w[0]; F2.I[1][j] = w[1];
      widen(F.D[0][j], kb - kb2, lane, w); F2.D[0][j] = w[0]; F2.D[1][j] = w[1];
    }
    done = affine_steps<2, XN, OEN, EN, B8>(S, qn, tn, kb2, fitting, max_step, lane, F2, used, t0,
                                            t1, st);
    if (!done) {  // wider than 128 diagonals: the ring limit of this tier
      st.step += 1;
      st.status = ST_OVERFLOW_WIDTH;
    }
  }
  st.used = used;
  flush_toff(S, t0, t1, st.status == ST_OK ? st.step : st.step - 1, lane);
  wave_sync<true>();
}

// Cost shapes with a register wavefront (reg_wavefront_shape); others run
// the LDS ring loop from step 1.
template <bool B8>
__device__ __forceinline__ bool affine_regs_any(const WaveStore<true>& S, int32_t qn, int32_t tn,
                                                int32_t m0, bool fitting, int32_t max_step,
                                                int32_t xn, int32_t oen, int32_t en, int lane,
                                                StepState& st) {
  if (xn == 2 && oen == 4 && en == 1)  // 4/6/2
    affine_regs<2, 4, 1, B8>(S, qn, tn, m0, fitting, max_step, lane, st);
  else if (xn == 2 && oen == 5 && en == 1)  // 2/4/1
    affine_regs<2, 5, 1, B8>(S, qn, tn, m0, fitting, max_step, lane, st);
  else if (xn == 1 && oen == 1 && en == 1)  // 1/0/1
    affine_regs<1, 1, 1, B8>(S, qn, tn, m0, fitting, max_step, lane, st);
  else if (xn == 3 && oen == 7 && en == 2)  // 3/5/2
    affine_regs<3, 7, 2, B8>(S, qn, tn, m0, fitting, max_step, lane, st);
  else
    return false;
  return true;
}

// --- register wavefront for workgroup tiers -------------------------------------
//
// A workgroup of NT threads (W = NT / 64 waves) keeps a pair's wavefront in
// registers over a window of C = W * J chunks of 64 diagonals: chunk c is
// diagonals kb + 64 c .. kb + 64 c + 63, lane L holding kb + 64 c + L, and
// chunk c belongs to wave c % W as its chunk j = c / W.  Frames as in the
// LDS tier (P = offset + 1, 0 = no cell); neighbours k - 1 / k + 1 come from
// the adjacent lanes by DPP, and at a chunk's ends from the neighbouring
// chunk's end lane, which the other wave published in LDS after the step
// that computed it (every value a step reads is from an earlier step, and a
// barrier ends each step).  The step range only grows, so a chunk not yet
// reached holds no cells and is skipped.  History cells go to the ring
// path's HBM layout, so the traceback is shared.  A pair whose range would
// leave the window stops with ST_OVERFLOW_WIDTH and is re-run on the ring
// path.
//
// The step body is written to keep scalar work low (the scalar unit, shared
// by the CU's waves, issues one instruction per cycle like the vector
// units): no lane-divergent branch on the common path, table reads without
// bounds tests, history stores whose disabled lanes the buffer unit drops.

// Published chunk-end values of the last MD + 1 steps, one entry of four
// ints {m at lane 0, d at lane 0, m at lane 63, i at lane 63} per (slot,
// chunk) with slot = step % (MD + 1); entries are indexed c + 1 and the
// entries of c = -1 and c = C stay zero, so the end chunks read "no cell"
// without a test.  Laid out after the staged sequences (MD + 1 = 6 slots
// for every shape with a register wavefront).

struct BlockEdges {
  lmem_t<int32_t>* tab;  // [slot][C + 2][4]
  int32_t C;
};

// Frames of J chunks, two 16-bit P values per register (P = offset + 1 <=
// qn + 1 < 65536: the host sends only queries of at most 65,534 bases):
// the m layer's MD frames in MR registers per chunk (frame d back, d = 1 ..
// MD, is half (d - 1) % 2 of register (d - 1) / 2, so a step shifts every
// register by 16 bits), then the gap layers (one frame back, EN = 1) in one
// register, i low, d high.  Register p of every chunk forms one vector of J
// elements, element j for chunk j.  The chunk loop is not unrolled (one copy
// of the step body keeps the kernel inside the instruction cache); its
// chunk index is wave-uniform, so an element is read and written by a
// relative register move, with no scratch.
template <int MD>
struct BlockShape {
  static constexpr int MR = (MD + 1) / 2;  // m-layer registers per chunk
};

template <int MD>
__device__ __forceinline__ int32_t frame_m(const uint32_t (&m)[(MD + 1) / 2], int d) {
  const uint32_t r = m[(d - 1) >> 1];
  return (int32_t)(((d - 1) & 1) ? r >> 16 : r & 0xFFFFu);
}

template <int MD>
__device__ __forceinline__ void push_m(uint32_t (&m)[(MD + 1) / 2], uint32_t v) {
#pragma unroll
  for (int p = (MD + 1) / 2 - 1; p >= 1; --p) m[p] = __builtin_amdgcn_alignbit(m[p], m[p - 1], 16);
  m[0] = (m[0] << 16) | v;
}

// extend_diag over sequences staged in LDS with a first 32-bit window (16
// bases, or 4 bytes) compared by every lane without a branch (a lane with
// nothing to extend reads word 0 and discards it): most cells stop within
// it, and only when some lane's match runs past it does the wave enter
// extend_diag's loop for the rest.
template <bool B8>
__device__ __forceinline__ int32_t extend_first(const lmem_t<const uint64_t>* qw, uint32_t qsh,
                                                const lmem_t<const uint64_t>* tw, uint32_t tsh, int32_t qn,
                                                int32_t tn, int32_t p, int32_t k) {
  constexpr uint32_t BITS = B8 ? 8 : 2;
  constexpr uint32_t CTZ_SHIFT = B8 ? 3 : 1;
  constexpr int32_t PER = B8 ? 4 : 16;
  const int32_t j = p - k;
  const bool live = (uint32_t)p < (uint32_t)qn && (uint32_t)j < (uint32_t)tn;
  const int32_t lim = min(qn - p, tn - j);
  const uint32_t qb = live ? (qsh + (uint32_t)p) * BITS : 0u, tb = live ? (tsh + (uint32_t)j) * BITS : 0u;
  const lmem_t<const uint32_t>* q32 = (const lmem_t<const uint32_t>*)qw + (qb >> 5);
  const lmem_t<const uint32_t>* t32 = (const lmem_t<const uint32_t>*)tw + (tb >> 5);
  const uint32_t qx = __builtin_amdgcn_alignbit(q32[1], q32[0], qb);  // shift = qb & 31
  const uint32_t tx = __builtin_amdgcn_alignbit(t32[1], t32[0], tb);
  const uint32_t z = (uint32_t)__builtin_ctzg(qx ^ tx, 32);  // 32 when all match
  const int32_t n = (int32_t)(z >> CTZ_SHIFT);
  const bool more = live && z == 32u && lim > PER;
  if (__builtin_amdgcn_ballot_w64(more) == 0) return live ? p + min(n, lim) : p;
  if (more) return extend_diag<B8>(qw, qsh, tw, tsh, qn, tn, p + PER, k);
  return live ? p + min(n, lim) : p;
}

template <int NT, int J, int XN, int OEN, int EN, bool B8>
__device__ void affine_block_regs(const WaveStore<false, true>& S, const BlockEdges& E, const Band& band,
                                  int32_t qn, int32_t tn, int32_t m0, bool fitting, int32_t max_step, int tid,
                                  PairShared* sh, bool hist_cap_ok, StepState& st) {
  constexpr int W = NT / 64, MD = XN > OEN ? XN : OEN, SL = MD + 1;
  constexpr int MR = BlockShape<MD>::MR;
  static_assert(EN == 1, "the gap layers keep one frame");
  using frames_t = uint32_t __attribute__((ext_vector_type(J)));
  constexpr int32_t NDIAG = 64 * W * J;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // With a bound the computed cells of every step lie within (bound + 1) /
  // en diagonals centred on kend / 2 (between the start diagonal 0 and the
  // end diagonal kend), so the window is centred there.
  const int32_t kend_g = qn - tn;
  const int32_t kb = band.bound > 0
                         ? min(0, max(-(NDIAG - 1), max(-tn, min(kend_g / 2 - NDIAG / 2, qn - (NDIAG - 1)))))
                         : reg_window_base(NDIAG, qn, tn);
  const int32_t C = E.C, CE = C + 2;  // table entries per slot
  lmem_t<int32_t>* const tab = E.tab;
  frames_t FM[MR], FID;
#pragma unroll
  for (int p = 0; p < MR; ++p) FM[p] = (frames_t)0u;
  FID = (frames_t)0u;
  const int32_t d0 = 0 - kb, c0 = d0 >> 6, l0 = d0 & 63;
  if (c0 % W == w) {  // step 0 (m0 on diagonal 0), one frame back from step 1
    const int32_t j0 = __builtin_amdgcn_readfirstlane(c0 / W);
    FM[0][j0] = lane == l0 ? (uint32_t)(m0 + 1) : 0u;
  }
  // The table starts empty, then holds step 0's chunk ends.
  for (int32_t i = tid; i < SL * CE * 4; i += NT) tab[i] = 0;
  __syncthreads();
  if (tid == 0) {
    if (l0 == 0) tab[(c0 + 1) * 4 + 0] = m0 + 1;
    if (l0 == 63) tab[(c0 + 1) * 4 + 2] = m0 + 1;
    sh->key[0] = sh->key[1] = sh->key[2] = INT32_MAX;
  }
  __syncthreads();
  st = StepState{0, 0, false, 0, 0, ST_OK};
  uint64_t used = 0;
  bool hist_on = hist_cap_ok;
  const uint32_t uq = (uint32_t)qn;
  const lmem_t<const uint64_t>* qw = S.qw;
  const lmem_t<const uint64_t>* tw = S.tw;
  // History stores through a buffer resource over the rest of the slot: a
  // lane with no cell to store gets an offset past its end, which the
  // buffer unit drops.
  constexpr uint32_t kDrop = 0x80000000u;
  for (;;) {
    const int32_t s = __builtin_amdgcn_readfirstlane(st.step + 1);
    int32_t lo, hi;  // the step's computed diagonals (its band)
    band_range(band, s, qn, tn, lo, hi);
    // The window is too narrow from here on, or the band is empty (the bound
    // is below the optimum): the pair is re-run on the ring path.
    if (lo > hi || lo < kb || hi > kb + NDIAG - 1) {
      st.status = ST_OVERFLOW_WIDTH;
      st.step = s;
      break;
    }
    st.step = s;
    if (s > max_step) { st.status = ST_ALIGNMENT_FAILED; break; }
    const int32_t width = hi - lo + 1;
    if (hist_on && !S.fits(s, used, width)) hist_on = false;  // go on to learn the step count
    if (hist_on && tid == 0) S.toff(s) = (uint32_t)used;
    if (tid == 0) sh->key[(s + 1) % 3] = INT32_MAX;
    const uint64_t hb = (uint64_t)(uintptr_t)(S.cells + used);
    const uint32_t hb_lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)hb);
    const uint32_t hb_hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)(hb >> 32));
    const __amdgpu_buffer_rsrc_t hrs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(uintptr_t)(((uint64_t)hb_hi << 32) | hb_lo), (short)0, width * 4, 0x00020000);
    // Table rows of steps s - OEN and s - EN (s >= 1; a step before 0 maps
    // to a row no step has written yet, which reads as no cell) and of s.
    const int32_t ro = ((s + SL - OEN) % SL) * CE, re = ((s + SL - EN) % SL) * CE, rw = (s % SL) * CE;
    // This wave's chunks c = w + W j that the step range reaches (both
    // ends lie in the window, checked above).  A wave with none runs one
    // chunk outside the range (no cell changes: every lane is out of
    // range), so the loop always runs: a loop that may run zero times makes
    // the compiler copy every frame vector each step.
    const int32_t c_lo = (lo - kb) >> 6, c_hi = (hi - kb) >> 6;
    // With a bound the band also shrinks, and a chunk it has left would
    // keep its earlier ends in the table: the chunks next to the range get
    // "no cell" ends in this step's row (entries c_lo and c_hi + 2, inside
    // the row: entries 0 and C + 1 are the row's permanent zero entries).
    // Their frames need nothing: a chunk's steps inside the band are one
    // interval (lo(s) falls then rises, hi(s) rises then falls), so a
    // chunk the band reaches again later does not exist, and one it has not
    // reached yet still holds zero frames.
    if (band.bound > 0 && w == 0 && lane < 8)
      tab[(rw + (lane < 4 ? c_lo : c_hi + 2)) * 4 + (lane & 3)] = 0;
    int32_t j = c_lo <= w ? 0 : (c_lo - w + W - 1) / W;
    int32_t j_end = c_hi < w ? -1 : (c_hi - w) / W;
    if (j > j_end) {
      j = min(j, J - 1);
      j_end = j;
    }
    const int32_t ke_lo = fitting ? max(kend_g, lo) : kend_g;
    const int32_t ke_hi = fitting ? hi : kend_g;
    do {
      const int32_t c = w + W * j;
      uint32_t m[MR];
#pragma unroll
      for (int p = 0; p < MR; ++p) m[p] = FM[p][j];
      const uint32_t idr = FID[j];
      const int32_t om = frame_m<MD>(m, OEN);
      const int32_t ie = (int32_t)(idr & 0xFFFFu), de = (int32_t)(idr >> 16);
      // Chunk ends: the neighbouring chunks' end lanes, published earlier
      // (entry c + 1 of the row; c - 1 and c + 1 are entries c and c + 2),
      // taken by lane 0 / lane 63 in the DPP shifts.
      const int32_t eb_m = tab[(ro + c) * 4 + 2], ea_m = tab[(ro + c + 2) * 4 + 0];
      const int32_t eb_i = tab[(re + c) * 4 + 3], ea_d = tab[(re + c + 2) * 4 + 1];
      const int32_t vol = below_or(eb_m, om), vor = above_or(ea_m, om);
      const int32_t vil = below_or(eb_i, ie), vdr = above_or(ea_d, de);
      const int32_t k = kb + 64 * c + lane;
      const uint32_t kl = (uint32_t)(k - lo);
      const bool inr = kl <= (uint32_t)(width - 1);
      const uint32_t utk = (uint32_t)(tn + k);
      const int32_t vx = frame_m<MD>(m, XN);
      int32_t ins = (uint32_t)(vol - 1) < uq ? vol + 1 : 0;
      uint32_t iop = 0;
      if ((uint32_t)(vil - 1) < uq && vil + 1 > ins) { ins = vil + 1; iop = 1; }
      int32_t del = (uint32_t)(vor - 1) <= utk ? vor : 0;
      uint32_t dop = 0;
      if ((uint32_t)(vdr - 1) <= utk && vdr > del) { del = vdr; dop = 1; }
      int32_t best = (uint32_t)(vx - 1) < min(uq, utk) ? vx + 1 : 0;
      uint32_t mop = 0;
      if (ins > best) { best = ins; mop = 1; }
      if (del > best) { best = del; mop = 2; }
      const int32_t mval = extend_first<B8>(qw, S.qsh, tw, S.tsh, qn, tn, inr ? best - 1 : -1, k);
      __builtin_amdgcn_raw_buffer_store_b32(((uint32_t)best << 4) | mop | (iop << 2) | (dop << 3), hrs,
                                            inr && hist_on ? kl * 4u : kDrop, 0, 0);
      // End test: the chunk holds a diagonal of [ke_lo, ke_hi] (wave-uniform).
      const int32_t kc = kb + 64 * c;
      if (kc <= ke_hi && kc + 63 >= ke_lo) {
        if (inr && k >= ke_lo && k <= ke_hi) {
          const int32_t layer = mval == qn ? 0 : ins == qn + 1 ? 1 : del == qn + 1 ? 2 : -1;
          if (layer >= 0) atomicMin(&sh->key[s % 3], ((k + tn) << 2) | layer);
        }
      }
      const int32_t nm = inr ? mval + 1 : 0, ni = inr ? ins : 0, nd = inr ? del : 0;
      push_m<MD>(m, (uint32_t)nm);
#pragma unroll
      for (int p = 0; p < MR; ++p) FM[p][j] = m[p];
      FID[j] = (uint32_t)ni | ((uint32_t)nd << 16);
      if (lane == 0 || lane == 63) {
        const bool top = lane == 63;
        lmem_t<uint64_t>* dst = (lmem_t<uint64_t>*)(tab + (rw + c + 1) * 4 + (top ? 2 : 0));
        *dst = (uint64_t)(uint32_t)nm | ((uint64_t)(uint32_t)(top ? ni : nd) << 32);
      }
    } while (++j <= j_end);
    used += (uint64_t)width;
    lds_barrier();  // this step's chunk ends and end key, for every wave
    const int32_t key = __builtin_amdgcn_readfirstlane(sh->key[s % 3]);
    if (key != INT32_MAX) {
      st.found = true;
      st.end_k = (key >> 2) - tn;
      st.end_layer = key & 3;
      break;
    }
  }
  if (!hist_on && st.status == ST_OK) st.status = ST_OVERFLOW_HIST;
  st.used = used;
  __syncthreads();  // every wave's history stores, for the traceback
}

// Workgroup register wavefront for the cost shapes that have one (shared
// with the LDS tier); false for others (the ring path runs).
template <int NT, int J, bool B8>
__device__ __forceinline__ bool affine_block_regs_any(const WaveStore<false, true>& S, const BlockEdges& E,
                                                      const Band& band, int32_t qn, int32_t tn, int32_t m0,
                                                      bool fitting, int32_t max_step, int32_t xn, int32_t oen,
                                                      int32_t en, int tid, PairShared* sh, StepState& st) {
  if (xn == 2 && oen == 4 && en == 1)  // 4/6/2
    affine_block_regs<NT, J, 2, 4, 1, B8>(S, E, band, qn, tn, m0, fitting, max_step, tid, sh, true, st);
  else if (xn == 2 && oen == 5 && en == 1)  // 2/4/1
    affine_block_regs<NT, J, 2, 5, 1, B8>(S, E, band, qn, tn, m0, fitting, max_step, tid, sh, true, st);
  else
    return false;
  return true;
}

// Gap-linear steps in registers over a window of CH chunks (one frame
// back, P = offset + 1 as above); recurrence, bounds and tie rules as in
// wave_linear's ring loop.
template <int CH, bool B8>
__device__ bool linear_steps(const WaveStore<true>& S, int32_t qn, int32_t tn, int32_t kb,
                             bool fitting, int32_t explore, int lane, int32_t (&Pf)[CH],
                             uint32_t& used, uint32_t& t0, uint32_t& t1, StepState& st,
                             bool& exhausted) {
  const int32_t kend_global = qn - tn;
  for (;;) {
    const int32_t s = st.step + 1;
    const int32_t lo = step_lo(s, tn), hi = step_hi(s, qn);
    if (lo < kb || hi > kb + 64 * CH - 1 || s > 64 * CH - 1) return false;
    st.step = s;
      if (s > explore) { exhausted = true; return true; }
    const uint32_t width = (uint32_t)(hi - lo + 1);
    if ((uint32_t)s >= S.steps_cap || used + width > (uint32_t)S.cap) {
      st.status = ST_OVERFLOW_HIST;
      return true;
    }
    if (s < 64) t0 = lane == s ? used : t0;
    else t1 = lane == s - 64 ? used : t1;
    int32_t below[CH], above[CH];
    neighbours<CH>(Pf, lane, below, above);
    const int32_t ke_lo = fitting ? max(qn - tn, lo) : kend_global;
    const int32_t ke_hi = fitting ? hi : kend_global;
    uint64_t fmask = 0;
    int32_t fc = 0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      const int32_t k = kb + 64 * c + lane;
      const uint32_t kl = (uint32_t)(k - lo);
      const bool inr = kl <= width - 1;
      const int32_t vm = Pf[c];
      const int32_t vl = k > -tn ? below[c] : 0;
      const int32_t vr = k < qn ? above[c] : 0;
      int32_t best = 0;
      uint32_t op = 0;
      if (vm > 0) { best = vm + 1; op = 1; }
      if (vl > 0 && vl + 1 >= best) { best = vl + 1; op = 2; }
      if (vr > 0 && vr >= best) { best = vr; op = 3; }
      const bool ok = op != 0 && best <= qn + 1 && (uint32_t)(best - 1 - k) <= (uint32_t)tn;
      const int32_t val =
          extend_diag<B8>(S.qw, S.qsh, S.tw, S.tsh, qn, tn, inr && ok ? best - 1 : -1, k);
      if (inr) elem(S.cells, used + kl) = (uint16_t)(ok ? (((uint32_t)best << 4) | op) : 0u);
      const bool hit = inr && ok && k >= ke_lo && k <= ke_hi && val >= qn;
      Pf[c] = inr ? val + 1 : 0;
      const uint64_t m = __ballot(hit);
      if (m && !fmask) {
        fmask = m;
        fc = c;
      }
    }
    used += width;
    if (fmask) {
      st.found = true;
      st.end_k = kb + 64 * fc + (int32_t)__builtin_ctzll(fmask);
      return true;
    }
  }
}

// The whole register wavefront of one gap-linear pair.
template <bool B8>
__device__ void linear_regs(const WaveStore<true>& S, int32_t qn, int32_t tn, int32_t w0,
                            bool fitting, int32_t explore, int lane, StepState& st, bool& exhausted) {
  st = StepState{0, 0, false, 0, 0, ST_OK};
  const int32_t kb = reg_window_base(64, qn, tn);
  int32_t Pf[1] = {kb + lane == 0 ? w0 + 1 : 0};
  uint32_t used = 0, t0 = 0, t1 = 0;
  bool done =
      linear_steps<1, B8>(S, qn, tn, kb, fitting, explore, lane, Pf, used, t0, t1, st, exhausted);
  if (!done) {
    const int32_t kb2 = reg_window_base(128, qn, tn);
    int32_t P2[2];
    widen(Pf[0], kb - kb2, lane, P2);
    done = linear_steps<2, B8>(S, qn, tn, kb2, fitting, explore, lane, P2, used, t0, t1, st,
                               exhausted);
    if (!done) {
      st.step += 1;
      st.status = ST_OVERFLOW_WIDTH;
    }
  }
  st.used = used;
  flush_toff(S, t0, t1, st.status == ST_OK && !exhausted ? st.step : st.step - 1, lane);
  wave_sync<true>();
}

// --- HBM ring without wrap-around ----------------------------------------------
//
// An HBM ring (wave and workgroup tiers whose ring is not in LDS) is wider
// than any pair it holds: W >= qn + tn + 2 * frames + 3 (host sizing), so
// diagonal k sits at index k + tn + M with no wrap, where M is at least every
// lookback + 1.  Each frame written at step f also gets -1 in the M entries
// beyond both ends of its range [lo(f), hi(f)]; a later step reads frame f at
// most lookback + 1 diagonals beyond that range, so out-of-range reads come
// out -1 by themselves and the per-cell range checks of the wrapped rings are
// not needed.  At pair start the neighbourhood of diagonal 0 of every frame
// is set to -1 (the frames of steps before 0 and step 0's margins).
template <int NT>
__device__ __forceinline__ void ring_fill_origin(int32_t* ring, uint32_t W, int32_t nframes,
                                                 int32_t span, int32_t koff, int32_t qn, int32_t tn,
                                                 int32_t M, int lane) {
  const int32_t klo = max(-span, -tn - M), khi = min(span, qn + M);
  const int32_t n = khi - klo + 1;
  for (int32_t t = lane; t < n * nframes; t += NT) {
    const int32_t fr = t / n, k = klo + t % n;
    ring[(size_t)fr * W + (uint32_t)(k + koff)] = -1;
  }
}

// --- affine --------------------------------------------------------------------

// J > 0 (workgroup tiers, RL layout without a ring): the register
// wavefront of J chunks per wave (affine_block_regs) instead of the ring loop;
// the host launches it only for cost shapes that have one.
template <bool LDS, bool B8, int NT, bool RL = false, int J = 0>
__device__ __forceinline__ void wave_affine(const WfaArgs& a, const WaveGeom& G, uint8_t* lds_base,
                            const PairDesc& pd, PairOut* po, int lane, PairShared* sh) {
  using S_t = WaveStore<LDS, RL>;
  const int32_t qn = pd.qn, tn = pd.tn;
  const bool forced_global = (pd.flags & PF_GLOBAL) != 0;
  if (qn == 0) {
    // Wrapper (wfa_simd.rs:925-927): empty query -> empty alignment.  Forced
    // global (fill_wfa_affine_generic with fitting = false, :594-601): every
    // target base is an I run costing gap_open + tn * gap_extend.
    if (lane == 0) {
      if (forced_global && tn > 0) {
        a.runs[pd.run_off + pd.run_cap - 1] = ((uint32_t)tn << 4) | OP_I;
        *po = PairOut{ST_OK, a.cp.o + tn * a.cp.e, tn, tn, 1, (int32_t)pd.run_cap - 1, 0, 0};
      } else {
        *po = PairOut{ST_OK, 0, 0, 0, 0, 0, 0, 0};
      }
    }
    return;
  }
  const S_t S = make_store<LDS, B8, NT, RL>(a, G, lds_base, pd, lane);
  const int32_t x = a.cp.x, o = a.cp.o, e = a.cp.e;
  const bool fitting = !forced_global && tn > qn + qn / 2 + 10;
  const int32_t g = a.cp.g, xn = a.cp.xn, oen = a.cp.oen, en = a.cp.en;
  // Band pruning: register-wavefront workgroup launches that ask for it
  // (the forward pass and the traceback use the same band).
  const Band band{J > 0 && a.band ? (int32_t)pd.bound : 0, qn - tn, en, fitting};
  // The m layer reads xn and oen steps back, the gap layers en steps back:
  // Rm and Rg frames of ring (the host sizes the ring as Rm + 2 Rg frames).
  const int32_t Rm = max(xn, oen) + 1, Rg = en + 1;
  const int64_t hard_max = (int64_t)(qn + tn) * max(x, 1) + o + e + 4;
  const int64_t explore_to = a.cp.cap >= 0 ? min((int64_t)a.cp.cap, hard_max) : hard_max;
  // A 32-bit division when the bound fits (it does for every pair an LDS or
  // HBM slot holds); the quotient is the last score step explored.
  const int32_t max_step = explore_to <= INT32_MAX ? (int32_t)explore_to / g
                                                   : (int32_t)min<int64_t>(explore_to / g, INT32_MAX);
  const uint32_t W = S.wmask + 1;
  typename S_t::ring_t* Mr = S.ring;
  typename S_t::ring_t* Ri = S.ring + (size_t)Rm * W;
  typename S_t::ring_t* Rd = Ri + (size_t)Rg * W;
  // HBM ring: unwrapped with -1 margins of M = Rm entries (ring_fill_origin).
  constexpr bool UNW = !LDS && !RL;
  const int32_t M = Rm, koff = tn + Rm;
  auto rix = [&](int32_t k) -> uint32_t {
    return UNW ? (uint32_t)(k + koff) : ((uint32_t)k & S.wmask);
  };
  if constexpr (UNW) {
    ring_fill_origin<NT>((int32_t*)S.ring, W, Rm + 2 * Rg, M + Rm + 1, koff, qn, tn, M, lane);
    pair_sync<LDS, NT, RL>();
  }
  // LDS tier or workgroup tier with a register wavefront: no ring.
  const bool regs = (LDS && NT == 64 && reg_wavefront_shape(xn, oen, en)) || J > 0;

  const int32_t m0 = extend_diag<B8>(S.qw, S.qsh, S.tw, S.tsh, qn, tn, 0, 0);
  if (lane == 0 && !regs) {
    // Step 0 holds only m on diagonal 0; its ins/del entries are unset.
    Mr[rix(0)] = (typename S_t::ring_v)m0;
    Ri[rix(0)] = -1;
    Rd[rix(0)] = -1;
    if (NT > 64) sh->key[0] = sh->key[1] = sh->key[2] = INT32_MAX;
  }
  pair_sync<LDS, NT, RL>();
  int32_t status = ST_OK;
  int32_t step = 0, slot = 0, gslot = 0;  // step mod Rm, step mod Rg
  uint64_t used = 0;
  bool hist_on = true;
  int32_t end_k = 0, end_layer = 0;
  bool found = fitting ? (m0 == qn) : (qn == tn && m0 == qn);  // s = 0 endpoint on diagonal 0
  const int32_t kend_g = qn - tn;
  if constexpr (LDS && NT == 64) {
    StepState st;
    if (!found && affine_regs_any<B8>(S, qn, tn, m0, fitting, max_step, xn, oen, en, lane, st)) {
      step = __builtin_amdgcn_readfirstlane(st.step);
      used = st.used;
      found = st.found;
      end_k = st.end_k;
      end_layer = st.end_layer;
      status = st.status;
      if (!found && status == ST_OK) {  // the ring loop goes on from this step
        slot = __builtin_amdgcn_readfirstlane(step % Rm);
        gslot = __builtin_amdgcn_readfirstlane(step % Rg);
      }
    }
  }
  if constexpr (J > 0) {
    if (!found) {
      constexpr int C = NT / 64 * J;
      lmem_t<int32_t>* eb = (lmem_t<int32_t>*)(lds_base + G.tab_off);
      const BlockEdges E{eb, C};
      StepState st;
      if (affine_block_regs_any<NT, J, B8>(S, E, band, qn, tn, m0, fitting, max_step, xn, oen, en, lane, sh,
                                           st)) {
        step = __builtin_amdgcn_readfirstlane(st.step);
        used = st.used;
        found = st.found;
        end_k = st.end_k;
        end_layer = st.end_layer;
        status = st.status;
      } else {
        status = ST_INTERNAL;
      }
    }
  }
  while (J == 0 && !found && status == ST_OK) {
    step = __builtin_amdgcn_readfirstlane(step + 1);
    if (step > max_step) { status = ST_ALIGNMENT_FAILED; break; }
    slot = __builtin_amdgcn_readfirstlane(slot + 1 == Rm ? 0 : slot + 1);
    gslot = __builtin_amdgcn_readfirstlane(gslot + 1 == Rg ? 0 : gslot + 1);
    const int32_t lo = step_lo(step, tn), hi = step_hi(step, qn);
    const int32_t width = hi - lo + 1;
    if ((uint32_t)width > W) { status = ST_OVERFLOW_WIDTH; break; }
    if (hist_on && !S.fits(step, used, width)) {
      if (LDS) { status = ST_OVERFLOW_HIST; break; }
      hist_on = false;  // keep going to learn the exact number of steps
    }
    if (hist_on && lane == 0) S.toff(step) = (typename S_t::cell_v)used;
    if (NT > 64 && lane == 0) sh->key[(step + 1) % 3] = INT32_MAX;
    const int32_t sx = step - xn, so = step - oen, se = step - en;
    int32_t rx = slot - xn, ro = slot - oen;
    rx += rx < 0 ? Rm : 0;
    ro += ro < 0 ? Rm : 0;
    const int32_t re = gslot + 1 == Rg ? 0 : gslot + 1;  // (step - en) mod Rg, en = Rg - 1
    const int32_t xlo = sx >= 0 ? step_lo(sx, tn) : 1, xhi = sx >= 0 ? step_hi(sx, qn) : 0;
    const int32_t olo = so >= 0 ? step_lo(so, tn) : 1, ohi = so >= 0 ? step_hi(so, qn) : 0;
    const int32_t elo = se >= 0 ? step_lo(se, tn) : 1, ehi = se >= 0 ? step_hi(se, qn) : 0;
    const int32_t ke_lo = fitting ? max(kend_g, lo) : kend_g;
    const int32_t ke_hi = fitting ? hi : kend_g;
    const typename S_t::ring_t* Rmx = Mr + (size_t)rx * W;
    const typename S_t::ring_t* Rmo = Mr + (size_t)ro * W;
    const typename S_t::ring_t* Rie = Ri + (size_t)re * W;
    const typename S_t::ring_t* Rde = Rd + (size_t)re * W;
    typename S_t::ring_t* Wm = Mr + (size_t)slot * W;
    typename S_t::ring_t* Wi = Ri + (size_t)gslot * W;
    typename S_t::ring_t* Wd = Rd + (size_t)gslot * W;
    typename S_t::cell_t* cells = S.cells + used;
    if constexpr (UNW) {  // this step's frames: -1 margins beyond [lo, hi]
      for (int32_t t = lane; t < 2 * M; t += NT) {
        const uint32_t ix = rix(t < M ? lo - M + t : hi + 1 + (t - M));
        elem(Wm, ix) = -1;
        elem(Wi, ix) = -1;
        elem(Wd, ix) = -1;
      }
    }
    const int32_t nch = (width + NT - 1) / NT;
    for (int32_t c = 0; c < nch; ++c) {
      const int32_t k = lo + c * NT + lane;
      // Lanes past hi compute nothing that is kept; their reads stay in the
      // frame (k clamped to hi + 1).
      const int32_t kr = UNW ? min(k, hi + 1) : k;
      const uint32_t ik = rix(kr), il = rix(kr - 1), ir = rix(kr + 1);
      const int32_t rvx = elem(Rmx, ik), rvol = elem(Rmo, il), rvor = elem(Rmo, ir),
                    rvil = elem(Rie, il), rvdr = elem(Rde, ir);
      const int32_t vx = (UNW || (k >= xlo && k <= xhi)) ? rvx : -1;
      const int32_t vol = (UNW || (k - 1 >= olo && k - 1 <= ohi)) ? rvol : -1;
      const int32_t vor = (UNW || (k + 1 >= olo && k + 1 <= ohi)) ? rvor : -1;
      const int32_t vil = (UNW || (k - 1 >= elo && k - 1 <= ehi)) ? rvil : -1;
      const int32_t vdr = (UNW || (k + 1 >= elo && k + 1 <= ehi)) ? rvdr : -1;
      // Stored offsets are valid on their own diagonal, so a candidate from
      // k-1 (+1 query) needs offset <= qn only, one from k+1 needs target <= tn
      // only, and the mismatch candidate needs both (affine_pos_valid).
      int32_t ins = (vol >= 0 && vol < qn) ? vol + 1 : -1;
      uint32_t iop = 0;
      if (vil >= 0 && vil < qn && vil + 1 > ins) { ins = vil + 1; iop = 1; }
      int32_t del = (vor >= 0 && vor - k <= tn) ? vor : -1;
      uint32_t dop = 0;
      if (vdr >= 0 && vdr - k <= tn && vdr > del) { del = vdr; dop = 1; }
      int32_t best = (vx >= 0 && vx < qn && vx + 1 - k <= tn) ? vx + 1 : -1;
      uint32_t mop = 0;
      if (ins > best) { best = ins; mop = 1; }
      if (del > best) { best = del; mop = 2; }
      const int32_t mval = extend_diag<B8>(S.qw, S.qsh, S.tw, S.tsh, qn, tn, best, k);
      int32_t layer = -1;
      if (k <= hi) {
        elem(Wm, ik) = (typename S_t::ring_v)mval;
        elem(Wi, ik) = (typename S_t::ring_v)ins;
        elem(Wd, ik) = (typename S_t::ring_v)del;
        if (hist_on)
          elem(cells, (uint32_t)(k - lo)) = (typename S_t::cell_v)(((uint32_t)(best + 1) << 4) | mop | (iop << 2) |
                                                 (dop << 3));
        if (k >= ke_lo && k <= ke_hi) layer = mval == qn ? 0 : ins == qn ? 1 : del == qn ? 2 : -1;
      }
      if (NT == 64) {
        const uint64_t mask = __ballot(layer >= 0);
        if (mask && !found) {
          const int32_t ln = __builtin_ctzll(mask);
          found = true;
          end_k = lo + c * 64 + ln;
          end_layer = readlane(layer, ln);
        }
      } else if (layer >= 0) {
        atomicMin(&sh->key[step % 3], ((k + tn) << 2) | layer);
      }
    }
    used += (uint64_t)width;
    pair_sync<LDS, NT, RL>();
    if (NT > 64) {
      const int32_t key = __builtin_amdgcn_readfirstlane(sh->key[step % 3]);
      if (key != INT32_MAX) {
        found = true;
        end_k = (key >> 2) - tn;
        end_layer = key & 3;
      }
    }
  }
  if (!hist_on && status == ST_OK) status = ST_OVERFLOW_HIST;
  if (NT > 64 && RL) __syncthreads();  // history stores of every wave, for wave 0

  // Traceback (affine_traceback, wfa_simd.rs:785-886), wave-uniform state;
  // gap runs are resolved 64 cells per ballot.
  PairOut r{status, 0, 0, 0, 0, 0, step, (int32_t)min<uint64_t>(used, INT32_MAX)};
  if (status == ST_OK && lane < 64) {
    if (step == 0) { end_k = 0; end_layer = 0; }
    const int32_t t_end = qn - end_k;
    RunWriter w;
    w.init(a.runs + pd.run_off, pd.run_cap, lane == 0);
    int32_t qi = qn, ti = t_end, k = end_k, s = step, layer = end_layer;
    while (qi > 0 || ti > 0) {
      if (s == 0) {
        w.emit(OP_M, min(qi, ti));
        break;
      }
      if (layer == 0) {
        int32_t lo, hi;
        band_range(band, s, qn, tn, lo, hi);
        if (k < lo || k > hi) { r.status = ST_INTERNAL; break; }
        const uint32_t cell = S.cells[(uint64_t)S.toff(s) + (uint64_t)(k - lo)];
        const int32_t pre = (int32_t)(cell >> 4) - 1;
        if (pre < 0) { r.status = ST_INTERNAL; break; }
        const uint32_t mop = cell & 3;
        const int32_t run = max(qi - pre, 0);
        w.emit(OP_M, run);
        qi -= run; ti -= run;
        if (mop == 0) { w.emit(OP_X, 1); --qi; --ti; s -= xn; }
        else layer = (int32_t)mop;
      } else {
        // Gap layer: cells extend (op bit 1) back to the one that opened the
        // gap (op bit 0) from the m layer o+e earlier.
        const int32_t dir = layer == 1 ? -1 : 1;
        const uint32_t bit = layer == 1 ? 2 : 3;
        bool opened = false;
        const int32_t n = gap_run(S, band, s, k, dir, en, bit, qn, tn, lane, &opened);
        w.emit(layer == 1 ? OP_D : OP_I, n);
        if (layer == 1) qi -= n; else ti -= n;
        k += dir * n;
        if (opened) { s -= (n - 1) * en + oen; layer = 0; }
        else s -= n * en;
      }
      if (s < 0) { r.status = ST_INTERNAL; break; }
    }
    w.flush();
    if (w.overflow && r.status == ST_OK) r.status = ST_OVERFLOW_RUNS;
    r.score = step * g;
    r.edit_distance = w.non_m;
    r.t_end = t_end;
    r.nruns = (int32_t)w.n;
    r.run_start = (int32_t)(pd.run_cap - w.n);
  }
  publish(a, po, r, lane);
}

// --- linear --------------------------------------------------------------------

template <bool LDS, bool B8, int NT, bool RL = false>
__device__ __forceinline__ void wave_linear(const WfaArgs& a, const WaveGeom& G, uint8_t* lds_base,
                            const PairDesc& pd, PairOut* po, int lane, PairShared* sh) {
  using S_t = WaveStore<LDS, RL>;
  const int32_t qn = pd.qn, tn = pd.tn;
  const bool forced_global = (pd.flags & PF_GLOBAL) != 0;
  if (qn == 0) {
    // Wrapper (wfa_simd.rs:194-203, :365-367): empty query -> empty
    // alignment.  Forced global (fill_wfa, :231-345): the only path is tn
    // target-only steps, distance tn.
    if (lane == 0) {
      if (forced_global && tn > 0) {
        a.runs[pd.run_off + pd.run_cap - 1] = ((uint32_t)tn << 4) | OP_I;
        *po = PairOut{ST_OK, tn, tn, tn, 1, (int32_t)pd.run_cap - 1, 0, 0};
      } else {
        *po = PairOut{ST_OK, 0, 0, 0, 0, 0, 0, 0};
      }
    }
    return;
  }
  const S_t S = make_store<LDS, B8, NT, RL>(a, G, lds_base, pd, lane);
  const bool fitting = !forced_global && tn > qn + qn / 2 + 10;
  const int32_t max_s = qn + tn;
  const int32_t explore = (fitting && a.cp.cap >= 0) ? min(a.cp.cap, max_s) : max_s;
  const int32_t kend_global = qn - tn;
  const uint32_t W = S.wmask + 1;
  typename S_t::ring_t* ring = S.ring;  // two slots
  // HBM ring: unwrapped with -1 margins of M = 2 entries (ring_fill_origin).
  constexpr bool UNW = !LDS && !RL;
  const int32_t M = 2, koff = tn + 2;
  auto rix = [&](int32_t k) -> uint32_t {
    return UNW ? (uint32_t)(k + koff) : ((uint32_t)k & S.wmask);
  };
  if constexpr (UNW) {
    ring_fill_origin<NT>((int32_t*)S.ring, W, 2, M + 3, koff, qn, tn, M, lane);
    pair_sync<LDS, NT, RL>();
  }

  const int32_t w0 = extend_diag<B8>(S.qw, S.qsh, S.tw, S.tsh, qn, tn, 0, 0);
  if (lane == 0) {
    if (!(LDS && NT == 64)) ring[rix(0)] = (typename S_t::ring_v)w0;  // LDS tier: no ring
    if (NT > 64) sh->key[0] = sh->key[1] = sh->key[2] = INT32_MAX;
  }
  pair_sync<LDS, NT, RL>();
  int32_t step = 0, status = ST_OK, end_k = 0;
  uint64_t used = 0;
  bool hist_on = true;
  bool exhausted = false;
  bool found = fitting ? (w0 >= qn) : (kend_global == 0 && w0 == qn);
  if constexpr (LDS && NT == 64) {
    if (!found) {
      StepState st;
      linear_regs<B8>(S, qn, tn, w0, fitting, explore, lane, st, exhausted);
      step = __builtin_amdgcn_readfirstlane(st.step);
      used = st.used;
      found = st.found;
      end_k = st.end_k;
      status = st.status;
    }
  }
  while (!found && !exhausted && status == ST_OK) {
    step = __builtin_amdgcn_readfirstlane(step + 1);
    if (step > explore) { exhausted = true; break; }
    const int32_t slot = step & 1, ps = slot ^ 1;
    // Upstream's loop range for step s (wfa_simd.rs:288-291); the previous
    // step's range is the same formula at s - 1.
    const int32_t lo = step_lo(step, tn), hi = step_hi(step, qn);
    const int32_t p_lo = step_lo(step - 1, tn), p_hi = step_hi(step - 1, qn);
    const int32_t width = hi - lo + 1;
    if ((uint32_t)width > W) { status = ST_OVERFLOW_WIDTH; break; }
    if (hist_on && !S.fits(step, used, width)) {
      if (LDS) { status = ST_OVERFLOW_HIST; break; }
      hist_on = false;
    }
    if (hist_on && lane == 0) S.toff(step) = (typename S_t::cell_v)used;
    if (NT > 64 && lane == 0) sh->key[(step + 1) % 3] = INT32_MAX;
    const typename S_t::ring_t* P = ring + (size_t)ps * W;
    typename S_t::ring_t* Wn = ring + (size_t)slot * W;
    typename S_t::cell_t* cells = S.cells + used;
    const int32_t ke_lo = fitting ? max(qn - tn, lo) : kend_global;
    const int32_t ke_hi = fitting ? hi : kend_global;
    if constexpr (UNW) {  // this step's frame: -1 margins beyond [lo, hi]
      for (int32_t t = lane; t < 2 * M; t += NT) elem(Wn, rix(t < M ? lo - M + t : hi + 1 + (t - M))) = -1;
    }
    const int32_t nch = (width + NT - 1) / NT;
    for (int32_t c = 0; c < nch; ++c) {
      const int32_t k = lo + c * NT + lane;
      const int32_t kr = UNW ? min(k, hi + 1) : k;  // lanes past hi: reads stay in the frame
      const uint32_t ik = rix(kr), il = rix(kr - 1), ir = rix(kr + 1);
      const int32_t rvm = elem(P, ik), rvl = elem(P, il), rvr = elem(P, ir);
      // k - 1 < -tn and k + 1 > qn fall in the margins, so with UNW the
      // upstream bounds (k > -tn, k < qn) come out of the reads as well.
      const int32_t vm = (UNW || (k >= p_lo && k <= p_hi)) ? rvm : -1;
      const int32_t vl = (UNW || (k > -tn && k - 1 >= p_lo && k - 1 <= p_hi)) ? rvl : -1;
      const int32_t vr = (UNW || (k < qn && k + 1 >= p_lo && k + 1 <= p_hi)) ? rvr : -1;
      // Candidates [mismatch, query-only, target-only]; the last maximum wins;
      // then the winner alone is bounds-checked (wfa_simd.rs:316-326).
      int32_t best = -1;
      uint32_t op = 0;
      if (vm >= 0) { best = vm + 1; op = 1; }
      if (vl >= 0 && vl + 1 >= best) { best = vl + 1; op = 2; }
      if (vr >= 0 && vr >= best) { best = vr; op = 3; }
      const bool ok = op != 0 && best <= qn && best - k >= 0 && best - k <= tn;
      const int32_t val = extend_diag<B8>(S.qw, S.qsh, S.tw, S.tsh, qn, tn, ok ? best : -1, k);
      bool hit = false;
      if (k <= hi) {
        elem(Wn, ik) = (typename S_t::ring_v)val;
        if (hist_on)
          elem(cells, (uint32_t)(k - lo)) = (typename S_t::cell_v)(ok ? (((uint32_t)(best + 1) << 4) | op) : 0u);
        hit = ok && k >= ke_lo && k <= ke_hi && val >= qn;
      }
      if (NT == 64) {
        const uint64_t mask = __ballot(hit);
        if (mask && !found) {
          found = true;
          end_k = lo + c * 64 + (int32_t)__builtin_ctzll(mask);
        }
      } else if (hit) {
        atomicMin(&sh->key[step % 3], k + tn);
      }
    }
    used += (uint64_t)width;
    pair_sync<LDS, NT, RL>();
    if (NT > 64) {
      const int32_t key = __builtin_amdgcn_readfirstlane(sh->key[step % 3]);
      if (key != INT32_MAX) {
        found = true;
        end_k = key - tn;
      }
    }
  }
  if (!hist_on && status == ST_OK && !exhausted) status = ST_OVERFLOW_HIST;
  if (NT > 64 && RL) __syncthreads();  // history stores of every wave, for wave 0

  PairOut r{status, 0, 0, 0, 0, 0, step, (int32_t)min<uint64_t>(used, INT32_MAX)};
  if (status == ST_OK && exhausted) {
    if (fitting) {
      r.status = ST_ALIGNMENT_FAILED;
    } else {
      r.t_end = tn;  // fill_wfa's fallthrough ("", 0) (wfa_simd.rs:343-344)
      r.run_start = (int32_t)pd.run_cap;
    }
  } else if (status == ST_OK && lane < 64) {
    const int32_t t_end = qn - end_k;
    const int32_t ed = step;
    if (!fitting && a.cp.cap >= 0 && ed > a.cp.cap) {
      r.status = ST_ALIGNMENT_FAILED;  // wfa_simd.rs:380-384
    } else {
      RunWriter w;
      w.init(a.runs + pd.run_off, pd.run_cap, lane == 0);
      int32_t qi = qn, ti = t_end, k = qi - ti, s = ed;
      while (qi > 0 || ti > 0) {
        if (s == 0) {
          w.emit(OP_M, min(qi, ti));
          break;
        }
        const int32_t lo = step_lo(s, tn), hi = step_hi(s, qn);
        if (k < lo || k > hi) { r.status = ST_INTERNAL; break; }
        const uint32_t cell = S.cells[(uint64_t)S.toff(s) + (uint64_t)(k - lo)];
        const uint32_t op = cell & 15;
        if (op == 0) { r.status = ST_INTERNAL; break; }
        const int32_t pre = (int32_t)(cell >> 4) - 1;
        const int32_t run = max(qi - pre, 0);
        w.emit(OP_M, run);
        qi -= run; ti -= run;
        if (op == 1) { w.emit(OP_X, 1); --qi; --ti; }
        else if (op == 2) { w.emit(OP_D, 1); --qi; --k; }
        else { w.emit(OP_I, 1); --ti; ++k; }
        --s;
      }
      w.flush();
      if (w.overflow && r.status == ST_OK) r.status = ST_OVERFLOW_RUNS;
      r.score = ed;
      r.edit_distance = ed;
      r.t_end = t_end;
      r.nruns = (int32_t)w.n;
      r.run_start = (int32_t)(pd.run_cap - w.n);
    }
  }
  publish(a, po, r, lane);
}

// --- banded rows: four short pairs per wave ---------------------------------------
//
// Global-branch gap-affine pairs whose cost shape has en = 1 (rows_shape):
// each row of RW lanes aligns a pair of its own.  At score step s, lane r of
// a row holds diagonal k = B0 + dir s + r: dir = -1, B0 = 0 when qn <= tn
// (the window [-s, RW - 1 - s] follows the lower end of the reachable range
// down) and dir = +1, B0 = 1 - RW when qn > tn (window [s - RW + 1, s]).
//
// Band: a cell (s, k) lies on an alignment of score <= B steps only if
// s + |k - kend| <= B, kend = qn - tn (each diagonal between k and the end
// costs at least one more gap-extend step).  With B = |kend| + RW - 1 every
// such cell is inside the row's window (going down: k <= kend + B - s =
// RW - 1 - s).  A candidate (s', k') of a cell (s, k) has s' + |k' - kend| <=
// s + |k - kend|, so the cells of the band, their values and their
// tie-breaks only depend on cells of the band: a pair whose optimum is at
// most B steps gets exactly the unbanded result (end, score, CIGAR), and the
// window cells outside the band compute values that nothing reads.  The end
// cell is on diagonal kend, inside the band at every step <= B.  Cells out
// of reach (|k| > s) need no test: their candidates are out of reach too and
// come out unset, as on the single-pair path.
//
// A wave first aligns four pairs in 16-lane rows; the pairs whose optimum
// is above their band (end not found by step B) are aligned again two at a
// time in 32-lane rows (B = |kend| + 31), and the rare ones beyond that by
// the whole wave on the single-pair path, all in the same launch.
//
// Frames keep each step's own lane layout (P = offset + 1, 0 = no cell); a
// read of step s - d on diagonal k + j is lane r + dir d + j of the same row:
// one DPP row shift with zero fill past the row's ends for 16-lane rows, a
// lane permute with the ends masked for 32-lane rows, so rows never mix.
// History: one u16 cell per (step, lane) of the row (the LDS tier's cell
// format), so the traceback finds diagonal k of step s at lane k - B0 - dir s.

// Value of lane r + N of the same row of RW lanes; 0 past either end of the
// row.  Call with every lane active.
template <int N, int RW>
__device__ __forceinline__ int32_t row_read(int32_t v, int lane) {
  static_assert(N > -RW && N < RW, "a row shift stays inside the row");
  if constexpr (N == 0) {
    return v;
  } else if constexpr (RW == 16) {
    if constexpr (N > 0) return __builtin_amdgcn_mov_dpp(v, 0x100 + N, 0xF, 0xF, true);  // row_shl:N
    else return __builtin_amdgcn_mov_dpp(v, 0x110 - N, 0xF, 0xF, true);                  // row_shr:-N
  } else {
    const int32_t x = __builtin_amdgcn_ds_bpermute(((lane + N) & 63) << 2, v);
    return (uint32_t)((lane & (RW - 1)) + N) < (uint32_t)RW ? x : 0;
  }
}

// The row's pair re-packed at bit 0 into its LDS region (query, then
// target, each followed by a zero word), the row's RW lanes striding.
template <bool B8, int RW>
__device__ __forceinline__ void stage_row(const PairDesc& pd, lmem_t<uint64_t>* seq, int r) {
  const uint32_t nq = seq_words<B8>(pd.qn), nt = seq_words<B8>(pd.tn);
  const uint32_t qb = pd.q_shift * (B8 ? 8u : 2u), tb = pd.t_shift * (B8 ? 8u : 2u);
  const gmem_t<const uint64_t>* gq = (const gmem_t<const uint64_t>*)pd.q;
  const gmem_t<const uint64_t>* gt = (const gmem_t<const uint64_t>*)pd.t;
  for (uint32_t i = (uint32_t)r; i < nq + nt; i += RW) {
    const bool isq = i < nq;
    const uint32_t j = isq ? i : i - nq, n = isq ? nq : nt, b = isq ? qb : tb;
    const gmem_t<const uint64_t>* src = isq ? gq : gt;
    uint64_t v = 0;
    if (j + 1 < n) {
      const uint64_t lo = src[j];
      v = b == 0 ? lo : (lo >> b) | (src[j + 1] << (64 - b));
    }
    seq[i] = v;
  }
}

// Match extension over sequences staged at bit 0 in LDS, for the banded
// rows: a first 32-bit window (16 bases, or 4 bytes) on every lane without a
// branch, then, while any lane's match runs on, 128-bit windows (64 bases,
// or 16 bytes): the ten dwords of a window are read together and compared
// with one wait, so a long match (the leading diagonal of a row runs to the
// next difference, about a hundred bases at 1% divergence) costs half the
// LDS round trips of 64-bit windows.  Reads run at most four dwords past a
// sequence's staged words, still inside the row's region (the target is
// followed by the history); bases past the suffix ends are cut by `lim`.
template <bool B8>
__device__ __forceinline__ int32_t extend_rows(const lmem_t<const uint64_t>* qw, const lmem_t<const uint64_t>* tw,
                                               int32_t qn, int32_t tn, int32_t p, int32_t k) {
  constexpr uint32_t BITS = B8 ? 8 : 2;
  constexpr uint32_t CTZ_SHIFT = B8 ? 3 : 1;
  constexpr int32_t PER = B8 ? 4 : 16;
  const int32_t j = p - k;
  const bool live = (uint32_t)p < (uint32_t)qn && (uint32_t)j < (uint32_t)tn;
  const int32_t lim = min(qn - p, tn - j);
  const uint32_t qb = live ? (uint32_t)p * BITS : 0u, tb = live ? (uint32_t)j * BITS : 0u;
  const lmem_t<const uint32_t>* q32 = (const lmem_t<const uint32_t>*)qw + (qb >> 5);
  const lmem_t<const uint32_t>* t32 = (const lmem_t<const uint32_t>*)tw + (tb >> 5);
  const uint32_t qx = __builtin_amdgcn_alignbit(q32[1], q32[0], qb);  // shift = qb & 31
  const uint32_t tx = __builtin_amdgcn_alignbit(t32[1], t32[0], tb);
  const uint32_t z0 = (uint32_t)__builtin_ctzg(qx ^ tx, 32);  // 32 when all match
  const bool more = live && z0 == 32u && lim > PER;
  if (__builtin_amdgcn_ballot_w64(more) == 0) return live ? p + min((int32_t)(z0 >> CTZ_SHIFT), lim) : p;
  int32_t n = (int32_t)(z0 >> CTZ_SHIFT);
  if (more) {
    q32 += 1;
    t32 += 1;
    for (;;) {
      uint32_t qa[5], ta[5];
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        qa[i] = q32[i];
        ta[i] = t32[i];
      }
      uint32_t z = 128u;
#pragma unroll
      for (int i = 3; i >= 0; --i) {
        const uint32_t x = __builtin_amdgcn_alignbit(qa[i + 1], qa[i], qb) ^ __builtin_amdgcn_alignbit(ta[i + 1], ta[i], tb);
        z = x ? 32u * (uint32_t)i + (uint32_t)__builtin_ctz(x) : z;
      }
      n += (int32_t)(z >> CTZ_SHIFT);
      if (z < 128u || n >= lim) break;
      q32 += 4;
      t32 += 4;
    }
  }
  return live ? p + min(n, lim) : p;
}

// Candidates of one cell (affine_step's recurrence and tie rules on P =
// offset + 1 values, 0 = no cell): gap-open / gap-extend into the
// query-only (ins) and target-only (del) layers, then the m layer's best of
// mismatch, ins and del; ops as the history cell stores them.
struct RowCell {
  int32_t ins, del, best;
  uint32_t ops;  // mop | iop << 2 | dop << 3
};

__device__ __forceinline__ RowCell row_cell(int32_t vx, int32_t vol, int32_t vor, int32_t vil, int32_t vdr,
                                            int32_t qn, int32_t tn, int32_t k) {
  const uint32_t uq = (uint32_t)qn, utk = (uint32_t)(tn + k);
  RowCell c;
  c.ins = (uint32_t)(vol - 1) < uq ? vol + 1 : 0;
  uint32_t iop = 0;
  if ((uint32_t)(vil - 1) < uq && vil + 1 > c.ins) { c.ins = vil + 1; iop = 1; }
  c.del = (uint32_t)(vor - 1) <= utk ? vor : 0;
  uint32_t dop = 0;
  if ((uint32_t)(vdr - 1) <= utk && vdr > c.del) { c.del = vdr; dop = 1; }
  c.best = (uint32_t)(vx - 1) < min(uq, utk) ? vx + 1 : 0;
  uint32_t mop = 0;
  if (c.ins > c.best) { c.best = c.ins; mop = 1; }
  if (c.del > c.best) { c.best = c.del; mop = 2; }
  c.ops = mop | (iop << 2) | (dop << 3);
  return c;
}

// One pass of the wave's 64 / RW rows (pi = ~0u: empty row).  XN, OEN:
// mismatch and open + extend in score steps (extend = 1); DIR: window
// direction of every pair.  Returns true on the lanes of a row whose pair's
// optimum lies beyond its band (nothing stored for it).
template <int XN, int OEN, int DIR, bool B8, int RW>
__device__ __forceinline__ bool rows_pass(const WfaArgs& a, const WaveGeom& G, uint8_t* base, int lane,
                                          uint32_t pi, uint32_t pos) {
  constexpr int MD = XN > OEN ? XN : OEN;
  constexpr int32_t B0 = DIR < 0 ? 0 : 1 - RW;
  constexpr uint64_t RMASK = RW == 64 ? ~0ull : (1ull << RW) - 1;
  const int r = lane & (RW - 1);
  const uint32_t rsh = (uint32_t)lane & (uint32_t)(64 - RW);  // the row's bits in a ballot
  const bool has = pi != 0xFFFFFFFFu;
  PairDesc pd{};
  if (has) {
    if (a.rows_desc) {  // staged at bit 0 in list order (shifts 0)
      const RowDesc rd = a.rows_desc[pos];
      pd.q = a.rows_seq + rd.q_word;
      pd.t = pd.q + seq_words<B8>(rd.qn);
      pd.qn = rd.qn;
      pd.tn = rd.tn;
      pd.run_off = rd.run_off;
      pd.run_cap = rd.run_cap;
    } else {
      pd = a.pairs[pi];
    }
  }
  const int32_t qn = pd.qn, tn = pd.tn;
  uint8_t* rb = base + (uint32_t)(lane / RW) * rows_region_bytes(G.rows_hist_off, G.rows_steps, RW);
  lmem_t<uint64_t>* seq = (lmem_t<uint64_t>*)rb;
  lmem_t<uint16_t>* hist = (lmem_t<uint16_t>*)(rb + G.rows_hist_off);
  if (has) stage_row<B8, RW>(pd, seq, r);
  wave_sync<true>();
  const lmem_t<const uint64_t>* qw = seq;
  const lmem_t<const uint64_t>* tw = seq + seq_words<B8>(qn);
  const int32_t kend = qn - tn;
  const int32_t B = (DIR < 0 ? -kend : kend) + (RW - 1);
  const int32_t g = a.cp.g;
  // The cap and the last score step explored, as on the single-pair path.
  const int64_t hard_max = (int64_t)(qn + tn) * max(a.cp.x, 1) + a.cp.o + a.cp.e + 4;
  const int64_t explore_to = a.cp.cap >= 0 ? min((int64_t)a.cp.cap, hard_max) : hard_max;
  const int32_t max_step = explore_to <= INT32_MAX ? (int32_t)explore_to / g
                                                   : (int32_t)min<int64_t>(explore_to / g, INT32_MAX);
  // Every live row's step stays below its region's step count: lim <= B <=
  // rows_steps + RW - 17 (the host sizes rows_steps for the largest |kend|).
  const int32_t lim = min(B, max_step);
  const uint32_t span = (uint32_t)(qn + tn);  // k + tn in [0, qn + tn]: inside the matrix

  const int32_t m0 = has ? extend_diag<B8>(qw, 0u, tw, 0u, qn, tn, 0, 0) : 0;
  int32_t M[MD];  // M[MD - d]: step s - d
#pragma unroll
  for (int j = 0; j < MD; ++j) M[j] = 0;
  M[MD - 1] = has && B0 + r == 0 ? m0 + 1 : 0;
  int32_t I1 = 0, D1 = 0;  // step s - 1
  bool live = has;
  int32_t end_s = -1, end_layer = 0;
  if (has && kend == 0 && m0 == qn) {  // s = 0 endpoint on diagonal 0
    end_s = 0;
    live = false;
  }
  // The end of a row: the first step whose end-diagonal cell reaches qn in
  // some layer (m, then ins, then del).
  auto end_found = [&](int32_t s, bool hit, int32_t mval, const RowCell& c) {
    const uint64_t hm = __builtin_amdgcn_ballot_w64(hit);
    if (hm) {
      const int32_t layer = mval == qn ? 0 : c.ins == qn + 1 ? 1 : 2;
      const uint64_t h1 = __builtin_amdgcn_ballot_w64(hit && layer == 1);
      const uint64_t h2 = __builtin_amdgcn_ballot_w64(hit && layer == 2);
      if (live && ((hm >> rsh) & RMASK)) {
        end_s = s;
        end_layer = ((h1 >> rsh) & RMASK) ? 1 : ((h2 >> rsh) & RMASK) ? 2 : 0;
        live = false;
      }
    }
  };
  // Every row stores each step (s stays inside every region of the launch
  // while any row is live); steps past a row's end are never read.
  for (int32_t s = 1;; ++s) {
    live = live && s <= lim;
    if (__builtin_amdgcn_ballot_w64(live) == 0) break;
    const int32_t k = B0 + DIR * s + r;
    const bool inr = live && (uint32_t)(k + tn) <= span;
    const RowCell c = row_cell(row_read<DIR * XN, RW>(M[MD - XN], lane),
                               row_read<DIR * OEN - 1, RW>(M[MD - OEN], lane),
                               row_read<DIR * OEN + 1, RW>(M[MD - OEN], lane),
                               row_read<DIR - 1, RW>(I1, lane), row_read<DIR + 1, RW>(D1, lane), qn, tn, k);
    const int32_t mval = extend_rows<B8>(qw, tw, qn, tn, inr ? c.best - 1 : -1, k);
    const uint32_t cell = ((uint32_t)c.best << 4) | c.ops;
    hist[(uint32_t)s * RW + (uint32_t)r] = (uint16_t)(inr ? cell : 0u);
    const bool hit = inr & (k == kend) & ((mval == qn) | (c.ins == qn + 1) | (c.del == qn + 1));
    end_found(s, hit, mval, c);
#pragma unroll
    for (int j = 0; j + 1 < MD; ++j) M[j] = M[j + 1];
    M[MD - 1] = inr ? mval + 1 : 0;
    I1 = inr ? c.ins : 0;
    D1 = inr ? c.del : 0;
  }
  // A row that did not reach its end: beyond the cap it fails as upstream's
  // search does, beyond the band alone it needs a wider pass.
  const bool failed = has && end_s < 0 && max_step <= B;
  const bool wider = has && end_s < 0 && max_step > B;

  // Traceback of each row that found its end (affine_traceback,
  // wfa_simd.rs:785-886), row-uniform state, gap runs resolved RW cells per
  // ballot.
  RunWriter w;
  w.init(a.runs + pd.run_off, pd.run_cap, r == 0);
  int32_t status = ST_OK;
  bool tb = end_s >= 0;
  int32_t qi = qn, ti = tn, k = kend, s = end_s, layer = end_layer;
  while (__builtin_amdgcn_ballot_w64(tb)) {
    if (tb) {
      if (s == 0) {
        w.emit(OP_M, min(qi, ti));
        tb = false;
      } else if (layer == 0) {
        const int32_t rr = k - (B0 + DIR * s);
        const uint32_t cell = (uint32_t)rr < (uint32_t)RW ? (uint32_t)hist[(uint32_t)s * RW + (uint32_t)rr] : 0u;
        const int32_t pre = (int32_t)(cell >> 4) - 1;
        if (pre < 0) {
          status = ST_INTERNAL;
          tb = false;
        } else {
          const uint32_t mop = cell & 3;
          const int32_t run = max(qi - pre, 0);
          w.emit(OP_M, run);
          qi -= run;
          ti -= run;
          if (mop == 0) { w.emit(OP_X, 1); --qi; --ti; s -= XN; }
          else layer = (int32_t)mop;
        }
      } else {
        // Lane j of the row looks at step s - j on diagonal k + dir j and
        // reports whether the gap was opened there (op bit 0).
        const int32_t dir = layer == 1 ? -1 : 1;
        const uint32_t bit = layer == 1 ? 2u : 3u;
        const int32_t sj = s - r, kj = k + dir * r;
        bool open = false;
        if (sj >= 1) {
          const int32_t rj = kj - (B0 + DIR * sj);
          open = (uint32_t)rj < (uint32_t)RW ? ((hist[(uint32_t)sj * RW + (uint32_t)rj] >> bit) & 1u) == 0 : true;
        }
        const uint64_t mr = (__builtin_amdgcn_ballot_w64(open) >> rsh) & RMASK;
        const int32_t n = mr ? (int32_t)__builtin_ctzll(mr) + 1 : RW;
        w.emit(layer == 1 ? OP_D : OP_I, n);
        if (layer == 1) qi -= n;
        else ti -= n;
        k += dir * n;
        if (mr) { s -= (n - 1) + OEN; layer = 0; }
        else s -= n;
      }
      if (tb && s < 0) { status = ST_INTERNAL; tb = false; }
      if (tb && qi <= 0 && ti <= 0) tb = false;
    }
  }
  w.flush();
  if (has && !wider && r == 0) {
    PairOut res{ST_ALIGNMENT_FAILED, 0, 0, 0, 0, 0, lim + 1, 0};
    if (!failed) {
      if (w.overflow && status == ST_OK) status = ST_OVERFLOW_RUNS;
      res = PairOut{status, end_s * g, w.non_m, tn, (int32_t)w.n, (int32_t)(pd.run_cap - w.n), end_s,
                    end_s * RW};
    }
    a.out[pi] = res;
    if (res.status >= ST_OVERFLOW_HIST && a.overflow)
      __hip_atomic_store(a.overflow, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  return wider;
}

// --- kernel --------------------------------------------------------------------

// Up to seven workgroups per CU (28 waves, 72 registers per lane): the
// gap-affine register wavefront needs about 71 registers (held to 64 it
// spilled a few values to scratch, about 1.5 KiB of writes per wave), and
// the tuning that fills a CU with short pairs gives each wave 5 KiB of LDS,
// which holds 28 waves anyway.  Kernels that need fewer registers (linear)
// still fit eight workgroups.
template <int MODE, bool LDS>
__global__ __launch_bounds__(64 * WV_WAVES, 7) void wfa_wave_kernel(WfaArgs a, WaveGeom G) {
  extern __shared__ __align__(16) uint8_t wave_lds_pool[];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int32_t)(threadIdx.x >> 6));
  uint8_t* base = LDS ? wave_lds_pool + (size_t)wv * G.per_wave_bytes : nullptr;
  // Work items: the first one per wave is the wave's own index.  With
  // static_order the wave then takes every nwaves-th item (the host orders
  // the list so that these strides balance); else the next item comes from
  // the queue head (offset by the number of waves) once the current one is
  // done, so no wave holds an item while another wave idles.  Claims on the
  // one queue head serialise (about 12 ns each), which would bound a queue
  // of short pairs; those use static_order.
  const uint32_t nwaves = gridDim.x * WV_WAVES;
  uint32_t w = blockIdx.x * WV_WAVES + (uint32_t)wv;
  while (w < a.nwork) {
    const uint32_t pi = a.work_list[w];
    const PairDesc pd = a.pairs[pi];
    PairOut* po = a.out + pi;
    if (MODE == ENG_AFFINE) {
      if (pd.flags & PF_BYTES) wave_affine<LDS, true, 64>(a, G, base, pd, po, lane, nullptr);
      else wave_affine<LDS, false, 64>(a, G, base, pd, po, lane, nullptr);
    } else {
      if (pd.flags & PF_BYTES) wave_linear<LDS, true, 64>(a, G, base, pd, po, lane, nullptr);
      else wave_linear<LDS, false, 64>(a, G, base, pd, po, lane, nullptr);
    }
    wave_sync<LDS>();  // the wave's storage is reused by its next pair
    if (a.static_order) {
      w += nwaves;
    } else {
      uint32_t nx = 0;
      if (lane == 0) nx = atomicAdd(&a.counters[0], 1u);
      w = (uint32_t)readlane((int32_t)nx, 0) + nwaves;
    }
  }
}

// Banded rows: one launch over a list of four segments (rows_seg): pairs of
// entries in 32-lane rows, then quads in 16-lane rows, each kind with the
// window going down (qn <= tn) and then up.  Wave v of the launch (v, v +
// waves, ... when the launch has fewer waves than sets) takes set v of the
// concatenated sets; the host orders each segment so that these strides
// balance, and the 32-lane sets (the pairs with the most steps) start first.
// A 16-lane row holds a pair only when the host has bounded its optimum by
// the row's band, so one that does not finish is reported as too wide (and
// re-run on the HBM tiers); a 32-lane row that does not finish goes to the
// single-pair path, the whole wave, one pair at a time.  Up to five
// workgroups per CU (102 registers per lane, no spills).
template <bool B8>
__global__ __launch_bounds__(64 * WV_WAVES, 5) void wfa_rows_kernel(WfaArgs a, WaveGeom G) {
  extern __shared__ __align__(16) uint8_t rows_lds_pool[];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int32_t)(threadIdx.x >> 6));
  uint8_t* base = rows_lds_pool + (size_t)wv * G.per_wave_bytes;
  const uint32_t nwaves = gridDim.x * WV_WAVES;
  uint32_t first[5], off[5];  // first set and first entry of each segment
  first[0] = off[0] = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint32_t per = j < 2 ? 2u : 4u;
    first[j + 1] = first[j] + (a.rows_seg[j] + per - 1) / per;
    off[j + 1] = off[j] + a.rows_seg[j];
  }
  const int32_t xn = a.cp.xn, oen = a.cp.oen;
  for (uint32_t v = blockIdx.x * WV_WAVES + (uint32_t)wv; v < first[4]; v += nwaves) {
    const int j = v < first[1] ? 0 : v < first[2] ? 1 : v < first[3] ? 2 : 3;
    const uint32_t per = j < 2 ? 2u : 4u;
    const uint32_t pos = off[j] + (v - first[j]) * per + (uint32_t)lane / (64u / per);
    const uint32_t pi = pos < off[j + 1] ? a.work_list[pos] : 0xFFFFFFFFu;
    bool wider;
    if (j == 0) {
      wider = xn == 2 && oen == 4 ? rows_pass<2, 4, -1, B8, 32>(a, G, base, lane, pi, pos)
            : xn == 2 && oen == 5 ? rows_pass<2, 5, -1, B8, 32>(a, G, base, lane, pi, pos)
                                  : rows_pass<1, 1, -1, B8, 32>(a, G, base, lane, pi, pos);
    } else if (j == 1) {
      wider = xn == 2 && oen == 4 ? rows_pass<2, 4, 1, B8, 32>(a, G, base, lane, pi, pos)
            : xn == 2 && oen == 5 ? rows_pass<2, 5, 1, B8, 32>(a, G, base, lane, pi, pos)
                                  : rows_pass<1, 1, 1, B8, 32>(a, G, base, lane, pi, pos);
    } else if (j == 2) {
      wider = xn == 2 && oen == 4 ? rows_pass<2, 4, -1, B8, 16>(a, G, base, lane, pi, pos)
            : xn == 2 && oen == 5 ? rows_pass<2, 5, -1, B8, 16>(a, G, base, lane, pi, pos)
                                  : rows_pass<1, 1, -1, B8, 16>(a, G, base, lane, pi, pos);
    } else {
      wider = xn == 2 && oen == 4 ? rows_pass<2, 4, 1, B8, 16>(a, G, base, lane, pi, pos)
            : xn == 2 && oen == 5 ? rows_pass<2, 5, 1, B8, 16>(a, G, base, lane, pi, pos)
                                  : rows_pass<1, 1, 1, B8, 16>(a, G, base, lane, pi, pos);
    }
    const uint32_t lead = (uint32_t)lane % (64u / per);  // 0 on the first lane of each row
    if (j >= 2) {
      if (wider && lead == 0) {
        a.out[pi] = PairOut{ST_OVERFLOW_WIDTH, 0, 0, 0, 0, 0, 0, 0};
        if (a.overflow) __hip_atomic_store(a.overflow, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
    } else {
      uint64_t f = __builtin_amdgcn_ballot_w64(wider && lead == 0);
      while (f) {
        const int32_t l = (int32_t)__builtin_ctzll(f);
        f &= f - 1;
        const uint32_t fpi = (uint32_t)readlane((int32_t)pi, l);
        wave_sync<true>();
        const PairDesc fpd = a.pairs[fpi];
        wave_affine<true, B8, 64>(a, G, base, fpd, a.out + fpi, lane, nullptr);
      }
    }
    wave_sync<true>();  // the wave's region is reused by its next set
  }
}

// Workgroup per pair (NT threads) for wide wavefronts; ring and history in
// the workgroup's HBM slot.
template <int MODE, int NT, bool RL, int J = 0>
__global__ __launch_bounds__(NT) void wfa_block_kernel(WfaArgs a, WaveGeom G) {
  __shared__ PairShared sh;
  extern __shared__ __align__(16) uint8_t block_lds[];
  uint8_t* base = RL ? block_lds : nullptr;
  const int tid = threadIdx.x;
  // First item: the workgroup's index; later ones from the queue head
  // (offset by the grid size), claimed by thread 0 when a pair is done and
  // read by all after the closing barrier.  The slot alternates so that a
  // claim never overwrites the value some wave is still reading for the
  // previous item (a pair with qn = 0 returns without a barrier).
  // The index is kept in a scalar register so that the loop exit is a scalar
  // branch: every wave leaves together.  With a per-lane exit the compiler
  // may give thread 0 (the only one with work before the next barrier) a loop
  // of its own, and the other lanes would spin on the old index.  For the
  // same reason no thread-0-only code follows the barrier at the end of the
  // body.
  uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)blockIdx.x);
  uint32_t par = 0;
  while (w < a.nwork) {
    const uint32_t pi = a.work_list[w];
    const PairDesc pd = a.pairs[pi];
    PairOut* po = a.out + pi;
    if (MODE == ENG_AFFINE) {
      if (pd.flags & PF_BYTES) wave_affine<false, true, NT, RL, J>(a, G, base, pd, po, tid, &sh);
      else wave_affine<false, false, NT, RL, J>(a, G, base, pd, po, tid, &sh);
    } else {
      if (pd.flags & PF_BYTES) wave_linear<false, true, NT, RL>(a, G, base, pd, po, tid, &sh);
      else wave_linear<false, false, NT, RL>(a, G, base, pd, po, tid, &sh);
    }
    if (tid == 0) sh.next[par] = atomicAdd(&a.counters[0], 1u);
    __syncthreads();  // the slot is reused by the next pair; sh.next[par] is set
    w = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)sh.next[par]) + gridDim.x;
    par ^= 1u;
  }
}

// Compacts CIGAR runs: pair i's count = off[i+1] - off[i] words move from
// runs[src[i] ..] to dst[off[i] ..]; a wave per pair.
__global__ __launch_bounds__(256) void gather_runs_kernel(const uint32_t* __restrict__ runs,
                                                          const uint64_t* __restrict__ src,
                                                          const uint64_t* __restrict__ off,
                                                          uint32_t* __restrict__ dst, uint32_t n) {
  const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (i >= n) return;
  const uint64_t a = off[i], cnt = off[i + 1] - a;
  const uint32_t* s = runs + src[i];
  for (uint64_t j = lane; j < cnt; j += 64) dst[a + j] = s[j];
}

hipError_t launch_gather_runs(const uint32_t* runs, const uint64_t* src, const uint64_t* off,
                              uint32_t* dst, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(gather_runs_kernel, dim3((n + 3) / 4), dim3(256), 0, st, runs, src, off, dst, n);
  return hipGetLastError();
}

// Row stage: for staged pair k (list position order), its query and target
// from wherever its descriptor points, shifted to bit 0 exactly as the row
// launch stages them into LDS (word j = bits of words j and j + 1 of the
// source for j + 1 < n, the last word 0), to the staging arena from word
// rd[k].q_word on (query, then target); a wave per pair.
__global__ __launch_bounds__(256) void stage_rows_kernel(const PairDesc* __restrict__ pairs,
                                                         const uint32_t* __restrict__ idx,
                                                         const RowDesc* __restrict__ rd, uint64_t* stage,
                                                         uint32_t n) {
  const uint32_t k = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (k >= n) return;
  const PairDesc pd = pairs[idx[k]];
  const bool b8 = (pd.flags & PF_BYTES) != 0;
  uint64_t* dst = stage + rd[k].q_word;
  for (int side = 0; side < 2; ++side) {
    const uint64_t* src = side ? pd.t : pd.q;
    const uint32_t len = (uint32_t)(side ? pd.tn : pd.qn);
    const uint32_t b = (side ? pd.t_shift : pd.q_shift) * (b8 ? 8u : 2u);
    const uint32_t nw = b8 ? seq_words<true>((int32_t)len) : seq_words<false>((int32_t)len);
    if (side) dst += b8 ? seq_words<true>(pd.qn) : seq_words<false>(pd.qn);
    for (uint32_t j = lane; j < nw; j += 64) {
      uint64_t v = 0;
      if (j + 1 < nw) {
        const uint64_t lo = src[j];
        v = b == 0 ? lo : (lo >> b) | (src[j + 1] << (64 - b));
      }
      dst[j] = v;
    }
  }
}

hipError_t launch_stage_rows(const PairDesc* pairs, const uint32_t* idx, const RowDesc* rd, uint64_t* stage,
                             uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(stage_rows_kernel, dim3((n + 3) / 4), dim3(256), 0, st, pairs, idx, rd, stage, n);
  return hipGetLastError();
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void set_bounds_kernel(PairDesc* pairs, const uint32_t* iv, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) pairs[iv[2 * k]].bound = iv[2 * k + 1];
}

hipError_t launch_set_bounds(PairDesc* pairs, const uint32_t* iv, uint32_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(set_bounds_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pairs, iv, n);
  return hipGetLastError();
}

hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_wfa_wave(int mode, bool lds, const WfaArgs& a, const WaveGeom& G,
                           uint32_t grid, hipStream_t st) {
  const size_t bytes = lds ? (size_t)G.per_wave_bytes * WV_WAVES : 0;
  const dim3 block(64 * WV_WAVES);
  if (mode == ENG_AFFINE) {
    if (lds) hipLaunchKernelGGL((wfa_wave_kernel<ENG_AFFINE, true>), dim3(grid), block, bytes, st, a, G);
    else hipLaunchKernelGGL((wfa_wave_kernel<ENG_AFFINE, false>), dim3(grid), block, 0, st, a, G);
  } else {
    if (lds) hipLaunchKernelGGL((wfa_wave_kernel<ENG_LINEAR, true>), dim3(grid), block, bytes, st, a, G);
    else hipLaunchKernelGGL((wfa_wave_kernel<ENG_LINEAR, false>), dim3(grid), block, 0, st, a, G);
  }
  return hipGetLastError();
}

hipError_t launch_wfa_rows(bool b8, const WfaArgs& a, const WaveGeom& G, uint32_t grid, hipStream_t st) {
  const dim3 block(64 * WV_WAVES);
  const size_t bytes = (size_t)G.per_wave_bytes * WV_WAVES;
  if (b8) hipLaunchKernelGGL((wfa_rows_kernel<true>), dim3(grid), block, bytes, st, a, G);
  else hipLaunchKernelGGL((wfa_rows_kernel<false>), dim3(grid), block, bytes, st, a, G);
  return hipGetLastError();
}

// nt = threads per pair: 64 or 512 (register wavefront only), 256 or 1024;
// storage 0: ring in HBM, 1: ring and sequences in LDS, 2: register
// wavefront, sequences and chunk ends in LDS (G gives the layout; dynamic
// LDS = G.per_wave_bytes).
template <int MODE>
hipError_t launch_block_mode(int nt, int storage, const WfaArgs& a, const WaveGeom& G, uint32_t grid,
                             hipStream_t st) {
  const bool rl = storage != 0;
  const size_t lds = rl ? G.per_wave_bytes : 0;
  if (storage == 2 && MODE == ENG_AFFINE) {
    // 16 chunks per wave: 16,384 diagonals for 1024 threads, 8,192 for 512,
    // 4,096 for 256, 1,024 for one wave.
    if (nt == 1024) hipLaunchKernelGGL((wfa_block_kernel<MODE, 1024, true, 16>), dim3(grid), dim3(1024), lds, st, a, G);
    else if (nt == 512) hipLaunchKernelGGL((wfa_block_kernel<MODE, 512, true, 16>), dim3(grid), dim3(512), lds, st, a, G);
    else if (nt == 256) hipLaunchKernelGGL((wfa_block_kernel<MODE, 256, true, 16>), dim3(grid), dim3(256), lds, st, a, G);
    else if (nt == 64) hipLaunchKernelGGL((wfa_block_kernel<MODE, 64, true, 16>), dim3(grid), dim3(64), lds, st, a, G);
    else return hipErrorInvalidValue;
    return hipGetLastError();
  }
  if (nt == 512 || nt == 64) return hipErrorInvalidValue;  // register wavefront only
  if (nt == 1024) {
    if (rl) hipLaunchKernelGGL((wfa_block_kernel<MODE, 1024, true>), dim3(grid), dim3(1024), lds, st, a, G);
    else hipLaunchKernelGGL((wfa_block_kernel<MODE, 1024, false>), dim3(grid), dim3(1024), 0, st, a, G);
  } else {
    if (rl) hipLaunchKernelGGL((wfa_block_kernel<MODE, 256, true>), dim3(grid), dim3(256), lds, st, a, G);
    else hipLaunchKernelGGL((wfa_block_kernel<MODE, 256, false>), dim3(grid), dim3(256), 0, st, a, G);
  }
  return hipGetLastError();
}

hipError_t launch_wfa_block(int mode, int nt, int storage, const WfaArgs& a, const WaveGeom& G,
                            uint32_t grid, hipStream_t st) {
  return mode == ENG_AFFINE ? launch_block_mode<ENG_AFFINE>(nt, storage, a, G, grid, st)
                            : launch_block_mode<ENG_LINEAR>(nt, storage, a, G, grid, st);
}

}  // namespace phx
