// wfa_wave.hip -- wave-resident WFA tier for short pairs (qn + tn <= 2048).
//
// Each wave64 of a 256-thread workgroup owns one pair at a time and pulls the
// next pair from the device work queue on its own; the four waves never
// synchronise with each other.  Everything a pair touches lives in the wave's
// LDS region: the packed query and target, a ring of the last R score frames
// (int16 offsets, 128 diagonals, two per lane), per-step cell offsets and
// 16-bit history cells (bits 15..4 = pre-extension offset + 1, bits 3..0 =
// ops).  Diagonal ranges are arithmetic functions of the step (see below).
// The fitting/global endpoint is found with a wave ballot (first chunk,
// lowest lane = smallest diagonal); the traceback runs wave-wide and resolves
// affine gap runs 64 cells per ballot.  A pair whose
// wavefront exceeds 128 diagonals or whose history exceeds the LDS budget
// reports an overflow status and is re-run on the workgroup-per-pair tier.
//
// Recurrences, candidate orders and tie rules are those of
// phraya-align/src/wfa_simd.rs (linear :283-330, fitting end :940-955,
// linear traceback :959-1130, affine :581-779 and :785-886), exactly as in
// wfa_kernels.hip.
#include <hip/hip_runtime.h>

#include "phx_device.h"
#include "wfa_common.h"

namespace phx {

constexpr int WV_W = 128;               // diagonals per ring frame
constexpr int WV_WAVES = 4;             // waves (pairs) per workgroup

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int32_t readlane(int32_t v, int32_t lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

struct WaveLds {
  int16_t* ring;
  uint64_t* seq;
  uint16_t* toff;   // first history cell of each step
  uint16_t* cells;
};

__device__ __forceinline__ WaveLds wave_lds(uint8_t* base, const WaveGeom& G) {
  WaveLds L;
  L.ring = reinterpret_cast<int16_t*>(base + G.ring_off);
  L.seq = reinterpret_cast<uint64_t*>(base + G.seq_off);
  L.toff = reinterpret_cast<uint16_t*>(base + G.tab_off);
  L.cells = reinterpret_cast<uint16_t*>(base + G.cell_off);
  return L;
}

// Copies the pair's packed words into LDS; returns the target's word offset.
template <bool B8>
__device__ __forceinline__ uint32_t stage_seqs(const WfaArgs& a, const PairDesc& pd, uint64_t* dst,
                                               int lane) {
  const uint32_t nq = B8 ? ((uint32_t)pd.qn + 7) / 8 + 1 : ((uint32_t)pd.qn + 31) / 32 + 1;
  const uint32_t nt = B8 ? ((uint32_t)pd.tn + 7) / 8 + 1 : ((uint32_t)pd.tn + 31) / 32 + 1;
  for (uint32_t i = (uint32_t)lane; i < nq + nt; i += 64)
    dst[i] = i < nq ? a.seq[pd.q_word + i] : a.seq[pd.t_word + (i - nq)];
  wave_sync();
  return nq;
}

// --- affine --------------------------------------------------------------------
//
// Diagonal ranges are arithmetic: every diagonal shift costs at least one
// gap-extend unit, so after u score steps (in units of g) all set cells lie in
// |k| <= u.  Step u computes every k in [max(-u, -tn), min(u, qn)] and writes
// all three layers there (-1 = unset), so a ring read is valid exactly when
// the predecessor step's range contains the diagonal; no per-slot metadata is
// kept.  Cells outside the reachable set simply come out unset, as upstream's
// full-width frames do.

__device__ __forceinline__ int32_t step_lo(int32_t u, int32_t tn) { return max(-u, -tn); }
__device__ __forceinline__ int32_t step_hi(int32_t u, int32_t qn) { return min(u, qn); }

// Parallel search along a gap: lane j looks at the cell of step s - j*en on
// diagonal k + dir*j and reports whether the gap was opened there (op bit 0).
// Returns the number of gap bases up to and including the opening cell, or
// 64 + (no opening within the window) as 64 with *opened = false.
__device__ __forceinline__ int32_t gap_run(const WaveLds& L, int32_t s, int32_t k, int32_t dir,
                                           int32_t en, uint32_t bit, int32_t qn, int32_t tn,
                                           int lane, bool* opened) {
  const int32_t sj = s - lane * en;
  const int32_t kj = k + dir * lane;
  bool open = false;
  if (sj >= 1) {
    const int32_t lo = step_lo(sj, tn), hi = step_hi(sj, qn);
    if (kj >= lo && kj <= hi) {
      const uint32_t cell = L.cells[(uint32_t)L.toff[sj] + (uint32_t)(kj - lo)];
      open = ((cell >> bit) & 1u) == 0;
    } else {
      open = true;  // unreachable for a valid path; ends the run (caught below)
    }
  }
  const uint64_t mask = __ballot(open);
  *opened = mask != 0;
  return mask ? (int32_t)__builtin_ctzll(mask) + 1 : 64;
}

template <bool B8>
__device__ void wave_affine(const WfaArgs& a, const WaveGeom& G, uint8_t* base,
                            const PairDesc& pd, PairOut* po, int lane) {
  const int32_t qn = pd.qn, tn = pd.tn;
  if (qn == 0) {  // wfa_simd.rs:925-927
    if (lane == 0) *po = PairOut{ST_OK, 0, 0, 0, 0, 0, 0, 0};
    return;
  }
  WaveLds L = wave_lds(base, G);
  const uint32_t tw_off = stage_seqs<B8>(a, pd, L.seq, lane);
  const uint64_t* qw = L.seq;
  const uint64_t* tw = L.seq + tw_off;
  const int32_t x = a.cp.x, o = a.cp.o, e = a.cp.e;
  const bool fitting = tn > qn + qn / 2 + 10;
  int32_t g = x;
  for (int32_t b = o; b != 0;) { const int32_t r = g % b; g = b; b = r; }
  for (int32_t b = e; b != 0;) { const int32_t r = g % b; g = b; b = r; }
  const int32_t xn = x / g, oen = (o + e) / g, en = e / g;
  const int32_t R = max(xn, max(oen, en)) + 1;
  const int64_t hard_max = (int64_t)(qn + tn) * max(x, 1) + o + e + 4;
  const int64_t explore_to = a.cp.cap >= 0 ? min((int64_t)a.cp.cap, hard_max) : hard_max;
  const int32_t max_step = (int32_t)(explore_to / g);
  int16_t* Rm = L.ring;
  int16_t* Ri = L.ring + R * WV_W;
  int16_t* Rd = L.ring + 2 * R * WV_W;

  const int32_t m0 = extend_diag<B8>(qw, tw, qn, tn, 0, 0);
  if (lane == 0) {
    // Step 0 holds only m on diagonal 0; its ins/del entries are unset.
    Rm[0] = (int16_t)m0;
    Ri[0] = -1;
    Rd[0] = -1;
  }
  wave_sync();
  int32_t status = ST_OK;
  int32_t step = 0, slot = 0;
  uint32_t used = 0;
  int32_t end_k = 0, end_layer = 0;
  bool found = fitting ? (m0 == qn) : (qn == tn && m0 == qn);  // s = 0 endpoint on diagonal 0
  const int32_t kend_g = qn - tn;
  while (!found) {
    step = __builtin_amdgcn_readfirstlane(step + 1);
    if (step > max_step) { status = ST_ALIGNMENT_FAILED; break; }
    slot = __builtin_amdgcn_readfirstlane(slot + 1 == R ? 0 : slot + 1);
    const int32_t lo = step_lo(step, tn), hi = step_hi(step, qn);
    const int32_t width = hi - lo + 1;
    if (width > WV_W) { status = ST_OVERFLOW_WIDTH; break; }
    if ((uint32_t)step >= G.steps_cap || used + (uint32_t)width > G.cells_cap) {
      status = ST_OVERFLOW_HIST;
      break;
    }
    if (lane == 0) L.toff[step] = (uint16_t)used;
    const int32_t sx = step - xn, so = step - oen, se = step - en;
    int32_t rx = slot - xn, ro = slot - oen, re = slot - en;
    rx += rx < 0 ? R : 0;
    ro += ro < 0 ? R : 0;
    re += re < 0 ? R : 0;
    // Predecessor ranges (empty when the step does not exist).
    const int32_t xlo = sx >= 0 ? step_lo(sx, tn) : 1, xhi = sx >= 0 ? step_hi(sx, qn) : 0;
    const int32_t olo = so >= 0 ? step_lo(so, tn) : 1, ohi = so >= 0 ? step_hi(so, qn) : 0;
    const int32_t elo = se >= 0 ? step_lo(se, tn) : 1, ehi = se >= 0 ? step_hi(se, qn) : 0;
    const int32_t ke_lo = fitting ? max(kend_g, lo) : kend_g;
    const int32_t ke_hi = fitting ? hi : kend_g;
    const int16_t* Rmx = Rm + rx * WV_W;
    const int16_t* Rmo = Rm + ro * WV_W;
    const int16_t* Rie = Ri + re * WV_W;
    const int16_t* Rde = Rd + re * WV_W;
    int16_t* Wm = Rm + slot * WV_W;
    int16_t* Wi = Ri + slot * WV_W;
    int16_t* Wd = Rd + slot * WV_W;
    uint16_t* cells = L.cells + used;
    const int32_t nch = (width + 63) >> 6;
    for (int32_t c = 0; c < nch; ++c) {
      const int32_t k = lo + c * 64 + lane;
      const uint32_t ik = (uint32_t)k & (WV_W - 1), il = (uint32_t)(k - 1) & (WV_W - 1),
                     ir = (uint32_t)(k + 1) & (WV_W - 1);
      const int32_t rvx = Rmx[ik], rvol = Rmo[il], rvor = Rmo[ir], rvil = Rie[il], rvdr = Rde[ir];
      const int32_t vx = (k >= xlo && k <= xhi) ? rvx : -1;
      const int32_t vol = (k - 1 >= olo && k - 1 <= ohi) ? rvol : -1;
      const int32_t vor = (k + 1 >= olo && k + 1 <= ohi) ? rvor : -1;
      const int32_t vil = (k - 1 >= elo && k - 1 <= ehi) ? rvil : -1;
      const int32_t vdr = (k + 1 >= elo && k + 1 <= ehi) ? rvdr : -1;
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
      int32_t mval = best;
      if (best >= 0) mval = extend_diag<B8>(qw, tw, qn, tn, best, k);
      int32_t layer = -1;
      if (k <= hi) {
        Wm[ik] = (int16_t)mval;
        Wi[ik] = (int16_t)ins;
        Wd[ik] = (int16_t)del;
        cells[k - lo] = (uint16_t)(((uint32_t)(best + 1) << 4) | mop | (iop << 2) | (dop << 3));
        if (k >= ke_lo && k <= ke_hi) layer = mval == qn ? 0 : ins == qn ? 1 : del == qn ? 2 : -1;
      }
      const uint64_t mask = __ballot(layer >= 0);
      if (mask && !found) {
        const int32_t ln = __builtin_ctzll(mask);
        found = true;
        end_k = lo + c * 64 + ln;
        end_layer = readlane(layer, ln);
      }
    }
    used += (uint32_t)width;
    wave_sync();
  }

  // Traceback (affine_traceback, wfa_simd.rs:785-886), wave-uniform state;
  // gap runs are resolved 64 cells at a time.
  PairOut r{status, 0, 0, 0, 0, 0, step, (int32_t)used};
  if (status == ST_OK) {
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
        const int32_t lo = step_lo(s, tn), hi = step_hi(s, qn);
        if (k < lo || k > hi) { r.status = ST_INTERNAL; break; }
        const uint32_t cell = L.cells[(uint32_t)L.toff[s] + (uint32_t)(k - lo)];
        const int32_t pre = (int32_t)(cell >> 4) - 1;
        if (pre < 0) { r.status = ST_INTERNAL; break; }
        const uint32_t mop = cell & 3;
        const int32_t run = max(qi - pre, 0);
        w.emit(OP_M, run);
        qi -= run; ti -= run;
        if (mop == 0) { w.emit(OP_X, 1); --qi; --ti; s -= xn; }
        else layer = (int32_t)mop;
      } else {
        // Gap layer: consecutive cells extend (op bit 1) until the one that
        // opened the gap (op bit 0) from the m layer o+e earlier.
        const int32_t dir = layer == 1 ? -1 : 1;
        const uint32_t bit = layer == 1 ? 2 : 3;
        bool opened = false;
        const int32_t n = gap_run(L, s, k, dir, en, bit, qn, tn, lane, &opened);
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
  if (lane == 0) *po = r;
}

// --- linear --------------------------------------------------------------------

template <bool B8>
__device__ void wave_linear(const WfaArgs& a, const WaveGeom& G, uint8_t* base,
                            const PairDesc& pd, PairOut* po, int lane) {
  const int32_t qn = pd.qn, tn = pd.tn;
  if (qn == 0) {  // wfa_simd.rs:194-203, :365-367
    if (lane == 0) *po = PairOut{ST_OK, 0, 0, 0, 0, 0, 0, 0};
    return;
  }
  WaveLds L = wave_lds(base, G);
  const uint32_t tw_off = stage_seqs<B8>(a, pd, L.seq, lane);
  const uint64_t* qw = L.seq;
  const uint64_t* tw = L.seq + tw_off;
  const bool fitting = tn > qn + qn / 2 + 10;
  const int32_t max_s = qn + tn;
  const int32_t explore = (fitting && a.cp.cap >= 0) ? min(a.cp.cap, max_s) : max_s;
  const int32_t kend_global = qn - tn;
  int16_t* ring = L.ring;  // two slots

  const int32_t w0 = extend_diag<B8>(qw, tw, qn, tn, 0, 0);
  if (lane == 0) ring[0] = (int16_t)w0;
  wave_sync();
  int32_t step = 0, status = ST_OK, end_k = 0;
  uint32_t used = 0;
  bool exhausted = false;
  bool found = fitting ? (w0 >= qn) : (kend_global == 0 && w0 == qn);
  while (!found) {
    step = __builtin_amdgcn_readfirstlane(step + 1);
    if (step > explore) { exhausted = true; break; }
    const int32_t slot = step & 1, ps = slot ^ 1;
    // Upstream's loop range for step s (wfa_simd.rs:288-291); the previous
    // step's range is the same formula at s - 1.
    const int32_t lo = step_lo(step, tn), hi = step_hi(step, qn);
    const int32_t p_lo = step_lo(step - 1, tn), p_hi = step_hi(step - 1, qn);
    const int32_t width = hi - lo + 1;
    if (width > WV_W) { status = ST_OVERFLOW_WIDTH; break; }
    if ((uint32_t)step >= G.steps_cap || used + (uint32_t)width > G.cells_cap) {
      status = ST_OVERFLOW_HIST;
      break;
    }
    if (lane == 0) L.toff[step] = (uint16_t)used;
    const int16_t* P = ring + ps * WV_W;
    int16_t* Wn = ring + slot * WV_W;
    uint16_t* cells = L.cells + used;
    const int32_t ke_lo = fitting ? max(qn - tn, lo) : kend_global;
    const int32_t ke_hi = fitting ? hi : kend_global;
    const int32_t nch = (width + 63) >> 6;
    for (int32_t c = 0; c < nch; ++c) {
      const int32_t k = lo + c * 64 + lane;
      const uint32_t ik = (uint32_t)k & (WV_W - 1), il = (uint32_t)(k - 1) & (WV_W - 1),
                     ir = (uint32_t)(k + 1) & (WV_W - 1);
      const int32_t rvm = P[ik], rvl = P[il], rvr = P[ir];
      const int32_t vm = (k >= p_lo && k <= p_hi) ? rvm : -1;
      const int32_t vl = (k > -tn && k - 1 >= p_lo && k - 1 <= p_hi) ? rvl : -1;
      const int32_t vr = (k < qn && k + 1 >= p_lo && k + 1 <= p_hi) ? rvr : -1;
      // Candidates [mismatch, query-only, target-only]; the last maximum wins;
      // then the winner alone is bounds-checked (wfa_simd.rs:316-326).
      int32_t best = -1;
      uint32_t op = 0;
      if (vm >= 0) { best = vm + 1; op = 1; }
      if (vl >= 0 && vl + 1 >= best) { best = vl + 1; op = 2; }
      if (vr >= 0 && vr >= best) { best = vr; op = 3; }
      const bool ok = op != 0 && best <= qn && best - k >= 0 && best - k <= tn;
      int32_t val = -1;
      if (ok) val = extend_diag<B8>(qw, tw, qn, tn, best, k);
      bool hit = false;
      if (k <= hi) {
        Wn[ik] = (int16_t)val;
        cells[k - lo] = (uint16_t)(ok ? (((uint32_t)(best + 1) << 4) | op) : 0u);
        hit = ok && k >= ke_lo && k <= ke_hi && val >= qn;
      }
      const uint64_t mask = __ballot(hit);
      if (mask && !found) {
        found = true;
        end_k = lo + c * 64 + (int32_t)__builtin_ctzll(mask);
      }
    }
    used += (uint32_t)width;
    wave_sync();
  }

  PairOut r{status, 0, 0, 0, 0, 0, step, (int32_t)used};
  if (status == ST_OK && exhausted) {
    if (fitting) {
      r.status = ST_ALIGNMENT_FAILED;
    } else {
      r.t_end = tn;  // fill_wfa's fallthrough ("", 0) (wfa_simd.rs:343-344)
      r.run_start = (int32_t)pd.run_cap;
    }
  } else if (status == ST_OK) {
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
        const uint32_t cell = L.cells[(uint32_t)L.toff[s] + (uint32_t)(k - lo)];
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
  if (lane == 0) *po = r;
}

// --- kernel --------------------------------------------------------------------

template <int MODE>
__global__ __launch_bounds__(64 * WV_WAVES) void wfa_wave_kernel(WfaArgs a, WaveGeom G) {
  extern __shared__ __align__(16) uint8_t wave_lds_pool[];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  uint8_t* base = wave_lds_pool + (size_t)wv * G.per_wave_bytes;
  for (;;) {
    uint32_t w = 0;
    if (lane == 0) w = atomicAdd(&a.counters[0], 1u);
    w = (uint32_t)readlane((int32_t)w, 0);
    if (w >= a.nwork) break;
    const uint32_t pi = a.work_list[w];
    const PairDesc pd = a.pairs[pi];
    PairOut* po = a.out + pi;
    if (MODE == ENG_AFFINE) {
      if (pd.flags & PF_BYTES) wave_affine<true>(a, G, base, pd, po, lane);
      else wave_affine<false>(a, G, base, pd, po, lane);
    } else {
      if (pd.flags & PF_BYTES) wave_linear<true>(a, G, base, pd, po, lane);
      else wave_linear<false>(a, G, base, pd, po, lane);
    }
    if (lane == 0 && po->status >= ST_OVERFLOW_HIST) atomicAdd(&a.counters[1], 1u);
    wave_sync();  // the wave's LDS region is reused by its next pair
  }
}

hipError_t launch_wfa_wave(int mode, const WfaArgs& a, const WaveGeom& G, uint32_t grid,
                           hipStream_t st) {
  const size_t lds = (size_t)G.per_wave_bytes * WV_WAVES;
  if (mode == ENG_AFFINE)
    hipLaunchKernelGGL(wfa_wave_kernel<ENG_AFFINE>, dim3(grid), dim3(64 * WV_WAVES), lds, st, a, G);
  else
    hipLaunchKernelGGL(wfa_wave_kernel<ENG_LINEAR>, dim3(grid), dim3(64 * WV_WAVES), lds, st, a, G);
  return hipGetLastError();
}

}  // namespace phx
