// myers_kernels.hip -- Myers/Hyyro bit-parallel edit distance on gfx950.
//
// One lane owns one pair; its query is split into NB 64-bit blocks kept in
// registers (NB <= 8, queries up to 512 symbols) or, for longer queries, in
// the lane's HBM scratch slot.  The per-column block update, the global top
// row (hp carry-in 1), the last-block mask and the score bit follow
// phraya-align/src/wfa_simd.rs:3552-3648; the fitting rule (abandon margin
// m/10 + 1, first strict minimum column) follows :3783-3814.
//
// With-CIGAR engines store (vp, vn) and the score per column and backtrace
// with the upstream priority D > I > diagonal (wfa_simd.rs:3702-3781).  The
// DP values the upstream rebuilds column by column are read here as
//   C_j[i] = score_j - popcount(vp_j >> i) + popcount(vn_j >> i)   (j >= 1)
//   C_0[i] = i
// which are the same integers (reconstruct_into, :3507-3516, sums the same
// vertical deltas from the bottom cell).
#include <hip/hip_runtime.h>

#include "phx_device.h"

namespace phx {

// 32-bit mask of the 2-bit fields of `w` equal to symbol c.
__device__ __forceinline__ uint32_t sym_mask32(uint64_t w, uint32_t c) {
  uint64_t x = w ^ (0x5555555555555555ull * c);
  x = ~(x | (x >> 1)) & 0x5555555555555555ull;
  x = (x | (x >> 1)) & 0x3333333333333333ull;
  x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x >> 4)) & 0x00FF00FF00FF00FFull;
  x = (x | (x >> 8)) & 0x0000FFFF0000FFFFull;
  x = (x | (x >> 16)) & 0x00000000FFFFFFFFull;
  return (uint32_t)x;
}

__device__ __forceinline__ uint32_t sym_at(const uint64_t* w, uint32_t i, bool b8) {
  return b8 ? (uint32_t)((w[i >> 3] >> ((i & 7) * 8)) & 0xFF) : (uint32_t)((w[i >> 5] >> ((i & 31) * 2)) & 3);
}

// Peq word of block b for 2-bit queries.
__device__ __forceinline__ uint64_t peq2(const uint64_t* qw, uint32_t b, uint32_t c, uint64_t mask) {
  return ((uint64_t)sym_mask32(qw[2 * b], c) | ((uint64_t)sym_mask32(qw[2 * b + 1], c) << 32)) & mask;
}

// Peq word of block b for byte queries (compared on the fly).
__device__ __forceinline__ uint64_t peq8(const uint64_t* qw, uint32_t b, uint32_t m, uint32_t tb) {
  uint64_t r = 0;
  const uint32_t base = b * 64;
  const uint32_t lim = min(64u, m - base);
  for (uint32_t i = 0; i < lim; ++i) r |= (uint64_t)(sym_at(qw, base + i, true) == tb) << i;
  return r;
}

// One Myers column over NB blocks held in arrays (registers when NB is a
// compile-time constant, HBM otherwise).  Returns the score delta.  Bits at
// or above the query length m in the last block are not cleared: every
// operation of the update moves information only upward (the addition's
// carries, the left shifts), so they never reach the score bit m - 1, and
// the traceback masks them out of its column sums (col_value).
template <typename Eq>
__device__ __forceinline__ int32_t myers_column(uint64_t* vp, uint64_t* vn, uint32_t nb,
                                                uint32_t score_bit, Eq eq_of) {
  unsigned long long add_carry = 0;
  uint64_t hp_in = 1, hn_in = 0, last_hp = 0, last_hn = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    const uint64_t eq = eq_of(b);
    const uint64_t v = vp[b], n = vn[b];
    const uint64_t xh = eq | n;
    // v + (xh & v) + carry in, carry out to the next block.
    unsigned long long cout;
    const uint64_t s2 = __builtin_addcll(v, xh & v, add_carry, &cout);
    add_carry = cout;
    const uint64_t d0 = (s2 ^ v) | xh;
    const uint64_t hn = v & d0;
    const uint64_t hp = n | ~(v | d0);
    last_hp = hp;
    last_hn = hn;
    const uint64_t x = (hp << 1) | hp_in;
    hp_in = hp >> 63;
    const uint64_t hs = (hn << 1) | hn_in;
    hn_in = hn >> 63;
    vn[b] = x & d0;
    vp[b] = hs | ~(d0 | x);
  }
  return (int32_t)((last_hp >> score_bit) & 1) - (int32_t)((last_hn >> score_bit) & 1);
}

// C_j[i] for j >= 1 from the stored column (vp, vn, score): the vertical
// deltas of rows i + 1 .. m (bits i .. m - 1; the last block's bits from m
// up are left over from the update and masked off with last_mask).
__device__ __forceinline__ int32_t col_value(const uint64_t* vp, const uint64_t* vn, uint32_t nb,
                                             uint64_t last_mask, int32_t score, uint32_t i) {
  int32_t pc = 0, nc = 0;
  const uint32_t b0 = i >> 6;
  for (uint32_t b = b0; b < nb; ++b) {
    uint64_t keep = b == b0 ? (~0ull << (i & 63)) : ~0ull;
    if (b + 1 == nb) keep &= last_mask;
    pc += __popcll(vp[b] & keep);
    nc += __popcll(vn[b] & keep);
  }
  return score - pc + nc;
}

struct MyersRunWriter {
  uint32_t* base;
  uint32_t cap, n, cur_op, cur_len;
  bool overflow;
  __device__ void init(uint32_t* b, uint32_t c) { base = b; cap = c; n = 0; cur_op = 0xFFFFFFFFu; cur_len = 0; overflow = false; }
  __device__ void flush() {
    if (cur_len == 0) return;
    if (n < cap) base[cap - 1 - n] = (cur_len << 4) | cur_op, ++n;
    else overflow = true;
    cur_len = 0;
  }
  __device__ void emit(uint32_t op) {
    if (op != cur_op) { flush(); cur_op = op; }
    ++cur_len;
  }
};

// NB > 0: compile-time block count (state in registers); NB == 0: runtime
// block count with state in the work item's scratch slot.
template <int NB>
__global__ __launch_bounds__(256) void myers_kernel(MyersArgs a) {
  // 2-bit queries of up to three blocks look their Peq words up in LDS,
  // [symbol][block][thread]: one read per block and column instead of a
  // four-way select of 64-bit registers; consecutive threads read
  // consecutive words, so the reads are free of bank conflicts.  Each
  // thread reads only its own entries (no barrier).
  constexpr bool PEQ_LDS = NB > 0 && NB <= 3;
  __shared__ uint64_t peq_lds[PEQ_LDS ? 4 * NB * 256 : 1];
  const uint32_t tid = threadIdx.x;
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= a.nwork) return;
  const uint32_t pi = a.work_list[w];
  const PairDesc pd = a.pairs[pi];
  const uint64_t* qw = pd.q;
  const uint64_t* tw = pd.t;
  const uint32_t tsh = pd.t_shift;  // the target may be a window of a resident reference
  const uint32_t m = (uint32_t)pd.qn, n = (uint32_t)pd.tn;
  const bool b8 = pd.flags & PF_BYTES;
  const bool with_cigar = a.engine != ENG_MYERS_DIST;
  const bool global_api = a.engine == ENG_MYERS_GLOBAL;
  PairOut r{ST_OK, 0, 0, 0, 0, 0, 0, 0};
  MyersRunWriter rw;
  rw.init(a.runs + pd.run_off, pd.run_cap);

  if (m == 0) {
    // global: (n, "nI"); fitting entry point: (0, "", 0) (wfa_simd.rs:3661-3664, :3787-3789)
    if (global_api) {
      r.score = r.edit_distance = (int32_t)n;
      r.t_end = (int32_t)n;
      if (with_cigar && n > 0) { rw.cur_op = OP_I; rw.cur_len = n; rw.flush(); }
    }
  } else if (n == 0) {
    r.score = r.edit_distance = (int32_t)m;  // (m, "mD"), consumed 0
    if (with_cigar) { rw.cur_op = OP_D; rw.cur_len = m; rw.flush(); }
  } else {
    const bool fitting = !global_api && n > m + m / 2 + 10;
    const uint32_t margin = m / 10 + 1;
    const uint32_t nb = NB > 0 ? (uint32_t)NB : (m + 63) / 64;
    const uint32_t last_bits = m % 64 == 0 ? 64 : m % 64;
    const bool last_partial = last_bits < 64;
    const uint64_t last_mask = last_partial ? ((1ull << last_bits) - 1) : ~0ull;
    const uint32_t score_bit = (m - 1) % 64;

    uint8_t* slot = a.scratch + (size_t)w * a.scratch_stride;
    // Scratch layout: [generic state: vp nb, vn nb, peq 4*nb] then, for CIGAR
    // engines, per column {vp nb, vn nb} and a separate i32 score array.
    uint64_t regs_vp[NB > 0 ? NB : 1], regs_vn[NB > 0 ? NB : 1], regs_peq[4][NB > 0 ? NB : 1];
    uint64_t* vp = NB > 0 ? regs_vp : reinterpret_cast<uint64_t*>(slot);
    uint64_t* vn = NB > 0 ? regs_vn : vp + nb;
    uint64_t* gpeq = vn + nb;  // generic 2-bit Peq, 4 * nb words
    const size_t state_bytes = NB > 0 ? 0 : (size_t)6 * nb * 8;
    uint64_t* colstore = reinterpret_cast<uint64_t*>(slot + state_bytes);
    int32_t* colscore = reinterpret_cast<int32_t*>(colstore + (size_t)2 * nb * n);

    for (uint32_t b = 0; b < nb; ++b) {
      vp[b] = (b + 1 == nb) ? last_mask : ~0ull;
      vn[b] = 0;
      if (!b8) {
        const uint64_t msk = (b + 1 == nb) ? last_mask : ~0ull;
        for (uint32_t c = 0; c < 4; ++c) {
          const uint64_t pe = peq2(qw, b, c, msk);
          if (PEQ_LDS) peq_lds[(c * NB + b) * 256 + tid] = pe;
          else if (NB > 0) regs_peq[c][b] = pe;
          else gpeq[c * nb + b] = pe;
        }
      }
    }

    int32_t score = (int32_t)m, best_seen = (int32_t)m;
    int32_t best_score = (int32_t)m, best_consumed = 0;
    uint32_t ncols = n;
    uint64_t tword = 0;
    for (uint32_t j = 0; j < n; ++j) {
      uint32_t tb;
      const uint32_t jt = j + tsh;
      if (b8) {
        tb = sym_at(tw, jt, true);
      } else {
        if (j == 0 || (jt & 31) == 0) tword = tw[jt >> 5];
        tb = (uint32_t)(tword >> ((jt & 31) * 2)) & 3;
      }
      int32_t delta;
      if (b8) {
        delta = myers_column(vp, vn, nb, score_bit,
                             [&](uint32_t b) { return peq8(qw, b, m, tb); });
      } else if (PEQ_LDS) {
        const uint64_t* row = peq_lds + tb * (NB * 256) + tid;
        delta = myers_column(vp, vn, nb, score_bit, [&](uint32_t b) { return row[b * 256]; });
      } else if (NB > 0) {
        delta = myers_column(vp, vn, nb, score_bit, [&](uint32_t b) {
          const uint64_t p0 = regs_peq[0][b], p1 = regs_peq[1][b], p2 = regs_peq[2][b], p3 = regs_peq[3][b];
          return tb == 0 ? p0 : tb == 1 ? p1 : tb == 2 ? p2 : p3;
        });
      } else {
        delta = myers_column(vp, vn, nb, score_bit,
                             [&](uint32_t b) { return gpeq[tb * nb + b]; });
      }
      score += delta;
      if (score < 0) score = 0;  // saturating_sub
      if (with_cigar) {
        uint64_t* cs = colstore + (size_t)2 * nb * j;
        for (uint32_t b = 0; b < nb; ++b) { cs[b] = vp[b]; cs[nb + b] = vn[b]; }
        colscore[j] = score;
      }
      if (score < best_seen) best_seen = score;
      if (score < best_score) { best_score = score; best_consumed = (int32_t)j + 1; }
      if (fitting && score > best_seen + (int32_t)margin) { ncols = j + 1; break; }
    }
    (void)ncols;
    int32_t ed, end_col;
    if (fitting) { ed = best_score; end_col = best_consumed; }
    else { ed = score; end_col = (int32_t)n; }
    r.score = r.edit_distance = ed;
    r.t_end = end_col;
    if (with_cigar) {
      // Backtrace from (m, end_col); priority D, then I, then diagonal.
      uint32_t qi = m, ti = (uint32_t)end_col;
      while (qi > 0 || ti > 0) {
        if (qi == 0) { rw.emit(OP_I); --ti; continue; }
        if (ti == 0) { rw.emit(OP_D); --qi; continue; }
        const uint64_t* cvp = colstore + (size_t)2 * nb * (ti - 1);
        const uint64_t* cvn = cvp + nb;
        const int32_t cur = col_value(cvp, cvn, nb, last_mask, colscore[ti - 1], qi);
        const bool vbit = (cvp[(qi - 1) >> 6] >> ((qi - 1) & 63)) & 1;  // C[qi] == C[qi-1] + 1
        if (vbit) { rw.emit(OP_D); --qi; continue; }
        int32_t left;
        if (ti - 1 == 0) {
          left = (int32_t)qi;
        } else {
          const uint64_t* pvp = colstore + (size_t)2 * nb * (ti - 2);
          left = col_value(pvp, pvp + nb, nb, last_mask, colscore[ti - 2], qi);
        }
        if (cur == left + 1) { rw.emit(OP_I); --ti; continue; }
        const bool eq = sym_at(qw, qi - 1, b8) == sym_at(tw, ti - 1 + tsh, b8);
        rw.emit(eq ? OP_M : OP_X);
        --qi; --ti;
      }
    }
  }
  if (with_cigar) {
    rw.flush();
    if (rw.overflow) r.status = ST_OVERFLOW_RUNS;
    r.nruns = (int32_t)rw.n;
    r.run_start = (int32_t)(pd.run_cap - rw.n);
  } else {
    r.run_start = (int32_t)pd.run_cap;
  }
  a.out[pi] = r;
}

hipError_t launch_myers(int nb_template, const MyersArgs& a, hipStream_t st) {
  const dim3 grid((a.nwork + 255) / 256), block(256);
  switch (nb_template) {
    case 1: hipLaunchKernelGGL(myers_kernel<1>, grid, block, 0, st, a); break;
    case 2: hipLaunchKernelGGL(myers_kernel<2>, grid, block, 0, st, a); break;
    case 3: hipLaunchKernelGGL(myers_kernel<3>, grid, block, 0, st, a); break;
    case 4: hipLaunchKernelGGL(myers_kernel<4>, grid, block, 0, st, a); break;
    case 5: hipLaunchKernelGGL(myers_kernel<5>, grid, block, 0, st, a); break;
    case 6: hipLaunchKernelGGL(myers_kernel<6>, grid, block, 0, st, a); break;
    case 7: hipLaunchKernelGGL(myers_kernel<7>, grid, block, 0, st, a); break;
    case 8: hipLaunchKernelGGL(myers_kernel<8>, grid, block, 0, st, a); break;
    default: hipLaunchKernelGGL(myers_kernel<0>, grid, block, 0, st, a); break;
  }
  return hipGetLastError();
}

}  // namespace phx
