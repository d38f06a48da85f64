// alncheck.cpp -- TEST INFRASTRUCTURE ONLY: full-size checks of alignments
// returned by the HIP engines, for pairs too long for the oracle (which keeps
// phraya-align's full-width score frames).  Nothing in the product links it.
//
// Per pair (global branch: the whole query against the whole target):
//   * CIGAR replay: M bases equal, X bases different, no run leaves either
//     sequence; the query is consumed fully and the target up to t_end;
//   * the CIGAR's cost equals the reported score (affine: X = mismatch, each
//     I / D run gap_open + len * gap_extend; linear: every non-M base 1);
//   * the reported edit distance equals the sum of non-M run lengths
//     (levenshtein_from_cigar, phraya-align/src/wfa_simd.rs:893-908);
//   * t_end equals the target length;
//   * the score equals the optimum of a banded Gotoh DP (the recurrence of
//     wfa_simd.rs:3118-3170 with the gap-state switch costing a new gap
//     open), over the band that holds every alignment of cost <= score: a
//     path reaching diagonal k = i - j needs gaps of total length at least
//     |k| + |k - kend| (kend = qn - tn), so one of cost <= S never leaves
//     [min(0, kend) - d, max(0, kend) + d] with d = (S - o - e |kend|) / (2e).
//     The banded optimum is therefore the true optimum whenever that is <= S,
//     and the CIGAR itself shows it is <= S.  Linear WFA is the same DP with
//     mismatch 1, gap_open 0, gap_extend 1.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace {

constexpr int64_t kInf = INT64_MAX / 4;

enum Flag : uint8_t {
  F_REPLAY = 1,    // an M base differs, an X base matches, or a run leaves a sequence
  F_CONSUME = 2,   // query not fully consumed or target consumed != t_end
  F_COST = 4,      // cost of the CIGAR != score
  F_ED = 8,        // edit distance != sum of non-M run lengths
  F_OPT = 16,      // score != banded DP optimum
  F_TEND = 32,     // t_end != target length (global branch)
};

// Banded global Gotoh optimum, or kInf when no alignment of cost <= S lies in
// the band (then the optimum exceeds S).  cells (may be null) gets the cells
// computed.  Scores are int32 with kBig as "unset": real scores stay below
// 2^27 (the callers' scores are int32 and pairs at most 2^27 bases), and an
// unset value drifts by at most max(x, o + e) per row and e * band below, so
// anything at or above kUnset is unset.
constexpr int32_t kBig = 1 << 30;
constexpr int32_t kUnset = 1 << 29;

// The band of diagonals k = i - j holding every global alignment of cost <= S
// (false when there is none: S below the unavoidable gap cost).
bool band_of(int64_t qn, int64_t tn, int64_t o, int64_t e, int64_t S, int64_t* klo, int64_t* khi) {
  if (S < 0 || e < 1 || S >= (1 << 27) || qn >= (1 << 27) || tn >= (1 << 27)) return false;
  const int64_t kend = qn - tn;
  const int64_t akend = kend < 0 ? -kend : kend;
  int64_t d = 0;
  if (kend != 0 || S >= o + 2 * e) {
    const int64_t extra = S - o - e * akend;
    if (extra < 0) {
      if (kend != 0) return false;  // any path needs a gap of |kend| bases
    } else {
      d = extra / (2 * e);
    }
  }
  *klo = std::max(std::min<int64_t>(0, kend) - d, -tn);
  *khi = std::min(std::max<int64_t>(0, kend) + d, qn);
  return true;
}

// Row by row (query position i), diagonals k of a row in one pass plus a
// suffix-minimum scan for the target-only state.  A second implementation of
// band_global, kept to cross-check it.
int64_t band_global_rows(const uint8_t* q, int64_t qn, const uint8_t* t, int64_t tn, int64_t x64,
                         int64_t o64, int64_t e64, int64_t S, uint64_t* cells) {
  if (cells) *cells = 0;
  int64_t klo = 0, khi = 0;
  if (!band_of(qn, tn, o64, e64, S, &klo, &khi)) return kInf;
  const int32_t x = (int32_t)x64, o = (int32_t)o64, e = (int32_t)e64;
  const int64_t kend = qn - tn;
  const int64_t W = khi - klo + 1;
  // Arrays indexed by a = k - klo + 1, with an unset sentinel at both ends.
  std::vector<int32_t> pm(W + 2, kBig), pv(W + 2, kBig), ph(W + 2, kBig);
  std::vector<int32_t> cm(W + 2, kBig), cv(W + 2, kBig), ch(W + 2, kBig), u(W + 2, kBig);
  // The target reversed: cell (i, k) compares q[i - 1] with t[i - 1 - k] =
  // tr[tn - i + k], ascending in k.
  std::vector<uint8_t> tr((size_t)tn + 1);
  for (int64_t j = 0; j < tn; ++j) tr[(size_t)(tn - 1 - j)] = t[j];
  const int32_t oe = o + e;
  // Row 0: (0, 0) in M; (0, j) for j > 0 is a target-only gap.
  for (int64_t k = klo; k <= khi; ++k) {
    const int64_t j = -k, a = k - klo + 1;
    if (j < 0 || j > tn) continue;
    if (j == 0) pm[a] = 0;
    else ph[a] = o + e * (int32_t)j;
  }
  uint64_t nc = (uint64_t)W;
  // Row i's cells are k in [lo, hi], both non-decreasing in i.  Every cell of
  // that range is written, and the entry just below it is cleared: the next
  // row reads k - 1 >= lo - 1, and entries above hi were never written in
  // either buffer.
  for (int64_t i = 1; i <= qn; ++i) {
    const int64_t lo = std::max(klo, i - tn), hi = std::min(khi, i);
    const int64_t alo = lo - klo + 1, ahi = hi - klo + 1;
    const uint8_t qc = q[i - 1];
    cm[alo - 1] = cv[alo - 1] = ch[alo - 1] = kBig;
    const int32_t* __restrict PM = pm.data();
    const int32_t* __restrict PV = pv.data();
    const int32_t* __restrict PH = ph.data();
    int32_t* __restrict CM = cm.data();
    int32_t* __restrict CV = cv.data();
    int32_t* __restrict U = u.data();
    const uint8_t* __restrict T = tr.data() + (tn - i) + (klo - 1);  // T[a] = tr[tn - i + k]
    // (i-1, j-1) is the same diagonal; (i-1, j) is diagonal k - 1 (a
    // query-only base, D).  U holds min(M, V) + o + e + e * (a - 1) for the
    // target-only scan below.
    for (int64_t a = alo; a <= ahi; ++a) {
      const int32_t best = std::min(PM[a], std::min(PV[a], PH[a]));
      const int32_t m = best + (qc == T[a] ? 0 : x);
      const int32_t v = std::min(std::min(PM[a - 1], PH[a - 1]) + oe, PV[a - 1] + e);
      CM[a] = m;
      CV[a] = v;
      U[a] = std::min(m, v) + oe + e * (int32_t)(a - 1);
    }
    if (hi == i) {  // j = 0: only a query-only gap reaches it
      cm[ahi] = kBig;
      u[ahi] = kBig;
    }
    // (i, j-1) is diagonal k + 1 of this row (a target-only base, I): a run
    // of I from cell a' > a costs o + e + e * (a' - 1 - a) more than min(M, V)
    // there, so H[a] = min over a' > a of U[a'] - e * a.
    int32_t acc = kBig;
    for (int64_t a = ahi; a >= alo; --a) {
      ch[a] = acc - e * (int32_t)a;
      acc = std::min(acc, u[a]);
    }
    std::swap(pm, cm);
    std::swap(pv, cv);
    std::swap(ph, ch);
    nc += (uint64_t)(hi - lo + 1);
  }
  if (cells) *cells = nc;
  if (kend < klo || kend > khi) return kInf;
  const int64_t a = kend - klo + 1;
  const int32_t r = std::min(pm[a], std::min(pv[a], ph[a]));
  return r >= kUnset ? kInf : r;
}

// Anti-diagonal by anti-diagonal (s = i + j): M reads s - 2, the query-only
// (V, from (i-1, j)) and target-only (H, from (i, j-1)) states read s - 1, so
// every cell of an anti-diagonal is independent of the others and the loop
// vectorises.  Cells of anti-diagonal s are the diagonals k = s (mod 2) of
// the band, stored at m = (k - base_s) / 2 with base_s = klo or klo + 1.
int64_t band_global(const uint8_t* q, int64_t qn, const uint8_t* t, int64_t tn, int64_t x64, int64_t o64,
                    int64_t e64, int64_t S, uint64_t* cells) {
  if (cells) *cells = 0;
  int64_t klo = 0, khi = 0;
  if (!band_of(qn, tn, o64, e64, S, &klo, &khi)) return kInf;
  const int32_t x = (int32_t)x64, o = (int32_t)o64, e = (int32_t)e64, oe = o + e;
  const int64_t kend = qn - tn;
  if (kend < klo || kend > khi) return kInf;
  const int64_t W2 = (khi - klo) / 2 + 1;
  const int64_t pad = 2;  // entries m = -2, -1 and W2, W2 + 1 exist and stay unset unless cleared
  std::vector<int32_t> buf[3][3];  // [s % 3][M, V, H]
  for (auto& b : buf)
    for (auto& a : b) a.assign((size_t)(W2 + 2 * pad), kBig);
  // qp[i] = q[i - 1] and tp[j] = t[j - 1] with a pad byte at index 0; the
  // target reversed so that j descending is an ascending address.
  std::vector<uint8_t> qp((size_t)qn + 2, 0), tr((size_t)tn + 2, 0);
  for (int64_t i = 0; i < qn; ++i) qp[(size_t)i + 1] = q[i];
  for (int64_t j = 0; j < tn; ++j) tr[(size_t)(tn - 1 - j)] = t[j];  // tr[tn - j] = t[j - 1]
  // s = 0: the origin, k = 0.
  {
    const int64_t base = (klo % 2 == 0) ? klo : klo + 1;
    buf[0][0][(size_t)(pad + (0 - base) / 2)] = 0;
  }
  uint64_t nc = 1;
  for (int64_t s = 1; s <= qn + tn; ++s) {
    const int64_t par = ((s - klo) % 2 + 2) % 2;  // base_s = klo + par
    const int64_t base = klo + par;
    // Valid diagonals of s: k in [klo, khi], 0 <= i = (s + k) / 2 <= qn,
    // 0 <= j = (s - k) / 2 <= tn.
    int64_t lo = std::max(std::max(klo, -s), s - 2 * tn);
    int64_t hi = std::min(std::min(khi, 2 * qn - s), s);
    if (((lo - base) % 2 + 2) % 2) ++lo;
    if (((hi - base) % 2 + 2) % 2) --hi;
    int32_t* __restrict CM = buf[s % 3][0].data() + pad;
    int32_t* __restrict CV = buf[s % 3][1].data() + pad;
    int32_t* __restrict CH = buf[s % 3][2].data() + pad;
    if (lo > hi) {  // nothing of s lies in the band: all unset
      std::fill(buf[s % 3][0].begin(), buf[s % 3][0].end(), kBig);
      std::fill(buf[s % 3][1].begin(), buf[s % 3][1].end(), kBig);
      std::fill(buf[s % 3][2].begin(), buf[s % 3][2].end(), kBig);
      continue;
    }
    const int64_t mlo = (lo - base) / 2, mhi = (hi - base) / 2;
    const int32_t* __restrict P2M = buf[(s + 1) % 3][0].data() + pad;  // s - 2
    const int32_t* __restrict P2V = buf[(s + 1) % 3][1].data() + pad;
    const int32_t* __restrict P2H = buf[(s + 1) % 3][2].data() + pad;
    // s - 1 has base klo + (1 - par): (s-1, k-1) is at m - 1 + par, (s-1, k+1)
    // at m + par.
    const int32_t* __restrict P1M = buf[(s + 2) % 3][0].data() + pad + par;
    const int32_t* __restrict P1V = buf[(s + 2) % 3][1].data() + pad + par;
    const int32_t* __restrict P1H = buf[(s + 2) % 3][2].data() + pad + par;
    // Cell m: i = (s + base) / 2 + m, j = (s - base) / 2 - m.
    const int64_t i0 = (s + base) / 2, j0 = (s - base) / 2;
    const uint8_t* __restrict QP = qp.data() + i0;            // QP[m] = qp[i]
    const uint8_t* __restrict TR = tr.data() + (tn - j0);     // TR[m] = tr[tn - j] = t[j - 1]
    // The three anti-diagonals live in distinct buffers.
#pragma GCC ivdep
    for (int64_t m = mlo; m <= mhi; ++m) {
      const int32_t best = std::min(P2M[m], std::min(P2V[m], P2H[m]));
      CM[m] = best + (QP[m] == TR[m] ? 0 : x);
      CV[m] = std::min(std::min(P1M[m - 1], P1H[m - 1]) + oe, P1V[m - 1] + e);
      CH[m] = std::min(std::min(P1M[m], P1V[m]) + oe, P1H[m] + e);
    }
    // Border cells: i = 0 has no M or V state, j = 0 no M or H state.
    if (i0 + mlo == 0) CM[mlo] = CV[mlo] = kBig;
    if (j0 - mhi == 0) CM[mhi] = CH[mhi] = kBig;
    // Entries just outside the range: the next two anti-diagonals read at
    // most one entry beyond it on each side.
    CM[mlo - 1] = CV[mlo - 1] = CH[mlo - 1] = kBig;
    CM[mhi + 1] = CV[mhi + 1] = CH[mhi + 1] = kBig;
    nc += (uint64_t)(mhi - mlo + 1);
  }
  if (cells) *cells = nc;
  const int64_t s = qn + tn;
  const int64_t base = klo + ((s - klo) % 2 + 2) % 2;
  const int64_t m = (kend - base) / 2;
  const int32_t r = std::min(buf[s % 3][0][(size_t)(pad + m)],
                             std::min(buf[s % 3][1][(size_t)(pad + m)], buf[s % 3][2][(size_t)(pad + m)]));
  return r >= kUnset ? kInf : r;
}

uint8_t check_one(int mode, int64_t x, int64_t o, int64_t e, const uint8_t* q, int64_t qn,
                  const uint8_t* t, int64_t tn, int64_t score, int64_t ed, int64_t t_end,
                  const uint32_t* runs, uint64_t nruns, int optimal, uint64_t* cells) {
  if (mode == 0) {
    x = 1;
    o = 0;
    e = 1;
  }
  uint8_t f = 0;
  int64_t qi = 0, ti = 0, cost = 0, nonm = 0;
  for (uint64_t r = 0; r < nruns; ++r) {
    const int64_t len = runs[r] >> 4;
    const uint32_t op = runs[r] & 15u;
    if (len <= 0 || op > 3) {
      f |= F_REPLAY;
      break;
    }
    if (op == 0 || op == 1) {
      if (qi + len > qn || ti + len > tn) {
        f |= F_REPLAY;
        break;
      }
      for (int64_t d = 0; d < len; ++d)
        if ((q[qi + d] == t[ti + d]) != (op == 0)) f |= F_REPLAY;
      qi += len;
      ti += len;
      if (op == 1) {
        cost += len * x;
        nonm += len;
      }
    } else if (op == 2) {  // I: target-only
      if (ti + len > tn) {
        f |= F_REPLAY;
        break;
      }
      ti += len;
      cost += o + len * e;
      nonm += len;
    } else {  // D: query-only
      if (qi + len > qn) {
        f |= F_REPLAY;
        break;
      }
      qi += len;
      cost += o + len * e;
      nonm += len;
    }
  }
  if (qi != qn || ti != t_end) f |= F_CONSUME;
  if (cost != score) f |= F_COST;
  if (nonm != ed) f |= F_ED;
  if (t_end != tn) f |= F_TEND;
  if (optimal && band_global(q, qn, t, tn, x, o, e, score, cells) != score) f |= F_OPT;
  return f;
}

}  // namespace

extern "C" {

// Single-pair banded optimum (mode 0 linear, 1 affine); -1 when no alignment
// of cost <= S lies in the band.
int64_t chk_band_global(int mode, int64_t x, int64_t o, int64_t e, const uint8_t* q, uint64_t qn,
                        const uint8_t* t, uint64_t tn, int64_t S, uint64_t* cells) {
  if (mode == 0) {
    x = 1;
    o = 0;
    e = 1;
  }
  const int64_t r = band_global(q, (int64_t)qn, t, (int64_t)tn, x, o, e, S, cells);
  return r >= kInf ? -1 : r;
}

int64_t chk_band_global_rows(int mode, int64_t x, int64_t o, int64_t e, const uint8_t* q, uint64_t qn,
                             const uint8_t* t, uint64_t tn, int64_t S, uint64_t* cells) {
  if (mode == 0) {
    x = 1;
    o = 0;
    e = 1;
  }
  const int64_t r = band_global_rows(q, (int64_t)qn, t, (int64_t)tn, x, o, e, S, cells);
  return r >= kInf ? -1 : r;
}

// All pairs of a batch on `threads` host threads (longest first); flags[i]
// gets pair i's F_* bits (0 = every check passed).  Returns the total DP cells.
uint64_t chk_alignments(int mode, int64_t x, int64_t o, int64_t e, uint32_t n, const uint8_t* q_seq,
                        const uint64_t* q_off, const uint8_t* t_seq, const uint64_t* t_off,
                        const int64_t* score, const int64_t* ed, const int64_t* t_end,
                        const uint32_t* runs, const uint64_t* run_off, int optimal, int threads,
                        uint8_t* flags) {
  std::vector<uint32_t> order(n);
  for (uint32_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    return (q_off[a + 1] - q_off[a]) + (t_off[a + 1] - t_off[a]) >
           (q_off[b + 1] - q_off[b]) + (t_off[b + 1] - t_off[b]);
  });
  std::atomic<uint32_t> next{0};
  std::atomic<uint64_t> total{0};
  auto work = [&] {
    uint64_t mine = 0;
    for (uint32_t k; (k = next.fetch_add(1)) < n;) {
      const uint32_t i = order[k];
      uint64_t c = 0;
      flags[i] = check_one(mode, x, o, e, q_seq + q_off[i], (int64_t)(q_off[i + 1] - q_off[i]),
                           t_seq + t_off[i], (int64_t)(t_off[i + 1] - t_off[i]), score[i], ed[i],
                           t_end[i], runs + run_off[i], run_off[i + 1] - run_off[i], optimal, &c);
      mine += c;
    }
    total += mine;
  };
  std::vector<std::thread> pool;
  for (int k = 0; k < std::max(1, threads); ++k) pool.emplace_back(work);
  for (auto& th : pool) th.join();
  return total.load();
}

}  // extern "C"
