// phraya_oracle.cpp -- CPU restatement of phraya-align's extension engines.
//
// TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the CPU
// baseline timed beside it.  The product never links or calls this file.
//
// Each function below follows the upstream Rust function named in its comment
// (paths relative to the upstream repository root).  The restatement keeps
// the upstream's data layout where it shapes cost (full-width frames pushed
// per score, an O(m) column rebuild per Myers backtrace step) so that the
// CPU baseline times the same algorithm the upstream runs.
#include "phraya_oracle.h"

#include <emmintrin.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Longest common prefix (phraya-align/src/wfa_simd.rs:52-155)
// ---------------------------------------------------------------------------

// Byte-by-byte tier (wfa_simd.rs:123-130).
inline size_t lcp_scalar(const uint8_t* a, size_t an, const uint8_t* b, size_t bn) {
  size_t n = std::min(an, bn), i = 0;
  while (i < n && a[i] == b[i]) ++i;
  return i;
}

// 16-byte SSE2 tier used in production on x86_64 (wfa_simd.rs:69-90): compare
// 16 bytes, movemask, first zero bit of the mask is the first mismatch; the
// sub-16-byte tail goes to the scalar tier.
inline size_t lcp_sse2(const uint8_t* a, size_t an, const uint8_t* b, size_t bn) {
  size_t n = std::min(an, bn), i = 0;
  while (i + 16 <= n) {
    __m128i va = _mm_loadu_si128(reinterpret_cast<const __m128i*>(a + i));
    __m128i vb = _mm_loadu_si128(reinterpret_cast<const __m128i*>(b + i));
    unsigned mask = static_cast<unsigned>(_mm_movemask_epi8(_mm_cmpeq_epi8(va, vb))) & 0xFFFFu;
    if (mask != 0xFFFFu) return i + static_cast<size_t>(__builtin_ctz(mask ^ 0xFFFFu));
    i += 16;
  }
  return i + lcp_scalar(a + i, n - i, b + i, n - i);
}

// 8-byte little-endian XOR tier (wfa_simd.rs:139-155).
inline size_t lcp_u64(const uint8_t* a, size_t an, const uint8_t* b, size_t bn) {
  size_t n = std::min(an, bn), i = 0;
  while (i + 8 <= n) {
    uint64_t aw, bw;
    std::memcpy(&aw, a + i, 8);
    std::memcpy(&bw, b + i, 8);
    uint64_t diff = aw ^ bw;
    if (diff) return i + static_cast<size_t>(__builtin_ctzll(diff)) / 8;
    i += 8;
  }
  while (i < n && a[i] == b[i]) ++i;
  return i;
}

inline size_t lcp(const uint8_t* a, size_t an, const uint8_t* b, size_t bn) {
  return lcp_sse2(a, an, b, bn);
}

// ---------------------------------------------------------------------------
// CIGAR helpers
// ---------------------------------------------------------------------------

// compact_cigar (wfa_simd.rs:1200-1222): run-length encode op characters.
std::string compact_cigar(const std::vector<char>& ops) {
  std::string out;
  if (ops.empty()) return out;
  char cur = ops[0];
  size_t count = 1;
  for (size_t i = 1; i < ops.size(); ++i) {
    if (ops[i] == cur) {
      ++count;
    } else {
      out += std::to_string(count);
      out.push_back(cur);
      cur = ops[i];
      count = 1;
    }
  }
  out += std::to_string(count);
  out.push_back(cur);
  return out;
}

// levenshtein_from_cigar (wfa_simd.rs:893-908): sum of all non-M run lengths.
size_t levenshtein_from_cigar(const std::string& cigar) {
  size_t total = 0, num = 0;
  for (char ch : cigar) {
    if (ch >= '0' && ch <= '9') {
      num = num * 10 + static_cast<size_t>(ch - '0');
    } else {
      if (ch != 'M') total += num;
      num = 0;
    }
  }
  return total;
}

// ---------------------------------------------------------------------------
// Linear (unit-cost) WFA (wfa_simd.rs:231-488, :940-1130)
// ---------------------------------------------------------------------------

constexpr int32_t UNSET = INT32_MIN;

struct LinFrame {
  std::vector<int32_t> wf;
  std::vector<uint8_t> op;  // 0 none, 1 X (same k), 2 query-only 'D' (from k-1), 3 target-only 'I' (from k+1)
};

struct Seqs {
  const uint8_t* q;
  int32_t qn;
  const uint8_t* t;
  int32_t tn;
  // Greedy diagonal extension (wfa_simd.rs:252-262): unchanged outside the
  // matrix, else position + LCP of the two suffixes.
  int32_t extend(int32_t qp, int32_t k) const {
    int32_t j = qp - k;
    if (qp < 0 || qp >= qn || j < 0 || j >= tn) return qp;
    return qp + static_cast<int32_t>(lcp(q + qp, static_cast<size_t>(qn - qp),
                                         t + j, static_cast<size_t>(tn - j)));
  }
};

// One score step of the linear recurrence, identical in the global
// (wfa_simd.rs:283-330) and fitting (:430-469) loops.
void linear_step(const Seqs& S, int32_t s, const std::vector<int32_t>& prev,
                 std::vector<int32_t>& next, std::vector<uint8_t>& ops) {
  const int32_t qn = S.qn, tn = S.tn;
  const int32_t lo = -std::min(s, tn);
  const int32_t hi = std::min(s, qn);
  for (int32_t k = lo; k <= hi; ++k) {
    const size_t ki = static_cast<size_t>(k + tn);
    // Candidates in upstream order [mismatch, ins, del]; max_by_key keeps the
    // LAST maximum, so a later candidate wins ties.
    int32_t best = UNSET;
    uint8_t best_op = 0;
    bool have = false;
    auto offer = [&](int32_t p, uint8_t op) {
      if (!have || p >= best) {
        best = p;
        best_op = op;
        have = true;
      }
    };
    if (prev[ki] != UNSET) offer(prev[ki] + 1, 1);
    if (k > -tn) {
      size_t pk = static_cast<size_t>(k - 1 + tn);
      if (prev[pk] != UNSET) offer(prev[pk] + 1, 2);
    }
    if (k < qn) {
      size_t pk = static_cast<size_t>(k + 1 + tn);
      if (prev[pk] != UNSET) offer(prev[pk], 3);
    }
    if (!have) continue;
    // Bounds check after the choice: an out-of-range winner leaves the cell
    // unset, with no fallback to another candidate (wfa_simd.rs:324-326).
    int32_t j = best - k;
    if (best < 0 || best > qn || j < 0 || j > tn) continue;
    next[ki] = S.extend(best, k);
    ops[ki] = best_op;
  }
}

// traceback_wfa / traceback_wfa_with_tend (wfa_simd.rs:959-1130).  The two
// upstream functions differ only in the starting target index.
std::string linear_traceback(const std::vector<LinFrame>& hist, const Seqs& S,
                             int32_t edit_dist, int32_t t_end) {
  const int32_t tn = S.tn;
  std::vector<char> ops;
  int32_t qi = S.qn, ti = t_end;
  int32_t k = qi - ti;
  int32_t s = edit_dist;
  auto prev_at = [&](const std::vector<int32_t>& wf, int64_t pk) -> int32_t {
    return (pk >= 0 && pk < static_cast<int64_t>(wf.size())) ? wf[static_cast<size_t>(pk)] : 0;
  };
  while (qi > 0 || ti > 0) {
    if (s < 0) break;
    int64_t ki = static_cast<int64_t>(k) + tn;
    if (s == 0) {
      while (qi > 0 && ti > 0) {
        ops.push_back('M');
        --qi;
        --ti;
      }
      break;
    }
    const LinFrame& cur = hist[static_cast<size_t>(s)];
    uint8_t op = (ki >= 0 && ki < static_cast<int64_t>(cur.op.size())) ? cur.op[static_cast<size_t>(ki)] : 0;
    const std::vector<int32_t>& pw = hist[static_cast<size_t>(s) - 1].wf;
    int32_t pred_pos, pred_k;
    switch (op) {
      case 1: pred_pos = prev_at(pw, ki); pred_k = k; break;
      case 2: pred_pos = prev_at(pw, ki - 1); pred_k = k - 1; break;
      case 3: pred_pos = prev_at(pw, ki + 1); pred_k = k + 1; break;
      default: pred_pos = prev_at(pw, ki); pred_k = k; break;
    }
    int32_t edit_start = (op == 1 || op == 2) ? pred_pos + 1 : pred_pos;
    int32_t match_count = qi - edit_start;
    for (int32_t i = 0; i < match_count; ++i) {
      ops.push_back('M');
      --qi;
      --ti;
    }
    if (op == 1) {
      ops.push_back('X');
      --qi;
      --ti;
    } else if (op == 2) {
      ops.push_back('D');
      --qi;
    } else if (op == 3) {
      ops.push_back('I');
      --ti;
    }
    k = pred_k;
    --s;
  }
  std::reverse(ops.begin(), ops.end());
  return compact_cigar(ops);
}

// fill_wfa (wfa_simd.rs:231-345): global unit-cost WFA.
std::pair<std::string, int32_t> fill_wfa(const Seqs& S) {
  const int32_t qn = S.qn, tn = S.tn;
  const size_t size = static_cast<size_t>(qn + tn + 1);
  std::vector<LinFrame> hist;
  LinFrame f0{std::vector<int32_t>(size, UNSET), std::vector<uint8_t>(size, 0)};
  f0.wf[static_cast<size_t>(tn)] = S.extend(0, 0);
  const size_t target_idx = static_cast<size_t>(qn);  // diagonal qn - tn, index (qn - tn) + tn
  if (f0.wf[target_idx] == qn) return {linear_traceback(hist, S, 0, tn), 0};
  hist.push_back(std::move(f0));
  const int32_t max_s = qn + tn;
  for (int32_t s = 1; s <= max_s; ++s) {
    LinFrame nf{std::vector<int32_t>(size, UNSET), std::vector<uint8_t>(size, 0)};
    linear_step(S, s, hist[static_cast<size_t>(s) - 1].wf, nf.wf, nf.op);
    bool done = nf.wf[target_idx] >= qn;
    hist.push_back(std::move(nf));
    if (done) return {linear_traceback(hist, S, s, tn), s};
  }
  return {std::string(), 0};  // upstream fallback (wfa_simd.rs:343-344)
}

// fitting_end_k (wfa_simd.rs:940-955): smallest k (largest t_end) with wf >= qn.
bool fitting_end_k(const std::vector<int32_t>& wf, int32_t qn, int32_t tn, int32_t* t_end) {
  bool found = false;
  int32_t best = 0;
  for (int32_t k = qn - tn; k <= qn; ++k) {
    int32_t te = qn - k;
    if (te < 0 || te > tn) continue;
    size_t ki = static_cast<size_t>(k + tn);
    if (ki >= wf.size()) continue;
    if (wf[ki] >= qn && (!found || te > best)) {
      best = te;
      found = true;
    }
  }
  *t_end = best;
  return found;
}

struct EngineOut {
  bool some;
  std::string cigar;
  int64_t score;          // search score
  int64_t edit_distance;  // reported distance
  int64_t t_consumed;
};

// fill_wfa_fitting_impl (wfa_simd.rs:360-482).  cap < 0 = None.
EngineOut fill_wfa_fitting_impl(const Seqs& S, int64_t cap) {
  const int32_t qn = S.qn, tn = S.tn;
  if (qn == 0) return {true, std::string(), 0, 0, 0};
  if (tn <= qn + qn / 2 + 10) {
    auto r = fill_wfa(S);
    if (cap >= 0 && static_cast<int64_t>(r.second) > cap) return {false, {}, 0, 0, 0};
    return {true, r.first, r.second, r.second, tn};
  }
  const size_t size = static_cast<size_t>(qn + tn + 1);
  std::vector<LinFrame> hist;
  LinFrame f0{std::vector<int32_t>(size, UNSET), std::vector<uint8_t>(size, 0)};
  f0.wf[static_cast<size_t>(tn)] = S.extend(0, 0);
  const int32_t max_s = qn + tn;
  // `(cap as i32).min(max_s)`: caps in this API are non-negative i32 values.
  const int32_t explore = cap >= 0 ? static_cast<int32_t>(std::min<int64_t>(cap, max_s)) : max_s;
  int32_t t_end;
  if (fitting_end_k(f0.wf, qn, tn, &t_end)) {
    hist.push_back(std::move(f0));
    return {true, linear_traceback(hist, S, 0, t_end), 0, 0, t_end};
  }
  hist.push_back(std::move(f0));
  for (int32_t s = 1; s <= explore; ++s) {
    LinFrame nf{std::vector<int32_t>(size, UNSET), std::vector<uint8_t>(size, 0)};
    linear_step(S, s, hist[static_cast<size_t>(s) - 1].wf, nf.wf, nf.op);
    bool done = fitting_end_k(nf.wf, qn, tn, &t_end);
    hist.push_back(std::move(nf));
    if (done) return {true, linear_traceback(hist, S, s, t_end), s, s, t_end};
  }
  return {false, {}, 0, 0, 0};
}

// ---------------------------------------------------------------------------
// Gap-affine WFA (wfa_simd.rs:542-935)
// ---------------------------------------------------------------------------

struct AffFrame {
  std::vector<int32_t> m, ins, del;
  std::vector<uint8_t> m_op, ins_op, del_op;
  explicit AffFrame(size_t n)
      : m(n, UNSET), ins(n, UNSET), del(n, UNSET), m_op(n, 0), ins_op(n, 0), del_op(n, 0) {}
};

struct Costs {
  int32_t x, o, e;
};

// affine_pos_valid (wfa_simd.rs:563-569).
inline bool pos_valid(int32_t pos, int32_t k, int32_t qn, int32_t tn) {
  if (pos == UNSET || pos < 0 || pos > qn) return false;
  int32_t tp = pos - k;
  return tp >= 0 && tp <= tn;
}

// affine_traceback (wfa_simd.rs:785-886).
std::string affine_traceback(const std::vector<AffFrame>& hist, const Seqs& S, int32_t s_end,
                             int32_t t_end, int32_t k_end, uint8_t start_layer, Costs c) {
  const int32_t tn = S.tn;
  std::vector<char> ops;
  int32_t qi = S.qn, ti = t_end, k = k_end, s = s_end;
  uint8_t layer = start_layer;
  auto emit_m = [&](int32_t n) {
    for (int32_t i = 0; i < n; ++i) {
      ops.push_back('M');
      --qi;
      --ti;
    }
  };
  while (qi > 0 || ti > 0) {
    if (s == 0) {
      while (qi > 0 && ti > 0) {
        ops.push_back('M');
        --qi;
        --ti;
      }
      break;
    }
    const size_t ki = static_cast<size_t>(k + tn);
    const AffFrame& f = hist[static_cast<size_t>(s)];
    if (layer == 0) {
      uint8_t op = f.m_op[ki];
      if (op == 0) {
        int32_t s_mm = s - c.x;
        int32_t raw = hist[static_cast<size_t>(s_mm)].m[ki] + 1;
        emit_m(std::max(qi - raw, 0));
        ops.push_back('X');
        --qi;
        --ti;
        s = s_mm;
      } else if (op == 1) {
        emit_m(std::max(qi - f.ins[ki], 0));
        layer = 1;
      } else {
        emit_m(std::max(qi - f.del[ki], 0));
        layer = 2;
      }
    } else if (layer == 1) {
      ops.push_back('D');  // query-extra
      --qi;
      if (f.ins_op[ki] == 0) {
        s -= c.o + c.e;
        layer = 0;
      } else {
        s -= c.e;
      }
      --k;
    } else {
      ops.push_back('I');  // target-extra
      --ti;
      if (f.del_op[ki] == 0) {
        s -= c.o + c.e;
        layer = 0;
      } else {
        s -= c.e;
      }
      ++k;
    }
  }
  std::reverse(ops.begin(), ops.end());
  return compact_cigar(ops);
}

// fill_wfa_affine_generic (wfa_simd.rs:581-779).  cap < 0 = None.
EngineOut fill_wfa_affine_generic(const Seqs& S, Costs c, bool fitting, int64_t cap) {
  const int32_t qn = S.qn, tn = S.tn;
  if (qn == 0 && tn == 0) return {true, std::string(), 0, 0, 0};
  if (qn == 0) {
    if (fitting) return {true, std::string(), 0, 0, 0};
    return {true, std::to_string(tn) + "I", c.o + static_cast<int64_t>(tn) * c.e, 0, tn};
  }
  const size_t size = static_cast<size_t>(qn + tn + 1);
  const int32_t diag_lo = -tn, diag_hi = qn;

  // Endpoint scan (wfa_simd.rs:618-650): layer order m, ins, del; fitting takes
  // the largest t_end, i.e. the smallest diagonal.
  auto endpoint = [&](const AffFrame& f, int32_t* t_end, int32_t* k_out, uint8_t* layer) -> bool {
    auto check = [&](size_t ki, uint8_t* l) -> bool {
      if (f.m[ki] == qn) { *l = 0; return true; }
      if (f.ins[ki] == qn) { *l = 1; return true; }
      if (f.del[ki] == qn) { *l = 2; return true; }
      return false;
    };
    if (!fitting) {
      int32_t k = qn - tn;
      uint8_t l;
      if (check(static_cast<size_t>(k + tn), &l)) {
        *t_end = tn;
        *k_out = k;
        *layer = l;
        return true;
      }
      return false;
    }
    bool found = false;
    for (int32_t k = qn - tn; k <= qn; ++k) {
      int32_t te = qn - k;
      if (te < 0 || te > tn) continue;
      uint8_t l;
      if (check(static_cast<size_t>(k + tn), &l) && (!found || te > *t_end)) {
        *t_end = te;
        *k_out = k;
        *layer = l;
        found = true;
      }
    }
    return found;
  };

  std::vector<AffFrame> hist;
  hist.emplace_back(size);
  hist[0].m[static_cast<size_t>(tn)] = S.extend(0, 0);
  int32_t t_end = 0, k_end = 0;
  uint8_t layer = 0;
  if (endpoint(hist[0], &t_end, &k_end, &layer)) {
    return {true, affine_traceback(hist, S, 0, t_end, k_end, layer, c), 0, 0, t_end};
  }
  const int64_t hard_max_s =
      static_cast<int64_t>(qn + tn) * std::max(c.x, 1) + c.o + c.e + 4;
  const int64_t explore_to = cap >= 0 ? std::min<int64_t>(cap, hard_max_s) : hard_max_s;

  for (int64_t s64 = 1; s64 <= explore_to; ++s64) {
    const int32_t s = static_cast<int32_t>(s64);
    AffFrame f(size);
    const int32_t s_open = s - c.o - c.e;
    const int32_t s_ext = s - c.e;
    for (int32_t k = diag_lo; k <= diag_hi; ++k) {
      const size_t ki = static_cast<size_t>(k + tn);
      // ins (query-extra): open from m[s_open][k-1]+1, extend from ins[s_ext][k-1]+1.
      {
        int32_t best = UNSET;
        uint8_t bop = 0;
        if (k - 1 >= diag_lo) {
          const size_t pk = ki - 1;
          if (s_open >= 0) {
            int32_t p = hist[static_cast<size_t>(s_open)].m[pk];
            if (p != UNSET) {
              int32_t cand = p + 1;
              if (pos_valid(cand, k, qn, tn) && cand > best) { best = cand; bop = 0; }
            }
          }
          if (s_ext >= 0) {
            int32_t p = hist[static_cast<size_t>(s_ext)].ins[pk];
            if (p != UNSET) {
              int32_t cand = p + 1;
              if (pos_valid(cand, k, qn, tn) && cand > best) { best = cand; bop = 1; }
            }
          }
        }
        if (best != UNSET) { f.ins[ki] = best; f.ins_op[ki] = bop; }
      }
      // del (target-extra): open from m[s_open][k+1], extend from del[s_ext][k+1].
      {
        int32_t best = UNSET;
        uint8_t bop = 0;
        if (k + 1 <= diag_hi) {
          const size_t pk = ki + 1;
          if (s_open >= 0) {
            int32_t p = hist[static_cast<size_t>(s_open)].m[pk];
            if (p != UNSET && pos_valid(p, k, qn, tn) && p > best) { best = p; bop = 0; }
          }
          if (s_ext >= 0) {
            int32_t p = hist[static_cast<size_t>(s_ext)].del[pk];
            if (p != UNSET && pos_valid(p, k, qn, tn) && p > best) { best = p; bop = 1; }
          }
        }
        if (best != UNSET) { f.del[ki] = best; f.del_op[ki] = bop; }
      }
    }
    const int32_t s_mm = s - c.x;
    for (int32_t k = diag_lo; k <= diag_hi; ++k) {
      const size_t ki = static_cast<size_t>(k + tn);
      int32_t best = UNSET;
      uint8_t bop = 0;
      if (s_mm >= 0) {
        int32_t p = hist[static_cast<size_t>(s_mm)].m[ki];
        if (p != UNSET) {
          int32_t cand = p + 1;
          if (pos_valid(cand, k, qn, tn) && cand > best) { best = cand; bop = 0; }
        }
      }
      if (f.ins[ki] != UNSET && pos_valid(f.ins[ki], k, qn, tn) && f.ins[ki] > best) { best = f.ins[ki]; bop = 1; }
      if (f.del[ki] != UNSET && pos_valid(f.del[ki], k, qn, tn) && f.del[ki] > best) { best = f.del[ki]; bop = 2; }
      if (best != UNSET) {
        f.m[ki] = S.extend(best, k);
        f.m_op[ki] = bop;
      }
    }
    hist.push_back(std::move(f));
    if (endpoint(hist[static_cast<size_t>(s)], &t_end, &k_end, &layer)) {
      return {true, affine_traceback(hist, S, s, t_end, k_end, layer, c), s, 0, t_end};
    }
  }
  return {false, {}, 0, 0, 0};
}

// fill_wfa_affine_fitting_impl (wfa_simd.rs:919-935).
EngineOut fill_wfa_affine_fitting_impl(const Seqs& S, Costs c, int64_t cap) {
  if (S.qn == 0) return {true, std::string(), 0, 0, 0};
  const bool fitting = S.tn > S.qn + S.qn / 2 + 10;
  EngineOut r = fill_wfa_affine_generic(S, c, fitting, cap);
  if (r.some) r.edit_distance = static_cast<int64_t>(levenshtein_from_cigar(r.cigar));
  return r;
}

// ---------------------------------------------------------------------------
// Myers / Hyyrö bit-parallel edit distance (wfa_simd.rs:3507-3814)
// ---------------------------------------------------------------------------

struct MyersColumns {
  std::vector<uint64_t> data;  // per column: vp[B] then vn[B]
  std::vector<size_t> scores;
  size_t nb = 0;
  size_t ncols = 0;
};

// myers_forward_flat_with_cutoff (wfa_simd.rs:3552-3648).  margin < 0 = None.
MyersColumns myers_forward(const uint8_t* q, size_t m, const uint8_t* t, size_t n, int64_t margin,
                           bool keep_columns) {
  constexpr size_t W = 64;
  MyersColumns out;
  const size_t nb = (m + W - 1) / W;
  const size_t last = nb - 1;
  const size_t last_bits = (m % W == 0) ? W : m % W;
  const uint64_t last_mask = last_bits == W ? ~0ull : ((1ull << last_bits) - 1);
  const size_t score_bit = (m - 1) % W;
  std::vector<uint64_t> pm(nb * 256, 0);
  for (size_t i = 0; i < m; ++i) pm[(i / W) * 256 + q[i]] |= 1ull << (i % W);
  std::vector<uint64_t> vp(nb, ~0ull), vn(nb, 0), nvp(nb, 0), nvn(nb, 0);
  vp[last] = last_mask;
  size_t score = m, best_seen = m;
  const size_t stride = nb * 2;
  out.nb = nb;
  if (keep_columns) out.data.assign(n * stride, 0);
  out.scores.reserve(n);
  for (size_t col = 0; col < n; ++col) {
    const uint8_t tb = t[col];
    uint64_t add_carry = 0, hp_carry = 1, hn_carry = 0, last_hp = 0, last_hn = 0;
    for (size_t b = 0; b < nb; ++b) {
      const uint64_t eq = pm[b * 256 + tb];
      const uint64_t xh = eq | vn[b];
      const uint64_t xhvp = xh & vp[b];
      uint64_t s1 = vp[b] + xhvp;
      uint64_t c1 = s1 < vp[b];
      uint64_t s2 = s1 + add_carry;
      uint64_t c2 = s2 < s1;
      add_carry = c1 | c2;
      uint64_t d0 = (s2 ^ vp[b]) | xh;
      uint64_t hn = vp[b] & d0;
      uint64_t hp = vn[b] | ~(vp[b] | d0);
      if (b == last && last_bits < W) {
        d0 &= last_mask;
        hn &= last_mask;
        hp &= last_mask;
      }
      if (b == last) {
        last_hp = hp;
        last_hn = hn;
      }
      const uint64_t nhp = hp >> (W - 1);
      const uint64_t x = (hp << 1) | hp_carry;
      hp_carry = nhp;
      const uint64_t nhn = hn >> (W - 1);
      const uint64_t hns = (hn << 1) | hn_carry;
      hn_carry = nhn;
      nvn[b] = x & d0;
      uint64_t v2 = hns | ~(d0 | x);
      nvp[b] = (b == last && last_bits < W) ? (v2 & last_mask) : v2;
    }
    if ((last_hp >> score_bit) & 1) score += 1;
    if ((last_hn >> score_bit) & 1) score = score > 0 ? score - 1 : 0;
    vp.swap(nvp);
    vn.swap(nvn);
    if (keep_columns) {
      std::copy(vp.begin(), vp.end(), out.data.begin() + static_cast<ptrdiff_t>(col * stride));
      std::copy(vn.begin(), vn.end(), out.data.begin() + static_cast<ptrdiff_t>(col * stride + nb));
    }
    out.scores.push_back(score);
    if (score < best_seen) best_seen = score;
    if (margin >= 0 && score > best_seen + static_cast<size_t>(margin)) {
      out.ncols = col + 1;
      if (keep_columns) out.data.resize(out.ncols * stride);
      return out;
    }
  }
  out.ncols = n;
  return out;
}

// fill_dp_column_flat + reconstruct_into (wfa_simd.rs:3507-3516, :3687-3699).
void dp_column(const MyersColumns& cols, size_t m, size_t j, std::vector<size_t>& buf) {
  if (j == 0) {
    for (size_t i = 0; i <= m; ++i) buf[i] = i;
    return;
  }
  const uint64_t* vp = &cols.data[(j - 1) * cols.nb * 2];
  const uint64_t* vn = vp + cols.nb;
  buf[m] = cols.scores[j - 1];
  for (size_t i = m; i-- > 0;) {
    const int64_t pb = static_cast<int64_t>((vp[i / 64] >> (i % 64)) & 1);
    const int64_t nb = static_cast<int64_t>((vn[i / 64] >> (i % 64)) & 1);
    buf[i] = static_cast<size_t>(static_cast<int64_t>(buf[i + 1]) - pb + nb);
  }
}

// myers_backtrace_flat (wfa_simd.rs:3702-3781): priority D, then I, then diagonal.
std::string myers_backtrace(const uint8_t* q, size_t m, const uint8_t* t, const MyersColumns& cols,
                            size_t end_ti) {
  std::vector<char> ops;
  ops.reserve(m + end_ti);
  size_t qi = m, ti = end_ti;
  std::vector<size_t> cur(m + 1, 0), prev(m + 1, 0);
  if (ti > 0) {
    dp_column(cols, m, ti, cur);
    dp_column(cols, m, ti - 1, prev);
  }
  while (qi > 0 || ti > 0) {
    if (qi == 0) {
      ops.push_back('I');
      --ti;
      continue;
    }
    if (ti == 0) {
      ops.push_back('D');
      --qi;
      continue;
    }
    const size_t c = cur[qi];
    const size_t diag_cost = q[qi - 1] == t[ti - 1] ? 0 : 1;
    if (c == cur[qi - 1] + 1) {
      ops.push_back('D');
      --qi;
    } else if (c == prev[qi] + 1) {
      ops.push_back('I');
      --ti;
      cur.swap(prev);
      if (ti > 0) dp_column(cols, m, ti - 1, prev);
    } else {
      ops.push_back(diag_cost == 0 ? 'M' : 'X');
      --qi;
      --ti;
      cur.swap(prev);
      if (ti > 0) dp_column(cols, m, ti - 1, prev);
    }
  }
  std::reverse(ops.begin(), ops.end());
  return compact_cigar(ops);
}

// myers_edit_distance_impl (wfa_simd.rs:3657-3673).
EngineOut myers_global(const uint8_t* q, size_t m, const uint8_t* t, size_t n, bool with_cigar) {
  if (m == 0) return {true, n > 0 ? std::to_string(n) + "I" : std::string(), (int64_t)n, (int64_t)n, (int64_t)n};
  if (n == 0) return {true, std::to_string(m) + "D", (int64_t)m, (int64_t)m, 0};
  MyersColumns cols = myers_forward(q, m, t, n, -1, with_cigar);
  const size_t ed = cols.ncols > 0 ? cols.scores[cols.ncols - 1] : m;
  std::string cigar = with_cigar ? myers_backtrace(q, m, t, cols, n) : std::string();
  return {true, cigar, (int64_t)ed, (int64_t)ed, (int64_t)n};
}

// myers_fitting_impl (wfa_simd.rs:3783-3814).
EngineOut myers_fitting(const uint8_t* q, size_t m, const uint8_t* t, size_t n, bool with_cigar) {
  if (m == 0) return {true, std::string(), 0, 0, 0};
  if (n == 0) return {true, std::to_string(m) + "D", (int64_t)m, (int64_t)m, 0};
  if (n <= m + m / 2 + 10) return myers_global(q, m, t, n, with_cigar);
  const int64_t margin = static_cast<int64_t>(m / 10 + 1);
  MyersColumns cols = myers_forward(q, m, t, n, margin, with_cigar);
  size_t best_consumed = 0, best_score = m;
  for (size_t c = 0; c < cols.ncols; ++c) {
    if (cols.scores[c] < best_score) {
      best_score = cols.scores[c];
      best_consumed = c + 1;
    }
  }
  std::string cigar = with_cigar ? myers_backtrace(q, m, t, cols, best_consumed) : std::string();
  return {true, cigar, (int64_t)best_score, (int64_t)best_score, (int64_t)best_consumed};
}

// ---------------------------------------------------------------------------
// Wrappers (phraya-align/src/lib.rs:63-167, wfa_simd.rs:173-223)
// ---------------------------------------------------------------------------

int finish(const EngineOut& e, size_t seed_q, size_t seed_t, size_t q_suffix_len, orc_result* r,
           char* cigar, size_t cap, std::string* keep) {
  if (!e.some) {
    r->status = ORC_ALIGNMENT_FAILED;
    r->cigar_len = 0;
    if (cigar && cap) cigar[0] = 0;
    return r->status;
  }
  r->status = ORC_OK;
  r->score = e.score;
  r->edit_distance = e.edit_distance;
  r->query_start = static_cast<int64_t>(seed_q);
  r->query_end = static_cast<int64_t>(seed_q + q_suffix_len);
  r->target_start = static_cast<int64_t>(seed_t);
  r->target_end = static_cast<int64_t>(seed_t) + e.t_consumed;
  r->cigar_len = static_cast<int64_t>(e.cigar.size());
  if (cigar && cap) {
    size_t nc = std::min(cap - 1, e.cigar.size());
    std::memcpy(cigar, e.cigar.data(), nc);
    cigar[nc] = 0;
  }
  if (keep) *keep = e.cigar;
  return ORC_OK;
}

int invalid(orc_result* r, char* cigar, size_t cap) {
  std::memset(r, 0, sizeof(*r));
  r->status = ORC_INVALID_INPUT;
  if (cigar && cap) cigar[0] = 0;
  return ORC_INVALID_INPUT;
}

int run_engine(int engine, const uint8_t* q, size_t qn, const uint8_t* t, size_t tn, size_t sq,
               size_t st, Costs c, int64_t cap, orc_result* r, char* cigar, size_t ccap,
               std::string* keep) {
  std::memset(r, 0, sizeof(*r));
  if (engine == ORC_ENGINE_MYERS_GLOBAL) {
    EngineOut e = myers_global(q, qn, t, tn, true);
    return finish(e, 0, 0, qn, r, cigar, ccap, keep);
  }
  if (sq > qn || st > tn) return invalid(r, cigar, ccap);
  const uint8_t* qs = q + sq;
  const uint8_t* ts = t + st;
  const size_t qsn = qn - sq, tsn = tn - st;
  if (qsn > static_cast<size_t>(INT32_MAX / 4) || tsn > static_cast<size_t>(INT32_MAX / 4))
    return invalid(r, cigar, ccap);
  Seqs S{qs, static_cast<int32_t>(qsn), ts, static_cast<int32_t>(tsn)};
  switch (engine) {
    case ORC_ENGINE_LINEAR: {
      if (qsn == 0 && tsn == 0) return finish({true, {}, 0, 0, 0}, sq, st, 0, r, cigar, ccap, keep);
      return finish(fill_wfa_fitting_impl(S, cap), sq, st, qsn, r, cigar, ccap, keep);
    }
    case ORC_ENGINE_AFFINE: {
      if (c.x < 1 || c.e < 1 || c.o < 0) {
        r->status = ORC_BAD_COSTS;
        return ORC_BAD_COSTS;
      }
      return finish(fill_wfa_affine_fitting_impl(S, c, cap), sq, st, qsn, r, cigar, ccap, keep);
    }
    case ORC_ENGINE_LINEAR_GLOBAL: {
      // fill_wfa (wfa_simd.rs:231-345) whatever the length ratio; no cap.
      auto g = fill_wfa(S);
      return finish({true, g.first, g.second, g.second, static_cast<int64_t>(tsn)}, sq, st, qsn, r,
                    cigar, ccap, keep);
    }
    case ORC_ENGINE_AFFINE_GLOBAL: {
      if (c.x < 1 || c.e < 1 || c.o < 0) {
        r->status = ORC_BAD_COSTS;
        return ORC_BAD_COSTS;
      }
      // fill_wfa_affine_generic(fitting = false) (wfa_simd.rs:581-779), the
      // reported distance from the CIGAR as the wrapper does (:932-934).
      EngineOut e = fill_wfa_affine_generic(S, c, false, cap);
      if (e.some) e.edit_distance = static_cast<int64_t>(levenshtein_from_cigar(e.cigar));
      return finish(e, sq, st, qsn, r, cigar, ccap, keep);
    }
    case ORC_ENGINE_MYERS:
      return finish(myers_fitting(qs, qsn, ts, tsn, true), sq, st, qsn, r, cigar, ccap, keep);
    case ORC_ENGINE_MYERS_FWD:
      return finish(myers_fitting(qs, qsn, ts, tsn, false), sq, st, qsn, r, cigar, ccap, keep);
    default:
      return invalid(r, cigar, ccap);
  }
}

}  // namespace

extern "C" {

size_t orc_lcp(const uint8_t* a, size_t an, const uint8_t* b, size_t bn) { return lcp_sse2(a, an, b, bn); }
size_t orc_lcp_scalar(const uint8_t* a, size_t an, const uint8_t* b, size_t bn) { return lcp_scalar(a, an, b, bn); }
size_t orc_lcp_u64(const uint8_t* a, size_t an, const uint8_t* b, size_t bn) { return lcp_u64(a, an, b, bn); }

int orc_wfa_extend(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn, size_t seed_q,
                   size_t seed_t, int64_t cap, orc_result* r, char* cigar, size_t cigar_cap) {
  return run_engine(ORC_ENGINE_LINEAR, q, qn, t, tn, seed_q, seed_t, {0, 0, 0}, cap, r, cigar,
                    cigar_cap, nullptr);
}

int orc_wfa_extend_affine(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn, size_t seed_q,
                          size_t seed_t, int32_t mismatch, int32_t gap_open, int32_t gap_extend,
                          int64_t cap, orc_result* r, char* cigar, size_t cigar_cap) {
  return run_engine(ORC_ENGINE_AFFINE, q, qn, t, tn, seed_q, seed_t,
                    {mismatch, gap_open, gap_extend}, cap, r, cigar, cigar_cap, nullptr);
}

int orc_wfa_global(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn, orc_result* r,
                   char* cigar, size_t cigar_cap) {
  return run_engine(ORC_ENGINE_LINEAR_GLOBAL, q, qn, t, tn, 0, 0, {0, 0, 0}, -1, r, cigar, cigar_cap,
                    nullptr);
}

int orc_wfa_affine_global(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                          int32_t mismatch, int32_t gap_open, int32_t gap_extend, orc_result* r,
                          char* cigar, size_t cigar_cap) {
  return run_engine(ORC_ENGINE_AFFINE_GLOBAL, q, qn, t, tn, 0, 0, {mismatch, gap_open, gap_extend},
                    -1, r, cigar, cigar_cap, nullptr);
}

int orc_myers_extend(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn, size_t seed_q,
                     size_t seed_t, orc_result* r, char* cigar, size_t cigar_cap) {
  return run_engine(ORC_ENGINE_MYERS, q, qn, t, tn, seed_q, seed_t, {0, 0, 0}, -1, r, cigar,
                    cigar_cap, nullptr);
}

int orc_myers_edit_distance(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                            orc_result* r, char* cigar, size_t cigar_cap) {
  return run_engine(ORC_ENGINE_MYERS_GLOBAL, q, qn, t, tn, 0, 0, {0, 0, 0}, -1, r, cigar,
                    cigar_cap, nullptr);
}

int orc_myers_extend_distance(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn,
                              size_t seed_q, size_t seed_t, orc_result* r) {
  return run_engine(ORC_ENGINE_MYERS_FWD, q, qn, t, tn, seed_q, seed_t, {0, 0, 0}, -1, r,
                    nullptr, 0, nullptr);
}

struct orc_batch {
  std::vector<std::string> cigars;
};

orc_batch* orc_batch_run(int engine, uint32_t n, const uint8_t* q_seq, const uint64_t* q_off,
                         const uint8_t* t_seq, const uint64_t* t_off, const uint64_t* seed_q,
                         const uint64_t* seed_t, int32_t mismatch, int32_t gap_open,
                         int32_t gap_extend, int64_t cap, int threads, orc_result* res) {
  orc_batch* b = new orc_batch;
  b->cigars.resize(n);
  std::atomic<uint32_t> next{0};
  auto worker = [&]() {
    for (;;) {
      uint32_t i = next.fetch_add(1, std::memory_order_relaxed);
      if (i >= n) break;
      const uint8_t* q = q_seq + q_off[i];
      const uint8_t* t = t_seq + t_off[i];
      size_t qn = static_cast<size_t>(q_off[i + 1] - q_off[i]);
      size_t tn = static_cast<size_t>(t_off[i + 1] - t_off[i]);
      run_engine(engine, q, qn, t, tn, seed_q ? seed_q[i] : 0, seed_t ? seed_t[i] : 0,
                 {mismatch, gap_open, gap_extend}, cap, &res[i], nullptr, 0, &b->cigars[i]);
    }
  };
  if (threads <= 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    pool.reserve(static_cast<size_t>(threads));
    for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
  }
  return b;
}

uint64_t orc_batch_cigar_bytes(const orc_batch* b) {
  uint64_t total = 0;
  for (const auto& s : b->cigars) total += s.size();
  return total;
}

void orc_batch_cigar_copy(const orc_batch* b, char* arena, uint64_t* off) {
  uint64_t pos = 0;
  for (size_t i = 0; i < b->cigars.size(); ++i) {
    off[i] = pos;
    std::memcpy(arena + pos, b->cigars[i].data(), b->cigars[i].size());
    pos += b->cigars[i].size();
  }
  off[b->cigars.size()] = pos;
}

void orc_batch_free(orc_batch* b) { delete b; }

int orc_hardware_threads(void) {
  unsigned h = std::thread::hardware_concurrency();
  return h ? static_cast<int>(h) : 1;
}

}  // extern "C"

// --- post-extension: variant sites from CIGAR runs ---------------------------
//
// extract_variants_from_cigar (phraya-align/src/executor.rs:814-1044) for a
// batch of alignments against one target, restated step by step: M advances
// both positions; each X base with qp < |query| and tp < |target| is an SNV
// (ref target[tp], alt query[qp]); an I run (target-only bases, upstream's
// WFA convention :912) with t_pos < |target| and q_pos < |query| is a
// deletion at t_pos (ref = first deleted base, alt '.'); a D run (query-only
// bases) with q_pos < |query| and t_pos < |target| is an insertion at t_pos
// (ref '.', alleles = the inserted bases query[q_pos .. min(q_pos + count,
// |query|)]).  Each variant carries its local coverage window
// [p - r (floored at 0), min(p + r + 1, |target|)) with, per position, the
// number of the alignment's coverage windows (compute_windowed_coverage,
// :1071-1090: [min(target_start, |target|), min(target_end, |target|)),
// counts 1, empty ones dropped) that cover it (get_abs_multi, :218-220), and
// the repeat / hotspot flags by binary search over sorted, non-overlapping
// [start, end) intervals (:795-811).  Runs are (length << 4) | op with op
// M = 0, X = 1, I = 2, D = 3 (compacted, as parse_cigar(compact_cigar(..))).
struct orc_variant {
  uint32_t alignment, position;
  uint8_t type, ref_base, alt_base, flags;
  uint32_t ins_query_pos, ins_len;
  uint32_t window_start, window_len, window_offset;
  uint64_t coverage_off;
};

namespace {

template <class T>
bool in_intervals(uint64_t pos, const T* iv, uint64_t n) {
  // partition_point(start <= pos), then pos < end of the interval before it
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if ((uint64_t)iv[2 * mid] <= pos) lo = mid + 1;
    else hi = mid;
  }
  return lo > 0 && pos < (uint64_t)iv[2 * (lo - 1) + 1];
}

struct VarIn {
  const uint32_t* runs;
  const uint64_t *run_off, *target_start;
  const uint8_t* q_seq;
  const uint64_t* q_off;
  const uint8_t* target;
  uint64_t target_len;
  const uint64_t *spans, *span_off, *repeats;
  uint64_t n_repeats;
  const uint32_t* hotspots;
  uint64_t n_hotspots;
  uint32_t radius;
};

// Walks alignment i; with out == nullptr only counts variants and coverage
// words.
void variants_of(const VarIn& in, uint32_t i, orc_variant* out, uint32_t* cov, uint64_t* nv,
                 uint64_t* nc) {
  const uint8_t* q = in.q_seq + in.q_off[i];
  const uint64_t qn = in.q_off[i + 1] - in.q_off[i], tn = in.target_len;
  uint64_t q_pos = 0, t_pos = in.target_start[i], v = 0, c = 0;
  auto emit = [&](uint64_t p, uint8_t type, uint8_t ref, uint8_t alt, uint64_t iq, uint64_t il) {
    const uint64_t ws = p >= in.radius ? p - in.radius : 0;
    const uint64_t we = std::min<uint64_t>(p + in.radius + 1, tn);
    if (out) {
      orc_variant& o = out[v];
      o.alignment = i;
      o.position = (uint32_t)p;
      o.type = type;
      o.ref_base = ref;
      o.alt_base = alt;
      o.flags = (uint8_t)((in_intervals(p, in.repeats, in.n_repeats) ? 1 : 0) |
                          (in_intervals(p, in.hotspots, in.n_hotspots) ? 2 : 0));
      o.ins_query_pos = (uint32_t)iq;
      o.ins_len = (uint32_t)il;
      o.window_start = (uint32_t)ws;
      o.window_len = (uint32_t)(we - ws);
      o.window_offset = (uint32_t)(p - ws);
      o.coverage_off = c;
      for (uint64_t pos = ws; pos < we; ++pos) {
        uint32_t n = 0;
        for (uint64_t s = in.span_off[i]; s < in.span_off[i + 1]; ++s) {
          const uint64_t a = std::min(in.spans[2 * s], tn), b = std::min(in.spans[2 * s + 1], tn);
          if (a < b && pos >= a && pos < b) ++n;
        }
        cov[c + (pos - ws)] = n;
      }
    }
    ++v;
    c += we - ws;
  };
  for (uint64_t r = in.run_off[i]; r < in.run_off[i + 1]; ++r) {
    const uint64_t count = in.runs[r] >> 4;
    switch (in.runs[r] & 15u) {
      case 0:  // M
        q_pos += count;
        t_pos += count;
        break;
      case 1:  // X
        for (uint64_t k = 0; k < count; ++k) {
          const uint64_t qp = q_pos + k, tp = t_pos + k;
          if (qp < qn && tp < tn) emit(tp, 0, in.target[tp], q[qp], 0, 0);
        }
        q_pos += count;
        t_pos += count;
        break;
      case 2:  // I: target-only bases -> deletion
        if (t_pos < tn && q_pos < qn) emit(t_pos, 1, in.target[t_pos], '.', 0, 0);
        t_pos += count;
        break;
      case 3:  // D: query-only bases -> insertion
        if (q_pos < qn && t_pos < tn)
          emit(t_pos, 2, '.', 0, q_pos, std::min(q_pos + count, qn) - q_pos);
        q_pos += count;
        break;
      default:
        break;
    }
  }
  *nv = v;
  *nc = c;
}

}  // namespace

extern "C" {

// Two passes over the alignments (counts, then the records at exclusive
// offsets), each split over `threads` threads.  var_off gets n + 1 entries;
// *n_var / *n_cov the totals.  Returns 0, or 1 when a capacity is too small
// (totals still set).
int orc_extract_variants(uint32_t n, const uint32_t* runs, const uint64_t* run_off,
                         const uint64_t* target_start, const uint8_t* q_seq, const uint64_t* q_off,
                         const uint8_t* target, uint64_t target_len, const uint64_t* spans,
                         const uint64_t* span_off, const uint64_t* repeats, uint64_t n_repeats,
                         const uint32_t* hotspots, uint64_t n_hotspots, uint32_t radius, int threads,
                         orc_variant* out, uint64_t out_cap, uint32_t* cov, uint64_t cov_cap,
                         uint64_t* var_off, uint64_t* n_var, uint64_t* n_cov) {
  const VarIn in{runs, run_off, target_start, q_seq, q_off, target, target_len, spans, span_off,
                 repeats, n_repeats, hotspots, n_hotspots, radius};
  std::vector<uint64_t> nv(n), nc(n), co(n + 1, 0);
  const int T = std::max(1, threads);
  auto par = [&](auto&& fn) {
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
      pool.emplace_back([&, t] {
        for (uint32_t i = (uint32_t)t; i < n; i += (uint32_t)T) fn(i);
      });
    for (auto& th : pool) th.join();
  };
  par([&](uint32_t i) { variants_of(in, i, nullptr, nullptr, &nv[i], &nc[i]); });
  var_off[0] = 0;
  for (uint32_t i = 0; i < n; ++i) {
    var_off[i + 1] = var_off[i] + nv[i];
    co[i + 1] = co[i] + nc[i];
  }
  *n_var = var_off[n];
  *n_cov = co[n];
  if (var_off[n] > out_cap || co[n] > cov_cap) return 1;
  par([&](uint32_t i) {
    uint64_t a, b;
    variants_of(in, i, out + var_off[i], cov + co[i], &a, &b);
    for (uint64_t k = var_off[i]; k < var_off[i + 1]; ++k) out[k].coverage_off += co[i];
  });
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Pre-extension step: seeding and chaining (phraya-align/src/seeding.rs,
// phraya-align/src/chaining.rs), restated with upstream's data structures and
// orders: a hash map value -> target positions in sketch order
// (build_minimizer_index, seeding.rs:29-35), seeds in query-sketch order then
// stable-sorted by query position (find_seeds_indexed_capped, :50-74), and
// minibwa-style chaining DP (chain_seeds, chaining.rs:131-253) with its
// stable sorts, backward scan, skip bookkeeping and backtrack.
// ---------------------------------------------------------------------------
extern "C" {
typedef struct orc_seed {
  uint32_t query_pos;
  uint32_t target_pos;
  uint64_t minimizer;
} orc_seed;

typedef struct orc_chain_params {
  int64_t seed_len, max_band, max_dist_x;
  uint64_t max_iter, max_skip;
  double gap_penalty;
  int64_t min_chain_score;
} orc_chain_params;

typedef struct orc_chain {
  int64_t score;
  uint64_t target_start;  // Chain::target_start (chaining.rs:31-35)
  uint64_t seed_off;      // its seeds at [seed_off, seed_off + n_seeds) of the seed output, by query position
  uint32_t n_seeds;
  uint32_t reserved;
} orc_chain;
}

namespace {

using SeedIndex = std::unordered_map<uint64_t, std::vector<uint32_t>>;

SeedIndex seed_index(const uint64_t* vals, const uint32_t* pos, uint64_t n) {
  SeedIndex ix;
  for (uint64_t i = 0; i < n; ++i) ix[vals[i]].push_back(pos[i]);  // seeding.rs:31-34
  return ix;
}

// find_seeds_indexed_capped (seeding.rs:56-74).
std::vector<orc_seed> find_seeds(const SeedIndex& ix, const uint64_t* qv, const uint32_t* qp, uint64_t n,
                                 uint64_t max_occ) {
  std::vector<orc_seed> seeds;
  for (uint64_t i = 0; i < n; ++i) {
    auto it = ix.find(qv[i]);
    if (it == ix.end()) continue;
    if (it->second.size() > max_occ) continue;  // :63-65
    for (uint32_t tp : it->second) seeds.push_back(orc_seed{qp[i], tp, qv[i]});
  }
  std::stable_sort(seeds.begin(), seeds.end(),
                   [](const orc_seed& a, const orc_seed& b) { return a.query_pos < b.query_pos; });
  return seeds;
}

// chain_score (chaining.rs:96-117).
bool chain_score(const orc_seed& si, const orc_seed& sj, const orc_chain_params& p, int64_t* sc) {
  const int64_t dq = (int64_t)si.query_pos - (int64_t)sj.query_pos;
  if (dq <= 0) return false;
  const int64_t dr = (int64_t)si.target_pos - (int64_t)sj.target_pos;
  if (dr <= 0 || dr > p.max_dist_x + p.seed_len) return false;
  const int64_t dd = dr > dq ? dr - dq : dq - dr;
  if (dd > p.max_band) return false;
  const int64_t dg = std::min(dr, dq);
  int64_t s = std::min(p.seed_len, dg);
  if (dd > 0) {
    const double lin = p.gap_penalty * (double)dd;
    const double lg = 0.5 * std::log2((double)dd + 1.0);
    s -= (int64_t)std::round(lin + lg);  // f64::round: half away from zero
  }
  *sc = s;
  return true;
}

struct OrcChain {
  std::vector<orc_seed> seeds;
  int64_t score;
};

// chain_seeds (chaining.rs:131-253).
std::vector<OrcChain> chain_seeds(const std::vector<orc_seed>& seeds, const orc_chain_params& p) {
  std::vector<OrcChain> chains;
  if (seeds.empty()) return chains;
  std::vector<orc_seed> s = seeds;
  std::stable_sort(s.begin(), s.end(), [](const orc_seed& a, const orc_seed& b) {
    return a.target_pos != b.target_pos ? a.target_pos < b.target_pos : a.query_pos < b.query_pos;
  });
  const size_t n = s.size();
  std::vector<int64_t> f(n, 0);
  std::vector<int64_t> pred(n, -1);
  std::vector<uint64_t> dominated(n, 0);
  for (size_t i = 0; i < n; ++i) {
    int64_t best = p.seed_len;
    int64_t best_pred = -1;
    uint64_t n_skip = 0;
    const size_t floor = i > p.max_iter ? i - p.max_iter : 0;  // saturating_sub
    for (size_t j = i; j-- > floor;) {
      if ((int64_t)s[i].target_pos - (int64_t)s[j].target_pos >= p.max_dist_x + p.seed_len) break;
      int64_t sc;
      if (!chain_score(s[i], s[j], p, &sc)) continue;
      const int64_t total = sc + f[j];
      if (total > best) {
        best = total;
        best_pred = (int64_t)j;
        if (n_skip > 0) --n_skip;
      } else {
        dominated[j] += 1;
        if (dominated[j] > p.max_skip) {
          n_skip += 1;
          if (n_skip > p.max_skip) break;
        }
      }
    }
    f[i] = best;
    pred[i] = best_pred;
  }
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return f[a] > f[b]; });
  std::vector<bool> used(n, false);
  for (size_t end : order) {
    if (used[end] || f[end] < p.min_chain_score) continue;
    std::vector<orc_seed> rev;
    int64_t cur = (int64_t)end;
    while (cur >= 0) {
      if (used[cur]) break;
      used[cur] = true;
      rev.push_back(s[cur]);
      cur = pred[cur];
    }
    if (rev.empty()) continue;
    std::reverse(rev.begin(), rev.end());
    std::stable_sort(rev.begin(), rev.end(),
                     [](const orc_seed& a, const orc_seed& b) { return a.query_pos < b.query_pos; });
    chains.push_back(OrcChain{std::move(rev), f[end]});
  }
  std::stable_sort(chains.begin(), chains.end(),
                   [](const OrcChain& a, const OrcChain& b) { return a.score > b.score; });
  return chains;
}

uint64_t chain_target_start(const OrcChain& c) {
  const int64_t v = (int64_t)c.seeds[0].target_pos - (int64_t)c.seeds[0].query_pos;
  return v > 0 ? (uint64_t)v : 0;
}

}  // namespace

extern "C" {

// seed_occurrence_cap (seeding.rs:90-97): max(floor, 8 x the median occurrence).
uint64_t orc_seed_occurrence_cap(const uint64_t* vals, const uint32_t* pos, uint64_t n, uint64_t floor) {
  const SeedIndex ix = seed_index(vals, pos, n);
  if (ix.empty()) return floor;
  std::vector<uint64_t> counts;
  for (const auto& kv : ix) counts.push_back(kv.second.size());
  std::sort(counts.begin(), counts.end());
  const uint64_t median = counts[counts.size() / 2];
  const uint64_t m8 = median > UINT64_MAX / 8 ? UINT64_MAX : median * 8;
  return std::max(floor, m8);
}

// seed_and_chain (executor.rs:511-519) for a batch of query sketches against
// one target sketch.  Query i's minimizers at q_off[i] .. q_off[i + 1].  Two
// calls: with chains_cap / seeds_cap too small it returns 1 and sets the
// totals; outputs are per query in order: n_seeds[i], chains at
// chain_off[i] .. chain_off[i + 1], each chain's seeds in out_seeds.
int orc_seed_and_chain(const uint64_t* t_vals, const uint32_t* t_pos, uint64_t tn, uint64_t max_occ,
                       const uint64_t* q_vals, const uint32_t* q_pos, const uint64_t* q_off, uint32_t nq,
                       const orc_chain_params* p, int threads, uint64_t* n_seeds, orc_chain* chains,
                       uint64_t chains_cap, uint64_t* chain_off, orc_seed* out_seeds, uint64_t seeds_cap,
                       uint64_t* total_chains, uint64_t* total_seeds) {
  const SeedIndex ix = seed_index(t_vals, t_pos, tn);
  std::vector<std::vector<OrcChain>> per(nq);
  std::vector<uint64_t> ns(nq, 0);
  const int T = std::max(1, threads);
  std::vector<std::thread> pool;
  for (int w = 0; w < T; ++w)
    pool.emplace_back([&, w]() {
      for (uint32_t i = (uint32_t)w; i < nq; i += (uint32_t)T) {
        const auto seeds = find_seeds(ix, q_vals + q_off[i], q_pos + q_off[i], q_off[i + 1] - q_off[i], max_occ);
        ns[i] = seeds.size();
        per[i] = chain_seeds(seeds, *p);
      }
    });
  for (auto& t : pool) t.join();
  uint64_t C = 0, S = 0;
  for (uint32_t i = 0; i < nq; ++i) {
    C += per[i].size();
    for (const auto& c : per[i]) S += c.seeds.size();
  }
  *total_chains = C;
  *total_seeds = S;
  for (uint32_t i = 0; i < nq; ++i) n_seeds[i] = ns[i];
  if (C > chains_cap || S > seeds_cap) return 1;
  uint64_t c = 0, s = 0;
  for (uint32_t i = 0; i < nq; ++i) {
    chain_off[i] = c;
    for (const auto& ch : per[i]) {
      chains[c] = orc_chain{ch.score, chain_target_start(ch), s, (uint32_t)ch.seeds.size(), 0};
      for (const auto& sd : ch.seeds) out_seeds[s++] = sd;
      ++c;
    }
  }
  chain_off[nq] = c;
  return 0;
}

// CPU baseline: the index built once (as upstream's TargetContext holds it),
// then `passes` rounds of seed_and_chain over every query; returns the
// seconds of the passes alone.
double orc_seed_and_chain_timed(const uint64_t* t_vals, const uint32_t* t_pos, uint64_t tn, uint64_t max_occ,
                                const uint64_t* q_vals, const uint32_t* q_pos, const uint64_t* q_off, uint32_t nq,
                                const orc_chain_params* p, int threads, int passes, uint64_t* checksum) {
  const SeedIndex ix = seed_index(t_vals, t_pos, tn);
  const int T = std::max(1, threads);
  std::atomic<uint64_t> sum{0};
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < passes; ++r) {
    std::vector<std::thread> pool;
    for (int w = 0; w < T; ++w)
      pool.emplace_back([&, w]() {
        uint64_t s = 0;
        for (uint32_t i = (uint32_t)w; i < nq; i += (uint32_t)T) {
          const auto seeds = find_seeds(ix, q_vals + q_off[i], q_pos + q_off[i], q_off[i + 1] - q_off[i], max_occ);
          const auto ch = chain_seeds(seeds, *p);
          s += seeds.size() + ch.size() + (ch.empty() ? 0 : chain_target_start(ch[0]));
        }
        sum += s;
      });
    for (auto& t : pool) t.join();
  }
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *checksum = sum.load();
  return el;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Minimizer sketch (sketch_alphabet, phraya-core/src/types.rs:1846-1867).
//
// Upstream takes it from the simd-minimizers crate (2.3, phraya-core/
// Cargo.toml:9), whose source is not in the reference; this restates the
// scheme that crate publishes for canonical and forward minimizers over an
// ASCII sequence, as far as it is fixed by its description:
//   * base code (b >> 1) & 3 (A 0, C 1, T 2, G 3; other bytes by the same
//     rule), complement code ^ 2;
//   * k-mer value: the 2-bit codes with the first base most significant; a
//     canonical sketch takes the smaller of that and the reverse
//     complement's value;
//   * order: a 64-bit hash of the value (here the SplitMix64 finalizer);
//   * each window of w consecutive k-mers selects its smallest hash;
//     forward sketches take the leftmost of equal hashes; canonical ones
//     take the leftmost when more than half of the window's l = k + w - 1
//     bases are T or G and the rightmost otherwise (l odd), so a sequence
//     and its reverse complement select mirrored positions;
//   * consecutive windows selecting the same k-mer report it once.
// Exact agreement with the crate's hash and values is not pinned (no
// fixture of its output exists in the reference); what is pinned is the
// window rule (brute force), the strand symmetry and, for forward sketches,
// agreement with a second implementation.
// ---------------------------------------------------------------------------

namespace {

inline uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int64_t orc_sketch(const uint8_t* s, uint64_t n, uint32_t k, uint32_t w, int canonical, uint64_t* vals,
                   uint32_t* pos) {
  if (k == 0 || k > 32 || w == 0) return -1;
  const uint64_t l = (uint64_t)k + w - 1;
  if (canonical && l % 2 == 0) return -1;
  if (n < l) return 0;
  const uint64_t nk = n - k + 1, nwin = n - l + 1;
  std::vector<uint64_t> V(nk), H(nk);
  for (uint64_t i = 0; i < nk; ++i) {
    uint64_t F = 0, R = 0;
    for (uint32_t j = 0; j < k; ++j) {
      const uint64_t c = (s[i + j] >> 1) & 3;
      F = (F << 2) | c;
      R |= (c ^ 2) << (2 * j);
    }
    V[i] = canonical ? std::min(F, R) : F;
    H[i] = orc_mix64(V[i]);
  }
  int64_t cnt = 0;
  uint64_t last = ~0ull;
  for (uint64_t x = 0; x < nwin; ++x) {
    bool left = true;
    if (canonical) {
      uint64_t tg = 0;
      for (uint64_t j = x; j < x + l; ++j) tg += ((s[j] >> 1) & 3) >> 1;
      left = 2 * tg > l;
    }
    uint64_t best = x;
    for (uint64_t i = x + 1; i < x + w; ++i)
      if (H[i] < H[best] || (!left && H[i] == H[best])) best = i;
    if (best != last) {
      vals[cnt] = V[best];
      pos[cnt] = (uint32_t)best;
      ++cnt;
      last = best;
    }
  }
  return cnt;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The same sketch computed the way a CPU implementation would (rolling
// k-mer values, a running T/G count, one scan of w hashes per window), for
// the CPU baseline; checked against orc_sketch by the tests.
// ---------------------------------------------------------------------------

namespace {

int64_t sketch_rolling(const uint8_t* s, uint64_t n, uint32_t k, uint32_t w, bool canonical, std::vector<uint64_t>& V,
                       std::vector<uint64_t>& H, uint64_t* vals, uint32_t* pos) {
  const uint64_t l = (uint64_t)k + w - 1;
  if (n < l) return 0;
  const uint64_t nk = n - k + 1;
  V.resize(nk);
  H.resize(nk);
  const uint64_t kmask = k == 32 ? ~0ull : ((1ull << (2 * k)) - 1);
  const uint32_t rsh = 2 * (k - 1);
  uint64_t F = 0, R = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t c = (s[i] >> 1) & 3;
    F = ((F << 2) | c) & kmask;
    R = (R >> 2) | ((c ^ 2) << rsh);
    if (i + 1 >= k) {
      const uint64_t v = canonical && R < F ? R : F;
      V[i + 1 - k] = v;
      H[i + 1 - k] = orc_mix64(v);
    }
  }
  uint64_t tg = 0;
  for (uint64_t j = 0; j < l; ++j) tg += ((s[j] >> 1) & 3) >> 1;
  int64_t cnt = 0;
  uint64_t last = ~0ull;
  for (uint64_t x = 0;; ++x) {
    const bool left = !canonical || 2 * tg > l;
    uint64_t best = x, hb = H[x];
    for (uint64_t i = x + 1; i < x + w; ++i)
      if (H[i] < hb || (!left && H[i] == hb)) {
        hb = H[i];
        best = i;
      }
    if (best != last) {
      vals[cnt] = V[best];
      pos[cnt] = (uint32_t)best;
      ++cnt;
      last = best;
    }
    if (x + l >= n) break;
    tg += (((s[x + l] >> 1) & 3) >> 1) - (((s[x] >> 1) & 3) >> 1);
  }
  return cnt;
}

}  // namespace

extern "C" {

int64_t orc_sketch_rolling(const uint8_t* s, uint64_t n, uint32_t k, uint32_t w, int canonical, uint64_t* vals,
                           uint32_t* pos) {
  if (k == 0 || k > 32 || w == 0) return -1;
  if (canonical && ((uint64_t)k + w - 1) % 2 == 0) return -1;
  std::vector<uint64_t> V, H;
  return sketch_rolling(s, n, k, w, canonical != 0, V, H, vals, pos);
}

// CPU baseline: `passes` rounds of sketching every sequence of the batch
// (sequences split over `threads`); returns the seconds of the passes.
double orc_sketch_timed(const uint8_t* seq, const uint64_t* off, uint32_t nseq, uint32_t k, uint32_t w, int canonical,
                        int threads, int passes, uint64_t* checksum) {
  const int T = std::max(1, threads);
  std::atomic<uint64_t> sum{0};
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < passes; ++r) {
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
      pool.emplace_back([&, t]() {
        std::vector<uint64_t> V, H, vals;
        std::vector<uint32_t> pos;
        uint64_t s = 0;
        for (uint32_t i = (uint32_t)t; i < nseq; i += (uint32_t)T) {
          const uint64_t n = off[i + 1] - off[i];
          vals.resize(n + 1);
          pos.resize(n + 1);
          const int64_t c = sketch_rolling(seq + off[i], n, k, w, canonical != 0, V, H, vals.data(), pos.data());
          s += (uint64_t)c + (c > 0 ? vals[c - 1] : 0);
        }
        sum += s;
      });
    for (auto& th : pool) th.join();
  }
  const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  *checksum = sum.load();
  return el;
}

}  // extern "C"
