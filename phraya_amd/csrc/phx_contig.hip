// phx_contig.hip -- contig mode: align pairs of long contig segments by
// anchoring them on shared minimizers, aligning every inter-anchor gap on the
// GPU in one batched plan, and stitching each contig's CIGAR.
//
// Anchoring (host, threads over contig pairs):
//  * (k, w) minimizers of both sequences, forward strand: for every window of
//    w consecutive k-mers without non-ACGT bytes, the k-mer of smallest
//    hash (leftmost on ties);
//  * anchors = minimizers whose k-mer occurs exactly once among the
//    minimizers of each sequence, paired by k-mer;
//  * the longest chain of anchors increasing in both sequences;
//  * spaced out: walking the chain, the next anchor kept is the first one at
//    least `spacing` bases past the end of the previous one in the query,
//    with spacing drawn log-uniform in [gap_min, gap_max] (seeded per pair).
// Each gap between kept anchors (and before the first / after the last) is a
// window pair of the resident query and target arenas; the plan aligns all of
// them at once with the chosen engine.  Stitching concatenates, per contig,
// gap 0's runs, the first anchor as a k-base M run, gap 1's runs, ..., and
// merges equal neighbouring runs at the joins.  Every gap is pinned at both
// ends, so it always takes the global branch (never the fitting switch of the
// per-pair wrappers) and consumes its whole window; the contig score (the sum
// of the gap scores) is then the cost of the stitched CIGAR.
#include <hip/hip_runtime.h>

#include <chrono>
#include <mutex>
#include <string>
#include <vector>

#include "phraya_hip.h"
#include "phx_device.h"
#include "phx_host.h"

using namespace phx;
using namespace phx::host;

namespace {

struct ContigGap {
  uint64_t q0, q1, t0, t1;  // window [q0, q1) x [t0, t1) inside the contig pair
};

uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

struct Minimizer {
  uint32_t code;
  uint32_t pos;
};

// (k, w) minimizers, forward strand, leftmost smallest hash per window.
std::vector<Minimizer> minimizers(const uint8_t* s, size_t n, uint32_t k, uint32_t w) {
  std::vector<Minimizer> out;
  if (n < k) return out;
  const uint8_t* code = acgt_table();
  const uint32_t mask = k >= 16 ? 0xFFFFFFFFu : ((1u << (2 * k)) - 1);
  struct Item {
    uint32_t h, code, pos;
  };
  std::vector<Item> dq(n + 1);  // monotonic queue storage
  size_t head = 0, tail = 0;
  uint32_t cur = 0, valid = 0, kmers = 0;  // kmers: consecutive valid k-mers seen
  uint32_t last_pos = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; ++i) {
    const uint8_t c = code[s[i]];
    if (c == 0xFF) {
      valid = 0;
      kmers = 0;
      head = tail = 0;
      continue;
    }
    cur = ((cur << 2) | c) & mask;
    if (++valid < k) continue;
    const uint32_t pos = (uint32_t)(i + 1 - k);
    const uint32_t h = mix32(cur);
    while (tail > head && dq[tail - 1].h > h) --tail;  // keep the leftmost of equal hashes
    dq[tail++] = Item{h, cur, pos};
    ++kmers;
    while (dq[head].pos + w <= pos) ++head;
    if (kmers >= w && dq[head].pos != last_pos) {
      last_pos = dq[head].pos;
      out.push_back(Minimizer{dq[head].code, dq[head].pos});
    }
  }
  return out;
}

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Kept anchors (query pos, target pos) of one contig pair, in order.
std::vector<std::pair<uint64_t, uint64_t>> anchor_pair(const uint8_t* q, size_t qn, const uint8_t* t,
                                                       size_t tn, const phx_contig_params& p,
                                                       uint64_t pair_index) {
  std::vector<Minimizer> mq = minimizers(q, qn, p.k, p.w), mt = minimizers(t, tn, p.k, p.w);
  auto by_code = [](const Minimizer& a, const Minimizer& b) {
    return a.code < b.code || (a.code == b.code && a.pos < b.pos);
  };
  std::sort(mq.begin(), mq.end(), by_code);
  std::sort(mt.begin(), mt.end(), by_code);
  // Codes that occur once on each side.
  std::vector<std::pair<uint32_t, uint32_t>> cand;  // (q pos, t pos)
  size_t i = 0, j = 0;
  while (i < mq.size() && j < mt.size()) {
    const uint32_t a = mq[i].code, b = mt[j].code;
    if (a < b) { ++i; continue; }
    if (b < a) { ++j; continue; }
    size_t i2 = i, j2 = j;
    while (i2 < mq.size() && mq[i2].code == a) ++i2;
    while (j2 < mt.size() && mt[j2].code == a) ++j2;
    if (i2 - i == 1 && j2 - j == 1) cand.emplace_back(mq[i].pos, mt[j].pos);
    i = i2;
    j = j2;
  }
  std::sort(cand.begin(), cand.end());
  // Longest chain increasing in both (patience sorting on target positions).
  std::vector<uint32_t> tails, tail_idx;
  std::vector<int64_t> prev(cand.size(), -1);
  for (size_t c = 0; c < cand.size(); ++c) {
    const uint32_t tp = cand[c].second;
    const size_t at = std::lower_bound(tails.begin(), tails.end(), tp) - tails.begin();
    if (at == tails.size()) {
      tails.push_back(tp);
      tail_idx.push_back((uint32_t)c);
    } else {
      tails[at] = tp;
      tail_idx[at] = (uint32_t)c;
    }
    prev[c] = at ? (int64_t)tail_idx[at - 1] : -1;
  }
  std::vector<std::pair<uint32_t, uint32_t>> chain;
  for (int64_t c = tails.empty() ? -1 : (int64_t)tail_idx.back(); c >= 0; c = prev[c])
    chain.push_back(cand[(size_t)c]);
  std::reverse(chain.begin(), chain.end());
  // Space the anchors out.
  SplitMix rng{p.seed ^ (0xA0761D6478BD642Full * (pair_index + 1))};
  auto draw = [&] {
    const double lo = std::log((double)p.gap_min), hi = std::log((double)p.gap_max);
    return (uint64_t)std::exp(lo + rng.uniform() * (hi - lo));
  };
  std::vector<std::pair<uint64_t, uint64_t>> kept;
  uint64_t qe = 0, te = 0, want = draw();
  for (const auto& a : chain) {
    if (a.first < qe + want || a.second < te) continue;
    if ((uint64_t)a.first + p.k > qn || (uint64_t)a.second + p.k > tn) continue;
    kept.emplace_back(a.first, a.second);
    qe = (uint64_t)a.first + p.k;
    te = (uint64_t)a.second + p.k;
    want = draw();
  }
  return kept;
}

phx_contig_params fill_defaults(const phx_contig_params* in) {
  phx_contig_params p{};
  if (in) p = *in;
  if (!p.k) p.k = 15;
  if (!p.w) p.w = 10;
  if (!p.gap_min) p.gap_min = 1000;
  if (!p.gap_max) p.gap_max = 20000;
  return p;
}

bool params_ok(const phx_contig_params& p) {
  return p.k >= 8 && p.k <= 16 && p.w >= 1 && p.w <= 256 && p.gap_min >= 1 && p.gap_max >= p.gap_min;
}

struct DevMem {
  void* p = nullptr;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

struct phx_contig_plan {
  phx_ctx* ctx = nullptr;
  int engine = PHX_ENGINE_AFFINE;
  phx_contig_params params{};
  uint32_t n = 0;                       // contig pairs
  std::vector<uint64_t> qn, tn;         // contig lengths
  std::vector<uint64_t> gap_first;      // gaps of contig c: [gap_first[c], gap_first[c + 1])
  std::vector<ContigGap> gaps;
  std::vector<std::pair<uint64_t, uint64_t>> anchors;  // anchor after gap g (none after a contig's last)
  DevMem d_q2, d_q8, d_t2, d_t8;        // resident arenas (2-bit; 8-bit when needed)
  phx_plan* plan = nullptr;             // all gaps, one batch
  // Results of the last run.
  std::vector<phx_alignment> aln;       // per contig
  std::vector<uint32_t> runs;
  std::vector<uint64_t> run_off;
  double last_ms = 0, last_gpu_ms = 0, last_stitch_ms = 0;
  bool have = false;
};

namespace {

// Upload concatenated sequences (offsets off[0..n]) as a 2-bit arena and,
// when some window must compare raw bytes (want8), an 8-bit arena too; two pad
// words each.
bool upload_arena(const uint8_t* seq, const uint64_t* off, uint32_t n, bool want8, DevMem& d2,
                  DevMem& d8) {
  const uint64_t total = off[n] - off[0];
  const uint8_t* base = seq + off[0];
  std::vector<uint64_t> w2((total + 31) / 32 + 2, 0);
  const uint8_t* code = acgt_table();
  for (uint64_t i = 0; i < total; ++i) {
    const uint8_t c = code[base[i]];
    w2[i >> 5] |= (uint64_t)(c == 0xFF ? 0 : c) << ((i & 31) * 2);
  }
  if (hipMalloc(&d2.p, w2.size() * 8) != hipSuccess ||
      hipMemcpy(d2.p, w2.data(), w2.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
    return false;
  if (want8) {
    std::vector<uint64_t> w8((total + 7) / 8 + 2, 0);
    std::memcpy(w8.data(), base, total);
    if (hipMalloc(&d8.p, w8.size() * 8) != hipSuccess ||
        hipMemcpy(d8.p, w8.data(), w8.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return false;
  }
  return true;
}

bool window_acgt(const uint8_t* s, uint64_t a, uint64_t b) { return is_acgt(s + a, (size_t)(b - a)); }

}  // namespace

extern "C" {

int64_t phx_contig_anchors(const uint8_t* q, uint64_t qn, const uint8_t* t, uint64_t tn,
                           const phx_contig_params* params, uint64_t pair_index, uint64_t* out,
                           uint64_t cap) {
  const phx_contig_params p = fill_defaults(params);
  if (!params_ok(p) || (qn && !q) || (tn && !t)) return -1;
  const auto kept = anchor_pair(q, (size_t)qn, t, (size_t)tn, p, pair_index);
  for (size_t i = 0; i < kept.size() && i < cap; ++i) {
    out[2 * i] = kept[i].first;
    out[2 * i + 1] = kept[i].second;
  }
  return (int64_t)kept.size();
}

phx_contig_plan* phx_contig_plan_create(phx_ctx* ctx, int32_t engine, phx_costs costs,
                                        const phx_batch* b, const phx_contig_params* params) {
  if (!ctx || !b) return nullptr;
  const phx_contig_params p = fill_defaults(params);
  if (!params_ok(p) || (engine != PHX_ENGINE_AFFINE && engine != PHX_ENGINE_LINEAR)) {
    set_error(ctx, "contig mode needs the affine or linear engine and 8 <= k <= 16");
    return nullptr;
  }
  if (hipSetDevice(ctx->device) != hipSuccess) {
    set_error(ctx, "hipSetDevice failed", PHX_E_DEVICE);
    return nullptr;
  }
  phx_contig_plan* C = new phx_contig_plan;
  C->ctx = ctx;
  C->engine = engine;
  C->params = p;
  C->n = b->n;
  const uint32_t n = b->n;
  C->qn.resize(n);
  C->tn.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    C->qn[i] = b->q_off[i + 1] - b->q_off[i];
    C->tn[i] = b->t_off[i + 1] - b->t_off[i];
  }
  // Anchors per contig pair (host threads).
  std::vector<std::vector<std::pair<uint64_t, uint64_t>>> kept(n);
  parallel_for(n, [&](size_t i) {
    kept[i] = anchor_pair(b->q_seq + b->q_off[i], (size_t)C->qn[i], b->t_seq + b->t_off[i],
                          (size_t)C->tn[i], p, i);
  });
  C->gap_first.assign(n + 1, 0);
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t qe = 0, te = 0;
    for (const auto& a : kept[i]) {
      C->gaps.push_back(ContigGap{qe, a.first, te, a.second});
      C->anchors.push_back(a);
      qe = a.first + p.k;
      te = a.second + p.k;
    }
    C->gaps.push_back(ContigGap{qe, C->qn[i], te, C->tn[i]});
    C->anchors.emplace_back(0, 0);  // no anchor after a contig's last gap
    C->gap_first[i + 1] = C->gaps.size();
  }
  // Every gap must fit the device format (history cells hold offset + 1 in 28
  // bits): a pair with no anchors, or a long anchor-free stretch, makes one
  // gap as long as the stretch.  Refused here, before anything is uploaded.
  for (size_t g = 0; g < C->gaps.size(); ++g) {
    const ContigGap& G = C->gaps[g];
    if (G.q1 - G.q0 > (uint64_t)kMaxLen || G.t1 - G.t0 > (uint64_t)kMaxLen) {
      set_error(ctx, "contig gap " + std::to_string(g) + " (" + std::to_string(G.q1 - G.q0) + " x " +
                         std::to_string(G.t1 - G.t0) +
                         " bases) is longer than the device format supports (2^27 - 2)");
      delete C;
      return nullptr;
    }
  }
  // Resident arenas and one window descriptor per gap.
  const bool qb = !is_acgt(b->q_seq + b->q_off[0], (size_t)(b->q_off[n] - b->q_off[0]));
  const bool tb = !is_acgt(b->t_seq + b->t_off[0], (size_t)(b->t_off[n] - b->t_off[0]));
  if (!upload_arena(b->q_seq, b->q_off, n, qb || tb, C->d_q2, C->d_q8) ||
      !upload_arena(b->t_seq, b->t_off, n, qb || tb, C->d_t2, C->d_t8)) {
    set_error(ctx, "contig arena upload failed", PHX_E_NOMEM);
    delete C;
    return nullptr;
  }
  const uint64_t q_origin = b->q_off[0], t_origin = b->t_off[0];
  std::vector<PairDesc> pairs(C->gaps.size());
  std::vector<const uint8_t*> qh(C->gaps.size()), th(C->gaps.size());
  for (uint32_t i = 0; i < n; ++i) {
    const uint8_t* qs = b->q_seq + b->q_off[i];
    const uint8_t* ts = b->t_seq + b->t_off[i];
    for (uint64_t g = C->gap_first[i]; g < C->gap_first[i + 1]; ++g) {
      const ContigGap& G = C->gaps[g];
      const uint64_t qa = b->q_off[i] - q_origin + G.q0, ta = b->t_off[i] - t_origin + G.t0;
      const bool b8 = (qb && !window_acgt(qs, G.q0, G.q1)) || (tb && !window_acgt(ts, G.t0, G.t1));
      PairDesc& pd = pairs[g];
      const uint32_t per = b8 ? 8 : 32;
      pd.q = (const uint64_t*)(b8 ? C->d_q8.p : C->d_q2.p) + qa / per;
      pd.q_shift = (uint32_t)(qa % per);
      pd.t = (const uint64_t*)(b8 ? C->d_t8.p : C->d_t2.p) + ta / per;
      pd.t_shift = (uint32_t)(ta % per);
      pd.qn = (int32_t)(G.q1 - G.q0);
      pd.tn = (int32_t)(G.t1 - G.t0);
      // A gap is pinned by anchors (or a contig end) on both sides: always
      // the global branch, so the contig score is the cost of its CIGAR.
      pd.flags = (b8 ? PF_BYTES : 0) | PF_GLOBAL;
      qh[g] = qs + G.q0;
      th[g] = ts + G.t0;
    }
  }
  C->plan = plan_create_windows(ctx, engine, costs, -1, std::move(pairs), qh, th);
  if (!C->plan) {
    delete C;
    return nullptr;
  }
  return C;
}

uint64_t phx_contig_plan_gaps(const phx_contig_plan* C) { return C ? C->gaps.size() : 0; }

phx_plan* phx_contig_plan_gap_plan(phx_contig_plan* C) { return C ? C->plan : nullptr; }

int64_t phx_contig_plan_layout(const phx_contig_plan* C, uint32_t contig, uint64_t* gaps,
                               uint64_t cap) {
  if (!C || contig >= C->n) return -1;
  const uint64_t a = C->gap_first[contig], z = C->gap_first[contig + 1];
  for (uint64_t g = a; g < z && g - a < cap; ++g) {
    const ContigGap& G = C->gaps[g];
    uint64_t* o = gaps + 4 * (g - a);
    o[0] = G.q0;
    o[1] = G.q1;
    o[2] = G.t0;
    o[3] = G.t1;
  }
  return (int64_t)(z - a);
}

int phx_contig_plan_run(phx_contig_plan* C) {
  if (!C) return PHX_E_ARG;
  using Clock = std::chrono::steady_clock;
  const auto t0 = Clock::now();
  C->have = false;
  int rc = phx_plan_run(C->plan);
  if (rc != PHX_OK) return rc;
  // Gap results.
  const uint64_t G = C->gaps.size();
  std::vector<phx_alignment> ga(G);
  std::vector<uint64_t> goff(G + 1);
  std::vector<uint32_t> gr;
  phx_results r{ga.data(), nullptr, 0, goff.data()};
  rc = phx_plan_results(C->plan, &r);  // sizes the run arena (run_off[G])
  if (rc == PHX_E_CAPACITY) {
    gr.resize(goff[G]);
    r.runs = gr.data();
    r.runs_cap = gr.size();
    rc = phx_plan_results(C->plan, &r);
  }
  const auto t1 = Clock::now();
  if (rc != PHX_OK) return rc;
  // Stitch per contig (host threads).
  const uint32_t k = C->params.k;
  std::vector<std::vector<uint32_t>> per(C->n);
  C->aln.assign(C->n, phx_alignment{});
  parallel_for(C->n, [&](size_t c) {
    std::vector<uint32_t>& out = per[c];
    phx_alignment& a = C->aln[c];
    auto put = [&](uint32_t op, uint64_t len) {
      if (!len) return;
      if (!out.empty() && (out.back() & 15u) == op) out.back() += (uint32_t)(len << 4);
      else out.push_back((uint32_t)(len << 4) | op);
    };
    int64_t score = 0;
    for (uint64_t g = C->gap_first[c]; g < C->gap_first[c + 1]; ++g) {
      const phx_alignment& x = ga[g];
      if (x.status != PHX_OK) {
        a.status = x.status;
        out.clear();
        return;
      }
      // Global gap alignments consume their whole target window.
      if ((uint64_t)x.target_end != C->gaps[g].t1 - C->gaps[g].t0) {
        a.status = PHX_E_INTERNAL;
        out.clear();
        return;
      }
      for (uint64_t i = goff[g]; i < goff[g + 1]; ++i) put(gr[i] & 15u, gr[i] >> 4);
      score += x.score;
      if (g + 1 < C->gap_first[c + 1]) put(OP_M, k);
    }
    int64_t ed = 0;
    for (uint32_t run : out)
      if ((run & 15u) != OP_M) ed += run >> 4;
    a.status = PHX_OK;
    a.score = score;
    a.edit_distance = ed;
    a.query_start = a.target_start = 0;
    a.query_end = (int64_t)C->qn[c];
    a.target_end = (int64_t)C->tn[c];
  });
  C->run_off.assign(C->n + 1, 0);
  for (uint32_t c = 0; c < C->n; ++c) C->run_off[c + 1] = C->run_off[c] + per[c].size();
  C->runs.resize(C->run_off[C->n]);
  parallel_for(C->n, [&](size_t c) {
    std::copy(per[c].begin(), per[c].end(), C->runs.begin() + (ptrdiff_t)C->run_off[c]);
  });
  const auto t2 = Clock::now();
  C->last_gpu_ms = phx_plan_last_ms(C->plan);
  C->last_stitch_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
  C->last_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
  C->have = true;
  return PHX_OK;
}

int phx_contig_plan_results(phx_contig_plan* C, phx_results* out) {
  if (!C || !out || !out->aln || !out->run_off) return PHX_E_ARG;
  if (!C->have) return PHX_E_ARG;
  for (uint32_t c = 0; c < C->n; ++c) out->aln[c] = C->aln[c];
  for (uint32_t c = 0; c <= C->n; ++c) out->run_off[c] = C->run_off[c];
  if (C->runs.size() > out->runs_cap || (!out->runs && !C->runs.empty())) return PHX_E_CAPACITY;
  if (!C->runs.empty()) std::memcpy(out->runs, C->runs.data(), C->runs.size() * 4);
  return PHX_OK;
}

double phx_contig_plan_last_ms(const phx_contig_plan* C) { return C ? C->last_ms : 0.0; }
double phx_contig_plan_last_gpu_ms(const phx_contig_plan* C) { return C ? C->last_gpu_ms : 0.0; }
double phx_contig_plan_last_stitch_ms(const phx_contig_plan* C) { return C ? C->last_stitch_ms : 0.0; }

void phx_contig_plan_destroy(phx_contig_plan* C) {
  if (!C) return;
  if (C->plan) phx_plan_destroy(C->plan);
  delete C;
}

}  // extern "C"
