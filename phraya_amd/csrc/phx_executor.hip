// phx_executor.hip -- batched form of phraya-align's per-read extension step
// (extend_chains, phraya-align/src/executor.rs:528-567) and its scoring
// (score_alignments, phraya-align/src/lib.rs:170-199).
//
// Upstream extends one read's anchors one call at a time on a rayon thread.
// Here every anchor of every read in a batch becomes one pair of one device
// batch per engine, so a 4096-read batch of the executor is one or two
// launches; the per-read selection of the primary and the alternatives runs
// on the host afterwards, in flat passes, with upstream's rules.  The target
// is either a phx_reference kept resident in HBM (each anchor's window is a
// descriptor into it: only the queries are packed and uploaded per call) or
// host bytes (the windows are packed with the queries).  The rules:
//  * anchors_from_chains (executor.rs:347-362): the first chain_cap(strategy)
//    chains (Fast 1, Balanced 2, Sensitive 50, :330-336) give anchors
//    (0, chain.target_start); a read with no chain gets the anchor (0, 0);
//  * the target window is target[.. min(anchor.target_pos + 2 * qlen, |target|)]
//    (executor.rs:551-553), so the pair aligned is query vs the window's
//    suffix from target_pos, with positions reported from the window start;
//  * extend_anchor (executor.rs:379-404): Sensitive + affine -> gap-affine WFA
//    with AffineCosts::default() (2/4/1, wfa_simd.rs:527-534), Sensitive +
//    linear -> linear WFA, Balanced / Fast -> Myers fitting for qlen <= 500
//    (MYERS_MAX_QUERY_LEN, :248) else linear WFA;
//  * failed anchors are skipped (:556-559); no successful anchor -> no result;
//  * score_alignments: primary = the first alignment with the smallest edit
//    distance; alternatives = every other alignment not equal to the primary
//    (all fields, CIGAR included) whose normalised score ratio
//    (1 - ed / qlen) / (1 - ed_primary / qlen) is >= 0.95, in anchor order.
// And align_read's strand handling (executor.rs:573-733) for a batch:
//  * reverse_complement (phraya-core/src/types.rs:91-107) of DNA reads;
//  * extend only the orientation with chains when the other has none,
//    otherwise both (:676-698; protein: no reverse chains, empty reverse
//    query, :632-646), all orientations of all reads in one batch;
//  * keep the reverse orientation only when its primary edit distance is
//    strictly smaller (:700-711); then the Fast cutoff (divergence
//    ed / max(qlen, 1) > 0.20, :253, :716-723) and the outcome: placed when
//    1 - ed / max(qlen, 1) >= 0.95 (:263), else below-threshold when the
//    read had seeds, else no-seed (:727-733).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "phraya_hip.h"
#include "phx_device.h"
#include "phx_host.h"

using namespace phx;
using namespace phx::host;

// A target resident in HBM: 2-bit words (other bytes packed as A; windows
// holding them use the 8-bit copy) and 8-bit words, each with pad words, and
// on the host the bytes and per-64-byte-block prefix counts of non-ACGT bytes
// (a window's encoding is then found in O(1)).
struct phx_reference {
  phx_ctx* ctx = nullptr;
  uint64_t len = 0;
  std::vector<uint8_t> host;
  std::vector<uint32_t> bad_before;  // non-ACGT bytes before block b (64 bytes each)
  void* d2 = nullptr;
  void* d8 = nullptr;
};

namespace {

uint32_t cap_of(int32_t strategy) {
  return strategy == PHX_STRATEGY_FAST ? 1u : strategy == PHX_STRATEGY_BALANCED ? 2u : 50u;
}

int32_t engine_of(int32_t strategy, int32_t gap_model, uint64_t qlen) {
  if (strategy == PHX_STRATEGY_SENSITIVE)
    return gap_model == PHX_GAP_LINEAR ? PHX_ENGINE_LINEAR : PHX_ENGINE_AFFINE;
  return qlen <= 500 ? PHX_ENGINE_MYERS : PHX_ENGINE_LINEAR;
}

bool same_alignment(const phx_alignment& a, const uint32_t* ra, uint64_t na, const phx_alignment& b,
                    const uint32_t* rb, uint64_t nb) {
  return a.edit_distance == b.edit_distance && a.query_start == b.query_start &&
         a.query_end == b.query_end && a.target_start == b.target_start &&
         a.target_end == b.target_end && na == nb &&
         (na == 0 || std::memcmp(ra, rb, na * sizeof(uint32_t)) == 0);
}

uint8_t complement(uint8_t b) {
  switch (b) {
    case 'A': return 'T';
    case 'T': return 'A';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'a': return 't';
    case 't': return 'a';
    case 'c': return 'g';
    case 'g': return 'c';
    default: return b;
  }
}

// Non-ACGT bytes in [a, b) of a reference.
bool window_acgt(const phx_reference* R, uint64_t a, uint64_t b) {
  if (a >= b) return true;
  const uint64_t ba = (a + 63) / 64, bb = b / 64;
  if (ba >= bb) return is_acgt(R->host.data() + a, (size_t)(b - a));
  if (R->bad_before[bb] != R->bad_before[ba]) return false;
  return is_acgt(R->host.data() + a, (size_t)(ba * 64 - a)) &&
         is_acgt(R->host.data() + bb * 64, (size_t)(b - bb * 64));
}

// The target of one call: a resident reference, or host bytes.
struct TargetSrc {
  const phx_reference* ref;
  const uint8_t* host;
  uint64_t len;
};

// Every anchor of every read through the device (one plan per engine), then
// the per-read selection.  See phx_extend_chains_batch.
int extend_core(phx_ctx* ctx, int32_t strategy, int32_t gap_model, const TargetSrc& T,
                const phx_read_chains* reads, uint32_t n_reads, phx_read_extension* out,
                phx_alignment* anchor_aln, uint8_t* anchor_role, uint64_t anchors_cap, uint32_t* runs,
                uint64_t runs_cap, uint64_t* run_off) {
  // Anchors, read by read (executor.rs:347-362).
  std::vector<uint64_t> first(n_reads + 1, 0);
  for (uint32_t r = 0; r < n_reads; ++r) {
    if (!reads[r].query && reads[r].query_len) return PHX_E_ARG;
    const uint32_t nc = reads[r].n_chains;
    if (nc && !reads[r].chain_target_start) return PHX_E_ARG;
    if (reads[r].query_len > (uint64_t)kMaxLen) {
      set_error(ctx, "read longer than the device format supports (2^27 - 2)");
      return PHX_E_ARG;
    }
    first[r + 1] = first[r] + (nc == 0 ? 1 : std::min(nc, cap_of(strategy)));
  }
  const uint64_t A = first[n_reads];
  uint64_t need_a = 0, need_r = 0;
  phx_extend_chains_bounds(strategy, T.len, reads, n_reads, &need_a, &need_r);
  if (!anchor_aln || !anchor_role || anchors_cap < A || !runs || runs_cap < need_r) return PHX_E_CAPACITY;
  // Per anchor: its read, window start and end, engine (5 = invalid seed).
  std::vector<uint32_t> read_of(A);
  std::vector<uint64_t> tpos(A), wend(A);
  std::vector<uint8_t> eng(A);
  uint64_t count[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t r = 0; r < n_reads; ++r) {
    const phx_read_chains& rd = reads[r];
    const int32_t e = engine_of(strategy, gap_model, rd.query_len);
    for (uint64_t a = first[r]; a < first[r + 1]; ++a) {
      read_of[a] = r;
      const uint64_t tp = rd.n_chains == 0 ? 0 : rd.chain_target_start[a - first[r]];
      tpos[a] = tp;
      // The seed must lie within the window target[..min(tp + 2 qlen, |t|)]
      // (lib.rs:145-149): it does exactly when tp <= |t|.
      wend[a] = tp > T.len ? tp : std::min(tp + 2 * rd.query_len, T.len);
      eng[a] = (uint8_t)(tp > T.len ? 5 : e);
      ++count[eng[a]];
    }
  }
  // Encodings: a pair is 2-bit when its query and window are pure ACGT.
  std::vector<uint8_t> q_acgt(n_reads), pair_b8(A, 0);
  parallel_for(n_reads, [&](size_t r) { q_acgt[r] = is_acgt(reads[r].query, reads[r].query_len); });
  parallel_for(A, [&](size_t a) {
    if (eng[a] == 5) return;
    const bool w = T.ref ? window_acgt(T.ref, tpos[a], wend[a]) : is_acgt(T.host + tpos[a], wend[a] - tpos[a]);
    pair_b8[a] = !(w && q_acgt[read_of[a]]);
  });
  // Host arena: each read's query in the encodings its pairs use (2-bit and /
  // or 8-bit, word aligned), and without a reference every window.
  std::vector<uint64_t> q2(n_reads, kNoWord), q8(n_reads, kNoWord), tw(A, kNoWord);
  std::vector<uint8_t> need(n_reads, 0);  // bit 0: 2-bit, bit 1: 8-bit
  for (uint64_t a = 0; a < A; ++a)
    if (eng[a] != 5) need[read_of[a]] |= pair_b8[a] ? 2 : 1;
  uint64_t words = 0;
  for (uint32_t r = 0; r < n_reads; ++r) {
    if (need[r] & 1) { q2[r] = words; words += words_for(reads[r].query_len, false); }
    if (need[r] & 2) { q8[r] = words; words += words_for(reads[r].query_len, true); }
  }
  if (!T.ref)
    for (uint64_t a = 0; a < A; ++a)
      if (eng[a] != 5) { tw[a] = words; words += words_for(wend[a] - tpos[a], pair_b8[a]); }
  std::vector<uint64_t> arena(std::max<uint64_t>(words, 1));
  parallel_for(n_reads, [&](size_t r) {
    if (q2[r] != kNoWord) pack_into(arena.data() + q2[r], reads[r].query, reads[r].query_len, false);
    if (q8[r] != kNoWord) pack_into(arena.data() + q8[r], reads[r].query, reads[r].query_len, true);
  });
  if (!T.ref)
    parallel_for(A, [&](size_t a) {
      if (tw[a] != kNoWord) pack_into(arena.data() + tw[a], T.host + tpos[a], wend[a] - tpos[a], pair_b8[a]);
    });
  // Results of each engine's plan: per-anchor fields, and its runs in a
  // block of the run arena (one engine: the caller's arena itself).
  int n_eng = 0;
  for (int e = 0; e < 5; ++e) n_eng += count[e] > 0;
  std::vector<uint64_t> slot(A, 0);  // anchor -> index in its engine's plan
  std::vector<std::vector<uint32_t>> eruns(5);
  std::vector<std::vector<uint64_t>> eoff(5);
  std::vector<phx_alignment> ealn(A);
  std::vector<uint64_t> ebase(6, 0);  // first ealn index of each engine
  for (int e = 0; e < 5; ++e) ebase[e + 1] = ebase[e] + count[e];
  for (int e = 0; e < 5; ++e) {
    if (!count[e]) continue;
    std::vector<PairDesc> pairs(count[e]);
    std::vector<uint64_t> qw(count[e]), twv(count[e]);
    std::vector<const uint8_t*> qh(count[e]), th(count[e]);
    uint64_t k = 0;
    for (uint64_t a = 0; a < A; ++a) {
      if (eng[a] != e) continue;
      slot[a] = k;
      const phx_read_chains& rd = reads[read_of[a]];
      const bool b8 = pair_b8[a];
      PairDesc& pd = pairs[k];
      pd.qn = (int32_t)rd.query_len;
      pd.tn = (int32_t)(wend[a] - tpos[a]);
      pd.flags = b8 ? PF_BYTES : 0;
      qw[k] = b8 ? q8[read_of[a]] : q2[read_of[a]];
      if (T.ref) {
        const uint32_t per = b8 ? 8 : 32;
        pd.t = (const uint64_t*)(b8 ? T.ref->d8 : T.ref->d2) + tpos[a] / per;
        pd.t_shift = (uint32_t)(tpos[a] % per);
        twv[k] = kNoWord;
      } else {
        twv[k] = tw[a];
      }
      qh[k] = rd.query;
      th[k] = (T.ref ? T.ref->host.data() : T.host) + tpos[a];
      ++k;
    }
    // AffineCosts::default() (wfa_simd.rs:527-534) for the affine engine.
    phx_plan* P = plan_create_windows(ctx, e, phx_costs{2, 4, 1}, -1, std::move(pairs), qh, th, true,
                                      &arena, &qw, &twv);
    if (!P) return last_error_code(ctx);
    int rc = phx_plan_run(P);
    uint32_t* rdst = nullptr;
    uint64_t rcap = 0;
    eoff[e].assign(count[e] + 1, 0);
    phx_results res{ealn.data() + ebase[e], nullptr, 0, eoff[e].data()};
    if (rc == PHX_OK) {
      rc = phx_plan_results(P, &res);  // sizes the runs (run_off[count])
      if (rc == PHX_E_CAPACITY) {
        if (n_eng == 1) {
          rdst = runs;
          rcap = runs_cap;
        } else {
          eruns[e].resize(eoff[e][count[e]]);
          rdst = eruns[e].data();
          rcap = eruns[e].size();
        }
        res.runs = rdst;
        res.runs_cap = rcap;
        rc = phx_plan_results(P, &res);
      }
    }
    phx_plan_destroy(P);
    if (rc != PHX_OK) return rc;
  }
  // Per anchor: fields (positions from the window start) and runs in anchor
  // order.
  uint64_t pos = 0;
  for (uint64_t a = 0; a < A; ++a) {
    anchor_role[a] = 0;
    run_off[a] = pos;
    if (eng[a] == 5) {
      std::memset(&anchor_aln[a], 0, sizeof(phx_alignment));
      anchor_aln[a].status = PHX_INVALID_INPUT;
      continue;
    }
    const int e = eng[a];
    phx_alignment x = ealn[ebase[e] + slot[a]];
    if (x.status == PHX_OK) {
      x.target_start += (int64_t)tpos[a];
      x.target_end += (int64_t)tpos[a];
    }
    anchor_aln[a] = x;
    const uint64_t r0 = eoff[e][slot[a]], r1 = eoff[e][slot[a] + 1];
    if (n_eng > 1 && r1 > r0) std::memcpy(runs + pos, eruns[e].data() + r0, (r1 - r0) * 4);
    pos += r1 - r0;
  }
  run_off[A] = pos;
  // score_alignments per read over its successful anchors in order
  // (lib.rs:170-199): primary = the first smallest edit distance; an
  // alternative is any other one not equal to the primary whose normalised
  // score ratio is >= 0.95.
  for (uint32_t r = 0; r < n_reads; ++r) {
    phx_read_extension& o = out[r];
    o.first_anchor = first[r];
    o.n_anchors = (uint32_t)(first[r + 1] - first[r]);
    o.primary = -1;
    o.n_alternatives = 0;
    int64_t p = -1;
    for (uint64_t a = first[r]; a < first[r + 1]; ++a)
      if (anchor_aln[a].status == PHX_OK && (p < 0 || anchor_aln[a].edit_distance < anchor_aln[p].edit_distance))
        p = (int64_t)a;
    o.status = p < 0 ? PHX_ALIGNMENT_FAILED : PHX_OK;
    if (p < 0) continue;
    o.primary = (int32_t)((uint64_t)p - first[r]);
    anchor_role[p] = 2;
    const double ql = (double)reads[r].query_len;
    const double pn = 1.0 - (double)anchor_aln[p].edit_distance / ql;
    for (uint64_t a = first[r]; a < first[r + 1]; ++a) {
      if ((int64_t)a == p || anchor_aln[a].status != PHX_OK) continue;
      anchor_role[a] = 1;
      if (same_alignment(anchor_aln[a], runs + run_off[a], run_off[a + 1] - run_off[a], anchor_aln[p],
                         runs + run_off[p], run_off[p + 1] - run_off[p]))
        continue;
      const double an = 1.0 - (double)anchor_aln[a].edit_distance / ql;
      if (an / pn >= 0.95) {
        anchor_role[a] = 3;
        ++o.n_alternatives;
      }
    }
  }
  return PHX_OK;
}

// The orientations align_read extends (executor.rs:632-698): reverse chains
// count only for DNA (protein has no reverse strand and its reverse query is
// empty); one orientation alone when only it has chains, otherwise both.
// Reverse complements live in one flat buffer.
struct Orientations {
  std::vector<phx_read_chains> entry;  // forward and / or reverse extension inputs
  std::vector<uint8_t> rc;             // reverse complements, read r's at rc_off[r]
  std::vector<uint64_t> rc_off;
  std::vector<int32_t> fwd, rev;       // entry of each orientation, -1 when not extended
};

bool orientations(const phx_read_strands* reads, uint32_t n, int32_t is_dna, Orientations* o) {
  o->fwd.assign(n, -1);
  o->rev.assign(n, -1);
  o->entry.clear();
  o->rc_off.assign(n + 1, 0);
  for (uint32_t r = 0; r < n; ++r) {
    const phx_read_strands& rd = reads[r];
    if ((!rd.query && rd.query_len) || (rd.n_fwd_chains && !rd.fwd_chain_target_start) ||
        (is_dna && rd.n_rev_chains && !rd.rev_chain_target_start))
      return false;
    const uint32_t nf = rd.n_fwd_chains, nr = is_dna ? rd.n_rev_chains : 0;
    const bool do_rev = !(nf != 0 && nr == 0);
    o->rc_off[r + 1] = o->rc_off[r] + (do_rev && is_dna ? rd.query_len : 0);
  }
  o->rc.resize(o->rc_off[n]);
  parallel_for(n, [&](size_t r) {
    const uint64_t len = o->rc_off[r + 1] - o->rc_off[r];
    if (len) phx_reverse_complement(reads[r].query, len, o->rc.data() + o->rc_off[r]);
  });
  for (uint32_t r = 0; r < n; ++r) {
    const phx_read_strands& rd = reads[r];
    const uint32_t nf = rd.n_fwd_chains, nr = is_dna ? rd.n_rev_chains : 0;
    const bool do_fwd = !(nf == 0 && nr != 0), do_rev = !(nf != 0 && nr == 0);
    if (do_fwd) {
      o->fwd[r] = (int32_t)o->entry.size();
      o->entry.push_back(phx_read_chains{rd.query, rd.query_len, rd.fwd_chain_target_start, nf});
    }
    if (do_rev) {
      const uint64_t len = o->rc_off[r + 1] - o->rc_off[r];
      o->rev[r] = (int32_t)o->entry.size();
      o->entry.push_back(phx_read_chains{len ? o->rc.data() + o->rc_off[r] : nullptr, len,
                                         nr ? rd.rev_chain_target_start : nullptr, nr});
    }
  }
  return true;
}

int align_reads_core(phx_ctx* ctx, int32_t strategy, int32_t gap_model, int32_t is_dna, const TargetSrc& T,
                     const phx_read_strands* reads, uint32_t n_reads, phx_read_placement* placement,
                     phx_read_extension* ext, uint64_t ext_cap, phx_alignment* anchor_aln,
                     uint8_t* anchor_role, uint64_t anchors_cap, uint32_t* runs, uint64_t runs_cap,
                     uint64_t* run_off) {
  if (!ctx || !reads || !placement || !ext) return PHX_E_ARG;
  Orientations o;
  if (!orientations(reads, n_reads, is_dna, &o)) return PHX_E_ARG;
  if (ext_cap < o.entry.size()) return PHX_E_CAPACITY;
  // Every orientation of every read in one batched extend_chains.
  const int rc = extend_core(ctx, strategy, gap_model, T, o.entry.data(), (uint32_t)o.entry.size(), ext,
                             anchor_aln, anchor_role, anchors_cap, runs, runs_cap, run_off);
  if (rc != PHX_OK) return rc;
  // The better orientation, ties to forward (executor.rs:700-711); then the
  // Fast divergence cutoff (:716-723) and the outcome (:727-733).
  for (uint32_t r = 0; r < n_reads; ++r) {
    phx_read_placement& pl = placement[r];
    pl.entry_fwd = o.fwd[r];
    pl.entry_rev = o.rev[r];
    auto primary_ed = [&](int32_t e) {
      const phx_read_extension& x = ext[e];
      return anchor_aln[x.first_anchor + (uint64_t)x.primary].edit_distance;
    };
    const bool f = o.fwd[r] >= 0 && ext[o.fwd[r]].status == PHX_OK;
    const bool v = o.rev[r] >= 0 && ext[o.rev[r]].status == PHX_OK;
    if (f && v && primary_ed(o.rev[r]) < primary_ed(o.fwd[r])) pl.strand = PHX_STRAND_REVERSE;
    else if (f) pl.strand = PHX_STRAND_FORWARD;
    else if (v) pl.strand = PHX_STRAND_REVERSE;
    else pl.strand = -1;
    if (pl.strand < 0) {
      pl.entry = -1;
      pl.outcome = PHX_OUTCOME_NO_ALIGNMENT;
      pl.primary_score = 0.0;
      continue;
    }
    pl.entry = pl.strand == PHX_STRAND_FORWARD ? o.fwd[r] : o.rev[r];
    const double ql = (double)std::max<uint64_t>(reads[r].query_len, 1);
    const double ed = (double)primary_ed(pl.entry);
    pl.primary_score = 1.0 - ed / ql;
    if (strategy == PHX_STRATEGY_FAST && ed / ql > 0.20) pl.outcome = PHX_OUTCOME_FAST_CUTOFF;
    else if (pl.primary_score >= 0.95) pl.outcome = PHX_OUTCOME_PLACED;
    else if (reads[r].had_seeds) pl.outcome = PHX_OUTCOME_BELOW_THRESHOLD;
    else pl.outcome = PHX_OUTCOME_NO_SEED;
  }
  return PHX_OK;
}

bool valid_modes(int32_t strategy, int32_t gap_model) {
  return strategy >= PHX_STRATEGY_FAST && strategy <= PHX_STRATEGY_SENSITIVE &&
         gap_model >= PHX_GAP_AFFINE && gap_model <= PHX_GAP_LINEAR;
}

}  // namespace

extern "C" {

void phx_reverse_complement(const uint8_t* seq, uint64_t n, uint8_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = complement(seq[n - 1 - i]);
}

phx_reference* phx_reference_create(phx_ctx* ctx, const uint8_t* target, uint64_t len) {
  if (!ctx || (!target && len)) return nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) {
    set_error(ctx, "hipSetDevice failed", PHX_E_DEVICE);
    return nullptr;
  }
  phx_reference* R = new phx_reference;
  R->ctx = ctx;
  R->len = len;
  R->host.assign(target, target + len);
  const uint64_t nb = len / 64 + 1;
  R->bad_before.assign(nb + 1, 0);
  for (uint64_t b = 0; b < nb; ++b) {
    uint32_t c = 0;
    for (uint64_t i = b * 64; i < std::min(len, b * 64 + 64); ++i) c += acgt_table()[target[i]] == 0xFF;
    R->bad_before[b + 1] = R->bad_before[b] + c;
  }
  // 2-bit words (a non-ACGT byte packs as A: such windows use the 8-bit
  // copy) and 8-bit words, each followed by pad words.
  const uint64_t w2 = words_for(len, false) + 1, w8 = words_for(len, true) + 1;
  std::vector<uint64_t> p2(w2, 0), p8(w8, 0);
  const uint8_t* code = acgt_table();
  parallel_for((size_t)(w2 - 2 + 1), [&](size_t w) {
    uint64_t v = 0;
    for (uint64_t i = w * 32; i < std::min(len, w * 32 + 32); ++i) {
      const uint8_t c = code[target[i]];
      v |= (uint64_t)(c == 0xFF ? 0 : c) << ((i & 31) * 2);
    }
    p2[w] = v;
  });
  if (len) std::memcpy(p8.data(), target, len);
  if (hipMalloc(&R->d2, w2 * 8) != hipSuccess || hipMalloc(&R->d8, w8 * 8) != hipSuccess ||
      hipMemcpy(R->d2, p2.data(), w2 * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(R->d8, p8.data(), w8 * 8, hipMemcpyHostToDevice) != hipSuccess) {
    set_error(ctx, "phx_reference_create: device allocation or upload failed", PHX_E_NOMEM);
    if (R->d2) (void)hipFree(R->d2);
    if (R->d8) (void)hipFree(R->d8);
    delete R;
    return nullptr;
  }
  return R;
}

void phx_reference_destroy(phx_reference* R) {
  if (!R) return;
  (void)hipStreamSynchronize(R->ctx->stream);
  if (R->d2) (void)hipFree(R->d2);
  if (R->d8) (void)hipFree(R->d8);
  delete R;
}

int phx_align_reads_bounds(int32_t strategy, int32_t is_dna, uint64_t target_len,
                           const phx_read_strands* reads, uint32_t n_reads, uint64_t* entries,
                           uint64_t* anchors, uint64_t* runs) {
  if (!reads || !entries || !anchors || !runs) return PHX_E_ARG;
  Orientations o;
  if (!orientations(reads, n_reads, is_dna, &o)) return PHX_E_ARG;
  *entries = o.entry.size();
  return phx_extend_chains_bounds(strategy, target_len, o.entry.data(), (uint32_t)o.entry.size(),
                                  anchors, runs);
}

int phx_align_reads_batch(phx_ctx* ctx, int32_t strategy, int32_t gap_model, int32_t is_dna,
                          const uint8_t* target, uint64_t target_len, const phx_read_strands* reads,
                          uint32_t n_reads, phx_read_placement* placement, phx_read_extension* ext,
                          uint64_t ext_cap, phx_alignment* anchor_aln, uint8_t* anchor_role,
                          uint64_t anchors_cap, uint32_t* runs, uint64_t runs_cap, uint64_t* run_off) {
  if ((!target && target_len) || !valid_modes(strategy, gap_model)) return PHX_E_ARG;
  return align_reads_core(ctx, strategy, gap_model, is_dna, TargetSrc{nullptr, target, target_len}, reads,
                          n_reads, placement, ext, ext_cap, anchor_aln, anchor_role, anchors_cap, runs,
                          runs_cap, run_off);
}

int phx_align_reads_batch_ref(phx_ctx* ctx, int32_t strategy, int32_t gap_model, int32_t is_dna,
                              const phx_reference* ref, const phx_read_strands* reads, uint32_t n_reads,
                              phx_read_placement* placement, phx_read_extension* ext, uint64_t ext_cap,
                              phx_alignment* anchor_aln, uint8_t* anchor_role, uint64_t anchors_cap,
                              uint32_t* runs, uint64_t runs_cap, uint64_t* run_off) {
  if (!ref || ref->ctx != ctx || !valid_modes(strategy, gap_model)) return PHX_E_ARG;
  return align_reads_core(ctx, strategy, gap_model, is_dna, TargetSrc{ref, nullptr, ref->len}, reads, n_reads,
                          placement, ext, ext_cap, anchor_aln, anchor_role, anchors_cap, runs, runs_cap,
                          run_off);
}

uint32_t phx_chain_cap(int32_t strategy) { return cap_of(strategy); }

int phx_extend_chains_bounds(int32_t strategy, uint64_t target_len, const phx_read_chains* reads,
                             uint32_t n_reads, uint64_t* anchors, uint64_t* runs) {
  if (!reads || !anchors || !runs || strategy < PHX_STRATEGY_FAST || strategy > PHX_STRATEGY_SENSITIVE)
    return PHX_E_ARG;
  uint64_t A = 0, R = 0;
  for (uint32_t r = 0; r < n_reads; ++r) {
    const uint32_t nc = reads[r].n_chains;
    if (nc && !reads[r].chain_target_start) return PHX_E_ARG;
    const uint32_t na = nc == 0 ? 1 : std::min(nc, cap_of(strategy));
    for (uint32_t i = 0; i < na; ++i) {
      const uint64_t tp = nc == 0 ? 0 : reads[r].chain_target_start[i];
      const uint64_t wl = tp > target_len ? 0 : std::min(2 * reads[r].query_len, target_len - tp);
      R += reads[r].query_len + wl + 1;  // a run per base at most, plus one
    }
    A += na;
  }
  *anchors = A;
  *runs = R;
  return PHX_OK;
}

int phx_score_alignments(const phx_alignment* aln, const uint32_t* runs, const uint64_t* run_off,
                         uint32_t n, uint64_t query_len, int32_t* primary, uint8_t* is_alternative) {
  if (!aln || !run_off || !primary || !is_alternative || n == 0) return PHX_E_ARG;
  uint32_t p = 0;
  for (uint32_t i = 1; i < n; ++i)
    if (aln[i].edit_distance < aln[p].edit_distance) p = i;  // first minimum
  *primary = (int32_t)p;
  const double ql = (double)query_len;
  const double pn = 1.0 - (double)aln[p].edit_distance / ql;
  for (uint32_t i = 0; i < n; ++i) {
    is_alternative[i] = 0;
    const uint64_t ni = run_off[i + 1] - run_off[i], np = run_off[p + 1] - run_off[p];
    if (same_alignment(aln[i], runs ? runs + run_off[i] : nullptr, ni, aln[p],
                       runs ? runs + run_off[p] : nullptr, np))
      continue;  // the primary itself and any copy of it
    const double an = 1.0 - (double)aln[i].edit_distance / ql;
    if (an / pn >= 0.95) is_alternative[i] = 1;
  }
  return PHX_OK;
}

int phx_extend_chains_batch(phx_ctx* ctx, int32_t strategy, int32_t gap_model,
                            const uint8_t* target, uint64_t target_len,
                            const phx_read_chains* reads, uint32_t n_reads,
                            phx_read_extension* out, phx_alignment* anchor_aln,
                            uint8_t* anchor_role, uint64_t anchors_cap, uint32_t* runs,
                            uint64_t runs_cap, uint64_t* run_off) {
  if (!ctx || !reads || !out || !run_off || (!target && target_len) || !valid_modes(strategy, gap_model))
    return PHX_E_ARG;
  return extend_core(ctx, strategy, gap_model, TargetSrc{nullptr, target, target_len}, reads, n_reads, out,
                     anchor_aln, anchor_role, anchors_cap, runs, runs_cap, run_off);
}

int phx_extend_chains_batch_ref(phx_ctx* ctx, int32_t strategy, int32_t gap_model, const phx_reference* ref,
                                const phx_read_chains* reads, uint32_t n_reads, phx_read_extension* out,
                                phx_alignment* anchor_aln, uint8_t* anchor_role, uint64_t anchors_cap,
                                uint32_t* runs, uint64_t runs_cap, uint64_t* run_off) {
  if (!ctx || !ref || ref->ctx != ctx || !reads || !out || !run_off || !valid_modes(strategy, gap_model))
    return PHX_E_ARG;
  return extend_core(ctx, strategy, gap_model, TargetSrc{ref, nullptr, ref->len}, reads, n_reads, out,
                     anchor_aln, anchor_role, anchors_cap, runs, runs_cap, run_off);
}

}  // extern "C"
