// phx_executor.hip -- batched form of phraya-align's per-read extension step
// (extend_chains, phraya-align/src/executor.rs:528-567) and its scoring
// (score_alignments, phraya-align/src/lib.rs:170-199).
//
// Upstream extends one read's anchors one call at a time on a rayon thread.
// Here every anchor of every read in a batch becomes one pair of one device
// batch per engine, so a 4096-read batch of the executor is one or two
// launches; the per-read selection of the primary and the alternatives runs
// on the host afterwards, with upstream's rules:
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
#include <vector>

#include "phraya_hip.h"

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

// The orientations align_read extends (executor.rs:632-698): reverse chains
// count only for DNA (protein has no reverse strand and its reverse query is
// empty); one orientation alone when only it has chains, otherwise both.
struct Orientations {
  std::vector<phx_read_chains> entry;     // forward and / or reverse extension inputs
  std::vector<std::vector<uint8_t>> rc;   // reverse complements, one per read
  std::vector<int32_t> fwd, rev;          // entry of each orientation, -1 when not extended
};

bool orientations(const phx_read_strands* reads, uint32_t n, int32_t is_dna, Orientations* o) {
  o->rc.assign(n, {});
  o->fwd.assign(n, -1);
  o->rev.assign(n, -1);
  o->entry.clear();
  for (uint32_t r = 0; r < n; ++r) {
    const phx_read_strands& rd = reads[r];
    if ((!rd.query && rd.query_len) || (rd.n_fwd_chains && !rd.fwd_chain_target_start) ||
        (is_dna && rd.n_rev_chains && !rd.rev_chain_target_start))
      return false;
    const uint32_t nf = rd.n_fwd_chains, nr = is_dna ? rd.n_rev_chains : 0;
    const bool do_fwd = !(nf == 0 && nr != 0), do_rev = !(nf != 0 && nr == 0);
    if (do_fwd) {
      o->fwd[r] = (int32_t)o->entry.size();
      o->entry.push_back(phx_read_chains{rd.query, rd.query_len, rd.fwd_chain_target_start, nf});
    }
    if (do_rev) {
      if (is_dna) {
        o->rc[r].resize(rd.query_len);
        phx_reverse_complement(rd.query, rd.query_len, o->rc[r].data());
      }
      o->rev[r] = (int32_t)o->entry.size();
      o->entry.push_back(phx_read_chains{o->rc[r].empty() ? nullptr : o->rc[r].data(),
                                         (uint64_t)o->rc[r].size(),
                                         nr ? rd.rev_chain_target_start : nullptr, nr});
    }
  }
  return true;
}

}  // namespace

extern "C" {

void phx_reverse_complement(const uint8_t* seq, uint64_t n, uint8_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = complement(seq[n - 1 - i]);
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
  if (!ctx || !reads || !placement || !ext) return PHX_E_ARG;
  Orientations o;
  if (!orientations(reads, n_reads, is_dna, &o)) return PHX_E_ARG;
  if (ext_cap < o.entry.size()) return PHX_E_CAPACITY;
  // Every orientation of every read in one batched extend_chains.
  const int rc = phx_extend_chains_batch(ctx, strategy, gap_model, target, target_len, o.entry.data(),
                                         (uint32_t)o.entry.size(), ext, anchor_aln, anchor_role,
                                         anchors_cap, runs, runs_cap, run_off);
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
  if (!ctx || !reads || !out || !run_off || (!target && target_len) ||
      strategy < PHX_STRATEGY_FAST || strategy > PHX_STRATEGY_SENSITIVE ||
      gap_model < PHX_GAP_AFFINE || gap_model > PHX_GAP_LINEAR)
    return PHX_E_ARG;
  // Anchors, read by read (executor.rs:347-362).
  std::vector<uint64_t> first(n_reads + 1, 0);
  for (uint32_t r = 0; r < n_reads; ++r) {
    if (!reads[r].query && reads[r].query_len) return PHX_E_ARG;
    const uint32_t nc = reads[r].n_chains;
    if (nc && !reads[r].chain_target_start) return PHX_E_ARG;
    first[r + 1] = first[r] + (nc == 0 ? 1 : std::min(nc, cap_of(strategy)));
  }
  const uint64_t A = first[n_reads];
  uint64_t need_a = 0, need_r = 0;
  phx_extend_chains_bounds(strategy, target_len, reads, n_reads, &need_a, &need_r);
  if (!anchor_aln || !anchor_role || anchors_cap < A || !runs || runs_cap < need_r)
    return PHX_E_CAPACITY;
  std::vector<uint64_t> tpos(A);
  for (uint32_t r = 0; r < n_reads; ++r)
    for (uint64_t a = first[r]; a < first[r + 1]; ++a)
      tpos[a] = reads[r].n_chains == 0 ? 0 : reads[r].chain_target_start[a - first[r]];
  // One device batch per engine: query vs target[tpos .. window end], seed (0, 0).
  std::vector<uint32_t> read_of(A);
  for (uint32_t r = 0; r < n_reads; ++r)
    for (uint64_t a = first[r]; a < first[r + 1]; ++a) read_of[a] = r;
  std::vector<std::vector<uint64_t>> members(5);
  for (uint64_t a = 0; a < A; ++a) {
    anchor_role[a] = 0;
    std::memset(&anchor_aln[a], 0, sizeof(phx_alignment));
    if (tpos[a] > target_len) {  // the seed lies past the window (lib.rs:145-149)
      anchor_aln[a].status = PHX_INVALID_INPUT;
      continue;
    }
    members[engine_of(strategy, gap_model, reads[read_of[a]].query_len)].push_back(a);
  }
  std::vector<std::vector<uint32_t>> a_runs(A);
  for (int eng = 0; eng < 5; ++eng) {
    const auto& m = members[eng];
    if (m.empty()) continue;
    std::vector<uint8_t> qs, ts;
    std::vector<uint64_t> qo(1, 0), to(1, 0);
    for (uint64_t a : m) {
      const phx_read_chains& rd = reads[read_of[a]];
      const uint64_t wend = std::min(tpos[a] + 2 * rd.query_len, target_len);
      qs.insert(qs.end(), rd.query, rd.query + rd.query_len);
      ts.insert(ts.end(), target + tpos[a], target + wend);
      qo.push_back(qs.size());
      to.push_back(ts.size());
    }
    phx_batch b{};
    b.n = (uint32_t)m.size();
    b.engine = eng;
    b.costs = phx_costs{2, 4, 1};  // AffineCosts::default()
    b.cap = -1;
    b.q_seq = qs.data();
    b.q_off = qo.data();
    b.t_seq = ts.data();
    b.t_off = to.data();
    std::vector<phx_alignment> aln(m.size());
    std::vector<uint64_t> off(m.size() + 1);
    std::vector<uint32_t> rr(std::max<uint64_t>(1, phx_cigar_bound(&b)));
    phx_results res{aln.data(), rr.data(), rr.size(), off.data()};
    const int rc = phx_align_batch(ctx, &b, &res);
    if (rc != PHX_OK) return rc;
    for (size_t i = 0; i < m.size(); ++i) {
      phx_alignment x = aln[i];
      if (x.status == PHX_OK) {  // positions from the window start
        x.target_start += (int64_t)tpos[m[i]];
        x.target_end += (int64_t)tpos[m[i]];
        a_runs[m[i]].assign(rr.begin() + off[i], rr.begin() + off[i + 1]);
      }
      anchor_aln[m[i]] = x;
    }
  }
  // Runs layout, then the per-read selection.
  uint64_t pos = 0;
  for (uint64_t a = 0; a < A; ++a) {
    run_off[a] = pos;
    pos += a_runs[a].size();
  }
  run_off[A] = pos;
  for (uint64_t a = 0; a < A; ++a)
    if (!a_runs[a].empty()) std::memcpy(runs + run_off[a], a_runs[a].data(), a_runs[a].size() * 4);
  for (uint32_t r = 0; r < n_reads; ++r) {
    phx_read_extension& o = out[r];
    o.first_anchor = first[r];
    o.n_anchors = (uint32_t)(first[r + 1] - first[r]);
    o.primary = -1;
    o.n_alternatives = 0;
    // Successful anchors in order: the vector score_alignments receives.
    std::vector<uint64_t> ok;
    for (uint64_t a = first[r]; a < first[r + 1]; ++a)
      if (anchor_aln[a].status == PHX_OK) ok.push_back(a);
    o.status = ok.empty() ? PHX_ALIGNMENT_FAILED : PHX_OK;
    if (ok.empty()) continue;
    std::vector<phx_alignment> sel(ok.size());
    std::vector<uint64_t> so(ok.size() + 1, 0);
    std::vector<uint32_t> sr;
    for (size_t i = 0; i < ok.size(); ++i) {
      sel[i] = anchor_aln[ok[i]];
      sr.insert(sr.end(), a_runs[ok[i]].begin(), a_runs[ok[i]].end());
      so[i + 1] = sr.size();
    }
    int32_t p = 0;
    std::vector<uint8_t> alt(ok.size());
    phx_score_alignments(sel.data(), sr.data(), so.data(), (uint32_t)ok.size(), reads[r].query_len,
                         &p, alt.data());
    o.primary = (int32_t)(ok[p] - first[r]);
    anchor_role[ok[p]] = 2;
    for (size_t i = 0; i < ok.size(); ++i) {
      if ((int32_t)i == p) continue;
      anchor_role[ok[i]] = alt[i] ? 3 : 1;
      o.n_alternatives += alt[i];
    }
  }
  return PHX_OK;
}

}  // extern "C"
