// phx_chain.hip -- pre-extension step: seeding and chaining for a batch.
//
// seed_and_chain (phraya-align/src/executor.rs:511-519) for many query
// sketches against one target sketch:
//  * build_minimizer_index (seeding.rs:29-35): value -> target positions in
//    sketch order.  Here an open-addressing table of the distinct values
//    (16-byte slots: value, first position, count) over the positions
//    grouped by value in sketch order, built on the host once per target
//    like upstream's TargetContext, so a lookup is usually one load;
//  * find_seeds_indexed_capped (seeding.rs:56-74): for each query minimizer
//    in sketch order, every target position of its value unless the value
//    occurs more than max_occ times; then a stable sort by query position;
//  * chain_seeds (chaining.rs:131-253): stable sort by (target, query)
//    position, the backward-scan DP with chain_score (:96-117), the
//    dominated / skip bookkeeping, endpoints by score (stable), backtrack
//    without reuse, each chain's seeds stable-sorted by query position,
//    chains by score (stable).
// Every sort here sorts by a key made unique by the element's earlier
// position, so the heap sorts give exactly the stable sorts' orders.  The
// gap penalty round(gap_penalty * dd + 0.5 * log2(dd + 1)) takes only
// max_band values; the host computes them once with the same double
// arithmetic as upstream and the kernel reads the table.
//
// Plan creation counts each query's seeds once (a wave per query) to lay
// out the working arrays; a run is one launch, a wave per query: lookups,
// seed count, seeds, sorts, DP and chains, in LDS up to 128 seeds, else in
// global memory at the query's offset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "phraya_hip.h"
#include "phx_host.h"

namespace {

constexpr uint32_t LDS_SEEDS = 128;   // seeds a query may have for the LDS path
constexpr uint32_t CW = 2;            // waves (queries) per workgroup of the chain kernel

// Open-addressing slot of the index: a distinct value with its positions at
// ix_pos[start .. start + count); count == 0 marks an empty slot.
struct IxSlot {
  uint64_t value;
  uint32_t start, count;
};

__host__ __device__ inline uint64_t mix64(uint64_t z) {  // SplitMix64 finalizer
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct ChainDev {
  const IxSlot* slots;  // linear probing, at most half full
  uint64_t mask;
  const uint32_t* ix_pos;
  uint64_t max_occ;
  const uint64_t* q_val;
  const uint32_t* q_pos;
  const uint64_t* q_off;
  uint32_t nq;
  int64_t seed_len, max_band, max_dist_x, min_chain_score;
  uint64_t max_iter, max_skip;
  const int64_t* pen;    // pen[dd], dd in [1, max_band]
  const int32_t* pen32;  // the same, for the 32-bit LDS path
  int32_t small;         // parameters and query positions fit the 32-bit LDS path
};

struct Work {  // per seed, at the query's seed offset (queries above LDS_SEEDS seeds)
  uint32_t *gq, *gt;   // seeds in generation order
  uint64_t* gv;
  uint32_t* ord;       // permutations
  uint32_t *sq, *st;   // seeds by (target, query) position
  uint64_t* sv;
  int64_t* f;
  int32_t* pred;
  uint32_t* dom;
  uint8_t* used;
  phx_chain* tmp;      // chains in backtrack order
};

// The value's seeds (start, count) if the target holds it at most max_occ
// times (find_seeds_indexed_capped's mask); count 0 otherwise.
__device__ inline void lookup(const ChainDev& c, uint64_t v, uint32_t* start, uint32_t* count) {
  uint64_t h = mix64(v) & c.mask;
  *start = 0;
  *count = 0;
  while (true) {
    const IxSlot s = c.slots[h];
    if (s.count == 0) return;
    if (s.value == v) {
      if (s.count <= c.max_occ) {
        *start = s.start;
        *count = s.count;
      }
      return;
    }
    h = (h + 1) & c.mask;
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint64_t wave_sum(uint64_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += (uint64_t)__shfl_xor((unsigned long long)x, d, 64);
  return x;
}

__device__ inline uint32_t wave_excl_scan32(uint32_t x, uint32_t lane, uint32_t* total) {
  uint32_t s = x;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(s, d, 64);
    if (lane >= d) s += y;
  }
  *total = __shfl(s, 63, 64);
  return s - x;
}

// Lane shifts within each row of 16 lanes (DPP row_shr:D): a lane whose
// source lies outside its row gets `old`.
template <int D>
__device__ __forceinline__ int32_t row_shr(int32_t old, int32_t x) {
  return __builtin_amdgcn_update_dpp(old, x, 0x110 + D, 0xf, 0xf, false);
}

// The DP's two scans over the lanes [0, span) of row 0, span <= 16, in DPP
// steps: an inclusive max, and the inclusive composition of the maps
// x -> max(x + A, B) (earlier lanes applied first; identity (0, neg)).
template <int D>
__device__ __forceinline__ void max_step(int32_t& v) {
  v = max(v, row_shr<D>(INT32_MIN, v));
}

template <int D>
__device__ __forceinline__ void map_step(int32_t& A, int32_t& B, int32_t neg) {
  const int32_t pa = row_shr<D>(0, A), pb = row_shr<D>(neg, B);
  B = max(pb + A, B);
  A = pa + A;
}

__device__ __forceinline__ void max_scan_row(int32_t& v, uint32_t span) {
  if (span > 1) max_step<1>(v);
  if (span > 2) max_step<2>(v);
  if (span > 4) max_step<4>(v);
  if (span > 8) max_step<8>(v);
}

__device__ __forceinline__ void map_scan_row(int32_t& A, int32_t& B, uint32_t span, int32_t neg) {
  if (span > 1) map_step<1>(A, B, neg);
  if (span > 2) map_step<2>(A, B, neg);
  if (span > 4) map_step<4>(A, B, neg);
  if (span > 8) map_step<8>(A, B, neg);
}

// chain_score (chaining.rs:96-117) with the gap penalty from the table;
// INT64_MIN when j cannot precede i.
__device__ inline int64_t link_score(const ChainDev& c, int64_t qi, int64_t ti, int64_t qj, int64_t tj) {
  const int64_t dq = qi - qj, dr = ti - tj;
  if (dq <= 0 || dr <= 0 || dr > c.max_dist_x + c.seed_len) return INT64_MIN;
  const int64_t dd = dr > dq ? dr - dq : dq - dr;
  if (dd > c.max_band) return INT64_MIN;
  return min(c.seed_len, min(dr, dq)) - (dd > 0 ? c.pen[dd] : 0);
}

template <class Less>
__device__ void sift_down(uint32_t* a, uint32_t root, uint32_t n, const Less& less) {
  while (2 * root + 1 < n) {
    uint32_t ch = 2 * root + 1;
    if (ch + 1 < n && less(a[ch], a[ch + 1])) ++ch;
    if (!less(a[root], a[ch])) return;
    const uint32_t x = a[root];
    a[root] = a[ch];
    a[ch] = x;
    root = ch;
  }
}

// Ascending by a strict total order (keys made unique by the caller).
template <class Less>
__device__ void heap_sort(uint32_t* a, uint32_t n, const Less& less) {
  for (uint32_t s = n / 2; s-- > 0;) sift_down(a, s, n, less);
  for (uint32_t e = n; e-- > 1;) {
    const uint32_t x = a[0];
    a[0] = a[e];
    a[e] = x;
    sift_down(a, 0, e, less);
  }
}

// A wave per query: its seed count (lanes over its minimizers); run once at
// plan creation to lay out the working arrays.
__global__ __launch_bounds__(256) void seed_count_kernel(ChainDev c, uint64_t* n_seeds) {
  const uint32_t q = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (q >= c.nq) return;  // whole waves
  uint64_t n = 0;
  for (uint64_t m = c.q_off[q] + lane; m < c.q_off[q + 1]; m += 64) {
    uint32_t st, cnt;
    lookup(c, c.q_val[m], &st, &cnt);
    n += cnt;
  }
  n = wave_sum(n);
  if (lane == 0) n_seeds[q] = n;
}

// One query on one thread with its working arrays in global memory: the
// path for queries with more than LDS_SEEDS seeds.
__device__ void chain_serial(const ChainDev& c, uint32_t q, uint64_t base, uint32_t n, const Work& w,
                             phx_chain* out, phx_seed* cs, uint32_t* n_chains) {
  uint32_t* gq = w.gq + base;
  uint32_t* gt = w.gt + base;
  uint64_t* gv = w.gv + base;
  uint32_t* ord = w.ord + base;
  uint32_t g = 0;  // seeds in find_seeds' generation order
  for (uint64_t m = c.q_off[q]; m < c.q_off[q + 1]; ++m) {
    uint32_t lo, cnt;
    lookup(c, c.q_val[m], &lo, &cnt);
    for (uint32_t k = 0; k < cnt; ++k) {
      gq[g] = c.q_pos[m];
      gt[g] = c.ix_pos[lo + k];
      gv[g] = c.q_val[m];
      ord[g] = g;
      ++g;
    }
  }
  // find_seeds' stable sort by query position followed by chain_seeds'
  // stable sort by (target, query): one sort by (target, query, generation).
  heap_sort(ord, n, [&](uint32_t a, uint32_t b) {
    if (gt[a] != gt[b]) return gt[a] < gt[b];
    if (gq[a] != gq[b]) return gq[a] < gq[b];
    return a < b;
  });
  uint32_t* sq = w.sq + base;
  uint32_t* st = w.st + base;
  uint64_t* sv = w.sv + base;
  int64_t* f = w.f + base;
  int32_t* pred = w.pred + base;
  uint32_t* dom = w.dom + base;
  uint8_t* used = w.used + base;
  for (uint32_t k = 0; k < n; ++k) {
    sq[k] = gq[ord[k]];
    st[k] = gt[ord[k]];
    sv[k] = gv[ord[k]];
    dom[k] = 0;
    used[k] = 0;
  }
  const int64_t reach = c.max_dist_x + c.seed_len;
  for (uint32_t i = 0; i < n; ++i) {
    int64_t best = c.seed_len;
    int32_t bp = -1;
    uint64_t skip = 0;
    const uint32_t first = (uint64_t)i > c.max_iter ? (uint32_t)(i - c.max_iter) : 0u;  // saturating
    const int64_t qi = sq[i], ti = st[i];
    for (uint32_t j = i; j-- > first;) {
      if (ti - (int64_t)st[j] >= reach) break;
      const int64_t sc = link_score(c, qi, ti, sq[j], st[j]);
      if (sc == INT64_MIN) continue;
      const int64_t total = sc + f[j];
      if (total > best) {
        best = total;
        bp = (int32_t)j;
        if (skip > 0) --skip;
      } else {
        dom[j] += 1;
        if (dom[j] > c.max_skip && ++skip > c.max_skip) break;
      }
    }
    f[i] = best;
    pred[i] = bp;
  }
  for (uint32_t k = 0; k < n; ++k) ord[k] = k;  // endpoints: key (-f, index)
  heap_sort(ord, n, [&](uint32_t a, uint32_t b) { return f[a] != f[b] ? f[a] > f[b] : a < b; });
  phx_chain* tmp = w.tmp + base;
  phx_seed* out_seeds = cs + base;
  uint32_t nc = 0, ns = 0;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t e = ord[k];
    if (used[e] || f[e] < c.min_chain_score) continue;
    const uint32_t start = ns;
    for (int32_t cur = (int32_t)e; cur >= 0 && !used[cur]; cur = pred[cur]) {
      used[cur] = 1;
      out_seeds[ns++] = phx_seed{sq[cur], st[cur], sv[cur]};
    }
    for (uint32_t a = start, b = ns - 1; a < b; ++a, --b) {  // back to target order
      const phx_seed x = out_seeds[a];
      out_seeds[a] = out_seeds[b];
      out_seeds[b] = x;
    }
    for (uint32_t a = start + 1; a < ns; ++a) {  // stable insertion sort by query position
      const phx_seed x = out_seeds[a];
      uint32_t b = a;
      while (b > start && out_seeds[b - 1].query_pos > x.query_pos) {
        out_seeds[b] = out_seeds[b - 1];
        --b;
      }
      out_seeds[b] = x;
    }
    const int64_t ts = (int64_t)out_seeds[start].target_pos - (int64_t)out_seeds[start].query_pos;
    tmp[nc++] = phx_chain{f[e], (uint64_t)(ts > 0 ? ts : 0), base + start, ns - start, 0};
  }
  for (uint32_t k = 0; k < nc; ++k) ord[k] = k;  // chains: key (-score, backtrack order)
  heap_sort(ord, nc, [&](uint32_t a, uint32_t b) {
    return tmp[a].score != tmp[b].score ? tmp[a].score > tmp[b].score : a < b;
  });
  for (uint32_t k = 0; k < nc; ++k) out[base + k] = tmp[ord[k]];
  n_chains[q] = nc;
}

// A query's working arrays in LDS (queries of up to LDS_SEEDS seeds).
struct WaveLds {
  uint32_t gq[LDS_SEEDS], gt[LDS_SEEDS], gm[LDS_SEEDS];  // generation order; later the chained seeds
  uint32_t sq[LDS_SEEDS], st[LDS_SEEDS], sm[LDS_SEEDS];  // by (target, query, generation)
  int32_t f[LDS_SEEDS];
  int32_t cscore[LDS_SEEDS];
  uint64_t cts[LDS_SEEDS];
  int16_t pred[LDS_SEEDS];
  uint16_t dom[LDS_SEEDS], ord[LDS_SEEDS], cstart[LDS_SEEDS], ccount[LDS_SEEDS];
  uint8_t used[LDS_SEEDS];
  uint32_t nc;
};

// Pass 3, a wave per query.  Up to LDS_SEEDS seeds: the lanes generate the
// seeds (prefix sums of the per-minimizer counts give each lane its slots),
// rank-sort them (each lane counts the keys below its own: keys are unique,
// so the ranks are the stable sort's positions), lane 0 runs the DP and the
// backtrack in LDS, the lanes rank the endpoints and the chains and write
// the results.  Larger queries take chain_serial on lane 0.
__global__ __launch_bounds__(64 * CW) void seed_chain_kernel(ChainDev c, const uint64_t* seed_base, Work w,
                                                             uint64_t* n_seeds, uint32_t* n_chains,
                                                             phx_chain* out, phx_seed* cs) {
  __shared__ WaveLds lds[CW];
  const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t q = blockIdx.x * CW + wv;
  if (q >= c.nq) return;  // whole waves; no workgroup barrier below
  WaveLds& S = lds[wv];
  const uint64_t base = seed_base[q];
  const uint64_t m0 = c.q_off[q], m1 = c.q_off[q + 1];
  // Lookups: the first 64 minimizers' stay in registers for the generation.
  uint32_t lo0 = 0, cnt0 = 0;
  if (m0 + lane < m1) lookup(c, c.q_val[m0 + lane], &lo0, &cnt0);
  uint64_t nn = cnt0;
  for (uint64_t m = m0 + 64 + lane; m < m1; m += 64) {
    uint32_t lo, cnt;
    lookup(c, c.q_val[m], &lo, &cnt);
    nn += cnt;
  }
  const uint32_t n = (uint32_t)wave_sum(nn);  // below 2^31 (checked when the plan was laid out)
  if (lane == 0) n_seeds[q] = n;
  if (n > LDS_SEEDS || !c.small) {
    if (lane == 0) chain_serial(c, q, base, n, w, out, cs, n_chains);
    return;
  }
  uint32_t g = 0;
  for (uint64_t mb = m0; mb < m1; mb += 64) {
    const uint64_t m = mb + lane;
    uint32_t lo = lo0, cnt = cnt0;
    if (mb != m0) {
      lo = cnt = 0;
      if (m < m1) lookup(c, c.q_val[m], &lo, &cnt);
    }
    uint32_t tot;
    const uint32_t at = g + wave_excl_scan32(cnt, lane, &tot);
    for (uint32_t k = 0; k < cnt; ++k) {
      S.gq[at + k] = c.q_pos[m];
      S.gt[at + k] = c.ix_pos[lo + k];
      S.gm[at + k] = (uint32_t)(m - m0);
    }
    g += tot;
  }
  wave_sync();
  for (uint32_t i = lane; i < n; i += 64) {  // rank by (target, query, generation)
    const uint32_t ti = S.gt[i], qi = S.gq[i];
    uint32_t r = 0;
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t tj = S.gt[j], qj = S.gq[j];
      r += tj < ti || (tj == ti && (qj < qi || (qj == qi && j < i)));
    }
    S.sq[r] = qi;
    S.st[r] = ti;
    S.sm[r] = S.gm[i];
    S.dom[r] = 0;
    S.used[r] = 0;
  }
  wave_sync();
  // The DP (chaining.rs:158-196), one seed i at a time, its candidate
  // predecessors j = i-1, i-2, ... 64 per round on the lanes in scan order.
  // Upstream's sequential scan is reproduced exactly: the positional break
  // cuts the candidates at the first lane too far back; j improves the
  // running best when its total beats every earlier total (an exclusive
  // prefix max); the skip counter (saturating -1 on an improvement, +1 on a
  // dominated j whose count passes max_skip) is a scan of the maps
  // x -> max(x + a, b), and the scan stops after the first lane where it
  // passes max_skip.  Only the lanes up to that stop count as scanned.
  // In 32 bits: plan creation checked seed_len, max_dist_x, max_band and
  // every gap penalty within 2^20 and query positions below 2^31, so the
  // target gap of a candidate (below reach after the break test), the query
  // gap, the scores and f (at most LDS_SEEDS x (seed_len + penalty)) fit.
  // Seeds are sorted by target position, so ti - tj needs no sign.
  const uint32_t reach = (uint32_t)(c.max_dist_x + c.seed_len);
  const int32_t seed_len = (int32_t)c.seed_len, band = (int32_t)c.max_band;
  const uint32_t mskip = c.max_skip > 0xffffffffull ? 0xffffffffu : (uint32_t)c.max_skip;
  constexpr int32_t NEG = INT32_MIN / 4;
  for (uint32_t i = 0; i < n; ++i) {
    const int32_t qi = (int32_t)S.sq[i];
    const uint32_t ti = S.st[i];
    int32_t best = seed_len, bp = -1, skip = 0;
    const int32_t first = (uint64_t)i > c.max_iter ? (int32_t)(i - c.max_iter) : 0;
    for (int32_t top = (int32_t)i - 1; top >= first; top -= 64) {
      const int32_t j = top - (int32_t)lane;
      const bool in = j >= first;
      const uint32_t tj = in ? S.st[j] : ti;
      const uint64_t brk = __ballot(in && ti - tj >= reach);
      const uint32_t cut = brk ? (uint32_t)__builtin_ctzll(brk) : 64u;
      // Lanes [0, span) take part; the scans need only log2(span) steps.
      const uint32_t span = min(cut, (uint32_t)(top - first + 1));
      const bool act = in && lane < cut;
      bool valid = false;
      int32_t total = INT32_MIN;
      if (act) {  // chain_score (chaining.rs:96-117)
        const int32_t dq = qi - (int32_t)S.sq[j], dr = (int32_t)(ti - tj);
        const int32_t dd = dr > dq ? dr - dq : dq - dr;
        if (dq > 0 && dr > 0 && dd <= band) {
          valid = true;
          total = min(seed_len, min(dr, dq)) - (dd > 0 ? c.pen32[dd] : 0) + S.f[j];
        }
      }
      int32_t incl = total, pm;  // inclusive / exclusive prefix max of the totals
      if (span <= 16) {          // one row: DPP steps
        max_scan_row(incl, span);
        pm = row_shr<1>(INT32_MIN, incl);
      } else {
        for (uint32_t d = 1; d < span; d <<= 1) {
          const int32_t y = __shfl_up(incl, d, 64);
          if (lane >= d) incl = max(incl, y);
        }
        pm = __shfl_up(incl, 1, 64);
        if (lane == 0) pm = INT32_MIN;
      }
      const bool imp = valid && total > max(best, pm);
      const bool dominated = valid && !imp;
      const uint32_t dnew = dominated ? S.dom[j] + 1u : 0u;
      const bool inc = dominated && dnew > mskip;
      int32_t A = imp ? -1 : inc ? 1 : 0, B = imp ? 0 : NEG;  // this lane's map x -> max(x + A, B)
      if (span <= 16) {
        map_scan_row(A, B, span, NEG);
      } else {
        for (uint32_t d = 1; d < span; d <<= 1) {  // inclusive composition, earlier lanes first
          const int32_t pa = __shfl_up(A, d, 64), pb = __shfl_up(B, d, 64);
          if (lane >= d) {
            B = max(pb + A, B);
            A = pa + A;
          }
        }
      }
      const int32_t after = max(skip + A, B);
      const uint64_t stopm = __ballot(inc && (uint32_t)after > mskip);
      const uint32_t stop = stopm ? (uint32_t)__builtin_ctzll(stopm) : (span ? span - 1 : 0u);
      const bool scanned = act && lane <= stop;
      if (scanned && dominated) S.dom[j] = (uint16_t)dnew;
      const uint64_t impm = __ballot(scanned && imp);
      if (impm) {  // uniform source lane: a register read, no LDS round trip
        const int src = 63 - __builtin_clzll(impm);
        best = __builtin_amdgcn_readlane(total, src);
        bp = __builtin_amdgcn_readlane(j, src);
      }
      skip = __builtin_amdgcn_readlane(after, (int)stop);
      if (stopm || brk) break;
    }
    if (lane == 0) {
      S.f[i] = best;
      S.pred[i] = (int16_t)bp;
    }
    wave_sync();
  }
  for (uint32_t i = lane; i < n; i += 64) {  // endpoints by (-f, index)
    const int32_t fi = S.f[i];
    uint32_t r = 0;
    for (uint32_t j = 0; j < n; ++j) r += S.f[j] > fi || (S.f[j] == fi && j < i);
    S.ord[r] = (uint16_t)i;
  }
  wave_sync();
  if (lane == 0) {  // backtrack (chaining.rs:198-241); chained seeds staged in gq / gt / gm
    uint32_t nc = 0, ns = 0;
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t e = S.ord[k];
      if (S.used[e] || S.f[e] < c.min_chain_score) continue;
      const uint32_t start = ns;
      for (int32_t cur = (int32_t)e; cur >= 0 && !S.used[cur]; cur = S.pred[cur]) {
        S.used[cur] = 1;
        S.gq[ns] = S.sq[cur];
        S.gt[ns] = S.st[cur];
        S.gm[ns] = S.sm[cur];
        ++ns;
      }
      for (uint32_t a = start, b = ns - 1; a < b; ++a, --b) {  // back to target order
        uint32_t x = S.gq[a]; S.gq[a] = S.gq[b]; S.gq[b] = x;
        x = S.gt[a]; S.gt[a] = S.gt[b]; S.gt[b] = x;
        x = S.gm[a]; S.gm[a] = S.gm[b]; S.gm[b] = x;
      }
      for (uint32_t a = start + 1; a < ns; ++a) {  // stable insertion sort by query position
        const uint32_t xq = S.gq[a], xt = S.gt[a], xm = S.gm[a];
        uint32_t b = a;
        while (b > start && S.gq[b - 1] > xq) {
          S.gq[b] = S.gq[b - 1];
          S.gt[b] = S.gt[b - 1];
          S.gm[b] = S.gm[b - 1];
          --b;
        }
        S.gq[b] = xq;
        S.gt[b] = xt;
        S.gm[b] = xm;
      }
      const int64_t ts = (int64_t)S.gt[start] - (int64_t)S.gq[start];
      S.cscore[nc] = S.f[e];
      S.cts[nc] = (uint64_t)(ts > 0 ? ts : 0);
      S.cstart[nc] = (uint16_t)start;
      S.ccount[nc] = (uint16_t)(ns - start);
      ++nc;
    }
    S.nc = nc;
  }
  wave_sync();
  const uint32_t nc = S.nc;
  uint32_t ns = 0;
  for (uint32_t k = 0; k < nc; ++k) ns += S.ccount[k];
  for (uint32_t k = lane; k < ns; k += 64)
    cs[base + k] = phx_seed{S.gq[k], S.gt[k], c.q_val[m0 + S.gm[k]]};
  for (uint32_t i = lane; i < nc; i += 64) {  // chains by (-score, backtrack order)
    const int32_t si = S.cscore[i];
    uint32_t r = 0;
    for (uint32_t j = 0; j < nc; ++j) r += S.cscore[j] > si || (S.cscore[j] == si && j < i);
    out[base + r] = phx_chain{si, S.cts[i], base + S.cstart[i], S.ccount[i], 0};
  }
  if (lane == 0) n_chains[q] = nc;
}

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

void free_buf(Buf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

bool alloc_copy(Buf& b, const void* src, size_t bytes) {
  b.bytes = std::max<size_t>(bytes, 8);
  if (hipMalloc(&b.p, b.bytes) != hipSuccess) {
    b.p = nullptr;
    return false;
  }
  return !src || bytes == 0 || hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
}

int cfail(phx_ctx* ctx, int code, const std::string& msg) {
  phx::host::set_error(ctx, msg, code);
  return code;
}

}  // namespace

struct phx_seed_index {
  phx_ctx* ctx = nullptr;
  Buf slots, pos;
  uint64_t mask = 0, n = 0;
  std::vector<uint64_t> counts;  // occurrences per distinct value (seed_occurrence_cap)
};

struct phx_chain_plan {
  phx_ctx* ctx = nullptr;
  const phx_seed_index* ix = nullptr;
  ChainDev c{};
  Buf q_val, q_pos, q_off, pen, pen32;
  Buf n_seeds, seed_base, n_chains;
  Buf gq, gt, gv, ord, sq, st, sv, f, pred, dom, used, tmp, out, cs;
  uint64_t total = 0;  // seeds of the batch (fixed by the inputs)
  bool ran = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

extern "C" {

phx_seed_index* phx_seed_index_create(phx_ctx* ctx, const uint64_t* values, const uint32_t* positions,
                                      uint64_t n) {
  if (!ctx || (n && (!values || !positions))) return nullptr;
  if (n >= (1ull << 32)) {
    phx::host::set_error(ctx, "phx_seed_index_create: more than 2^32 - 1 minimizers");
    return nullptr;
  }
  // Distinct values; within a value, positions in sketch order
  // (build_minimizer_index pushes them in that order).
  std::vector<uint64_t> idx(n);
  for (uint64_t i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) { return values[a] < values[b]; });
  std::vector<uint32_t> pos(n);
  std::vector<uint64_t> counts;
  uint64_t slots_n = 16;
  while (slots_n < 2 * n) slots_n <<= 1;  // at most half full
  std::vector<IxSlot> slots(slots_n, IxSlot{0, 0, 0});
  const uint64_t mask = slots_n - 1;
  for (uint64_t k = 0; k < n;) {
    uint64_t e = k;
    while (e < n && values[idx[e]] == values[idx[k]]) {
      pos[e] = positions[idx[e]];
      ++e;
    }
    const uint64_t v = values[idx[k]];
    uint64_t h = mix64(v) & mask;
    while (slots[h].count) h = (h + 1) & mask;
    slots[h] = IxSlot{v, (uint32_t)k, (uint32_t)(e - k)};
    counts.push_back(e - k);
    k = e;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return nullptr;
  phx_seed_index* X = new phx_seed_index();
  X->ctx = ctx;
  X->n = n;
  X->mask = mask;
  X->counts = std::move(counts);
  if (!alloc_copy(X->slots, slots.data(), slots.size() * sizeof(IxSlot)) || !alloc_copy(X->pos, pos.data(), n * 4)) {
    phx::host::set_error(ctx, "phx_seed_index_create: device allocation or upload failed", PHX_E_NOMEM);
    phx_seed_index_destroy(X);
    return nullptr;
  }
  return X;
}

uint64_t phx_seed_occurrence_cap(const phx_seed_index* X, uint64_t floor) {
  if (!X || X->counts.empty()) return floor;
  std::vector<uint64_t> c = X->counts;
  std::sort(c.begin(), c.end());
  const uint64_t median = c[c.size() / 2];
  return std::max(floor, median > UINT64_MAX / 8 ? UINT64_MAX : median * 8);
}

void phx_seed_index_destroy(phx_seed_index* X) {
  if (!X) return;
  free_buf(X->slots);
  free_buf(X->pos);
  delete X;
}

phx_chain_plan* phx_chain_plan_create(phx_ctx* ctx, const phx_seed_index* X, uint64_t max_occ,
                                      const phx_chain_params* p, const uint64_t* q_values,
                                      const uint32_t* q_positions, const uint64_t* q_off, uint32_t nq) {
  if (!ctx || !X || !p || !q_off) return nullptr;
  const uint64_t nm = q_off[nq];
  if (nm && (!q_values || !q_positions)) {
    phx::host::set_error(ctx, "phx_chain_plan_create: null query sketch arrays");
    return nullptr;
  }
  if (p->max_band < 0 || p->max_band > (1 << 20) || p->seed_len < 0 || p->max_dist_x < 0) {
    phx::host::set_error(ctx, "phx_chain_plan_create: max_band must lie in [0, 2^20], seed_len and max_dist_x >= 0");
    return nullptr;
  }
  // chain_score's gap penalty per diagonal drift (chaining.rs:111-115).
  std::vector<int64_t> pen((size_t)p->max_band + 1, 0);
  std::vector<int32_t> pen32(pen.size(), 0);
  constexpr int64_t SMALL = 1 << 20;
  bool small = p->seed_len <= SMALL && p->max_dist_x <= SMALL;
  for (int64_t dd = 1; dd <= p->max_band; ++dd) {
    pen[dd] = (int64_t)std::round(p->gap_penalty * (double)dd + 0.5 * std::log2((double)dd + 1.0));
    small = small && pen[dd] >= -SMALL && pen[dd] <= SMALL;
    pen32[dd] = small ? (int32_t)pen[dd] : 0;
  }
  for (uint64_t m = 0; small && m < nm; ++m) small = q_positions[m] < (1u << 31);
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return nullptr;
  phx_chain_plan* P = new phx_chain_plan();
  P->ctx = ctx;
  P->ix = X;
  bool ok = alloc_copy(P->q_val, q_values, nm * 8) && alloc_copy(P->q_pos, q_positions, nm * 4) &&
            alloc_copy(P->q_off, q_off, ((size_t)nq + 1) * 8) && alloc_copy(P->pen, pen.data(), pen.size() * 8) &&
            alloc_copy(P->pen32, pen32.data(), pen32.size() * 4) &&
            alloc_copy(P->n_seeds, nullptr, (size_t)nq * 8) && alloc_copy(P->n_chains, nullptr, (size_t)nq * 4) &&
            hipEventCreate(&P->ev0) == hipSuccess && hipEventCreate(&P->ev1) == hipSuccess;
  ChainDev& c = P->c;
  c.slots = (const IxSlot*)X->slots.p;
  c.mask = X->mask;
  c.ix_pos = (const uint32_t*)X->pos.p;
  c.max_occ = max_occ;
  c.q_val = (const uint64_t*)P->q_val.p;
  c.q_pos = (const uint32_t*)P->q_pos.p;
  c.q_off = (const uint64_t*)P->q_off.p;
  c.nq = nq;
  c.seed_len = p->seed_len;
  c.max_band = p->max_band;
  c.max_dist_x = p->max_dist_x;
  c.min_chain_score = p->min_chain_score;
  c.max_iter = p->max_iter;
  c.max_skip = p->max_skip;
  c.pen = (const int64_t*)P->pen.p;
  c.pen32 = (const int32_t*)P->pen32.p;
  c.small = small ? 1 : 0;
  // Size the working arrays: the seed count is fixed by the inputs.
  if (ok && nq) {
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(seed_count_kernel, dim3((nq + 3) / 4), dim3(256), 0, s, c, (uint64_t*)P->n_seeds.p);
    std::vector<uint64_t> ns(nq);
    ok = hipGetLastError() == hipSuccess &&
         hipMemcpyAsync(ns.data(), P->n_seeds.p, (size_t)nq * 8, hipMemcpyDeviceToHost, s) == hipSuccess &&
         hipStreamSynchronize(s) == hipSuccess;
    std::vector<uint64_t> base(nq);
    for (uint32_t i = 0; ok && i < nq; ++i) {
      if (ns[i] >= (1ull << 31)) ok = false;  // per-query seed indices are 31-bit
      base[i] = P->total;
      P->total += ns[i];
    }
    ok = ok && alloc_copy(P->seed_base, base.data(), (size_t)nq * 8);
  }
  ok = ok && (P->seed_base.p || alloc_copy(P->seed_base, nullptr, 8));
  const size_t S = (size_t)P->total;
  ok = ok && alloc_copy(P->gq, nullptr, S * 4) && alloc_copy(P->gt, nullptr, S * 4) &&
       alloc_copy(P->gv, nullptr, S * 8) && alloc_copy(P->ord, nullptr, S * 4) && alloc_copy(P->sq, nullptr, S * 4) &&
       alloc_copy(P->st, nullptr, S * 4) && alloc_copy(P->sv, nullptr, S * 8) && alloc_copy(P->f, nullptr, S * 8) &&
       alloc_copy(P->pred, nullptr, S * 4) && alloc_copy(P->dom, nullptr, S * 4) &&
       alloc_copy(P->used, nullptr, S) && alloc_copy(P->tmp, nullptr, S * sizeof(phx_chain)) &&
       alloc_copy(P->out, nullptr, S * sizeof(phx_chain)) && alloc_copy(P->cs, nullptr, S * sizeof(phx_seed));
  if (!ok) {
    phx::host::set_error(ctx, "phx_chain_plan_create: device allocation, upload or seed count failed", PHX_E_NOMEM);
    phx_chain_plan_destroy(P);
    return nullptr;
  }
  return P;
}

int phx_chain_plan_run(phx_chain_plan* P) {
  if (!P) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return cfail(ctx, PHX_E_DEVICE, "hipSetDevice");
  hipStream_t s = ctx->stream;
  const uint32_t nq = P->c.nq;
  P->ran = false;
  if (hipEventRecord(P->ev0, s) != hipSuccess) return cfail(ctx, PHX_E_DEVICE, "event");
  if (nq) {
    Work w;
    w.gq = (uint32_t*)P->gq.p;
    w.gt = (uint32_t*)P->gt.p;
    w.gv = (uint64_t*)P->gv.p;
    w.ord = (uint32_t*)P->ord.p;
    w.sq = (uint32_t*)P->sq.p;
    w.st = (uint32_t*)P->st.p;
    w.sv = (uint64_t*)P->sv.p;
    w.f = (int64_t*)P->f.p;
    w.pred = (int32_t*)P->pred.p;
    w.dom = (uint32_t*)P->dom.p;
    w.used = (uint8_t*)P->used.p;
    w.tmp = (phx_chain*)P->tmp.p;
    hipLaunchKernelGGL(seed_chain_kernel, dim3((nq + CW - 1) / CW), dim3(64 * CW), 0, s, P->c,
                       (const uint64_t*)P->seed_base.p, w, (uint64_t*)P->n_seeds.p, (uint32_t*)P->n_chains.p,
                       (phx_chain*)P->out.p, (phx_seed*)P->cs.p);
    if (hipGetLastError() != hipSuccess) return cfail(ctx, PHX_E_DEVICE, "seed / chain launch");
  }
  if (hipEventRecord(P->ev1, s) != hipSuccess || hipEventSynchronize(P->ev1) != hipSuccess)
    return cfail(ctx, PHX_E_DEVICE, "seed / chain run");
  P->ran = true;
  return PHX_OK;
}

int phx_chain_plan_results(phx_chain_plan* P, phx_chain_results* out) {
  if (!P || !out) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  if (!P->ran) return cfail(ctx, PHX_E_ARG, "phx_chain_plan_results before a run");
  std::lock_guard<std::mutex> lk(ctx->mu);
  const uint32_t nq = P->c.nq;
  const size_t S = (size_t)P->total;
  std::vector<uint64_t> ns(nq), base(nq);
  std::vector<uint32_t> nc(nq);
  std::vector<phx_chain> ch(S);
  std::vector<phx_seed> cs(S);
  if (hipSetDevice(ctx->device) != hipSuccess ||
      (nq && (hipMemcpy(ns.data(), P->n_seeds.p, (size_t)nq * 8, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(base.data(), P->seed_base.p, (size_t)nq * 8, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(nc.data(), P->n_chains.p, (size_t)nq * 4, hipMemcpyDeviceToHost) != hipSuccess)) ||
      (S && (hipMemcpy(ch.data(), P->out.p, S * sizeof(phx_chain), hipMemcpyDeviceToHost) != hipSuccess ||
             hipMemcpy(cs.data(), P->cs.p, S * sizeof(phx_seed), hipMemcpyDeviceToHost) != hipSuccess)))
    return cfail(ctx, PHX_E_DEVICE, "seed / chain download");
  uint64_t C = 0, CS = 0;
  for (uint32_t i = 0; i < nq; ++i) {
    C += nc[i];
    for (uint32_t k = 0; k < nc[i]; ++k) CS += ch[base[i] + k].n_seeds;
  }
  out->n_chains = C;
  out->n_chain_seeds = CS;
  if (!out->chain_off || (nq && !out->n_seeds) || out->chains_cap < C || out->seeds_cap < CS ||
      (C && !out->chains) || (CS && !out->seeds))
    return PHX_E_CAPACITY;
  uint64_t c = 0, s = 0;
  for (uint32_t i = 0; i < nq; ++i) {
    out->n_seeds[i] = ns[i];
    out->chain_off[i] = c;
    for (uint32_t k = 0; k < nc[i]; ++k) {
      phx_chain x = ch[base[i] + k];
      std::copy(cs.begin() + x.seed_off, cs.begin() + x.seed_off + x.n_seeds, out->seeds + s);
      x.seed_off = s;
      s += x.n_seeds;
      out->chains[c++] = x;
    }
  }
  out->chain_off[nq] = c;
  return PHX_OK;
}

double phx_chain_plan_last_kernel_ms(const phx_chain_plan* P) {
  if (!P || !P->ran) return 0.0;
  float ms = 0;
  if (hipEventElapsedTime(&ms, P->ev0, P->ev1) != hipSuccess) return 0.0;
  return ms;
}

uint64_t phx_chain_plan_seeds(const phx_chain_plan* P) { return P ? P->total : 0; }

void phx_chain_plan_destroy(phx_chain_plan* P) {
  if (!P) return;
  (void)hipStreamSynchronize(P->ctx->stream);
  for (Buf* b : {&P->q_val, &P->q_pos, &P->q_off, &P->pen, &P->pen32, &P->n_seeds, &P->seed_base,
                 &P->n_chains, &P->gq, &P->gt, &P->gv, &P->ord, &P->sq, &P->st, &P->sv, &P->f, &P->pred, &P->dom,
                 &P->used, &P->tmp, &P->out, &P->cs})
    free_buf(*b);
  if (P->ev0) (void)hipEventDestroy(P->ev0);
  if (P->ev1) (void)hipEventDestroy(P->ev1);
  delete P;
}

int phx_seed_and_chain_batch(phx_ctx* ctx, const phx_seed_index* X, uint64_t max_occ, const phx_chain_params* p,
                             const uint64_t* q_values, const uint32_t* q_positions, const uint64_t* q_off,
                             uint32_t nq, phx_chain_results* out) {
  if (!ctx || !X || !p || !q_off || !out) return PHX_E_ARG;
  phx_chain_plan* P = phx_chain_plan_create(ctx, X, max_occ, p, q_values, q_positions, q_off, nq);
  if (!P) return phx::host::last_error_code(ctx);
  int rc = phx_chain_plan_run(P);
  if (rc == PHX_OK) rc = phx_chain_plan_results(P, out);
  phx_chain_plan_destroy(P);
  return rc;
}

}  // extern "C"
