// phx_variants.hip -- post-extension step: variant sites from CIGAR runs.
//
// Batched form of extract_variants_from_cigar (phraya-align/src/executor.rs:
// 814-1044) with the coverage windows of compute_windowed_coverage
// (:1071-1090, summed as get_abs_multi, :218-220) and the repeat / hotspot
// lookups (is_in_repeat_region / is_in_hotspot, :795-811).  Upstream walks
// one read's CIGAR string per call on a rayon thread; here a wave takes one
// alignment, a lane per CIGAR run: wave prefix sums of the runs' advances
// give each run's query / target position at once, instead of a walk.  Plan
// creation lays out the outputs once (counts per alignment, then exclusive
// offsets: a scan per 1024-alignment tile plus the totals of the tiles
// before); a run is one pass that writes the records and coverage windows
// at those offsets, so the records come out in alignment order and, per
// alignment, in CIGAR order, exactly as upstream's Vec would be
// concatenated.
//
// Rules per run (ops M = 0, X = 1, I = 2, D = 3; upstream's WFA convention,
// :912: I = target-only bases, D = query-only bases):
//   M: q_pos += n, t_pos += n
//   X: for each base k, qp = q_pos + k < |query| and tp = t_pos + k < |target|
//      -> SNV at tp, ref target[tp], alt query[qp]; then both += n
//   I: t_pos < |target| and q_pos < |query| -> deletion at t_pos, ref
//      target[t_pos] (the first deleted base), alt '.'; t_pos += n
//   D: q_pos < |query| and t_pos < |target| -> insertion at t_pos, ref '.',
//      alleles query[q_pos .. min(q_pos + n, |query|)); q_pos += n
// Coverage window of a variant at p: [p - r floored at 0, min(p + r + 1,
// |target|)), each value the number of the alignment's windows [min(start,
// |target|), min(end, |target|)) (empty ones dropped) that hold it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "phraya_hip.h"
#include "phx_device.h"
#include "phx_host.h"

namespace {

struct VarDev {
  uint32_t n;
  const uint32_t* runs;
  const uint64_t* run_off;
  const uint64_t* target_start;
  const uint8_t* q_seq;
  const uint64_t* q_off;
  const uint8_t* target;
  uint64_t target_len;
  const uint64_t* spans;
  const uint64_t* span_off;
  const uint64_t* repeats;
  uint64_t n_repeats;
  const uint32_t* hotspots;
  uint64_t n_hotspots;
  uint32_t radius;
  const uint32_t* rep_dir;  // interval directories (interval_flags)
  const uint32_t* hot_dir;
  uint32_t rep_shift, hot_shift;
};

constexpr uint32_t TILE = 1024;  // alignments per block of the offset scan

// Membership of [start, end) intervals sorted by start, as upstream's
// partition_point lookup: k = the number of intervals starting at or before
// pos, and pos must lie before the end of interval k - 1.  A directory with
// one entry per 2^shift target positions (dir[b] = intervals starting before
// b << shift) narrows the search to the intervals starting in pos's bucket,
// usually none or one, so a lookup is two or three dependent loads instead
// of a binary search over every interval.
// Both interval sets are looked up at once: the four directory entries are
// loaded before either search, so the two lookups overlap.
__device__ inline uint8_t interval_flags(const VarDev& a, uint64_t pos) {
  uint64_t rlo = 0, rhi = 0, hlo = 0, hhi = 0;
  if (a.n_repeats) {
    rlo = a.rep_dir[pos >> a.rep_shift];
    rhi = a.rep_dir[(pos >> a.rep_shift) + 1];
  }
  if (a.n_hotspots) {
    hlo = a.hot_dir[pos >> a.hot_shift];
    hhi = a.hot_dir[(pos >> a.hot_shift) + 1];
  }
  while (rlo < rhi) {
    const uint64_t mid = (rlo + rhi) >> 1;
    if (a.repeats[2 * mid] <= pos) rlo = mid + 1;
    else rhi = mid;
  }
  while (hlo < hhi) {
    const uint64_t mid = (hlo + hhi) >> 1;
    if ((uint64_t)a.hotspots[2 * mid] <= pos) hlo = mid + 1;
    else hhi = mid;
  }
  const bool rep = rlo > 0 && pos < a.repeats[2 * rlo - 1];
  const bool hot = hlo > 0 && pos < (uint64_t)a.hotspots[2 * hlo - 1];
  return (uint8_t)((rep ? PHX_VAR_IN_REPEAT : 0) | (hot ? PHX_VAR_IN_HOTSPOT : 0));
}

__device__ inline uint64_t shfl_up64(uint64_t x, uint32_t d) {
  return (uint64_t)__shfl_up((unsigned long long)x, d, 64);
}

__device__ inline uint64_t shfl64(uint64_t x, int src) {
  return (uint64_t)__shfl((unsigned long long)x, src, 64);
}

// Exclusive prefix sum over the wave's 64 lanes (all active); *total = sum.
__device__ inline uint64_t wave_excl_scan(uint64_t x, uint32_t lane, uint64_t* total) {
  uint64_t s = x;
#pragma unroll
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const uint64_t y = shfl_up64(s, d);
    if (lane >= d) s += y;
  }
  *total = shfl64(s, 63);
  return s - x;
}

__device__ inline uint64_t wave_sum(uint64_t x) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) x += (uint64_t)__shfl_xor((unsigned long long)x, d, 64);
  return x;
}

// Sum of the window lengths min(p + r + 1, tn) - max(p - r, 0) over p in
// [a, b) (closed forms of the two arithmetic series on either side of the
// clip points).
__device__ inline uint64_t window_words(uint64_t a, uint64_t b, uint64_t r, uint64_t tn) {
  if (b <= a) return 0;
  const uint64_t c = r + 1;
  const uint64_t s = tn > c ? min(max(tn - c, a), b) : a;  // p < s: p + c < tn
  const uint64_t upper = (s - a) * (a + s - 1 + 2 * c) / 2 + (b - s) * tn;
  const uint64_t e = min(max(r, a), b);  // p >= e: p - r >= 0
  const uint64_t lower = (b - e) * (e + b - 1 - 2 * r) / 2;
  return upper - lower;
}

// One lane's run in a 64-run chunk of an alignment: where it starts (wave
// prefix sums of the query / target advances), how many variants it yields
// (m, at target positions t_pos .. t_pos + m - 1 for X, t_pos for I / D) and
// how many coverage words their windows take (w).
struct LaneRun {
  uint32_t op;
  uint64_t len, q_pos, t_pos, m, w;
};

struct Walk {
  uint64_t r0, r1, qn, tn, q, t;  // run range; query / target length; positions at the next chunk

  __device__ Walk(const VarDev& a, uint32_t i)
      : r0(a.run_off[i]), r1(a.run_off[i + 1]), qn(a.q_off[i + 1] - a.q_off[i]), tn(a.target_len), q(0),
        t(a.target_start[i]) {}

  __device__ LaneRun chunk(const VarDev& a, uint64_t r, uint32_t lane) {
    LaneRun L;
    const uint32_t run = r + lane < r1 ? a.runs[r + lane] : 0u;  // past the end: M of length 0
    L.op = run & 15u;
    L.len = run >> 4;
    const bool adv_q = L.op == phx::OP_M || L.op == phx::OP_X || L.op == phx::OP_D;
    const bool adv_t = L.op == phx::OP_M || L.op == phx::OP_X || L.op == phx::OP_I;
    uint64_t tq, tt;
    L.q_pos = q + wave_excl_scan(adv_q ? L.len : 0, lane, &tq);
    L.t_pos = t + wave_excl_scan(adv_t ? L.len : 0, lane, &tt);
    q += tq;
    t += tt;
    const bool inside = L.q_pos < qn && L.t_pos < tn;
    if (L.op == phx::OP_X) L.m = inside ? min(L.len, min(qn - L.q_pos, tn - L.t_pos)) : 0;
    else L.m = inside && (L.op == phx::OP_I || L.op == phx::OP_D) ? 1 : 0;
    L.w = window_words(L.t_pos, L.t_pos + L.m, a.radius, tn);
    return L;
  }
};

// Pass 1, one wave per alignment: its variant and coverage-word counts.
__global__ __launch_bounds__(256) void variant_count_kernel(VarDev a, uint64_t* nvar, uint64_t* ncov) {
  const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= a.n) return;  // whole waves
  Walk w(a, i);
  uint64_t v = 0, c = 0;
  for (uint64_t r = w.r0; r < w.r1; r += 64) {
    const LaneRun L = w.chunk(a, r, lane);
    v += L.m;
    c += L.w;
  }
  v = wave_sum(v);
  c = wave_sum(c);
  if (lane == 0) {
    nvar[i] = v;
    ncov[i] = c;
  }
}

// Pass 2, one workgroup per TILE alignments: the counts become exclusive
// prefix sums within the tile (in place), and the tile's totals go to
// tv / tc.
__global__ __launch_bounds__(256) void tile_scan_kernel(uint64_t* nvar, uint64_t* ncov, uint64_t* tv,
                                                        uint64_t* tc, uint32_t n) {
  __shared__ uint64_t wv[4], wc[4];
  const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint64_t base = (uint64_t)blockIdx.x * TILE + 4ull * t;
  uint64_t xv[4], xc[4], sv = 0, sc = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    xv[k] = base + k < n ? nvar[base + k] : 0;
    xc[k] = base + k < n ? ncov[base + k] : 0;
    sv += xv[k];
    sc += xc[k];
  }
  uint64_t totv, totc;
  uint64_t ev = wave_excl_scan(sv, lane, &totv), ec = wave_excl_scan(sc, lane, &totc);
  if (lane == 0) {
    wv[wave] = totv;
    wc[wave] = totc;
  }
  __syncthreads();
  for (uint32_t j = 0; j < wave; ++j) {
    ev += wv[j];
    ec += wc[j];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) {
      nvar[base + k] = ev;
      ncov[base + k] = ec;
    }
    ev += xv[k];
    ec += xc[k];
  }
  if (t == 255) {
    tv[blockIdx.x] = ev;
    tc[blockIdx.x] = ec;
  }
}

// Pass 3, one wave per alignment: its offsets (tile prefix + the totals of
// the tiles before), its records (each lane the variants of its run: record
// order is CIGAR order), the repeat / hotspot flags, and then, variant by
// variant, the coverage windows with consecutive lanes on consecutive words:
// for each target position, how many of the alignment's windows hold it.
// The last alignment's wave stores the batch totals: var_off[n] on the
// device and both totals in host memory (system-scope stores).
__global__ __launch_bounds__(256) void variant_write_kernel(VarDev a, const uint64_t* vpre, const uint64_t* cpre,
                                                            const uint64_t* tv, const uint64_t* tc,
                                                            uint64_t* var_off, phx_variant* out, uint32_t* cov,
                                                            uint64_t* totals) {
  const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= a.n) return;  // whole waves
  uint64_t pv = 0, pc = 0;
  for (uint32_t j = lane; j < i / TILE; j += 64) {
    pv += tv[j];
    pc += tc[j];
  }
  const uint64_t vb = wave_sum(pv) + vpre[i], cb = wave_sum(pc) + cpre[i];
  if (lane == 0) var_off[i] = vb;
  const uint8_t* q = a.q_seq + a.q_off[i];
  const uint64_t s0 = a.span_off[i], s1 = a.span_off[i + 1], r = a.radius;
  // The first two coverage windows, loaded now (they do not depend on the
  // runs); further windows are read where they are used.
  const uint64_t w0a = s1 > s0 ? a.spans[2 * s0] : 0, w0b = s1 > s0 ? a.spans[2 * s0 + 1] : 0;
  const uint64_t w1a = s1 > s0 + 1 ? a.spans[2 * s0 + 2] : 0, w1b = s1 > s0 + 1 ? a.spans[2 * s0 + 3] : 0;
  Walk w(a, i);
  uint64_t v = vb, c = cb;  // the next chunk's first record / coverage word
  for (uint64_t rr = w.r0; rr < w.r1; rr += 64) {
    const LaneRun L = w.chunk(a, rr, lane);
    uint64_t tm, tw;
    const uint64_t vi = v + wave_excl_scan(L.m, lane, &tm);
    const uint64_t ci0 = c + wave_excl_scan(L.w, lane, &tw);
    uint64_t ci = ci0;
    for (uint64_t k = 0; k < L.m; ++k) {
      const uint64_t p = L.t_pos + k;
      const uint64_t ws = p >= r ? p - r : 0, we = min(p + r + 1, w.tn);
      phx_variant rec;
      rec.alignment = i;
      rec.position = (uint32_t)p;
      rec.type = L.op == phx::OP_X ? PHX_VAR_SNV : L.op == phx::OP_I ? PHX_VAR_DELETION : PHX_VAR_INSERTION;
      rec.ref_base = L.op == phx::OP_D ? (uint8_t)'.' : a.target[p];
      rec.alt_base = L.op == phx::OP_X ? q[L.q_pos + k] : L.op == phx::OP_I ? (uint8_t)'.' : 0;
      rec.flags = interval_flags(a, p);
      rec.ins_query_pos = L.op == phx::OP_D ? (uint32_t)L.q_pos : 0;
      rec.ins_len = L.op == phx::OP_D ? (uint32_t)(min(L.q_pos + L.len, w.qn) - L.q_pos) : 0;
      rec.window_start = (uint32_t)ws;
      rec.window_len = (uint32_t)(we - ws);
      rec.window_offset = (uint32_t)(p - ws);
      rec.coverage_off = ci;
      out[vi + k] = rec;
      ci += we - ws;
    }
    // Coverage words, the whole wave on one variant at a time.  A window
    // lies inside the target, so a span [lo, hi) holds pos exactly when
    // lo <= pos < hi, clipped to the target or not.
    for (uint64_t has = __ballot(L.m > 0); has; has &= has - 1) {
      const int src = __builtin_ctzll(has);
      const uint64_t tp = shfl64(L.t_pos, src), m = shfl64(L.m, src);
      uint64_t cw = shfl64(ci0, src);
      for (uint64_t k = 0; k < m; ++k) {
        const uint64_t p = tp + k;
        const uint64_t ws = p >= r ? p - r : 0, wl = min(p + r + 1, w.tn) - ws;
        for (uint64_t j = lane; j < wl; j += 64) {
          const uint64_t pos = ws + j;
          uint32_t hold = 0;
          hold += (w0a <= pos && pos < w0b) + (w1a <= pos && pos < w1b);  // empty when absent
          for (uint64_t s = s0 + 2; s < s1; ++s) hold += a.spans[2 * s] <= pos && pos < a.spans[2 * s + 1];
          cov[cw + j] = hold;
        }
        cw += wl;
      }
    }
    v += tm;
    c += tw;
  }
  if (i == a.n - 1 && lane == 0) {
    var_off[a.n] = v;
    __hip_atomic_store(&totals[0], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&totals[1], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct phx_variant_plan {
  phx_ctx* ctx = nullptr;
  VarDev a{};
  Buf runs, run_off, target_start, q_seq, q_off, target, spans, span_off, repeats, hotspots;
  Buf rep_dir, hot_dir;     // interval directories
  Buf nvar, ncov;           // per alignment: counts, then exclusive prefixes within the tile
  Buf tile_v, tile_c;       // per tile: totals
  Buf var_off, out, cov;    // outputs
  uint32_t n_tiles = 0;
  uint64_t cap_v = 0, cap_c = 0;  // device capacities (upper bounds from the runs)
  uint64_t n_var = 0, n_cov = 0;  // totals of the last run
  bool ran = false;
  uint64_t* tot_host = nullptr;  // pinned, mapped: the batch totals from the writing pass
  uint64_t* tot_dev = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  mutable bool kms_ready = true;
  mutable double kms = 0;
};

namespace {

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

int vfail(phx_ctx* ctx, int code, const std::string& msg) {
  phx::host::set_error(ctx, msg, code);
  return code;
}

// Checks of a batch that need no device.  Interval starts must be sorted,
// as upstream's lookups require; the directories depend on it.
int check_batch(const phx_variant_batch* b, std::string* err) {
  const uint32_t n = b->n;
  if (n && (!b->run_off || !b->target_start || !b->q_off || !b->span_off)) {
    *err = "missing per-alignment arrays";
    return PHX_E_ARG;
  }
  const uint64_t nruns = n ? b->run_off[n] : 0, nq = n ? b->q_off[n] : 0, nspans = n ? b->span_off[n] : 0;
  if ((nruns && !b->runs) || (nq && !b->q_seq) || (b->target_len && !b->target) || (nspans && !b->spans) ||
      (b->n_repeats && !b->repeats) || (b->n_hotspots && !b->hotspots)) {
    *err = "null array with a non-zero length";
    return PHX_E_ARG;
  }
  if (b->n_repeats >= (1ull << 32) || b->n_hotspots >= (1ull << 32)) {
    *err = "more than 2^32 - 1 intervals";
    return PHX_E_ARG;
  }
  for (uint64_t j = 1; j < b->n_repeats; ++j)
    if (b->repeats[2 * j] < b->repeats[2 * j - 2]) {
      *err = "repeat intervals not sorted by start";
      return PHX_E_ARG;
    }
  for (uint64_t j = 1; j < b->n_hotspots; ++j)
    if (b->hotspots[2 * j] < b->hotspots[2 * j - 2]) {
      *err = "hotspot intervals not sorted by start";
      return PHX_E_ARG;
    }
  return PHX_OK;
}

// Directory of intervals sorted by start over target positions [0, tn): the
// bucket width 2^shift (at least 16 positions) gives about one interval per
// bucket; dir[b] = intervals starting before b << shift, dir[nb] = all.
template <class T>
std::vector<uint32_t> interval_dir(const T* iv, uint64_t n, uint64_t tn, uint32_t* shift) {
  uint32_t s = 4;
  while (s < 63 && (tn >> s) > std::max<uint64_t>(n, 1)) ++s;
  const uint64_t nb = (tn >> s) + 1;
  std::vector<uint32_t> dir(nb + 1);
  uint64_t j = 0;
  for (uint64_t bkt = 0; bkt < nb; ++bkt) {
    while (j < n && (uint64_t)iv[2 * j] < (bkt << s)) ++j;
    dir[bkt] = (uint32_t)j;
  }
  dir[nb] = (uint32_t)n;
  *shift = s;
  return dir;
}

}  // namespace

extern "C" {

phx_variant_plan* phx_variant_plan_create(phx_ctx* ctx, const phx_variant_batch* b) {
  if (!ctx || !b) return nullptr;
  std::string err;
  if (check_batch(b, &err) != PHX_OK) {
    phx::host::set_error(ctx, "phx_variant_plan_create: " + err);
    return nullptr;
  }
  const uint32_t n = b->n;
  const uint64_t nruns = n ? b->run_off[n] : 0, nq = n ? b->q_off[n] : 0,
                 nspans = n ? b->span_off[n] : 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return nullptr;
  phx_variant_plan* P = new phx_variant_plan();
  P->ctx = ctx;
  // Upper bounds of the outputs: a variant per X base and per I / D run.
  uint64_t maxv = 0;
  for (uint64_t r = 0; r < nruns; ++r) {
    const uint32_t op = b->runs[r] & 15u;
    maxv += op == phx::OP_X ? (b->runs[r] >> 4) : (op == phx::OP_I || op == phx::OP_D) ? 1 : 0;
  }
  P->cap_v = std::max<uint64_t>(maxv, 1);
  P->cap_c = std::max<uint64_t>(maxv * std::min<uint64_t>(2ull * b->radius + 1, std::max<uint64_t>(b->target_len, 1)), 1);
  uint32_t rep_shift = 0, hot_shift = 0;
  const std::vector<uint32_t> rep_dir = interval_dir(b->repeats, b->n_repeats, b->target_len, &rep_shift);
  const std::vector<uint32_t> hot_dir = interval_dir(b->hotspots, b->n_hotspots, b->target_len, &hot_shift);
  P->n_tiles = (n + TILE - 1) / TILE;
  bool ok = alloc_copy(P->runs, b->runs, nruns * 4) &&
            alloc_copy(P->run_off, b->run_off, (size_t)(n + 1) * 8 * (n ? 1 : 0)) &&
            alloc_copy(P->target_start, b->target_start, (size_t)n * 8) &&
            alloc_copy(P->q_seq, b->q_seq, nq) &&
            alloc_copy(P->q_off, b->q_off, (size_t)(n + 1) * 8 * (n ? 1 : 0)) &&
            alloc_copy(P->target, b->target, b->target_len) &&
            alloc_copy(P->spans, b->spans, nspans * 16) &&
            alloc_copy(P->span_off, b->span_off, (size_t)(n + 1) * 8 * (n ? 1 : 0)) &&
            alloc_copy(P->repeats, b->repeats, b->n_repeats * 16) &&
            alloc_copy(P->hotspots, b->hotspots, b->n_hotspots * 8) &&
            alloc_copy(P->rep_dir, rep_dir.data(), rep_dir.size() * 4) &&
            alloc_copy(P->hot_dir, hot_dir.data(), hot_dir.size() * 4) &&
            alloc_copy(P->nvar, nullptr, (size_t)n * 8) &&
            alloc_copy(P->ncov, nullptr, (size_t)n * 8) &&
            alloc_copy(P->tile_v, nullptr, (size_t)P->n_tiles * 8) &&
            alloc_copy(P->tile_c, nullptr, (size_t)P->n_tiles * 8) &&
            alloc_copy(P->var_off, nullptr, (size_t)(n + 1) * 8) &&
            alloc_copy(P->out, nullptr, P->cap_v * sizeof(phx_variant)) &&
            alloc_copy(P->cov, nullptr, P->cap_c * 4);
  ok = ok && hipEventCreate(&P->ev0) == hipSuccess && hipEventCreate(&P->ev1) == hipSuccess &&
       hipHostMalloc((void**)&P->tot_host, 16, hipHostMallocMapped) == hipSuccess &&
       hipHostGetDevicePointer((void**)&P->tot_dev, P->tot_host, 0) == hipSuccess;
  if (!ok) {
    phx::host::set_error(ctx, "phx_variant_plan_create: device allocation or upload failed", PHX_E_NOMEM);
    phx_variant_plan_destroy(P);
    return nullptr;
  }
  VarDev& a = P->a;
  a.n = n;
  a.runs = (const uint32_t*)P->runs.p;
  a.run_off = (const uint64_t*)P->run_off.p;
  a.target_start = (const uint64_t*)P->target_start.p;
  a.q_seq = (const uint8_t*)P->q_seq.p;
  a.q_off = (const uint64_t*)P->q_off.p;
  a.target = (const uint8_t*)P->target.p;
  a.target_len = b->target_len;
  a.spans = (const uint64_t*)P->spans.p;
  a.span_off = (const uint64_t*)P->span_off.p;
  a.repeats = (const uint64_t*)P->repeats.p;
  a.n_repeats = b->n_repeats;
  a.hotspots = (const uint32_t*)P->hotspots.p;
  a.n_hotspots = b->n_hotspots;
  a.radius = b->radius;
  a.rep_dir = (const uint32_t*)P->rep_dir.p;
  a.hot_dir = (const uint32_t*)P->hot_dir.p;
  a.rep_shift = rep_shift;
  a.hot_shift = hot_shift;
  // Layout: the per-alignment counts and their offsets depend only on the
  // plan's inputs, so they are found once here; every run recomputes the
  // variants themselves (positions, bases, flags, windows, coverage).
  if (n) {
    hipStream_t st = ctx->stream;
    const dim3 block(256), waves((n + 3) / 4);
    hipLaunchKernelGGL(variant_count_kernel, waves, block, 0, st, P->a, (uint64_t*)P->nvar.p,
                       (uint64_t*)P->ncov.p);
    hipLaunchKernelGGL(tile_scan_kernel, dim3(P->n_tiles), block, 0, st, (uint64_t*)P->nvar.p,
                       (uint64_t*)P->ncov.p, (uint64_t*)P->tile_v.p, (uint64_t*)P->tile_c.p, n);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
      phx::host::set_error(ctx, "phx_variant_plan_create: layout passes failed", PHX_E_DEVICE);
      phx_variant_plan_destroy(P);
      return nullptr;
    }
  }
  return P;
}

int phx_variant_plan_run(phx_variant_plan* P) {
  if (!P) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return vfail(ctx, PHX_E_DEVICE, "hipSetDevice");
  hipStream_t st = ctx->stream;
  const uint32_t n = P->a.n;
  P->ran = false;
  P->kms_ready = false;
  const dim3 block(256), waves((n + 3) / 4);
  P->tot_host[0] = P->tot_host[1] = 0;  // an empty batch launches nothing
  if (hipEventRecord(P->ev0, st) != hipSuccess) return vfail(ctx, PHX_E_DEVICE, "event");
  if (n) {  // one launch: offsets from the plan's layout, everything else computed here
    hipLaunchKernelGGL(variant_write_kernel, waves, block, 0, st, P->a, (const uint64_t*)P->nvar.p,
                       (const uint64_t*)P->ncov.p, (const uint64_t*)P->tile_v.p, (const uint64_t*)P->tile_c.p,
                       (uint64_t*)P->var_off.p, (phx_variant*)P->out.p, (uint32_t*)P->cov.p, P->tot_dev);
    if (hipGetLastError() != hipSuccess) return vfail(ctx, PHX_E_DEVICE, "variant pass");
  }
  if (hipEventRecord(P->ev1, st) != hipSuccess || hipEventSynchronize(P->ev1) != hipSuccess)
    return vfail(ctx, PHX_E_DEVICE, "variant totals");
  P->n_var = ((volatile uint64_t*)P->tot_host)[0];
  P->n_cov = ((volatile uint64_t*)P->tot_host)[1];
  if (P->n_var > P->cap_v || P->n_cov > P->cap_c)  // the bounds hold by construction
    return vfail(ctx, PHX_E_INTERNAL, "variant totals above their bounds");
  P->ran = true;
  return PHX_OK;
}

int phx_variant_plan_results(phx_variant_plan* P, phx_variant_results* out) {
  if (!P || !out) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  if (!P->ran) return vfail(ctx, PHX_E_ARG, "phx_variant_plan_results before a run");
  std::lock_guard<std::mutex> lk(ctx->mu);
  out->n_variants = P->n_var;
  out->n_coverage = P->n_cov;
  if ((P->n_var && !out->variants) || (P->n_cov && !out->coverage) || !out->var_off ||
      out->variants_cap < P->n_var || out->coverage_cap < P->n_cov)
    return PHX_E_CAPACITY;
  if (hipSetDevice(ctx->device) != hipSuccess ||
      (P->n_var && hipMemcpy(out->variants, P->out.p, P->n_var * sizeof(phx_variant),
                             hipMemcpyDeviceToHost) != hipSuccess) ||
      (P->n_cov && hipMemcpy(out->coverage, P->cov.p, P->n_cov * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
      (P->a.n && hipMemcpy(out->var_off, P->var_off.p, ((size_t)P->a.n + 1) * 8, hipMemcpyDeviceToHost) !=
                     hipSuccess))
    return vfail(ctx, PHX_E_DEVICE, "variant download");
  if (!P->a.n) out->var_off[0] = 0;
  return PHX_OK;
}

double phx_variant_plan_last_kernel_ms(const phx_variant_plan* P) {
  if (!P || !P->ran) return 0.0;
  if (!P->kms_ready) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, P->ev0, P->ev1) != hipSuccess) ms = 0;
    P->kms = ms;
    P->kms_ready = true;
  }
  return P->kms;
}

void phx_variant_plan_destroy(phx_variant_plan* P) {
  if (!P) return;
  (void)hipStreamSynchronize(P->ctx->stream);
  for (Buf* b : {&P->runs, &P->run_off, &P->target_start, &P->q_seq, &P->q_off, &P->target, &P->spans,
                 &P->span_off, &P->repeats, &P->hotspots, &P->rep_dir, &P->hot_dir, &P->nvar, &P->ncov,
                 &P->tile_v, &P->tile_c, &P->var_off, &P->out, &P->cov})
    free_buf(*b);
  if (P->ev0) (void)hipEventDestroy(P->ev0);
  if (P->ev1) (void)hipEventDestroy(P->ev1);
  if (P->tot_host) (void)hipHostFree(P->tot_host);
  delete P;
}

int phx_extract_variants(phx_ctx* ctx, const phx_variant_batch* b, phx_variant_results* out) {
  if (!ctx || !b || !out) return PHX_E_ARG;
  std::string err;
  if (check_batch(b, &err) != PHX_OK) return vfail(ctx, PHX_E_ARG, "phx_extract_variants: " + err);
  phx_variant_plan* P = phx_variant_plan_create(ctx, b);
  if (!P) return phx::host::last_error_code(ctx);
  int rc = phx_variant_plan_run(P);
  if (rc == PHX_OK) rc = phx_variant_plan_results(P, out);
  phx_variant_plan_destroy(P);
  return rc;
}

}  // extern "C"
