// phx_api.hip -- host runtime behind include/phraya_hip.h.
//
// Packs sequences (2-bit when a pair is pure uppercase ACGT, raw bytes
// otherwise), uploads them into one HBM word arena, routes pairs to kernel
// tiers, re-routes the few pairs a tier cannot hold (wider wavefront or longer
// history than its slots) to the next tier, and returns results in the
// upstream's units.  Seed handling and the InvalidInput rule follow
// phraya-align/src/lib.rs:107-111 and :145-149.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "phraya_hip.h"
#include "phx_device.h"

#include "phx_host.h"

using namespace phx;
using namespace phx::host;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct phx_plan {
  phx_ctx* ctx = nullptr;
  int engine = 0;
  phx_costs costs{};
  int64_t cap = -1;
  uint32_t n = 0;
  std::vector<uint64_t> seed_q, seed_t, qlen, tlen;  // suffix lengths
  std::vector<int32_t> host_status;                   // final status of pairs never sent to the GPU
  std::vector<PairDesc> pairs;
  std::vector<uint32_t> tier_of;                      // WFA: initial tier per pair
  uint64_t seq_words = 0, runs_total = 0;
  uint64_t input_bytes = 0;                            // packed algorithmic input bytes
  int32_t ring_slots = 2;
  DevBuf d_seq, d_pairs, d_out, d_runs, d_work, d_work2, d_ctr, d_scratch;
  DevBuf d_bounds;  // (pair, bound) pairs of a relaxed-bound re-run
  DevBuf d_gather, d_gidx;  // compacted CIGAR runs and their source / target offsets
  DevBuf d_arena;  // HBM-tier scratch (history + ring slots); each group of a run gets its own slice
  // Row stage (banded-row groups): RowDesc records (one per list position of
  // every row group, in list order), their pair indices, then the staged
  // sequence words; group_rdesc[g] = first record of group g (~0: none).
  DevBuf d_rstage;
  std::vector<uint64_t> group_rdesc;
  uint64_t rstage_recs = 0, rstage_words = 0;          // its capacity (0 records: no row stage)
  uint32_t* ctr_host = nullptr;  // pinned overflow flags, one per group (kernels write them)
  uint32_t* ctr_dev = nullptr;   // the same flags' device address
  std::vector<std::vector<uint32_t>> work;             // initial work per launch group / Myers group
  std::vector<int> group_tier;                         // WFA: 0 = LDS tier, 1 = HBM tier
  std::vector<uint64_t> group_slot;                    // WFA: HBM slot bytes of each group
  std::vector<double> group_share;                     // WFA: share of the HBM budget
  int wide_group = -1;                                 // WFA: the merged 1024-thread register group, if any
  std::vector<double> group_load;                      // WFA: duration model of each group
  std::vector<uint8_t> group_rl;                       // WFA block tiers: 0 HBM ring, 1 ring + sequences
                                                       // in LDS, 2 register wavefront
  std::vector<uint64_t> group_maxsum;                  // WFA: largest qn + tn of each group
  std::vector<uint32_t> group_seqw;                    // WFA: largest staged sequence words of each group
  std::vector<uint32_t> group_rows;                    // WFA banded-row groups: history steps per row
  std::vector<std::array<uint32_t, 4>> group_rowseg;   // WFA banded-row groups: segment sizes (WfaArgs::rows_seg)
  // hipGraph of a run's launch sequence (WFA): replayed by later runs.
  bool use_graph = false;
  bool graph_wanted = false;                           // build one after the current run
  hipGraphExec_t gexec = nullptr;
  void* g_arena = nullptr;                             // arena the graph's kernels point into
  std::vector<int> g_launched;
  uint32_t g_launches = 0;
  phx_tuning tun{};                                    // this plan's tuning (from the context)
  std::vector<uint32_t> order;                         // pending pairs, longest first
  std::vector<int32_t> est;                            // WFA: estimated score steps per pair
  std::vector<uint64_t> band_words;                    // WFA: history words within the pair's bound (0 = none)
  bool banded = false;                                 // WFA: the cost shape has a workgroup register wavefront
  int32_t en_steps = 1;                                //   and its gap extension in score steps
  std::vector<uint8_t> rows_ok;                        // WFA: LDS-tier pair a banded row can take
  std::vector<uint8_t> rows_cert;                      // ... whose optimum is bounded by its 16-lane band
  // Allocation-free plans: every device buffer is a carve of a fixed region
  // (the device's scratch array) instead of an allocation.  Buffers that live
  // for the whole plan are carved from its start, up to fixed_transient; the
  // rest is one region that the run's HBM-tier slots (or Myers scratch) use
  // and, once the run has finished, the result gather.
  uint8_t* fixed_base = nullptr;
  uint64_t fixed_cap = 0, fixed_used = 0, fixed_transient = 0;
  // One-shot plans (phx_align_batch, the executor's) borrow their device
  // buffers, pinned flags and events from the context's cache and give them
  // back when destroyed, so back-to-back calls allocate nothing.
  bool cached = false;
  uint8_t* pin = nullptr;  // the cache's pinned staging while borrowed
  size_t pin_bytes = 0;
  std::vector<uint32_t> flag_shadow;                   // host copy of the overflow flags (fixed plans)
  bool fixed() const { return fixed_base != nullptr; }
  double share_norm = 1.0;                             // sum of group shares, when above 1
  // HBM scratch a group may use (its share of the budget); re-runs (-1) run
  // alone after everything else and may use the whole budget.
  uint64_t group_budget(int g) const {
    const uint64_t total = fixed() ? fixed_cap - fixed_transient : tun.hbm_budget_mib << 20;
    if (g >= 0 && g < (int)group_share.size())
      return std::min<uint64_t>(total, std::max<uint64_t>((uint64_t)(total * group_share[g] / share_norm),
                                                          fixed() ? 0 : 64ull << 20));
    return total;
  }
  std::vector<PairOut> out_h;
  bool out_valid = false;
  uint32_t last_launches = 0;
  // What each kernel launch of the last run processed (for per-launch
  // timings): its tier, workgroups, and pairs list[off, off + count).
  struct LaunchRec {
    int32_t tier = 0;
    uint32_t grid = 0;
    uint32_t storage = 0;  // workgroup tiers: 0 HBM ring, 1 LDS ring, 2 register wavefront
    const std::vector<uint32_t>* list = nullptr;
    uint64_t off = 0, count = 0;
  };
  std::vector<LaunchRec> launch_rec, g_launch_rec;  // eager runs / the recorded graph
  std::vector<std::vector<uint32_t>> rerun_lists;   // re-run lists of the last run
  // Timing of the last run, from the plan's own events (so that runs of other
  // plans on the context leave it alone): a pair of events around every
  // kernel launch, the run's start (null: the first launch's start event)
  // and its end.  Elapsed times are read from the events only when asked
  // for, not on every run.
  std::vector<hipEvent_t> kev;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
  hipEvent_t start_ev = nullptr, end_ev = nullptr;
  mutable bool ms_ready = true, kernel_ms_ready = true;
  mutable double last_ms = 0, last_kernel_ms = 0;
};

struct PlanCache {
  bool busy = false;
  uint8_t* pin = nullptr;  // pinned staging for the plan's uploads and downloads
  size_t pin_bytes = 0;
  DevBuf seq, pairs, out, runs, work, work2, ctr, scratch, gather, gidx, arena, rstage;
  uint32_t* ctr_host = nullptr;
  uint32_t* ctr_dev = nullptr;
  std::vector<hipEvent_t> kev;
  hipEvent_t ev_start = nullptr, ev_end = nullptr;
};

namespace {

// Buffers above this size are freed when a one-shot plan returns them, so a
// context does not hold on to one unusually large batch's memory.
constexpr size_t kCacheKeepBytes = 256ull << 20;

void swap_cache(PlanCache& C, phx_plan* P) {
  std::swap(C.seq, P->d_seq);
  std::swap(C.pairs, P->d_pairs);
  std::swap(C.out, P->d_out);
  std::swap(C.runs, P->d_runs);
  std::swap(C.work, P->d_work);
  std::swap(C.work2, P->d_work2);
  std::swap(C.ctr, P->d_ctr);
  std::swap(C.scratch, P->d_scratch);
  std::swap(C.gather, P->d_gather);
  std::swap(C.gidx, P->d_gidx);
  std::swap(C.arena, P->d_arena);
  std::swap(C.rstage, P->d_rstage);
  std::swap(C.ctr_host, P->ctr_host);
  std::swap(C.ctr_dev, P->ctr_dev);
  std::swap(C.kev, P->kev);
  std::swap(C.ev_start, P->ev_start);
  std::swap(C.ev_end, P->ev_end);
  std::swap(C.pin, P->pin);
  std::swap(C.pin_bytes, P->pin_bytes);
}

// Borrows the context's cached buffers for a one-shot plan (false when
// another one-shot plan of the context holds them).
bool cache_take(phx_plan* P) {
  phx_ctx* ctx = P->ctx;
  std::lock_guard<std::mutex> lk(ctx->cache_mu);
  if (!ctx->plan_cache) ctx->plan_cache = new PlanCache;
  if (ctx->plan_cache->busy) return false;
  ctx->plan_cache->busy = true;
  swap_cache(*ctx->plan_cache, P);
  P->cached = true;
  return true;
}

void dev_free(DevBuf& b);

// Gives a one-shot plan's buffers back to the context's cache.
void cache_return(phx_plan* P) {
  phx_ctx* ctx = P->ctx;
  std::lock_guard<std::mutex> lk(ctx->cache_mu);
  for (DevBuf* b : {&P->d_seq, &P->d_pairs, &P->d_out, &P->d_runs, &P->d_work, &P->d_work2, &P->d_ctr,
                    &P->d_scratch, &P->d_gather, &P->d_gidx, &P->d_arena, &P->d_rstage})
    if (b->bytes > kCacheKeepBytes) dev_free(*b);
  swap_cache(*ctx->plan_cache, P);
  ctx->plan_cache->busy = false;
  P->cached = false;
}

int fail(phx_ctx* c, int code, const std::string& msg) {
  set_error(c, msg, code);
  return code;
}

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(ctx, e_ == hipErrorOutOfMemory ? PHX_E_NOMEM : PHX_E_DEVICE,          \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                   \
  } while (0)

int dev_alloc(phx_ctx* ctx, DevBuf& b, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return PHX_OK;
  if (b.p) {
    HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.bytes = 0;
  }
  HIPCHK(hipMalloc(&b.p, bytes));
  b.bytes = bytes;
  return PHX_OK;
}

void dev_free(DevBuf& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
}

// A buffer the plan keeps for its whole life: allocated, or on fixed plans
// carved (256-byte aligned) from the fixed region; a carve never grows.
int plan_alloc(phx_plan* P, DevBuf& b, size_t bytes) {
  if (!P->fixed()) return dev_alloc(P->ctx, b, bytes);
  if (bytes == 0) bytes = 8;
  if (b.bytes >= bytes) return PHX_OK;
  const uint64_t at = (P->fixed_used + 255) & ~(uint64_t)255;
  if (b.p || at + bytes > P->fixed_cap)
    return fail(P->ctx, PHX_E_NOMEM,
                "batch needs more device memory than the allocation-free scratch holds (" +
                    std::to_string(P->fixed_cap >> 20) + " MiB)");
  b.p = P->fixed_base + at;
  b.bytes = bytes;
  P->fixed_used = at + bytes;
  return PHX_OK;
}

// A buffer of one phase (a run's HBM slots, Myers scratch, the result
// gather): allocated (grow-only), or on fixed plans the transient region at
// `offset`, which every phase reuses.
int plan_transient(phx_plan* P, DevBuf& b, size_t bytes, uint64_t offset = 0) {
  if (!P->fixed()) return dev_alloc(P->ctx, b, bytes);
  if (bytes == 0) bytes = 8;
  const uint64_t at = P->fixed_transient + ((offset + 255) & ~(uint64_t)255);
  if (at + bytes > P->fixed_cap)
    return fail(P->ctx, PHX_E_NOMEM,
                "pairs need more wavefront / result memory than the allocation-free scratch holds (" +
                    std::to_string(P->fixed_cap >> 20) + " MiB)");
  b.p = P->fixed_base + at;
  b.bytes = bytes;
  return PHX_OK;
}

// Overflow flags (one per launch group): pinned host memory the kernels
// store to, or on fixed plans words of the fixed region cleared by a fill
// kernel and copied back after the run.
int clear_flags(phx_plan* P, uint32_t n) {
  phx_ctx* ctx = P->ctx;
  if (P->fixed()) {
    HIPCHK(launch_zero(P->ctr_dev, n, ctx->stream));
    return PHX_OK;
  }
  for (uint32_t i = 0; i < n; ++i) ((volatile uint32_t*)P->ctr_host)[i] = 0;
  return PHX_OK;
}
// Pinned staging of at least `bytes` for a cached plan's copies (grow-only,
// kept by the context's cache), or null: copies then go through pageable
// memory.  Staging stays below 256 MiB.
uint8_t* staging(phx_plan* P, size_t bytes) {
  if (!P->cached || bytes > kCacheKeepBytes) return nullptr;
  if (P->pin_bytes < bytes) {
    if (P->pin) (void)hipHostFree(P->pin);
    P->pin = nullptr;
    P->pin_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 1u << 20);
    if (hipHostMalloc((void**)&P->pin, want, hipHostMallocDefault) != hipSuccess) {
      P->pin = nullptr;
      return nullptr;
    }
    P->pin_bytes = want;
  }
  return P->pin;
}

// Device-to-host copy on the context's stream (ordered after the work
// queued there; its streams do not synchronise with the null stream),
// through the plan's staging when it has one.
int d2h(phx_plan* P, void* dst, const void* src, size_t bytes) {
  phx_ctx* ctx = P->ctx;
  if (!bytes) return PHX_OK;
  uint8_t* pin = staging(P, bytes);
  HIPCHK(hipMemcpyAsync(pin ? pin : dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (pin) std::memcpy(dst, pin, bytes);
  return PHX_OK;
}

int read_flags(phx_plan* P, uint32_t n) {
  phx_ctx* ctx = P->ctx;
  if (P->fixed() && n) HIPCHK(hipMemcpy(P->ctr_host, P->ctr_dev, n * 4, hipMemcpyDeviceToHost));
  return PHX_OK;
}

// Event pair number i of the plan (created on first use).
int kernel_events(phx_plan* P, uint32_t i, hipEvent_t* a, hipEvent_t* b) {
  phx_ctx* ctx = P->ctx;
  while (P->kev.size() < 2 * (size_t)(i + 1)) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    P->kev.push_back(e);
  }
  *a = P->kev[2 * i];
  *b = P->kev[2 * i + 1];
  return PHX_OK;
}

// The plan's run start / end event (created on first use).
int plan_event(phx_plan* P, hipEvent_t* e) {
  phx_ctx* ctx = P->ctx;
  if (!*e) HIPCHK(hipEventCreate(e));
  return PHX_OK;
}

bool is_wfa(int engine) { return engine == PHX_ENGINE_LINEAR || engine == PHX_ENGINE_AFFINE; }

// base with every non-zero field of t applied.
phx_tuning merge_tuning(phx_tuning d, const phx_tuning* t) {
  if (!t) return d;
  if (t->wave_lds_kib) d.wave_lds_kib = std::max<uint32_t>(1, std::min<uint32_t>(t->wave_lds_kib, 38));
  if (t->wave_slots) d.wave_slots = t->wave_slots;
  if (t->hbm_slots) d.hbm_slots = t->hbm_slots;
  if (t->hbm_slot_mib) d.hbm_slot_mib = t->hbm_slot_mib;
  if (t->hbm_budget_mib) d.hbm_budget_mib = t->hbm_budget_mib;
  if (t->block256_min_len) d.block256_min_len = t->block256_min_len;
  if (t->block1024_min_len) d.block1024_min_len = t->block1024_min_len;
  if (t->heavy_permille) d.heavy_permille = std::min<uint32_t>(t->heavy_permille, 1000);
  if (t->block_lds_ring) d.block_lds_ring = t->block_lds_ring == 1 ? 1 : next_pow2(t->block_lds_ring);
  if (t->block_flight) d.block_flight = std::min<uint32_t>(t->block_flight, 4);
  if (t->block_regs) d.block_regs = std::min<uint32_t>(t->block_regs, 2);
  if (t->lds_rows) d.lds_rows = std::min<uint32_t>(t->lds_rows, 3);
  if (t->band_permille) d.band_permille = std::min<uint32_t>(t->band_permille, 1000);
  if (t->regs_narrow) d.regs_narrow = std::min<uint32_t>(t->regs_narrow, 3);
  return d;
}
bool has_cigar(int engine) { return engine != PHX_ENGINE_MYERS_DIST; }

// HBM slots of one group: slot bytes (history + ring), ring geometry and how
// many pairs it keeps in flight within `budget` bytes.
struct GroupSlots {
  uint64_t hist = 0, ring_words = 0, per = 0, n = 0;
  uint32_t wcap = 0;
};

uint64_t max_sum_of(const phx_plan* P, const std::vector<uint32_t>& work) {
  uint64_t m = 1;
  for (uint32_t i : work) m = std::max<uint64_t>(m, (uint64_t)P->pairs[i].qn + P->pairs[i].tn);
  return m;
}

uint32_t seq_words_of(const phx_plan* P, const std::vector<uint32_t>& work) {
  uint32_t w = 2;
  for (uint32_t i : work) w = std::max(w, lds_seq_words(P->pairs[i]));
  return w;
}

// History steps per row of a banded-row group: B + 1 = |qn - tn| + 16.
uint32_t rows_steps_of(const phx_plan* P, const std::vector<uint32_t>& work) {
  uint32_t m = 16;
  for (uint32_t i : work)
    m = std::max<uint32_t>(m, (uint32_t)std::abs(P->pairs[i].qn - P->pairs[i].tn) + 16u);
  return m;
}

// max_sum / seq_words: the group's cached maxima (0 = compute from work).
GroupSlots group_slots(const phx_plan* P, int t, const std::vector<uint32_t>& work,
                       uint64_t slot_bytes, uint64_t budget, bool rl = false, uint64_t max_sum = 0) {
  GroupSlots g;
  if (max_sum == 0) max_sum = max_sum_of(P, work);
  g.ring_words = ring_words_for(ring_frames(P->engine, P->costs), max_sum, &g.wcap);
  if (rl) {  // the ring lives in LDS
    g.ring_words = 0;
    g.wcap = P->tun.block_lds_ring;
  }
  g.hist = slot_bytes ? slot_bytes : ((uint64_t)P->tun.hbm_slot_mib << 20);
  g.per = g.hist + ((g.ring_words * 4 + 255) & ~(uint64_t)255);
  // Pairs in flight: one per wave (wave tier) or per workgroup (block tiers).
  const uint64_t max_flight = t == TIER_WAVE ? P->tun.hbm_slots
                                              : (uint64_t)P->ctx->cu_count * tier_per_cu(t) * P->tun.block_flight;
  g.n = std::max<uint64_t>(1, budget / g.per);
  g.n = std::min<uint64_t>(std::min<uint64_t>(g.n, max_flight), work.size());
  // Whole workgroups of waves; fixed plans round down where that keeps one
  // (their region is exactly the budget).
  if (t == TIER_WAVE)
    g.n = P->fixed() && g.n >= (uint64_t)kWaveWaves ? g.n / kWaveWaves * kWaveWaves
                                                     : (g.n + kWaveWaves - 1) / kWaveWaves * kWaveWaves;
  return g;
}

// Launches one group: tier t over `work` on stream `st`; HBM tiers use the
// scratch slice [slice, slice + gs.n * gs.per).
// Timing events around a launch: both, or none (graph capture: event nodes
// are added around the kernel node).
enum LaunchTiming { TIME_BOTH, TIME_NONE };

int launch_wfa_tier(phx_plan* P, int t, const std::vector<uint32_t>& work, uint32_t* d_list,
                    uint32_t* d_counters, uint32_t* overflow, const GroupSlots& gs, uint8_t* slice,
                    hipStream_t st, int storage = 0, LaunchTiming timing = TIME_BOTH,
                    uint32_t seq_words = 0, uint32_t row_steps = 0, const uint32_t* rows_seg = nullptr,
                    const RowDesc* rows_desc = nullptr, const uint64_t* rows_seq = nullptr) {
  const bool rl = storage != 0;
  phx_ctx* ctx = P->ctx;
  WfaArgs a{};
  a.pairs = (const PairDesc*)P->d_pairs.p;
  a.work_list = d_list;
  a.nwork = (uint32_t)work.size();
  a.counters = d_counters;
  a.overflow = overflow;
  a.cp = cost_params(P->costs, P->cap < 0 ? -1 : (int32_t)std::min<int64_t>(P->cap, INT32_MAX));
  a.out = (PairOut*)P->d_out.p;
  a.runs = (uint32_t*)P->d_runs.p;
  const int mode = P->engine == PHX_ENGINE_AFFINE ? ENG_AFFINE : ENG_LINEAR;
  WaveGeom G{};
  uint32_t slots = 0;
  if ((t == TIER_LDS || rl) && seq_words == 0) seq_words = seq_words_of(P, work);
  if (t == TIER_LDS) {
    G = lds_geometry(lds_tier_ring(P->engine, P->costs) ? ring_frames(P->engine, P->costs) : 0,
                     seq_words, P->tun.wave_lds_kib);
    const bool rows = storage >= kRowsStorage;
    uint64_t sets = work.size();
    if (rows) {
      // The group's segment sizes (32-lane down / up, 16-lane down / up):
      // one wave per two or four of their pairs.
      if (!rows_seg) return fail(ctx, PHX_E_INTERNAL, "banded-row launch without its segments");
      const uint32_t* seg = rows_seg;
      G = rows_geometry(G, seq_words, row_steps ? row_steps : rows_steps_of(P, work), seg[0] + seg[1] > 0);
      sets = 0;
      for (int j = 0; j < 4; ++j) {
        a.rows_seg[j] = seg[j];
        sets += (seg[j] + (j < 2 ? 1u : 3u)) / (j < 2 ? 2u : 4u);
      }
      a.rows_desc = rows_desc;
      a.rows_seq = rows_seq;
    }
    slots = lds_tier_waves(sets, P->tun.wave_slots);
    a.static_order = 1;  // the list is stride-balanced for exactly these waves
  } else {
    a.hist = slice;
    a.hist_stride = gs.hist;
    a.ring_g = (int32_t*)(slice + gs.n * gs.hist);
    a.ring_stride = (gs.per - gs.hist) / 4;
    a.wcap_log2 = log2u(gs.wcap);
    slots = (uint32_t)gs.n;
    if (storage == 1) G = block_lds_geometry(ring_frames(P->engine, P->costs), gs.wcap, seq_words);
    if (storage == 2) {
      G = block_regs_geometry(tier_threads(t), seq_words);
      a.band = 1;
    }
  }
  phx_plan::LaunchRec rec;
  rec.tier = t;
  rec.storage = (uint32_t)storage;
  rec.grid = t <= TIER_WAVE ? std::max<uint32_t>(1, (slots + kWaveWaves - 1) / kWaveWaves)
                            : std::max<uint32_t>(1, slots);
  rec.list = &work;
  rec.count = work.size();
  P->launch_rec.push_back(rec);
  if (timing == TIME_NONE) {
    HIPCHK(launch_tier(t, mode, a, G, slots, st, storage));
    return PHX_OK;
  }
  hipEvent_t e0, e1;
  if (kernel_events(P, P->last_launches, &e0, &e1) != PHX_OK) return PHX_E_DEVICE;
  HIPCHK(hipEventRecord(e0, st));
  HIPCHK(launch_tier(t, mode, a, G, slots, st, storage));
  HIPCHK(hipEventRecord(e1, st));
  ++P->last_launches;
  return PHX_OK;
}

// Initial group lists live in d_work (uploaded once); lists of pairs that a
// group could not hold are staged per run in d_work2.  Counters: two words per
// launch (queue head, overflow count).  Groups run concurrently on the
// context's streams and join on the main stream.
// Phase 1 of a WFA run on the context's streams: counter reset, every group
// launched (longest first, each on the stream with the least work queued;
// groups on one stream run one after another), join, counter readback.
// capture: the streams are being captured into a graph; no timing events are
// recorded and each kernel node is returned in *knodes.
uint64_t rstage_seq_off(const phx_plan* P);

int issue_groups(phx_plan* P, bool capture, std::vector<int>* launched,
                 std::vector<hipGraphNode_t>* knodes) {
  phx_ctx* ctx = P->ctx;
  const int ngroups = (int)P->work.size();
  if (ngroups > (int)kMaxGroups) return fail(ctx, PHX_E_INTERNAL, "too many launch groups");
  uint32_t* d_ctr = (uint32_t*)P->d_ctr.p;
  // Queue heads of the groups that use one (the LDS group strides over its
  // list statically); overflow flags live in pinned host memory and are
  // cleared by the host before each run.
  bool queued_group = false;
  for (int gi = 0; gi < ngroups; ++gi)
    queued_group |= !P->work[gi].empty() && P->group_tier[gi] != TIER_LDS;
  if (queued_group) HIPCHK(launch_zero(d_ctr, 2 * (uint32_t)ngroups, ctx->stream));
  int nonempty = 0;
  for (int gi = 0; gi < ngroups; ++gi) nonempty += !P->work[gi].empty();
  // One launch and nothing else on the stream: its events start and end the
  // run.
  const bool single = !capture && !queued_group && nonempty == 1;
  // Fixed plans record no events (no timing): they create nothing.
  const bool timed = !capture && !P->fixed();
  if (!single && timed) {
    if (plan_event(P, &P->ev_start) != PHX_OK) return PHX_E_DEVICE;
    HIPCHK(hipEventRecord(P->ev_start, ctx->stream));
    P->start_ev = P->ev_start;
  }
  if (!single) HIPCHK(hipEventRecord(ctx->join[0], ctx->stream));
  // Group gi's list sits at offset group_off[gi] of d_work.
  std::vector<uint64_t> group_off(ngroups + 1, 0);
  for (int gi = 0; gi < ngroups; ++gi) group_off[gi + 1] = group_off[gi] + P->work[gi].size();
  std::vector<int> order(ngroups);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return P->group_load[a] > P->group_load[b];
  });
  // The wide register group runs alone on stream 0.  Its workgroups each
  // need a whole CU, which the persistent workgroups of the other groups
  // free only as those groups finish.  When it holds at most one pair per
  // CU (the longest pairs of the batch, a few rounds of steps each) it is
  // launched first, so those pairs start at once; a larger one is launched
  // last: its workgroups take the CUs the other groups leave, and the rest
  // start as the other groups' workgroups finish.
  const int wide = P->wide_group >= 0 && P->wide_group < ngroups && !P->work[P->wide_group].empty()
                       ? P->wide_group : -1;
  if (wide >= 0) {
    order.erase(std::find(order.begin(), order.end(), wide));
    if (P->work[wide].size() <= (size_t)ctx->cu_count) order.insert(order.begin(), wide);
    else order.push_back(wide);
  }
  // Scratch: one arena (grow-only), a disjoint slice per HBM group.
  std::vector<GroupSlots> gs(ngroups);
  std::vector<uint64_t> slice_off(ngroups, 0);
  std::vector<int> tier(P->group_tier.begin(), P->group_tier.end());
  uint64_t arena = 0;
  for (int gi = 0; gi < ngroups; ++gi) {
    if (P->work[gi].empty() || tier[gi] == TIER_LDS) continue;
    const uint64_t budget = P->group_budget(gi);
    gs[gi] = group_slots(P, tier[gi], P->work[gi], P->group_slot[gi], budget, P->group_rl[gi] != 0,
                         P->group_maxsum[gi]);
    // A fixed region holds exactly the budgets: a wave-tier group whose
    // budget is short of one workgroup's four slots runs on the 256-thread
    // workgroup tier instead (one slot per pair; results never depend on
    // the tier).
    if (P->fixed() && tier[gi] == TIER_WAVE && gs[gi].n * gs[gi].per > budget) {
      tier[gi] = TIER_BLOCK256;
      gs[gi] = group_slots(P, tier[gi], P->work[gi], P->group_slot[gi], budget, false, P->group_maxsum[gi]);
    }
    slice_off[gi] = arena;
    arena += gs[gi].n * gs[gi].per;
  }
  if (arena > P->d_arena.bytes && capture) return PHX_E_INTERNAL;  // never allocate while capturing
  if (arena) {
    const int rc = plan_transient(P, P->d_arena, (size_t)arena);
    if (rc != PHX_OK) return rc;
  }
  double queued[kStreams] = {0, 0, 0, 0};
  bool used_stream[kStreams] = {true, false, false, false};
  for (int gi : order) {
    const auto& wl = P->work[gi];
    if (wl.empty()) continue;
    const int si = gi == wide ? 0
                 : wide >= 0 ? 1 + (int)(std::min_element(queued + 1, queued + kStreams) - (queued + 1))
                             : (int)(std::min_element(queued, queued + kStreams) - queued);
    queued[si] += std::max(P->group_load[gi], 1.0);
    hipStream_t st = ctx->streams[si];
    if (si != 0 && !used_stream[si]) {
      HIPCHK(hipStreamWaitEvent(st, ctx->join[0], 0));
      used_stream[si] = true;
    }
    const bool staged = gi < (int)P->group_rdesc.size() && P->group_rdesc[gi] != ~0ull;
    uint8_t* const rbase = (uint8_t*)P->d_rstage.p;
    int rc = launch_wfa_tier(P, tier[gi], wl, (uint32_t*)P->d_work.p + group_off[gi],
                             d_ctr + 2 * gi, P->ctr_dev + gi, gs[gi],
                             (uint8_t*)P->d_arena.p + slice_off[gi], st, P->group_rl[gi],
                             timed ? TIME_BOTH : TIME_NONE, P->group_seqw[gi], P->group_rows[gi],
                             P->group_rowseg[gi].data(),
                             staged ? (const RowDesc*)rbase + P->group_rdesc[gi] : nullptr,
                             staged ? (const uint64_t*)(rbase + rstage_seq_off(P)) : nullptr);
    if (rc != PHX_OK) return rc;
    if (capture) {
      hipStreamCaptureStatus cs;
      unsigned long long id = 0;
      hipGraph_t cg = nullptr;
      const hipGraphNode_t* deps = nullptr;
      size_t nd = 0;
      HIPCHK(hipStreamGetCaptureInfo_v2(st, &cs, &id, &cg, &deps, &nd));
      if (nd != 1) return PHX_E_INTERNAL;
      knodes->push_back(deps[0]);
    }
    launched->push_back(gi);
  }
  if (launched->empty()) return PHX_OK;
  for (int si = 1; si < kStreams; ++si) {
    if (!used_stream[si]) continue;
    HIPCHK(hipEventRecord(ctx->join[si], ctx->streams[si]));
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->join[si], 0));
  }
  if (timed) {
    hipEvent_t e0, e1;
    if (single && kernel_events(P, 0, &e0, &e1) == PHX_OK) {
      P->end_ev = e1;
    } else {
      if (plan_event(P, &P->ev_end) != PHX_OK) return PHX_E_DEVICE;
      HIPCHK(hipEventRecord(P->ev_end, ctx->stream));
      P->end_ev = P->ev_end;
    }
  }
  return PHX_OK;
}

// Captures phase 1 into a graph (nothing runs), with event-record nodes
// around every kernel node for the per-launch timings.  Called right after a
// normal run, so every buffer it points to already exists.
int build_graph(phx_plan* P) {
  phx_ctx* ctx = P->ctx;
  if (P->gexec) {
    (void)hipGraphExecDestroy(P->gexec);
    P->gexec = nullptr;
  }
  std::vector<int> launched;
  std::vector<hipGraphNode_t> knodes;
  const std::vector<phx_plan::LaunchRec> eager = P->launch_rec;
  P->launch_rec.clear();
  HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
  const int rc = issue_groups(P, true, &launched, &knodes);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
  if (rc != PHX_OK || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return rc != PHX_OK ? rc : fail(ctx, PHX_E_DEVICE, "graph capture failed");
  }
  int ret = PHX_OK;
  auto deps_of = [](hipGraphNode_t n) {
    size_t nd = 0;
    std::vector<hipGraphNode_t> d;
    if (hipGraphNodeGetDependencies(n, nullptr, &nd) == hipSuccess && nd) {
      d.resize(nd);
      (void)hipGraphNodeGetDependencies(n, d.data(), &nd);
    }
    return d;
  };
  for (size_t i = 0; i < knodes.size() && ret == PHX_OK; ++i) {
    hipEvent_t e0, e1;
    if (kernel_events(P, (uint32_t)i, &e0, &e1) != PHX_OK) ret = PHX_E_DEVICE;
    hipGraphNode_t first = knodes[i];
    const std::vector<hipGraphNode_t> deps = ret == PHX_OK ? deps_of(first) : std::vector<hipGraphNode_t>{};
    hipGraphNode_t n0 = nullptr, n1 = nullptr;
    if (ret == PHX_OK &&
        (hipGraphAddEventRecordNode(&n0, g, deps.data(), deps.size(), e0) != hipSuccess ||
         hipGraphAddDependencies(g, &n0, &first, 1) != hipSuccess ||
         hipGraphAddEventRecordNode(&n1, g, &knodes[i], 1, e1) != hipSuccess))
      ret = fail(ctx, PHX_E_DEVICE, "graph timing nodes");
  }
  if (ret == PHX_OK && hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
    P->gexec = nullptr;
    ret = fail(ctx, PHX_E_DEVICE, "graph instantiation failed");
  }
  (void)hipGraphDestroy(g);
  std::swap(P->launch_rec, P->g_launch_rec);
  P->launch_rec = eager;  // the run that was just timed
  if (ret != PHX_OK) return ret;
  P->g_launched = launched;
  P->g_launches = (uint32_t)knodes.size();
  P->g_launch_rec = P->launch_rec;
  P->g_arena = P->d_arena.p;
  return PHX_OK;
}

int rerun_relaxed(phx_plan* P, std::vector<uint32_t>& wl, std::vector<PairOut>& outs, uint32_t* d_ctr,
                  std::vector<uint32_t>& exact);

int run_wfa(phx_plan* P) {
  phx_ctx* ctx = P->ctx;
  uint32_t* d_ctr = (uint32_t*)P->d_ctr.p;
  volatile uint32_t* ctr_h = P->ctr_host;
  const uint32_t ngroups = (uint32_t)P->work.size();
  if (clear_flags(P, ngroups) != PHX_OK) return PHX_E_DEVICE;
  std::vector<int> launched;
  if (P->use_graph && P->gexec && P->g_arena == P->d_arena.p) {
    HIPCHK(hipGraphLaunch(P->gexec, ctx->stream));
    if (plan_event(P, &P->ev_end) != PHX_OK) return PHX_E_DEVICE;
    HIPCHK(hipEventRecord(P->ev_end, ctx->stream));
    P->end_ev = P->ev_end;
    launched = P->g_launched;
    P->last_launches = P->g_launches;
    P->launch_rec = P->g_launch_rec;
  } else {
    const int rc = issue_groups(P, false, &launched, nullptr);
    if (rc != PHX_OK) return rc;
    P->graph_wanted = P->use_graph && !launched.empty();
  }
  if (launched.empty()) return PHX_OK;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (read_flags(P, ngroups) != PHX_OK) return PHX_E_DEVICE;
  bool any = false;
  for (int gi : launched) any |= ctr_h[gi] != 0;
  if (!any) return PHX_OK;
  // Re-route pairs that overflowed their group (rare path): LDS-tier pairs go
  // to an HBM slot sized by estimate, HBM pairs to a slot sized exactly from
  // the step count they reported.
  std::vector<PairOut> outs(P->n);
  HIPCHK(hipMemcpy(outs.data(), P->d_out.p, sizeof(PairOut) * P->n, hipMemcpyDeviceToHost));
  auto overflowed = [&](uint32_t i) {
    return outs[i].status >= ST_OVERFLOW_HIST && outs[i].status <= ST_OVERFLOW_RUNS;
  };
  // A register-wavefront pair whose band emptied before its end (status
  // WIDTH while the band fits the group's window: the band spans at most
  // bound + 1 diagonals around kend / 2, so only an optimum above the bound
  // stops it) is re-run with the whole estimate as bound when that is
  // larger and fits the 1024-thread window.
  auto relaxable = [&](int gi, uint32_t i) {
    const PairDesc& pd = P->pairs[i];
    const int nt = tier_threads(P->group_tier[gi]);
    return !P->fixed() && P->group_rl[gi] == 2 && outs[i].status == ST_OVERFLOW_WIDTH && pd.bound > 0 &&
           (uint64_t)pd.bound + 4 <= reg_window(nt) && P->est[i] > (int32_t)pd.bound &&
           (uint64_t)P->est[i] + 4 <= reg_window(1024);
  };
  std::vector<uint32_t> from_lds, exact, relaxed;
  for (int gi : launched)
    for (uint32_t i : P->work[gi])
      if (overflowed(i)) (P->group_tier[gi] == 0 ? from_lds : relaxable(gi, i) ? relaxed : exact).push_back(i);
  if (!relaxed.empty()) {
    const int rc = rerun_relaxed(P, relaxed, outs, d_ctr, exact);
    if (rc != PHX_OK) return rc;
  }
  for (int round = 0; round < 8 && (!from_lds.empty() || !exact.empty()); ++round) {
    std::vector<uint32_t>& wl = !from_lds.empty() ? from_lds : exact;
    const bool est = &wl == &from_lds;
    std::stable_sort(wl.begin(), wl.end(), [&](uint32_t a, uint32_t b) {
      return (uint64_t)P->pairs[a].qn + P->pairs[a].tn > (uint64_t)P->pairs[b].qn + P->pairs[b].tn;
    });
    uint64_t words = 0;
    for (uint32_t i : wl) {
      // A width overflow (LDS ring) stopped early: its step count is partial.
      const int32_t steps = est ? estimate_steps(P->engine, P->costs, P->pairs[i].qn, P->pairs[i].tn)
                          : outs[i].status == ST_OVERFLOW_WIDTH ? std::max(outs[i].steps, P->est[i])
                                                                 : outs[i].steps;
      words = std::max(words, exact_hist_words(steps, P->pairs[i].qn, P->pairs[i].tn));
    }
    const uint64_t slot_bytes = (words * 4 + 255) & ~(uint64_t)255;
    if (plan_alloc(P, P->d_work2, (size_t)P->n * 4) != PHX_OK) return PHX_E_NOMEM;
    HIPCHK(hipMemcpy(P->d_work2.p, wl.data(), wl.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(launch_zero(d_ctr, 2, ctx->stream));
    if (clear_flags(P, 1) != PHX_OK) return PHX_E_DEVICE;
    uint64_t maxL = 0;
    for (uint32_t i : wl) maxL = std::max<uint64_t>(maxL, (uint64_t)P->pairs[i].qn + P->pairs[i].tn);
    // Re-runs are few and run after everything else: give them wide tiers.
    const int rt = maxL <= 8192 ? TIER_BLOCK256 : TIER_BLOCK1024;
    // Every group has finished: the re-run may use the whole arena.
    const GroupSlots rg = group_slots(P, rt, wl, slot_bytes, P->group_budget(-1));
    int rc = plan_transient(P, P->d_arena, (size_t)(rg.n * rg.per));
    if (rc != PHX_OK) return rc;
    P->rerun_lists.push_back(wl);
    rc = launch_wfa_tier(P, rt, P->rerun_lists.back(), (uint32_t*)P->d_work2.p, d_ctr, P->ctr_dev, rg,
                         (uint8_t*)P->d_arena.p, ctx->stream, false, P->fixed() ? TIME_NONE : TIME_BOTH);
    if (rc != PHX_OK) return rc;
    if (!P->fixed()) {
      if (plan_event(P, &P->ev_end) != PHX_OK) return PHX_E_DEVICE;
      HIPCHK(hipEventRecord(P->ev_end, ctx->stream));
      P->end_ev = P->ev_end;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (read_flags(P, 1) != PHX_OK) return PHX_E_DEVICE;
    std::vector<uint32_t> again;
    if (ctr_h[0] != 0) {
      HIPCHK(hipMemcpy(outs.data(), P->d_out.p, sizeof(PairOut) * P->n, hipMemcpyDeviceToHost));
      for (uint32_t i : wl)
        if (overflowed(i)) again.push_back(i);
    }
    wl.clear();
    exact.insert(exact.end(), again.begin(), again.end());
  }
  return PHX_OK;
}

// The relaxed-bound round: pairs whose optimum exceeded their band bound
// run again on the 1024-thread register tier with the whole estimate as
// bound (set on the device copies of their descriptors for this launch
// and restored after it); those that overflow again join `exact`.
int rerun_relaxed(phx_plan* P, std::vector<uint32_t>& wl, std::vector<PairOut>& outs, uint32_t* d_ctr,
                  std::vector<uint32_t>& exact) {
  phx_ctx* ctx = P->ctx;
  std::stable_sort(wl.begin(), wl.end(), [&](uint32_t a, uint32_t b) { return P->est[a] > P->est[b]; });
  std::vector<uint32_t> set(2 * wl.size()), restore(2 * wl.size());
  uint64_t words = 0;
  for (size_t k = 0; k < wl.size(); ++k) {
    const uint32_t i = wl[k];
    const PairDesc& pd = P->pairs[i];
    const bool fitting = !(pd.flags & PF_GLOBAL) && pd.tn > pd.qn + pd.qn / 2 + 10;
    set[2 * k] = restore[2 * k] = i;
    set[2 * k + 1] = (uint32_t)P->est[i];
    restore[2 * k + 1] = pd.bound;
    words = std::max(words, band_hist_words(P->est[i], pd.qn, pd.tn, P->en_steps, fitting));
  }
  if (plan_alloc(P, P->d_work2, (size_t)P->n * 4) != PHX_OK || plan_alloc(P, P->d_bounds, set.size() * 4) != PHX_OK)
    return PHX_E_NOMEM;
  HIPCHK(hipMemcpy(P->d_work2.p, wl.data(), wl.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(P->d_bounds.p, set.data(), set.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(launch_set_bounds((PairDesc*)P->d_pairs.p, (const uint32_t*)P->d_bounds.p, (uint32_t)wl.size(),
                           ctx->stream));
  HIPCHK(launch_zero(d_ctr, 2, ctx->stream));
  if (clear_flags(P, 1) != PHX_OK) return PHX_E_DEVICE;
  const uint64_t slot_bytes = (words * 4 + 255) & ~(uint64_t)255;
  const GroupSlots rg = group_slots(P, TIER_BLOCK1024, wl, slot_bytes, P->group_budget(-1), true);
  int rc = plan_transient(P, P->d_arena, (size_t)(rg.n * rg.per));
  if (rc != PHX_OK) return rc;
  P->rerun_lists.push_back(wl);
  rc = launch_wfa_tier(P, TIER_BLOCK1024, P->rerun_lists.back(), (uint32_t*)P->d_work2.p, d_ctr, P->ctr_dev, rg,
                       (uint8_t*)P->d_arena.p, ctx->stream, 2, TIME_BOTH);
  if (rc != PHX_OK) return rc;
  if (plan_event(P, &P->ev_end) != PHX_OK) return PHX_E_DEVICE;
  HIPCHK(hipEventRecord(P->ev_end, ctx->stream));
  P->end_ev = P->ev_end;
  // The descriptors get their own bounds back (the upload of `restore`
  // waits for the launch: same stream, synchronous copy).
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(P->d_bounds.p, restore.data(), restore.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(launch_set_bounds((PairDesc*)P->d_pairs.p, (const uint32_t*)P->d_bounds.p, (uint32_t)wl.size(),
                           ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (read_flags(P, 1) != PHX_OK) return PHX_E_DEVICE;
  if (((volatile uint32_t*)P->ctr_host)[0] != 0) {
    HIPCHK(hipMemcpy(outs.data(), P->d_out.p, sizeof(PairOut) * P->n, hipMemcpyDeviceToHost));
    for (uint32_t i : wl)
      if (outs[i].status >= ST_OVERFLOW_HIST && outs[i].status <= ST_OVERFLOW_RUNS) exact.push_back(i);
  }
  return PHX_OK;
}

// WFA launch groups from the plan's tuning: the LDS group, then HBM length
// buckets, each split into regular and heavy groups, with slot sizes, budget
// shares and a duration model for the launch order.
void plan_group(phx_plan* P) {
  phx_ctx* ctx = P->ctx;
  // Group 0: LDS tier.  Then HBM length buckets with edges 4096, 8192,
  // 16384, 32768 (qn + tn), each split by estimated history into a regular
  // group (slot = the bucket's 95th-percentile estimate) and a heavy group
  // (the rest, slot = its largest estimate), so a few outliers do not
  // shrink the slots in flight of the whole bucket.  A pair that outgrows
  // its slot still completes (exact re-run).
  // A global-branch LDS-tier pair that no banded row takes and whose
  // estimated history (or step count) the LDS tier cannot hold goes to the
  // workgroup tiers (register shapes only; the estimate runs ~1.4x high, so
  // a pair near the limit may have fitted).  Banded-row pairs stay (a row
  // holds up to |kend| + 31 steps, which the estimate's constant terms
  // overstate for short pairs), and so do fitting pairs (the estimate
  // prices the whole length difference, which the fitting end does not pay).
  auto lds_ok = [&](uint32_t i) {
    if (P->qlen[i] + P->tlen[i] > kWaveMaxLen || P->ring_slots > kWaveRingMax) return false;
    if (!P->banded || P->est.size() != P->n || P->est[i] <= 0) return true;
    if (P->rows_ok.size() == P->n && P->rows_ok[i]) return true;
    const PairDesc& pd = P->pairs[i];
    if (!(pd.flags & PF_GLOBAL) && pd.tn > pd.qn + pd.qn / 2 + 10) return true;
    const int64_t e = P->est[i];
    const uint32_t cells = lds_geometry(0, lds_seq_words(pd), P->tun.wave_lds_kib).cells_cap;
    return e < (int64_t)kWaveSteps && exact_hist_words((int32_t)e, pd.qn, pd.tn) <= cells;
  };
  auto words_of = [&](uint32_t i) {
    return exact_hist_words(P->est[i], P->pairs[i].qn, P->pairs[i].tn);
  };
  std::vector<uint32_t> bucket[6];
  for (uint32_t i : P->order) {
    const uint64_t L = P->qlen[i] + P->tlen[i];
    const int b = lds_ok(i) ? 0 : L <= 4096 ? 1 : L <= 8192 ? 2 : L <= 16384 ? 3 : L <= 32768 ? 4 : 5;
    bucket[b].push_back(i);
  }
  // Tier of each bucket from its upper length edge and the tuning.
  static const uint64_t kBucketEdge[6] = {2048, 4096, 8192, 16384, 32768, ~0ull};
  // 256-thread edge 0 = by cost shape: with a workgroup register wavefront
  // pairs past 2,048 take it (several pairs per CU, bands in registers);
  // otherwise pairs up to 8,192 stay on the HBM wave tier (a wave each).
  const uint64_t edge256 = P->tun.block256_min_len ? P->tun.block256_min_len : P->banded ? 2048 : 8192;
  int kBucketTier[6];
  for (int b = 0; b < 6; ++b)
    kBucketTier[b] = b == 0 ? TIER_LDS
                   : kBucketEdge[b] > P->tun.block1024_min_len ? TIER_BLOCK1024
                   : kBucketEdge[b] > edge256 ? TIER_BLOCK256 : TIER_WAVE;
  P->work.clear();
  P->group_tier.clear();
  P->group_slot.clear();
  P->group_rl.clear();
  // LDS tier: no work queue; the waves stride over a cost-balanced list.
  // Global-branch affine pairs of a banded-row cost shape run four to a
  // wave (one launch group per encoding and window direction); the rest,
  // and every pair when rows are tuned off, one pair per wave.
  // A pair whose optimum is bounded by its 16-lane band (rows_cert, from
  // plan creation) runs in a 16-lane row, four to a wave; the others in
  // 32-lane rows, two to a wave (or all of them: tuning lds_rows = 3).
  std::vector<uint32_t> rows[8];  // [bytes * 4 + segment (32 down, 32 up, 16 down, 16 up)]
  {
    std::vector<uint32_t> single;
    for (uint32_t i : bucket[0]) {
      const PairDesc& pd = P->pairs[i];
      if (P->tun.lds_rows != 2 && P->rows_ok[i])
        rows[((pd.flags & PF_BYTES) ? 4 : 0) + (P->rows_cert[i] && P->tun.lds_rows != 3 ? 2 : 0) +
             (pd.qn > pd.tn ? 1 : 0)].push_back(i);
      else
        single.push_back(i);
    }
    bucket[0].swap(single);
  }
  stride_balance(bucket[0], P->est, lds_tier_waves(bucket[0].size(), P->tun.wave_slots));
  P->work.push_back(bucket[0]);
  P->group_tier.push_back(TIER_LDS);
  P->group_rl.push_back(0);
  P->group_slot.push_back(0);
  std::vector<double> vol(1, 0.0);  // estimated history words per group
  std::vector<std::array<uint32_t, 4>> rowseg(1, {0, 0, 0, 0});
  for (int b8 = 0; b8 < 2; ++b8) {
    std::vector<uint32_t> list;
    std::array<uint32_t, 4> seg{0, 0, 0, 0};
    for (int j = 0; j < 4; ++j) {
      std::vector<uint32_t>& r = rows[b8 * 4 + j];
      const uint32_t per = j < 2 ? 2u : 4u;
      quad_balance(r, P->est, lds_tier_waves((r.size() + per - 1) / per, P->tun.wave_slots), per);
      seg[j] = (uint32_t)r.size();
      list.insert(list.end(), r.begin(), r.end());
    }
    if (list.empty()) continue;
    P->work.push_back(std::move(list));
    P->group_tier.push_back(TIER_LDS);
    P->group_rl.push_back((uint8_t)(kRowsStorage + b8));
    P->group_slot.push_back(0);
    vol.push_back(0.0);
    rowseg.push_back(seg);
  }
  for (int b = 1; b < 6; ++b) {
    if (bucket[b].empty()) continue;
    std::vector<uint64_t> w;
    for (uint32_t i : bucket[b]) w.push_back(words_of(i));
    std::vector<uint64_t> sorted = w;
    std::sort(sorted.begin(), sorted.end());
    const uint64_t keep = 1000 - std::min<uint32_t>(P->tun.heavy_permille, 1000);
    const uint64_t cut = keep >= 1000 ? sorted.back() : sorted[(size_t)((sorted.size() - 1) * keep / 1000)];
    std::vector<uint32_t> reg, heavy;
    for (size_t k = 0; k < w.size(); ++k) (w[k] <= cut ? reg : heavy).push_back(bucket[b][k]);
    for (int h = 0; h < 2; ++h) {
      std::vector<uint32_t>& list = h ? heavy : reg;
      if (list.empty()) continue;
      // Heavy pairs are few and long-running: one tier wider, so they do
      // not trail the rest of the launch.
      const int tier = h ? std::min<int>(kBucketTier[b] + 1, TIER_BLOCK1024) : kBucketTier[b];
      // Workgroup tiers: pairs whose estimated wavefront and sequences fit
      // keep ring and sequences in LDS; cost shapes with a register
      // wavefront keep it in registers while the range fits its window
      // (the step estimate runs about 1.25x high; a pair that outgrows the
      // window is re-run on the ring path); the rest use the HBM ring.
      std::vector<uint32_t> parts[3];  // [0] HBM ring, [1] LDS ring, [2] registers
      const uint32_t W = P->tun.block_lds_ring;
      const uint64_t ring_bytes = (uint64_t)ring_frames(P->engine, P->costs) * W * 4;
      const int nt = tier == TIER_BLOCK256 ? 256 : 1024;
      const bool regs_ok = tier >= TIER_BLOCK256 && P->tun.block_regs != 2 && block_regs_shape(P->engine, P->costs);
      const WaveGeom RG = block_regs_geometry(nt, 0);
      for (uint32_t i : list) {
        int where = 0;
        const uint64_t width = std::min<uint64_t>(2 * (uint64_t)P->est[i] + 1, P->qlen[i] + P->tlen[i] + 1);
        const uint64_t seq_b = 16 + 8ull * lds_seq_words(P->pairs[i]);
        // A bounded pair's band spans at most bound + 1 diagonals (en = 1);
        // an unbounded one spans its whole range, about 2 steps + 1.
        const bool fits_window = P->pairs[i].bound ? (uint64_t)P->pairs[i].bound + 4 <= reg_window(nt)
                                                   : (uint64_t)P->est[i] * 4 / 5 < reg_window(nt) / 2;
        // Register frames hold offset + 1 in 16 bits: queries up to 65,534 bases.
        if (regs_ok && (P->tun.block_regs == 1 || fits_window) && P->qlen[i] <= kRegMaxQuery &&
            RG.per_wave_bytes + seq_b <= kBlockLdsMax)
          where = 2;
        else if (tier >= TIER_BLOCK256 && W > 1 && ring_bytes + 1024 <= kBlockLdsMax &&
                 width <= W && ring_bytes + seq_b <= kBlockLdsMax)
          where = 1;
        parts[where].push_back(i);
      }
      // Register pairs run on the narrowest register tier whose window holds
      // their band: one wave (1,024 diagonals, 16 pairs per CU in flight,
      // no cross-wave barrier), 256 threads (4,096 diagonals, four per CU)
      // or 512 (8,192, two per CU), whose step barriers and latencies
      // overlap, instead of one pair per CU.
      std::vector<uint32_t> narrow, mid, tiny;
      if (P->tun.regs_narrow != 2) {
        std::vector<uint32_t> keep;
        const bool wave_tier = P->tun.regs_narrow != 3;
        for (uint32_t i : parts[2]) {
          const uint64_t b = P->pairs[i].bound;
          (b && wave_tier && b + 4 <= reg_window(64)                     ? tiny
           : tier == TIER_BLOCK1024 && b && b + 4 <= reg_window(256)     ? narrow
           : tier == TIER_BLOCK1024 && b && b + 4 <= reg_window(512)     ? mid
                                                                         : keep).push_back(i);
        }
        parts[2].swap(keep);
        for (auto* l : {&mid, &tiny})
          std::stable_sort(l->begin(), l->end(), [&](uint32_t a, uint32_t b) { return P->est[a] > P->est[b]; });
      }
      for (int r = 0; r < 6; ++r) {
        std::vector<uint32_t>& part = r < 3 ? parts[r] : r == 3 ? narrow : r == 4 ? mid : tiny;
        if (part.empty()) continue;
        uint64_t mx = 0;
        double v = 0;
        for (uint32_t i : part) {
          // Register pairs store only their band's cells.
          const uint64_t wi = r >= 2 && P->pairs[i].bound ? P->band_words[i] : words_of(i);
          mx = std::max<uint64_t>(mx, wi);
          v += (double)wi;
        }
        P->work.push_back(part);
        P->group_tier.push_back(r < 3 ? tier : r == 3 ? TIER_BLOCK256 : r == 4 ? TIER_BLOCK512 : TIER_BLOCK64);
        P->group_rl.push_back((uint8_t)(r < 3 ? r : 2));
        P->group_slot.push_back((mx * 4 + 255) & ~(uint64_t)255);
        vol.push_back(v);
      }
    }
  }
  // The register-wavefront groups of the 1024-thread tier become one group,
  // longest estimate first: its workgroups each fill a CU and take pairs
  // from one queue, so the longest pairs start first and a CU freed by any
  // of its pairs goes to the next one (concurrent groups of fixed sizes
  // leave CUs idle when the smaller ones finish first).  Its history slots
  // are sized for its largest pair, so it is given its pairs in flight (one
  // per CU) from the budget before the other groups share the rest.
  // Likewise the 512-thread and the one-wave register groups of the length
  // buckets become one group each, longest estimate first.
  for (const int mt : {TIER_BLOCK512, TIER_BLOCK64}) {
    std::vector<uint32_t> merged;
    uint64_t slot = 0;
    double v = 0;
    size_t first = SIZE_MAX;
    std::vector<size_t> drop;
    for (size_t g = 0; g < P->work.size(); ++g) {
      if (P->group_tier[g] != mt) continue;
      merged.insert(merged.end(), P->work[g].begin(), P->work[g].end());
      slot = std::max(slot, P->group_slot[g]);
      v += vol[g];
      if (first == SIZE_MAX) first = g;
      else drop.push_back(g);
    }
    if (!drop.empty()) {
      std::stable_sort(merged.begin(), merged.end(), [&](uint32_t a, uint32_t b) { return P->est[a] > P->est[b]; });
      P->work[first] = std::move(merged);
      P->group_slot[first] = slot;
      vol[first] = v;
      for (size_t k = drop.size(); k-- > 0;) {
        const size_t g = drop[k];
        P->work.erase(P->work.begin() + g);
        P->group_tier.erase(P->group_tier.begin() + g);
        P->group_rl.erase(P->group_rl.begin() + g);
        P->group_slot.erase(P->group_slot.begin() + g);
        vol.erase(vol.begin() + g);  // (row-group segments index only the LDS groups in front)
      }
    }
  }
  int wide = -1;
  P->wide_group = -1;
  {
    std::vector<uint32_t> merged;
    uint64_t slot = 0;
    double v = 0;
    std::vector<size_t> keep;
    for (size_t g = 0; g < P->work.size(); ++g) {
      if (P->group_tier[g] == TIER_BLOCK1024 && P->group_rl[g] == 2) {
        merged.insert(merged.end(), P->work[g].begin(), P->work[g].end());
        slot = std::max(slot, P->group_slot[g]);
        v += vol[g];
      } else {
        keep.push_back(g);
      }
    }
    if (!merged.empty()) {
      std::stable_sort(merged.begin(), merged.end(), [&](uint32_t a, uint32_t b) { return P->est[a] > P->est[b]; });
      std::vector<std::vector<uint32_t>> work;
      std::vector<int> tier;
      std::vector<uint8_t> rl;
      std::vector<uint64_t> gslot;
      std::vector<double> gvol;
      for (size_t g : keep) {
        work.push_back(std::move(P->work[g]));
        tier.push_back(P->group_tier[g]);
        rl.push_back(P->group_rl[g]);
        gslot.push_back(P->group_slot[g]);
        gvol.push_back(vol[g]);
      }
      wide = P->wide_group = (int)work.size();
      work.push_back(std::move(merged));
      tier.push_back(TIER_BLOCK1024);
      rl.push_back(2);
      gslot.push_back(slot);
      gvol.push_back(v);
      P->work.swap(work);
      P->group_tier.swap(tier);
      P->group_rl.swap(rl);
      P->group_slot.swap(gslot);
      vol.swap(gvol);
    }
  }
  // HBM budget: each group asks for its pairs in flight x (slot + ring);
  // when the asks exceed the budget they are scaled down together.
  const uint64_t total = P->tun.hbm_budget_mib << 20;
  auto tier_flight = [&](int t) -> uint64_t {
    return t == TIER_WAVE ? P->tun.hbm_slots : (uint64_t)ctx->cu_count * tier_per_cu(t) * P->tun.block_flight;
  };
  std::vector<double> ask(P->work.size(), 0.0);
  double ask_sum = 0;
  for (size_t g = 1; g < P->work.size(); ++g) {
    if (P->group_tier[g] == TIER_LDS) continue;
    uint64_t max_sum = 1;
    for (uint32_t i : P->work[g]) max_sum = std::max<uint64_t>(max_sum, P->qlen[i] + P->tlen[i]);
    const uint64_t ring = P->group_rl[g] ? 0 : ring_words_for(ring_frames(P->engine, P->costs), max_sum, nullptr) * 4;
    const int t = P->group_tier[g];
    const uint64_t flight = tier_flight(t);
    ask[g] = (double)std::min<uint64_t>(flight, P->work[g].size()) * (double)(P->group_slot[g] + ring);
    ask_sum += ask[g];
  }
  P->group_share.assign(P->work.size(), 0.0);
  double share_sum = 0;
  // Over budget: the wide register group first (up to three quarters of
  // the budget), the others in proportion to their asks.
  const bool over = !P->fixed() && ask_sum > (double)total;
  const double wide_share = over && wide >= 0 ? std::min(ask[wide] / (double)total, 0.75) : 0.0;
  const double rest_sum = over && wide >= 0 ? ask_sum - ask[wide] : ask_sum;
  for (size_t g = 1; g < P->work.size(); ++g) {
    if (P->group_tier[g] == TIER_LDS) continue;
    // Fixed plans split their region (sized only after the persistent
    // buffers are carved) in proportion to the asks.
    P->group_share[g] = P->fixed() ? (ask_sum > 0 ? ask[g] / ask_sum : 0.0)
                      : !over ? ask[g] / (double)total
                      : (int)g == wide ? wide_share
                                       : std::max(0.01, (1.0 - wide_share) * ask[g] / std::max(rest_sum, 1.0));
    share_sum += P->group_share[g];
  }
  // The groups' slices together never exceed the budget.
  P->share_norm = std::max(1.0, share_sum);
  // Duration model per group for the launch order: a pair's time grows as
  // steps^2 / threads (cells) + steps (per-step latency); a group takes at
  // least its longest pair and at least its total over its pairs in flight.
  P->group_load.assign(P->work.size(), 0.0);
  for (size_t g = 0; g < P->work.size(); ++g) {
    const int t = P->group_tier[g];
    const double threads = t >= TIER_BLOCK256 ? tier_threads(t) : 64;
    const double flight = t == TIER_LDS ? (double)P->tun.wave_slots
                        : (double)tier_flight(t);
    double sum = 0, longest = 0;
    for (uint32_t i : P->work[g]) {
      const double st = t == TIER_LDS ? 32.0 : (double)P->est[i];
      const double d = st * st / threads + 8.0 * st;
      sum += d;
      longest = std::max(longest, d);
    }
    P->group_load[g] = std::max(longest, sum / std::max(1.0, std::min(flight, (double)P->work[g].size())));
  }
  // Segment sizes of the banded-row groups (they keep their places: only
  // workgroup-tier groups are merged above).
  P->group_rowseg.assign(P->work.size(), {0, 0, 0, 0});
  for (size_t g = 0; g < rowseg.size() && g < P->work.size(); ++g) P->group_rowseg[g] = rowseg[g];
  // Per-group maxima the launches need, computed once per grouping.
  P->group_maxsum.assign(P->work.size(), 0);
  P->group_seqw.assign(P->work.size(), 0);
  P->group_rows.assign(P->work.size(), 0);
  for (size_t g = 0; g < P->work.size(); ++g) {
    P->group_maxsum[g] = max_sum_of(P, P->work[g]);
    if (P->group_tier[g] == TIER_LDS || P->group_rl[g] != 0) P->group_seqw[g] = seq_words_of(P, P->work[g]);
    if (P->group_tier[g] == TIER_LDS && P->group_rl[g] >= kRowsStorage)
      P->group_rows[g] = rows_steps_of(P, P->work[g]);
  }
}

// Uploads the flattened work lists (group by group).
// Band bounds of the pairs that have a step estimate (cost shapes with a
// workgroup register wavefront): estimate x band_permille / 1000, and the
// history words of that band.  A pair whose optimum exceeds its bound is
// re-run with the whole estimate as bound (run_plan's re-run rounds).
int32_t band_bound(const phx_plan* P, uint32_t i) {
  // Plans on a fixed region keep the whole estimate (their re-runs take the
  // ring path, and a fixed region has no room for the relaxed round).
  const int64_t pm = P->fixed() ? 1000 : P->tun.band_permille;
  return std::max<int32_t>(1, (int32_t)((int64_t)P->est[i] * pm / 1000));
}

void apply_bounds(phx_plan* P) {
  if (!P->banded) return;
  for (uint32_t i = 0; i < P->n; ++i) {
    if (P->est[i] <= 0) continue;
    PairDesc& pd = P->pairs[i];
    const bool fitting = !(pd.flags & PF_GLOBAL) && pd.tn > pd.qn + pd.qn / 2 + 10;
    pd.bound = (uint32_t)band_bound(P, i);
    P->band_words[i] = band_hist_words((int32_t)pd.bound, pd.qn, pd.tn, P->en_steps, fitting);
  }
}

// Row stage capacity: a record and the staged words of every pair a banded
// row may take (whatever the grouping); none when a pair's run slots lie
// past a record's 32-bit offsets (its launches then read PairDesc).
void row_stage_caps(phx_plan* P) {
  P->rstage_recs = P->rstage_words = 0;
  if (P->rows_ok.size() != P->n) return;
  uint64_t recs = 0, words = 0;
  for (uint32_t i = 0; i < P->n; ++i) {
    if (!P->rows_ok[i]) continue;
    const PairDesc& pd = P->pairs[i];
    if (pd.run_off + pd.run_cap > UINT32_MAX) return;
    const bool b8 = (pd.flags & PF_BYTES) != 0;
    ++recs;
    words += words_for(pd.qn, b8) + words_for(pd.tn, b8);
  }
  if (words > UINT32_MAX) return;
  P->rstage_recs = recs;
  P->rstage_words = words;
}

// d_rstage layout: records, pair indices (16-byte aligned), staged words.
uint64_t rstage_idx_off(const phx_plan* P) { return P->rstage_recs * sizeof(RowDesc); }
uint64_t rstage_seq_off(const phx_plan* P) { return rstage_idx_off(P) + ((P->rstage_recs * 4 + 15) & ~15ull); }
uint64_t rstage_bytes(const phx_plan* P) {
  return P->rstage_recs ? rstage_seq_off(P) + P->rstage_words * 8 : 0;
}

// Records and pair indices of the current banded-row groups, in list order
// (group_rdesc: each group's first record).
void row_records(phx_plan* P, std::vector<RowDesc>& rd, std::vector<uint32_t>& idx) {
  P->group_rdesc.assign(P->work.size(), ~0ull);
  rd.clear();
  idx.clear();
  if (!P->rstage_recs) return;
  uint64_t words = 0;
  for (size_t g = 0; g < P->work.size(); ++g) {
    if (P->group_tier[g] != TIER_LDS || P->group_rl[g] < kRowsStorage || P->work[g].empty()) continue;
    P->group_rdesc[g] = rd.size();
    for (uint32_t i : P->work[g]) {
      const PairDesc& pd = P->pairs[i];
      const bool b8 = (pd.flags & PF_BYTES) != 0;
      rd.push_back(RowDesc{(uint32_t)words, (uint32_t)pd.run_off, pd.run_cap, (uint16_t)pd.qn, (uint16_t)pd.tn});
      idx.push_back(i);
      words += words_for(pd.qn, b8) + words_for(pd.tn, b8);
    }
  }
  if (rd.size() > P->rstage_recs || words > P->rstage_words) {  // (row groups hold only rows_ok pairs)
    P->group_rdesc.assign(P->work.size(), ~0ull);
    rd.clear();
    idx.clear();
  }
}

// Copies the row pairs' sequences into the stage (device descriptors and
// records already uploaded), on the context's stream ahead of every run.
int stage_rows(phx_plan* P, uint32_t nrec) {
  phx_ctx* ctx = P->ctx;
  uint8_t* base = (uint8_t*)P->d_rstage.p;
  HIPCHK(launch_stage_rows((const PairDesc*)P->d_pairs.p, (const uint32_t*)(base + rstage_idx_off(P)),
                           (const RowDesc*)base, (uint64_t*)(base + rstage_seq_off(P)), nrec, ctx->stream));
  return PHX_OK;
}

// After a regrouping: new records for the new lists, and their stage.
int restage_rows(phx_plan* P) {
  phx_ctx* ctx = P->ctx;
  std::vector<RowDesc> rd;
  std::vector<uint32_t> idx;
  row_records(P, rd, idx);
  if (rd.empty()) return PHX_OK;
  uint8_t* base = (uint8_t*)P->d_rstage.p;
  HIPCHK(hipMemcpy(base, rd.data(), rd.size() * sizeof(RowDesc), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(base + rstage_idx_off(P), idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
  return stage_rows(P, (uint32_t)rd.size());
}

int upload_work(phx_plan* P) {
  phx_ctx* ctx = P->ctx;
  std::vector<uint32_t> flat;
  for (auto& w : P->work) flat.insert(flat.end(), w.begin(), w.end());
  if (!flat.empty()) HIPCHK(hipMemcpy(P->d_work.p, flat.data(), flat.size() * 4, hipMemcpyHostToDevice));
  return PHX_OK;
}

// Work lists (length groups and tiers), memory estimates, device buffers and
// descriptor upload; shared by packed-batch plans and window plans.  qh / th:
// host bytes of each pair's query and target (memory estimates).  packed /
// q_word / t_word: the plan's own sequence arena and each pair's word offsets
// in it, or null when the descriptors already point into resident memory.
phx_plan* plan_setup(phx_plan* P, const std::vector<const uint8_t*>& qh,
                     const std::vector<const uint8_t*>& th, const std::vector<uint64_t>* packed,
                     const std::vector<uint64_t>* q_word, const std::vector<uint64_t>* t_word) {
  phx_ctx* ctx = P->ctx;
  const uint32_t n = P->n;
  // Work lists.
  std::vector<uint32_t>& order = P->order;
  order.clear();
  order.reserve(n);
  for (uint32_t i = 0; i < n; ++i)
    if (P->host_status[i] == ST_PENDING) order.push_back(i);
  // Longest first (qn + tn < 2^28 for every pair that reaches a kernel).
  radix_sort_desc(order, [&](uint32_t i) { return (uint32_t)(P->qlen[i] + P->tlen[i]); });
  if (is_wfa(P->engine)) {
    auto lds_ok = [&](uint32_t i) {
      return P->qlen[i] + P->tlen[i] <= kWaveMaxLen && P->ring_slots <= kWaveRingMax;
    };
    // Step estimates from the sequences (k-mer event rate) for the pairs that
    // need HBM slots, on host threads.
    P->est.assign(n, 0);
    P->band_words.assign(n, 0);
    std::vector<uint32_t> wide;
    for (uint32_t i : P->order)
      if (!lds_ok(i)) wide.push_back(i);
    // Cost shapes with a workgroup register wavefront: the estimate is also
    // the pair's bound for band pruning there (a pair whose optimum exceeds
    // it is re-run without one).
    const bool banded = block_regs_shape(P->engine, P->costs);
    const int32_t en_steps = banded ? P->costs.gap_extend / gcd3(P->costs.mismatch, P->costs.gap_open,
                                                                 P->costs.gap_extend) : 1;
    // Banded rows: global-branch affine pairs of a rows cost shape with
    // |qn - tn| <= kRowsMaxKend, and whether an alignment within the 16-lane
    // band exists (rows_upper_bound).
    P->rows_ok.assign(n, 0);
    P->rows_cert.assign(n, 0);
    const CostParams rcp = cost_params(P->costs, -1);
    if (P->engine == PHX_ENGINE_AFFINE && rows_shape(rcp.xn, rcp.oen, rcp.en)) {
      std::vector<uint32_t> narrow;
      for (uint32_t i : P->order) {
        const PairDesc& pd = P->pairs[i];
        const bool global = (pd.flags & PF_GLOBAL) || pd.tn <= pd.qn + pd.qn / 2 + 10;
        if (lds_ok(i) && global && pd.qn >= 1 && pd.tn >= 1 && std::abs(pd.qn - pd.tn) <= kRowsMaxKend)
          narrow.push_back(i);
      }
      parallel_for(narrow.size(), [&](size_t k) {
        const uint32_t i = narrow[k];
        P->rows_ok[i] = 1;
        P->rows_cert[i] = rows_upper_bound(qh[i], P->qlen[i], th[i], P->tlen[i], rcp.xn, rcp.oen) <=
                          (int64_t)std::abs((int64_t)P->qlen[i] - (int64_t)P->tlen[i]) + 15;
      });
    }
    // With a workgroup register wavefront, the global-branch LDS-tier pairs
    // that no banded row takes get an estimate too: plan_group
    // sends those whose estimated history outgrows the LDS tier to the
    // register tiers instead of finding that out by a failed attempt and a
    // serial re-run.
    if (banded)
      for (uint32_t i : P->order) {
        const PairDesc& pd = P->pairs[i];
        const bool fitting = !(pd.flags & PF_GLOBAL) && pd.tn > pd.qn + pd.qn / 2 + 10;
        if (lds_ok(i) && !P->rows_ok[i] && !fitting && pd.qn && pd.tn) wide.push_back(i);
      }
    parallel_for(wide.size(), [&](size_t k) {
      const uint32_t i = wide[k];
      P->est[i] = estimate_steps_seq(P->engine, P->costs, qh[i], P->qlen[i], th[i], P->tlen[i]);
    });
    P->en_steps = en_steps;
    P->banded = banded;
    apply_bounds(P);
    plan_group(P);
  } else {
    P->work.assign(9, {});
    for (uint32_t i : order) {
      const uint64_t m = P->qlen[i];
      const uint64_t nb = m == 0 ? 1 : (m + 63) / 64;
      P->work[nb <= 8 ? nb : 0].push_back(i);
    }
  }
  // Device buffers.
  auto fail_plan = [&](const char* what, int code = PHX_E_NOMEM) -> phx_plan* {
    // A fixed plan's carve failure has already said what did not fit.
    if (!(P->fixed() && code == PHX_E_NOMEM))
      set_error(ctx, std::string("device allocation failed: ") + what, code);
    phx_plan_destroy(P);
    return nullptr;
  };
  const uint64_t words = packed ? packed->size() : 0;
  if (plan_alloc(P, P->d_seq, words * 8 + 8) != PHX_OK) return fail_plan("seq");
  if (plan_alloc(P, P->d_pairs, sizeof(PairDesc) * (size_t)std::max<uint32_t>(n, 1)) != PHX_OK) return fail_plan("pairs");
  if (plan_alloc(P, P->d_out, sizeof(PairOut) * (size_t)std::max<uint32_t>(n, 1)) != PHX_OK) return fail_plan("out");
  if (plan_alloc(P, P->d_runs, P->runs_total * 4 + 4) != PHX_OK) return fail_plan("runs");
  if (plan_alloc(P, P->d_work, (size_t)std::max<uint32_t>(n, 1) * 4) != PHX_OK) return fail_plan("work");
  if (plan_alloc(P, P->d_ctr, 2 * kMaxGroups * sizeof(uint32_t)) != PHX_OK) return fail_plan("counters");
  // The row stage pays off over repeated runs of one plan; a one-shot plan
  // (it borrows the context's cached buffers) runs once, so it skips the
  // extra upload and copy kernel and its row launches read PairDesc.
  if (!P->cached) row_stage_caps(P);
  // The stage is an optimisation: a fixed region with no room for it keeps
  // the batch (its row launches read PairDesc).
  if (P->fixed() && P->rstage_recs &&
      ((P->fixed_used + 255) & ~(uint64_t)255) + rstage_bytes(P) + (64u << 10) > P->fixed_cap)
    P->rstage_recs = P->rstage_words = 0;
  if (P->rstage_recs && plan_alloc(P, P->d_rstage, rstage_bytes(P)) != PHX_OK) return fail_plan("row stage");
  if (P->fixed()) {
    // Re-run lists and the overflow flags (words of the region, read back
    // after each run) are carved up front: a run allocates nothing.
    DevBuf flags;
    if (plan_alloc(P, P->d_work2, (size_t)std::max<uint32_t>(n, 1) * 4) != PHX_OK ||
        plan_alloc(P, flags, kMaxGroups * sizeof(uint32_t)) != PHX_OK)
      return fail_plan("flags");
    P->ctr_dev = (uint32_t*)flags.p;
    P->flag_shadow.assign(kMaxGroups, 0);
    P->ctr_host = P->flag_shadow.data();
    P->fixed_transient = (P->fixed_used + 255) & ~(uint64_t)255;
    if (P->fixed_transient >= P->fixed_cap) return fail_plan("transient region");
  }
  // Descriptors address their sequences by device pointer: into the plan's
  // own arena at the word given (packed batches; window plans with an arena
  // for some sequences), else the pointer the descriptor came with.
  if (packed) {
    for (uint32_t i = 0; i < n; ++i) {
      if (q_word && (*q_word)[i] != kNoWord) P->pairs[i].q = (const uint64_t*)P->d_seq.p + (*q_word)[i];
      if (t_word && (*t_word)[i] != kNoWord) P->pairs[i].t = (const uint64_t*)P->d_seq.p + (*t_word)[i];
    }
  }
  // Sequences, descriptors and work lists: through one pinned staging block
  // (cached plans; async copies on the context's stream, which every run
  // orders after), else straight from pageable memory.
  size_t wl_words = 0;
  for (auto& w : P->work) wl_words += w.size();
  // Row stage records of the banded-row groups (their sequences are copied
  // into the stage by a kernel once the descriptors are on the device).
  std::vector<RowDesc> rd;
  std::vector<uint32_t> ridx;
  row_records(P, rd, ridx);
  const size_t seq_b = words * 8, pd_b = sizeof(PairDesc) * (size_t)n, wl_b = wl_words * 4;
  const size_t rd_b = rd.size() * sizeof(RowDesc), ri_b = ridx.size() * 4;
  uint8_t* const rbase = (uint8_t*)P->d_rstage.p;
  if (uint8_t* pin = staging(P, seq_b + pd_b + wl_b + rd_b + ri_b)) {
    if (seq_b) std::memcpy(pin, packed->data(), seq_b);
    if (pd_b) std::memcpy(pin + seq_b, P->pairs.data(), pd_b);
    uint32_t* wl = (uint32_t*)(pin + seq_b + pd_b);
    for (auto& w : P->work) wl = std::copy(w.begin(), w.end(), wl);
    uint8_t* const pr = pin + seq_b + pd_b + wl_b;
    if (rd_b) std::memcpy(pr, rd.data(), rd_b);
    if (ri_b) std::memcpy(pr + rd_b, ridx.data(), ri_b);
    if ((seq_b && hipMemcpyAsync(P->d_seq.p, pin, seq_b, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        (pd_b && hipMemcpyAsync(P->d_pairs.p, pin + seq_b, pd_b, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        (wl_b && hipMemcpyAsync(P->d_work.p, pin + seq_b + pd_b, wl_b, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        (rd_b && hipMemcpyAsync(rbase, pr, rd_b, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) ||
        (ri_b && hipMemcpyAsync(rbase + rstage_idx_off(P), pr + rd_b, ri_b, hipMemcpyHostToDevice,
                                ctx->stream) != hipSuccess))
      return fail_plan("upload", PHX_E_DEVICE);
    if (!rd.empty() && stage_rows(P, (uint32_t)rd.size()) != PHX_OK) return fail_plan("row stage", PHX_E_DEVICE);
    if (!P->fixed() && !P->ctr_host &&
        (hipHostMalloc((void**)&P->ctr_host, kMaxGroups * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
         hipHostGetDevicePointer((void**)&P->ctr_dev, P->ctr_host, 0) != hipSuccess))
      return fail_plan("pinned counters");
    // The staging block is reused by the results: let the uploads finish.
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail_plan("upload", PHX_E_DEVICE);
    return P;
  }
  if ((words && hipMemcpy(P->d_seq.p, packed->data(), words * 8, hipMemcpyHostToDevice) != hipSuccess) ||
      (n && hipMemcpy(P->d_pairs.p, P->pairs.data(), sizeof(PairDesc) * n, hipMemcpyHostToDevice) != hipSuccess))
    return fail_plan("upload", PHX_E_DEVICE);
  if (!P->fixed() && !P->ctr_host &&
      (hipHostMalloc((void**)&P->ctr_host, kMaxGroups * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
       hipHostGetDevicePointer((void**)&P->ctr_dev, P->ctr_host, 0) != hipSuccess))
    return fail_plan("pinned counters");
  // Work lists are static between tunings: upload once.
  if (upload_work(P) != PHX_OK) return fail_plan("work upload", PHX_E_DEVICE);
  if (!rd.empty() &&
      (hipMemcpy(rbase, rd.data(), rd_b, hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(rbase + rstage_idx_off(P), ridx.data(), ri_b, hipMemcpyHostToDevice) != hipSuccess ||
       stage_rows(P, (uint32_t)rd.size()) != PHX_OK))
    return fail_plan("row stage", PHX_E_DEVICE);
  return P;
}

int run_myers(phx_plan* P) {
  phx_ctx* ctx = P->ctx;
  uint32_t* d_list_base = (uint32_t*)P->d_work.p;
  const uint64_t budget = P->fixed() ? P->fixed_cap - P->fixed_transient : P->tun.hbm_budget_mib << 20;
  uint64_t list_off = 0;
  for (size_t g = 0; g < P->work.size(); ++g) {
    const auto& wl = P->work[g];
    if (wl.empty()) continue;
    const int nbt = g <= 8 ? (int)g : 0;  // group index = block count (0: generic)
    uint64_t stride = 0;
    for (uint32_t i : wl) {
      const uint64_t m = P->pairs[i].qn, n = P->pairs[i].tn;
      const uint64_t nb = (m + 63) / 64;
      uint64_t s = nbt == 0 ? 6 * nb * 8 : 0;
      if (has_cigar(P->engine)) s += n * (16 * nb + 4);
      stride = std::max<uint64_t>(stride, (s + 15) & ~(uint64_t)15);
    }
    uint64_t chunk = stride ? std::max<uint64_t>(1, budget / stride) : wl.size();
    chunk = std::min<uint64_t>(chunk, wl.size());
    if (stride) {
      const int rc = plan_transient(P, P->d_scratch, (size_t)(chunk * stride));
      if (rc != PHX_OK) return rc;
    }
    for (uint64_t c0 = 0; c0 < wl.size(); c0 += chunk) {
      const uint64_t cn = std::min<uint64_t>(chunk, wl.size() - c0);
      MyersArgs a{};
      a.pairs = (const PairDesc*)P->d_pairs.p;
      a.work_list = d_list_base + list_off + c0;
      a.nwork = (uint32_t)cn;
      a.out = (PairOut*)P->d_out.p;
      a.runs = (uint32_t*)P->d_runs.p;
      a.scratch = (uint8_t*)P->d_scratch.p;
      a.scratch_stride = stride;
      a.engine = P->engine == PHX_ENGINE_MYERS ? ENG_MYERS
               : P->engine == PHX_ENGINE_MYERS_GLOBAL ? ENG_MYERS_GLOBAL : ENG_MYERS_DIST;
      phx_plan::LaunchRec rec;
      rec.tier = 8;
      rec.grid = (uint32_t)((cn + 255) / 256);
      rec.list = &wl;
      rec.off = c0;
      rec.count = cn;
      P->launch_rec.push_back(rec);
      if (P->fixed()) {
        HIPCHK(launch_myers(nbt, a, ctx->stream));
      } else {
        hipEvent_t e0, e1;
        if (kernel_events(P, P->last_launches, &e0, &e1) != PHX_OK) return PHX_E_DEVICE;
        HIPCHK(hipEventRecord(e0, ctx->stream));
        HIPCHK(launch_myers(nbt, a, ctx->stream));
        HIPCHK(hipEventRecord(e1, ctx->stream));
      }
      ++P->last_launches;
      if (c0 + cn < wl.size()) HIPCHK(hipStreamSynchronize(ctx->stream));  // scratch reuse
    }
    list_off += wl.size();
  }
  return PHX_OK;
}

}  // namespace

namespace phx {
namespace host {

phx_plan* plan_create_windows(phx_ctx* ctx, int engine, phx_costs costs, int64_t cap,
                              std::vector<PairDesc> pairs, const std::vector<const uint8_t*>& qh,
                              const std::vector<const uint8_t*>& th, bool oneshot,
                              const std::vector<uint64_t>* arena, const std::vector<uint64_t>* q_word,
                              const std::vector<uint64_t>* t_word) {
  if (!ctx || engine < PHX_ENGINE_LINEAR || engine > PHX_ENGINE_MYERS_DIST) return nullptr;
  if (engine == PHX_ENGINE_AFFINE &&
      (costs.mismatch < 1 || costs.gap_extend < 1 || costs.gap_open < 0 ||
       ring_slots(engine, costs) > kRingSlotsMax)) {
    set_error(ctx, "affine costs need mismatch >= 1, gap_extend >= 1, gap_open >= 0");
    return nullptr;
  }
  for (size_t i = 0; i < pairs.size(); ++i)
    if (pairs[i].qn < 0 || pairs[i].tn < 0 || pairs[i].qn > kMaxLen || pairs[i].tn > kMaxLen) {
      set_error(ctx, "window " + std::to_string(i) + " is longer than the device format supports (2^27 - 2)");
      return nullptr;
    }
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) {
    set_error(ctx, "hipSetDevice failed", PHX_E_DEVICE);
    return nullptr;
  }
  phx_plan* P = new phx_plan;
  P->ctx = ctx;
  if (oneshot) cache_take(P);
  P->tun = ctx->tun;
  P->engine = engine;
  P->costs = engine == PHX_ENGINE_AFFINE ? costs : phx_costs{0, 0, 0};
  P->cap = cap;
  P->n = (uint32_t)pairs.size();
  P->ring_slots = ring_slots(engine, P->costs);
  P->seed_q.assign(P->n, 0);
  P->seed_t.assign(P->n, 0);
  P->qlen.assign(P->n, 0);
  P->tlen.assign(P->n, 0);
  P->host_status.assign(P->n, ST_PENDING);
  uint64_t runs = 0;
  for (uint32_t i = 0; i < P->n; ++i) {
    PairDesc& pd = pairs[i];
    P->qlen[i] = (uint64_t)pd.qn;
    P->tlen[i] = (uint64_t)pd.tn;
    pd.run_cap = has_cigar(engine) ? (uint32_t)(pd.qn + pd.tn + 1) : 0;
    pd.run_off = runs;
    runs += pd.run_cap;
    const bool b8 = pd.flags & PF_BYTES;
    P->input_bytes += b8 ? (uint64_t)pd.qn + pd.tn : ((uint64_t)pd.qn + 3) / 4 + ((uint64_t)pd.tn + 3) / 4;
  }
  P->runs_total = runs;
  P->pairs = std::move(pairs);
  return plan_setup(P, qh, th, arena, q_word, t_word);
}

}  // namespace host
}  // namespace phx

// Device scratch of phx_align_batch_noalloc: one array per GPU, held from the
// library's load until the process ends.  In host code this name is only a
// stand-in for the symbol; the device address of a GPU's copy comes from
// hipGetSymbolAddress with that device current (noalloc_scratch_of), and only
// that address is ever passed to kernels or copies.
static __device__ __align__(256) uint8_t scratch[PHX_NOALLOC_SCRATCH_BYTES];

namespace {

struct NoAllocDevice {
  bool looked_up = false;
  uint8_t* ptr = nullptr;   // this device's copy of `scratch`
  std::mutex busy;          // one allocation-free call per device at a time
};

NoAllocDevice& noalloc_device(int device) {
  static std::mutex mu;
  static std::vector<std::unique_ptr<NoAllocDevice>> devs;
  std::lock_guard<std::mutex> lk(mu);
  if (device >= (int)devs.size()) devs.resize((size_t)device + 1);
  if (!devs[(size_t)device]) devs[(size_t)device].reset(new NoAllocDevice);
  return *devs[(size_t)device];
}

// Address of `scratch` on `device`, looked up once per device (the device
// must be current); null when the lookup fails.
uint8_t* noalloc_scratch_of(int device) {
  static std::mutex mu;
  NoAllocDevice& d = noalloc_device(device);
  std::lock_guard<std::mutex> lk(mu);
  if (!d.looked_up) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    d.ptr = e == hipSuccess ? (uint8_t*)p : nullptr;
    d.looked_up = true;
  }
  return d.ptr;
}

}  // namespace

extern "C" {

const char* phx_version(void) { return "phraya_hip 0.1 gfx950"; }

int64_t phx_estimate_request(int32_t engine, phx_costs costs, const uint8_t* q, uint64_t qn,
                             const uint8_t* t, uint64_t tn, int32_t* steps) {
  if ((engine != PHX_ENGINE_AFFINE && engine != PHX_ENGINE_LINEAR) || (qn && !q) || (tn && !t) ||
      qn > (uint64_t)kMaxLen || tn > (uint64_t)kMaxLen)
    return -1;
  if (engine == PHX_ENGINE_AFFINE && (costs.mismatch < 1 || costs.gap_extend < 1 || costs.gap_open < 0))
    return -1;
  const int32_t s = estimate_steps_seq(engine, costs, q, qn, t, tn);
  if (steps) *steps = s;
  uint32_t wcap = 0;
  const uint64_t ring = ring_words_for(ring_frames(engine, costs), qn + tn, &wcap) * 4;
  return (int64_t)(hist_slot_bytes(s, (int32_t)qn, (int32_t)tn) + ((ring + 255) & ~(uint64_t)255));
}

phx_ctx* phx_ctx_create(int device, uint32_t flags, const phx_tuning* tuning) {
  (void)flags;
  phx_ctx* c = new phx_ctx;
  if (device >= 0) {
    if (hipSetDevice(device) != hipSuccess) { delete c; return nullptr; }
    c->device = device;
  } else {
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return nullptr; }
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) {
    c->cu_count = prop.multiProcessorCount;
    // "hip-gfx950" from "gfx950:sramecc+:xnack-".
    std::string arch(prop.gcnArchName);
    arch = arch.substr(0, arch.find(':'));
    std::snprintf(c->impl_name, sizeof c->impl_name, "hip-%s", arch.c_str());
  }
  // The allocation-free entry point's scratch on this device (its address is
  // looked up once per device, here, with the device current).
  c->noalloc_scratch = noalloc_scratch_of(c->device);
  for (int i = 0; i < kStreams; ++i) {
    if (hipStreamCreateWithFlags(&c->streams[i], hipStreamNonBlocking) != hipSuccess) { delete c; return nullptr; }
  }
  c->stream = c->streams[0];
  for (hipEvent_t& e : c->join) {
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { delete c; return nullptr; }
  }
  phx_tuning d{};
  d.wave_lds_kib = 9;  // 4 workgroups x 4 waves x 9 KiB fit a CU's 160 KiB
  d.wave_slots = (uint32_t)c->cu_count * 16;
  d.hbm_slots = (uint32_t)c->cu_count * 8;
  d.hbm_slot_mib = 32;
  size_t free_b = 0, total_b = 0;
  d.hbm_budget_mib = 65536;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
    d.hbm_budget_mib = std::max<uint64_t>(1024, (uint64_t)(0.6 * (double)free_b) >> 20);
  d.block256_min_len = 0;  // by cost shape (plan_group)
  d.block1024_min_len = 16384;
  d.heavy_permille = 50;
  d.block_lds_ring = 1024;
  d.block_flight = 1;
  d.band_permille = 800;
  c->tun = merge_tuning(d, tuning);
  return c;
}

void phx_ctx_destroy(phx_ctx* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  if (PlanCache* C = c->plan_cache) {
    for (DevBuf* b : {&C->seq, &C->pairs, &C->out, &C->runs, &C->work, &C->work2, &C->ctr, &C->scratch,
                      &C->gather, &C->gidx, &C->arena, &C->rstage})
      dev_free(*b);
    if (C->ctr_host) (void)hipHostFree(C->ctr_host);
    if (C->pin) (void)hipHostFree(C->pin);
    for (hipEvent_t e : C->kev) (void)hipEventDestroy(e);
    for (hipEvent_t e : {C->ev_start, C->ev_end})
      if (e) (void)hipEventDestroy(e);
    delete C;
  }
  for (hipEvent_t e : c->join) (void)hipEventDestroy(e);
  for (int i = 0; i < kStreams; ++i) (void)hipStreamDestroy(c->streams[i]);
  delete c;
}

const char* phx_ctx_last_error(const phx_ctx* c) {
  if (!c) return "no context";
  const LastError& e = last_error();
  return e.ctx == c ? e.msg.c_str() : "";
}

uint64_t phx_cigar_bound(const phx_batch* b) {
  if (!b || b->engine == PHX_ENGINE_MYERS_DIST) return 0;
  uint64_t s = 0;
  for (uint32_t i = 0; i < b->n; ++i) s += (b->q_off[i + 1] - b->q_off[i]) + (b->t_off[i + 1] - b->t_off[i]) + 1;
  return s;
}

}  // extern "C"

namespace {

// phx_plan_create, optionally on a fixed device region (fixed_base /
// fixed_cap bytes) from which every device buffer is carved.
phx_plan* plan_create_batch(phx_ctx* ctx, const phx_batch* b, uint8_t* fixed_base, uint64_t fixed_cap,
                            bool oneshot = false) {
  if (!ctx || !b || b->engine < 0 || b->engine > PHX_ENGINE_MYERS_DIST) {
    set_error(ctx, "bad batch");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) { set_error(ctx, "hipSetDevice failed", PHX_E_DEVICE); return nullptr; }
  phx_plan* P = new phx_plan;
  P->ctx = ctx;
  P->fixed_base = fixed_base;
  P->fixed_cap = fixed_base ? fixed_cap : 0;
  if (oneshot && !fixed_base) cache_take(P);
  P->tun = ctx->tun;
  P->engine = b->engine;
  P->costs = b->costs;
  P->cap = b->cap;
  P->n = b->n;
  const uint32_t n = b->n;
  if (b->engine == PHX_ENGINE_AFFINE) {
    const phx_costs& c = b->costs;
    if (c.mismatch < 1 || c.gap_extend < 1 || c.gap_open < 0) {
      set_error(ctx, "affine costs need mismatch >= 1, gap_extend >= 1, gap_open >= 0");
      delete P;
      return nullptr;
    }
    P->ring_slots = ring_slots(b->engine, c);
    if (P->ring_slots > kRingSlotsMax) {
      set_error(ctx, "affine costs span more score frames than the device ring holds");
      delete P;
      return nullptr;
    }
  }
  P->seed_q.assign(n, 0);
  P->seed_t.assign(n, 0);
  P->qlen.assign(n, 0);
  P->tlen.assign(n, 0);
  P->host_status.assign(n, ST_PENDING);
  P->pairs.assign(n, PairDesc{});
  // Pass 1 (host threads): seeds, lengths, encodings; then word and run
  // offsets in one serial prefix pass; pass 2 (host threads): pack.
  std::vector<uint8_t> b8(n, 0), too_long(n, 0);
  std::vector<uint64_t> q_word(n, 0), t_word(n, 0);
  parallel_for(n, [&](size_t i) {
    const uint64_t qfull = b->q_off[i + 1] - b->q_off[i], tfull = b->t_off[i + 1] - b->t_off[i];
    uint64_t sq = 0, st = 0;
    if (b->engine != PHX_ENGINE_MYERS_GLOBAL) {
      sq = b->seed_q ? b->seed_q[i] : 0;
      st = b->seed_t ? b->seed_t[i] : 0;
    }
    P->seed_q[i] = sq;
    P->seed_t[i] = st;
    if (sq > qfull || st > tfull) {
      P->host_status[i] = ST_INVALID_INPUT;  // lib.rs:107-111
      return;
    }
    const uint64_t qn = qfull - sq, tn = tfull - st;
    if (qn > (uint64_t)kMaxLen || tn > (uint64_t)kMaxLen) {
      too_long[i] = 1;
      return;
    }
    P->qlen[i] = qn;
    P->tlen[i] = tn;
    b8[i] = !(is_acgt(b->q_seq + b->q_off[i] + sq, qn) && is_acgt(b->t_seq + b->t_off[i] + st, tn));
  });
  uint64_t words = 0, runs = 0;
  for (uint32_t i = 0; i < n; ++i) {
    if (too_long[i]) {
      set_error(ctx, "sequence longer than the device format supports (2^27 - 2)");
      delete P;
      return nullptr;
    }
    if (P->host_status[i] == ST_INVALID_INPUT) continue;
    const uint64_t qn = P->qlen[i], tn = P->tlen[i];
    PairDesc& pd = P->pairs[i];
    pd.qn = (int32_t)qn;
    pd.tn = (int32_t)tn;
    pd.flags = b8[i] ? PF_BYTES : 0;
    q_word[i] = words;
    words += words_for(qn, b8[i]);
    t_word[i] = words;
    words += words_for(tn, b8[i]);
    pd.run_cap = has_cigar(b->engine) ? (uint32_t)(qn + tn + 1) : 0;
    pd.run_off = runs;
    runs += pd.run_cap;
    P->input_bytes += b8[i] ? qn + tn : (qn + 3) / 4 + (tn + 3) / 4;
  }
  P->seq_words = words;
  P->runs_total = runs;
  std::vector<uint64_t> packed(words ? words : 1);
  parallel_for(n, [&](size_t i) {
    if (P->host_status[i] == ST_INVALID_INPUT) return;
    pack_into(packed.data() + q_word[i], b->q_seq + b->q_off[i] + P->seed_q[i], P->qlen[i], b8[i]);
    pack_into(packed.data() + t_word[i], b->t_seq + b->t_off[i] + P->seed_t[i], P->tlen[i], b8[i]);
  });
  std::vector<const uint8_t*> qh(n, nullptr), th(n, nullptr);
  for (uint32_t i = 0; i < n; ++i) {
    if (P->host_status[i] != ST_PENDING) continue;
    qh[i] = b->q_seq + b->q_off[i] + P->seed_q[i];
    th[i] = b->t_seq + b->t_off[i] + P->seed_t[i];
  }
  return plan_setup(P, qh, th, &packed, &q_word, &t_word);
}

}  // namespace

extern "C" {

phx_plan* phx_plan_create(phx_ctx* ctx, const phx_batch* b) { return plan_create_batch(ctx, b, nullptr, 0); }

int phx_plan_run(phx_plan* P) {
  if (!P) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  P->out_valid = false;
  P->last_launches = 0;
  P->launch_rec.clear();
  P->rerun_lists.clear();
  P->rerun_lists.reserve(10);  // records point into it: no reallocation within a run (8 rounds + 1)
  P->start_ev = nullptr;
  P->end_ev = nullptr;
  P->ms_ready = P->kernel_ms_ready = false;
  int rc = is_wfa(P->engine) ? run_wfa(P) : run_myers(P);
  if (rc != PHX_OK) return rc;
  if (!P->end_ev && P->last_launches > 0 && !P->fixed()) {
    if (plan_event(P, &P->ev_end) != PHX_OK) return PHX_E_DEVICE;
    HIPCHK(hipEventRecord(P->ev_end, ctx->stream));
    P->end_ev = P->ev_end;
  }
  if (P->end_ev) HIPCHK(hipEventSynchronize(P->end_ev));
  else HIPCHK(hipStreamSynchronize(ctx->stream));
  if (P->graph_wanted) {  // outside the timed region
    P->graph_wanted = false;
    return build_graph(P);
  }
  return PHX_OK;
}

int phx_plan_use_graph(phx_plan* P, int enable) {
  if (!P || enable < 0 || enable > 1) return PHX_E_ARG;
  std::lock_guard<std::mutex> lk(P->ctx->mu);
  P->use_graph = enable != 0 && is_wfa(P->engine);
  if (!P->use_graph && P->gexec) {
    (void)hipGraphExecDestroy(P->gexec);
    P->gexec = nullptr;
  }
  return PHX_OK;
}

// Device time of the last run: from its start (its first launch's start
// event when it is one launch) to its end.
double phx_plan_last_ms(const phx_plan* P) {
  if (!P) return 0.0;
  if (!P->ms_ready) {
    float ms = 0;
    const hipEvent_t start = P->start_ev ? P->start_ev : P->kev.empty() ? nullptr : P->kev[0];
    if (P->last_launches > 0 && start && P->end_ev &&
        hipEventElapsedTime(&ms, start, P->end_ev) != hipSuccess)
      ms = 0;
    P->last_ms = ms;
    P->ms_ready = true;
  }
  return P->last_ms;
}

// Sum of the kernel times of the last run (events around each launch).
double phx_plan_last_kernel_ms(const phx_plan* P) {
  if (!P) return 0.0;
  if (!P->kernel_ms_ready) {
    double kms = 0;
    for (uint32_t i = 0; i < P->last_launches && 2 * (size_t)i + 1 < P->kev.size(); ++i) {
      float k = 0;
      if (hipEventElapsedTime(&k, P->kev[2 * i], P->kev[2 * i + 1]) == hipSuccess) kms += k;
    }
    P->last_kernel_ms = kms;
    P->kernel_ms_ready = true;
  }
  return P->last_kernel_ms;
}
uint32_t phx_plan_last_launches(const phx_plan* P) { return P ? P->last_launches : 0; }

static int fetch_out(phx_plan* P);

int phx_plan_last_launch(const phx_plan* Pc, uint32_t i, phx_launch_info* out) {
  phx_plan* P = const_cast<phx_plan*>(Pc);
  if (!P || !out || i >= P->launch_rec.size()) return PHX_E_ARG;
  std::lock_guard<std::mutex> lk(P->ctx->mu);
  if (fetch_out(P) != PHX_OK) return PHX_E_DEVICE;
  const phx_plan::LaunchRec& r = P->launch_rec[i];
  std::memset(out, 0, sizeof(*out));
  out->tier = r.tier;
  out->pairs = (uint32_t)r.count;
  out->grid = r.grid;
  out->storage = r.storage;
  if (i < P->last_launches && 2 * (size_t)i + 1 < P->kev.size()) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, P->kev[2 * i], P->kev[2 * i + 1]) == hipSuccess) out->ms = ms;
  }
  uint64_t bytes = 0, cells = 0, steps = 0;
  for (uint64_t j = r.off; j < r.off + r.count; ++j) {
    const uint32_t pi = (*r.list)[j];
    const PairDesc& pd = P->pairs[pi];
    const bool b8 = pd.flags & PF_BYTES;
    bytes += b8 ? (uint64_t)pd.qn + pd.tn : ((uint64_t)pd.qn + 3) / 4 + ((uint64_t)pd.tn + 3) / 4;
    bytes += P->engine == PHX_ENGINE_MYERS_DIST ? 16 : 12 + 4ull * (uint64_t)std::max(P->out_h[pi].nruns, 0);
    cells += (uint64_t)std::max(P->out_h[pi].cells, 0);
    steps += (uint64_t)std::max(P->out_h[pi].steps, 0);
  }
  out->algorithmic_bytes = bytes;
  out->cells = cells;
  out->steps = steps;
  return PHX_OK;
}

int phx_plan_set_tuning(phx_plan* P, const phx_tuning* t) {
  if (!P || !t) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  const uint32_t old_pm = P->tun.band_permille;
  P->tun = merge_tuning(P->tun, t);
  if (!is_wfa(P->engine)) return PHX_OK;
  if (P->gexec) {  // the launch sequence changes with the grouping
    (void)hipGraphExecDestroy(P->gexec);
    P->gexec = nullptr;
  }
  if (P->banded && P->tun.band_permille != old_pm && !P->fixed()) {
    apply_bounds(P);
    if (P->n) HIPCHK(hipMemcpy(P->d_pairs.p, P->pairs.data(), sizeof(PairDesc) * P->n, hipMemcpyHostToDevice));
  }
  plan_group(P);
  const int rc = upload_work(P);
  return rc != PHX_OK ? rc : restage_rows(P);
}

int phx_plan_get_tuning(const phx_plan* P, phx_tuning* t) {
  if (!P || !t) return PHX_E_ARG;
  *t = P->tun;
  return PHX_OK;
}

int phx_plan_autotune(phx_plan* P, uint32_t max_trials, phx_tuning* best_out, double* best_ms_out) {
  if (!P) return PHX_E_ARG;
  if (!is_wfa(P->engine)) {  // Myers has no tier choices
    if (best_out) *best_out = P->tun;
    if (best_ms_out) *best_ms_out = 0;
    return PHX_OK;
  }
  const uint32_t cu = (uint32_t)P->ctx->cu_count;
  uint32_t trials = 0;
  // Time one tuning on the whole plan.  Short runs take the best of several;
  // a run during which the scratch arena had to grow is repeated (the
  // allocation would otherwise be timed).
  auto measure = [&](const phx_tuning& t, double* ms) -> int {
    int rc = phx_plan_set_tuning(P, &t);
    if (rc != PHX_OK) return rc;
    double best = 1e300;
    for (int r = 0; r < 6; ++r) {
      const size_t arena = P->d_arena.bytes;
      rc = phx_plan_run(P);
      if (rc != PHX_OK) return rc;
      if (P->d_arena.bytes != arena) continue;
      best = std::min(best, phx_plan_last_ms(P));
      if (best > 20.0 || r >= 3) break;
    }
    *ms = best;
    ++trials;
    return PHX_OK;
  };
  phx_tuning best = P->tun;
  double best_ms = 0;
  int rc = measure(best, &best_ms);
  if (rc != PHX_OK) return rc;
  auto consider = [&](phx_tuning t) -> int {
    if (max_trials && trials >= max_trials) return PHX_OK;
    double ms = 0;
    const int r = measure(t, &ms);
    if (r != PHX_OK) return r;
    if (ms < best_ms) {
      best_ms = ms;
      best = P->tun;
    }
    return PHX_OK;
  };
  auto pairs_in = [&](int tier) {
    size_t c = 0;
    for (size_t g = 0; g < P->work.size(); ++g)
      if (P->group_tier[g] == tier) c += P->work[g].size();
    return c;
  };
  const size_t lds_pairs = pairs_in(TIER_LDS);
  const bool has_hbm = lds_pairs < P->order.size();
  // Stage 1: tier edges (threads per pair by length), then the heavy share.
  if (has_hbm) {
    static const uint32_t kEdges[5][2] = {{8192, 16384}, {8192, 32768}, {16384, 32768},
                                          {16384, 65536}, {32768, 65536}};
    for (const auto& e : kEdges) {
      phx_tuning t = best;
      t.block256_min_len = e[0];
      t.block1024_min_len = e[1];
      if ((rc = consider(t)) != PHX_OK) return rc;
    }
    for (uint32_t heavy : {20u, 100u}) {
      phx_tuning t = best;
      t.heavy_permille = heavy;
      if ((rc = consider(t)) != PHX_OK) return rc;
    }
  }
  rc = phx_plan_set_tuning(P, &best);
  if (rc != PHX_OK) return rc;
  // Stage 2: pairs in flight of the workgroup tiers (a second workgroup per
  // CU hides memory latency when a group has pairs to spare), then of the
  // HBM wave tier.
  if (pairs_in(TIER_BLOCK256) + pairs_in(TIER_BLOCK512) + pairs_in(TIER_BLOCK1024) + pairs_in(TIER_BLOCK64) >
      2 * (size_t)cu) {
    phx_tuning t = best;
    t.block_flight = 2;
    if ((rc = consider(t)) != PHX_OK) return rc;
  }
  if (pairs_in(TIER_WAVE))
    for (uint32_t k : {4u, 16u}) {
      phx_tuning t = best;
      t.hbm_slots = cu * k;
      if ((rc = consider(t)) != PHX_OK) return rc;
    }
  // Stage 3: LDS per wave vs waves in flight, when LDS-tier pairs dominate.
  if (2 * lds_pairs >= P->order.size()) {
    static const uint32_t kLds[5][2] = {{7, 20}, {6, 24}, {5, 28}, {4, 32}, {12, 12}};
    for (const auto& l : kLds) {
      phx_tuning t = best;
      t.wave_lds_kib = l[0];
      t.wave_slots = cu * l[1];
      if ((rc = consider(t)) != PHX_OK) return rc;
    }
  }
  rc = phx_plan_set_tuning(P, &best);
  if (rc != PHX_OK) return rc;
  if (best_out) *best_out = best;
  if (best_ms_out) *best_ms_out = best_ms;
  return PHX_OK;
}

static int fetch_out(phx_plan* P) {
  if (P->out_valid) return PHX_OK;
  P->out_h.resize(P->n);
  const int rc = d2h(P, P->out_h.data(), P->d_out.p, sizeof(PairOut) * P->n);
  if (rc != PHX_OK) return rc;
  P->out_valid = true;
  return PHX_OK;
}

uint64_t phx_plan_algorithmic_bytes(const phx_plan* Pc) {
  phx_plan* P = const_cast<phx_plan*>(Pc);
  if (!P) return 0;
  std::lock_guard<std::mutex> lk(P->ctx->mu);
  if (fetch_out(P) != PHX_OK) return 0;
  uint64_t bytes = P->input_bytes;
  for (uint32_t i = 0; i < P->n; ++i) {
    if (P->host_status[i] != ST_PENDING) continue;
    bytes += P->engine == PHX_ENGINE_MYERS_DIST ? 16 : 12 + 4ull * (uint64_t)P->out_h[i].nruns;
  }
  return bytes;
}

int phx_plan_results(phx_plan* P, phx_results* r) {
  if (!P || !r || !r->aln || !r->run_off) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  int rc = fetch_out(P);
  if (rc != PHX_OK) return rc;
  // Per-pair fields and the compacted run layout.
  std::vector<uint64_t> src(P->n, 0);
  uint64_t pos = 0;
  int ret = PHX_OK;
  for (uint32_t i = 0; i < P->n; ++i) {
    phx_alignment& a = r->aln[i];
    std::memset(&a, 0, sizeof(a));
    r->run_off[i] = pos;
    if (P->host_status[i] != ST_PENDING) {
      a.status = P->host_status[i];
      continue;
    }
    const PairOut& o = P->out_h[i];
    if (o.status != ST_OK && o.status != ST_ALIGNMENT_FAILED) {
      a.status = PHX_E_INTERNAL;
      ret = PHX_E_INTERNAL;
      set_error(ctx, "pair " + std::to_string(i) + " ended with device status " + std::to_string(o.status));
      continue;
    }
    a.status = o.status;
    if (o.status != ST_OK) continue;
    a.score = o.score;
    a.edit_distance = o.edit_distance;
    a.query_start = (int64_t)P->seed_q[i];
    a.query_end = (int64_t)(P->seed_q[i] + P->qlen[i]);
    a.target_start = (int64_t)P->seed_t[i];
    a.target_end = (int64_t)P->seed_t[i] + o.t_end;
    if (P->engine == PHX_ENGINE_MYERS_GLOBAL) {
      a.query_start = a.target_start = 0;
      a.query_end = (int64_t)P->qlen[i];
      a.target_end = (int64_t)P->tlen[i];
    }
    if (has_cigar(P->engine) && o.nruns > 0) {
      src[i] = P->pairs[i].run_off + (uint64_t)o.run_start;
      pos += (uint64_t)o.nruns;
    }
  }
  r->run_off[P->n] = pos;
  if (ret != PHX_OK) return ret;
  if (!has_cigar(P->engine) || pos == 0) return PHX_OK;
  if (pos > r->runs_cap || !r->runs) {
    set_error(ctx, "CIGAR arena too small");
    return PHX_E_CAPACITY;
  }
  // Gather the used runs on the device, then one copy of exactly those words.
  std::vector<uint64_t> idx(2 * (size_t)P->n + 1);
  std::copy(src.begin(), src.end(), idx.begin());
  std::copy(r->run_off, r->run_off + P->n + 1, idx.begin() + P->n);
  // (Fixed plans: the run has finished, so its transient region holds these.)
  rc = plan_transient(P, P->d_gidx, idx.size() * 8);
  if (rc == PHX_OK) rc = plan_transient(P, P->d_gather, pos * 4, idx.size() * 8);
  if (rc != PHX_OK) return rc;
  // Staging sized for both directions first, so the download below cannot
  // reallocate it under the pending upload.
  if (uint8_t* pin = staging(P, std::max(idx.size() * 8, (size_t)pos * 4))) {
    std::memcpy(pin, idx.data(), idx.size() * 8);
    HIPCHK(hipMemcpyAsync(P->d_gidx.p, pin, idx.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  } else {
    HIPCHK(hipMemcpyAsync(P->d_gidx.p, idx.data(), idx.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  }
  const uint64_t* d_src = (const uint64_t*)P->d_gidx.p;
  HIPCHK(launch_gather_runs((const uint32_t*)P->d_runs.p, d_src, d_src + P->n,
                            (uint32_t*)P->d_gather.p, P->n, ctx->stream));
  return d2h(P, r->runs, P->d_gather.p, pos * 4);
}

void phx_plan_destroy(phx_plan* P) {
  if (!P) return;
  (void)hipStreamSynchronize(P->ctx->stream);
  if (P->gexec) {
    (void)hipGraphExecDestroy(P->gexec);
    P->gexec = nullptr;
  }
  if (P->cached) cache_return(P);  // leaves the plan with nothing to free
  if (!P->fixed()) {  // a fixed plan's buffers are carves of the scratch array
    for (DevBuf* b : {&P->d_seq, &P->d_pairs, &P->d_out, &P->d_runs, &P->d_work, &P->d_work2, &P->d_ctr,
                      &P->d_scratch, &P->d_gather, &P->d_gidx, &P->d_rstage})
      dev_free(*b);
    dev_free(P->d_arena);
    dev_free(P->d_bounds);
  }
  if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
  for (hipEvent_t e : P->kev) (void)hipEventDestroy(e);
  for (hipEvent_t e : {P->ev_start, P->ev_end})
    if (e) (void)hipEventDestroy(e);
  if (P->ctr_host && !P->fixed()) (void)hipHostFree(P->ctr_host);
  delete P;
}

int phx_align_batch(phx_ctx* ctx, const phx_batch* b, phx_results* out) {
  if (!ctx || !b || !out) return PHX_E_ARG;
  phx_plan* P = plan_create_batch(ctx, b, nullptr, 0, true);
  if (!P) return last_error_code(ctx);
  int rc = phx_plan_run(P);
  if (rc == PHX_OK) rc = phx_plan_results(P, out);
  phx_plan_destroy(P);
  return rc;
}

int phx_align_batch_noalloc(phx_ctx* ctx, const phx_batch* b, phx_results* out) {
  if (!ctx || !b || !out) return PHX_E_ARG;
  if (!ctx->noalloc_scratch)
    return fail(ctx, PHX_E_DEVICE, "the allocation-free scratch has no address on this device");
  // Every call on this GPU shares its scratch array: calls are serialised.
  NoAllocDevice& d = noalloc_device(ctx->device);
  std::lock_guard<std::mutex> busy(d.busy);
  phx_plan* P = plan_create_batch(ctx, b, ctx->noalloc_scratch, PHX_NOALLOC_SCRATCH_BYTES);
  if (!P) return last_error_code(ctx);
  int rc = phx_plan_run(P);
  if (rc == PHX_OK) rc = phx_plan_results(P, out);
  phx_plan_destroy(P);
  return rc;
}

size_t phx_format_cigar(const uint32_t* runs, uint64_t nruns, char* buf, size_t cap) {
  static const char kOps[4] = {'M', 'X', 'I', 'D'};
  size_t len = 0;
  char tmp[16];
  for (uint64_t i = 0; i < nruns; ++i) {
    const int k = snprintf(tmp, sizeof(tmp), "%u%c", runs[i] >> 4, kOps[runs[i] & 3]);
    for (int j = 0; j < k; ++j) {
      if (buf && len + 1 < cap) buf[len] = tmp[j];
      ++len;
    }
  }
  if (buf && cap) buf[std::min(len, cap - 1)] = 0;
  return len;
}

uint64_t phx_format_cigars(const uint32_t* runs, const uint64_t* run_off, uint64_t n, char* buf,
                           uint64_t cap, uint64_t* text_off) {
  static const char kOps[4] = {'M', 'X', 'I', 'D'};
  if (!run_off || (!runs && n && run_off[n] > run_off[0])) return 0;
  uint64_t len = 0;
  char digits[12];
  for (uint64_t i = 0; i < n; ++i) {
    if (text_off) text_off[i] = len;
    for (uint64_t r = run_off[i]; r < run_off[i + 1]; ++r) {
      uint32_t v = runs[r] >> 4;
      int nd = 0;
      do {
        digits[nd++] = (char)('0' + v % 10);
        v /= 10;
      } while (v);
      if (buf && len + (uint64_t)nd + 1 <= cap) {
        for (int d = nd - 1; d >= 0; --d) buf[len++] = digits[d];
        buf[len++] = kOps[runs[r] & 3];
      } else {
        len += (uint64_t)nd + 1;
      }
    }
  }
  if (text_off) text_off[n] = len;
  return len;
}

// Implementation the calling thread's last per-pair call ran (upstream keeps
// the same per thread in LAST_IMPL).
static thread_local const char* t_last_impl = "";

static int one_pair(phx_ctx* ctx, int engine, const uint8_t* q, size_t qlen, const uint8_t* t,
                    size_t tlen, uint64_t sq, uint64_t st, phx_costs costs, phx_alignment* out,
                    char* cigar, size_t cigar_cap, size_t* cigar_len) {
  if (!ctx || !out) return PHX_E_ARG;
  t_last_impl = ctx->impl_name;
  static const uint8_t kEmpty = 0;
  const uint64_t qo[2] = {0, qlen}, to[2] = {0, tlen};
  phx_batch b{};
  b.n = 1;
  b.engine = engine;
  b.costs = costs;
  b.cap = -1;
  b.q_seq = q ? q : &kEmpty;
  b.q_off = qo;
  b.t_seq = t ? t : &kEmpty;
  b.t_off = to;
  b.seed_q = &sq;
  b.seed_t = &st;
  std::vector<uint32_t> runs(qlen + tlen + 2);
  uint64_t roff[2] = {0, 0};
  phx_results r{out, runs.data(), runs.size(), roff};
  const int rc = phx_align_batch(ctx, &b, &r);
  if (rc != PHX_OK) return rc;
  const size_t len = phx_format_cigar(runs.data(), roff[1], cigar, cigar_cap);
  if (cigar_len) *cigar_len = len;
  return out->status;
}

const char* phx_last_impl(void) { return t_last_impl; }

int phx_wfa_extend(phx_ctx* ctx, const uint8_t* q, size_t qlen, const uint8_t* t, size_t tlen,
                   uint64_t seed_q, uint64_t seed_t, phx_alignment* out, char* cigar,
                   size_t cigar_cap, size_t* cigar_len) {
  return one_pair(ctx, PHX_ENGINE_LINEAR, q, qlen, t, tlen, seed_q, seed_t, phx_costs{0, 0, 0}, out,
                  cigar, cigar_cap, cigar_len);
}

int phx_wfa_extend_affine(phx_ctx* ctx, const uint8_t* q, size_t qlen, const uint8_t* t,
                          size_t tlen, uint64_t seed_q, uint64_t seed_t, phx_costs costs,
                          phx_alignment* out, char* cigar, size_t cigar_cap, size_t* cigar_len) {
  return one_pair(ctx, PHX_ENGINE_AFFINE, q, qlen, t, tlen, seed_q, seed_t, costs, out, cigar,
                  cigar_cap, cigar_len);
}

int phx_myers_extend(phx_ctx* ctx, const uint8_t* q, size_t qlen, const uint8_t* t, size_t tlen,
                     uint64_t seed_q, uint64_t seed_t, phx_alignment* out, char* cigar,
                     size_t cigar_cap, size_t* cigar_len) {
  return one_pair(ctx, PHX_ENGINE_MYERS, q, qlen, t, tlen, seed_q, seed_t, phx_costs{0, 0, 0}, out,
                  cigar, cigar_cap, cigar_len);
}

int phx_myers_edit_distance(phx_ctx* ctx, const uint8_t* q, size_t qlen, const uint8_t* t,
                            size_t tlen, int64_t* edit_distance, char* cigar, size_t cigar_cap,
                            size_t* cigar_len) {
  phx_alignment a{};
  const int rc = one_pair(ctx, PHX_ENGINE_MYERS_GLOBAL, q, qlen, t, tlen, 0, 0, phx_costs{0, 0, 0},
                          &a, cigar, cigar_cap, cigar_len);
  if (edit_distance) *edit_distance = a.edit_distance;
  return rc;
}

}  // extern "C"
