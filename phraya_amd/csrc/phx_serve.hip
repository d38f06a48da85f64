// phx_serve.hip -- request server: continuous batching of WFA extension
// requests on one GPU.
//
// A scheduler thread owns everything device-side.  Submitted requests wait in
// one FIFO per kernel tier (LDS tier for qn + tn <= 2048, then wave / 256 /
// 1024 threads per pair).  Each tier has its own launch lane(s): a stream, an
// event, device buffers and pinned staging.  Whenever a lane is free the
// scheduler admits the oldest waiting requests of its tier, reserving for
// each HBM-tier request its own history + ring region in a resident arena
// (first fit; the size is the request's estimated step count turned into
// history cells), packs the queries (targets are windows of the resident
// reference or explicit), and launches: upload, kernel (which writes results
// and CIGAR runs into the lane's pinned host buffers) and an event, all on the
// lane's stream.  Lanes complete independently, so short
// reads keep flowing while long ones run.  A request that outgrows its tier
// (LDS overflow, or history larger than estimated) returns to the front of
// the next tier's queue with its exact step count; results do not depend on
// the route.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "phraya_hip.h"
#include "phx_device.h"
#include "phx_host.h"

using namespace phx;
using namespace phx::host;

namespace {

using Clock = std::chrono::steady_clock;

double now_ms() {
  return std::chrono::duration<double, std::milli>(Clock::now().time_since_epoch()).count();
}

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct HBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

bool grow_dev(DBuf& b, size_t need) {
  if (need == 0) need = 8;
  if (b.bytes >= need) return true;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  const size_t sz = std::max(need, need + need / 2);
  if (hipMalloc(&b.p, sz) != hipSuccess) return false;
  b.bytes = sz;
  return true;
}

bool grow_host(HBuf& b, size_t need) {
  if (need == 0) need = 8;
  if (b.bytes >= need) return true;
  if (b.p) (void)hipHostFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  const size_t sz = std::max(need, need + need / 2);
  if (hipHostMalloc(&b.p, sz, hipHostMallocDefault) != hipSuccess) return false;
  b.bytes = sz;
  return true;
}

// First-fit allocator over [0, cap) with 256-byte granules; neighbours merge
// on release.
class Arena {
 public:
  void init(uint64_t cap) {
    cap_ = cap;
    free_.clear();
    if (cap) free_[0] = cap;
    used_ = peak_ = 0;
  }
  bool alloc(uint64_t bytes, uint64_t* off) {
    bytes = (bytes + 255) & ~(uint64_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const uint64_t rest = it->second - bytes, at = it->first + bytes;
      free_.erase(it);
      if (rest) free_[at] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(uint64_t off, uint64_t bytes) {
    bytes = (bytes + 255) & ~(uint64_t)255;
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  uint64_t cap() const { return cap_; }
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }

 private:
  uint64_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<uint64_t, uint64_t> free_;
};

struct Req {
  uint64_t tag = 0;
  std::vector<uint8_t> q, t;  // suffixes after the seed (t empty for a reference window)
  bool ref = false;
  uint64_t ref_pos = 0;       // window start after the seed
  int32_t qn = 0, tn = 0;
  uint64_t seed_q = 0, seed_t = 0;
  bool b8 = false;
  double t_submit = 0;
  int tier = 0;
  int32_t est = 0;    // estimated score steps (HBM tiers)
  int32_t steps = 0;  // exact score steps once known
  int attempts = 0;
  // Arena reservation while in flight.
  uint64_t res_off = 0, res_bytes = 0, hist_bytes = 0;
  uint32_t ring_log2 = 0;
};

struct Done {
  uint64_t tag = 0;
  phx_alignment aln{};
  std::vector<uint32_t> runs;
  double latency_ms = 0;
};

struct Lane {
  int tier = 0;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Req> reqs;
  std::vector<uint64_t> run_off;
  uint64_t runs_total = 0;
  DBuf d_seq, d_pairs, d_work, d_ctr;
  // Pinned staging; results and CIGAR runs are written by the kernel straight
  // into h_out / h_runs (device-visible host memory), so only the used runs
  // ever cross the bus.
  HBuf h_seq, h_pairs, h_out, h_runs, h_work;
};

}  // namespace

struct phx_server {
  phx_ctx* ctx = nullptr;
  int device = 0;
  int engine = PHX_ENGINE_AFFINE;
  int mode = ENG_AFFINE;
  phx_costs costs{};
  int32_t cap = -1;
  int32_t R = 2;
  uint32_t max_group = 8192;
  uint32_t wave_lds_kib = 9, wave_slots = 4096;
  // Resident reference: 2-bit and 8-bit copies, positions of non-ACGT bytes.
  DBuf d_ref2, d_ref8;
  std::vector<uint8_t> ref_host;  // for the memory estimates of reference windows
  uint64_t ref_len = 0;
  std::vector<uint64_t> ref_bad;
  // Wavefront state of HBM-tier requests.
  DBuf d_arena;
  Arena arena;
  // Queues.  incoming / done / stop are shared with the API calls (mu);
  // waiting and lanes belong to the scheduler thread.
  std::mutex mu;
  std::condition_variable cv_sched, cv_done;
  std::deque<Req> incoming;
  std::deque<Done> done;
  bool stop = false;
  std::deque<Req> waiting[4];
  std::vector<Lane> lanes;
  std::thread thr;
  // Statistics.
  std::atomic<uint64_t> submitted{0}, completed{0}, launches{0}, reruns{0};
  std::atomic<uint64_t> n_waiting{0}, n_flight{0}, arena_used{0}, arena_peak{0};
};

namespace {

bool ref_window_acgt(const phx_server* S, uint64_t pos, uint64_t len) {
  auto it = std::lower_bound(S->ref_bad.begin(), S->ref_bad.end(), pos);
  return it == S->ref_bad.end() || *it >= pos + len;
}

int initial_tier(const phx_server* S, int32_t qn, int32_t tn) {
  const uint64_t L = (uint64_t)qn + (uint64_t)tn;
  if (L <= kWaveMaxLen && S->R <= kWaveRingMax) return TIER_LDS;
  return hbm_tier_for(L);
}

// Steps the request may take in its tier: the LDS tier's step table, else the
// estimate its history slot is sized for (or the exact count once known).
int32_t step_bound(const Req& r, int tier, const phx_server* S) {
  if (tier == TIER_LDS) return (int32_t)kWaveSteps;
  return r.steps > 0 ? r.steps : r.est > 0 ? r.est : estimate_steps(S->engine, S->costs, r.qn, r.tn);
}

// CIGAR run slots: every edit run costs at least one score step and M runs
// only separate edit runs, so s steps give at most 2s + 1 runs.  A pair that
// needs more (it ran past the estimate) reports an overflow and is re-run
// with its exact step count.
uint64_t run_cap_of(const Req& r, int tier, const phx_server* S) {
  const uint64_t full = (uint64_t)r.qn + (uint64_t)r.tn + 1;
  return std::min<uint64_t>(full, 2 * (uint64_t)step_bound(r, tier, S) + 1);
}

Done make_error(const Req& r, int32_t status, double t) {
  Done d;
  d.tag = r.tag;
  d.aln.status = status;
  d.latency_ms = t - r.t_submit;
  return d;
}

void deliver(phx_server* S, std::vector<Done>& fin) {
  if (fin.empty()) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Done& d : fin) S->done.push_back(std::move(d));
  }
  S->completed += fin.size();
  fin.clear();
  S->cv_done.notify_all();
}

// Admit the oldest waiting requests of the lane's tier and launch them.
// Returns false when nothing could be launched.
bool launch(phx_server* S, Lane& L, std::vector<Done>& fin) {
  std::deque<Req>& W = S->waiting[L.tier];
  std::vector<Req> g;
  const double t = now_ms();
  while (!W.empty() && g.size() < S->max_group) {
    Req& r = W.front();
    if (L.tier != TIER_LDS) {
      const int32_t steps = step_bound(r, L.tier, S);
      uint32_t wcap = 0;
      const uint64_t ring_w = ring_words_for(ring_frames(S->engine, S->costs), (uint64_t)r.qn + r.tn, &wcap);
      r.hist_bytes = hist_slot_bytes(steps, r.qn, r.tn);
      r.res_bytes = r.hist_bytes + ((ring_w * 4 + 255) & ~(uint64_t)255);
      r.ring_log2 = log2u(wcap);
      if (r.res_bytes > S->arena.cap()) {  // can never be admitted
        fin.push_back(make_error(r, PHX_E_NOMEM, t));
        W.pop_front();
        continue;
      }
      if (!S->arena.alloc(r.res_bytes, &r.res_off)) break;  // wait for memory
    }
    g.push_back(std::move(r));
    W.pop_front();
  }
  if (g.empty()) return false;
  const uint32_t n = (uint32_t)g.size();
  // Pack queries (and explicit targets).
  std::vector<uint64_t> qw(n), tw(n);
  uint64_t words = 0, runs = 0;
  L.run_off.assign(n, 0);
  for (uint32_t i = 0; i < n; ++i) {
    qw[i] = words;
    words += words_for(g[i].qn, g[i].b8);
    if (!g[i].ref) {
      tw[i] = words;
      words += words_for(g[i].tn, g[i].b8);
    }
    L.run_off[i] = runs;
    runs += run_cap_of(g[i], L.tier, S);
  }
  L.runs_total = runs;
  bool ok = grow_dev(L.d_seq, words * 8) && grow_dev(L.d_pairs, n * sizeof(PairDesc)) &&
            grow_dev(L.d_work, n * 4) && grow_dev(L.d_ctr, 64) && grow_host(L.h_seq, words * 8) &&
            grow_host(L.h_pairs, n * sizeof(PairDesc)) && grow_host(L.h_out, n * sizeof(PairOut)) &&
            grow_host(L.h_runs, runs * 4) && grow_host(L.h_work, n * 4);
  void* out_dev = nullptr;
  void* runs_dev = nullptr;
  ok = ok && hipHostGetDevicePointer(&out_dev, L.h_out.p, 0) == hipSuccess &&
       hipHostGetDevicePointer(&runs_dev, L.h_runs.p, 0) == hipSuccess;
  if (!ok) {
    for (Req& r : g) {
      if (L.tier != TIER_LDS) S->arena.release(r.res_off, r.res_bytes);
      fin.push_back(make_error(r, PHX_E_NOMEM, t));
    }
    return true;
  }
  uint64_t* hs = (uint64_t*)L.h_seq.p;
  PairDesc* hp = (PairDesc*)L.h_pairs.p;
  uint32_t* hw = (uint32_t*)L.h_work.p;
  const uint64_t* dseq = (const uint64_t*)L.d_seq.p;
  uint32_t max_seq_words = 2;
  for (uint32_t i = 0; i < n; ++i) {
    Req& r = g[i];
    pack_into(hs + qw[i], r.q.data(), r.q.size(), r.b8);
    PairDesc pd{};
    pd.q = dseq + qw[i];
    pd.q_shift = 0;
    if (r.ref) {
      const uint32_t per = r.b8 ? 8 : 32;
      pd.t = (const uint64_t*)(r.b8 ? S->d_ref8.p : S->d_ref2.p) + r.ref_pos / per;
      pd.t_shift = (uint32_t)(r.ref_pos % per);
    } else {
      pack_into(hs + tw[i], r.t.data(), r.t.size(), r.b8);
      pd.t = dseq + tw[i];
      pd.t_shift = 0;
    }
    pd.qn = r.qn;
    pd.tn = r.tn;
    pd.flags = r.b8 ? PF_BYTES : 0;
    pd.run_cap = (uint32_t)run_cap_of(r, L.tier, S);
    pd.run_off = L.run_off[i];
    if (L.tier != TIER_LDS) {
      uint8_t* base = (uint8_t*)S->d_arena.p + r.res_off;
      pd.hist = base;
      pd.hist_bytes = r.hist_bytes;
      pd.ring = (int32_t*)(base + r.hist_bytes);
      pd.ring_log2 = r.ring_log2;
    }
    hp[i] = pd;
    hw[i] = i;
    max_seq_words = std::max(max_seq_words, lds_seq_words(pd));
  }
  hipStream_t st = L.stream;
  WfaArgs a{};
  a.pairs = (const PairDesc*)L.d_pairs.p;
  a.work_list = (const uint32_t*)L.d_work.p;
  a.nwork = n;
  a.counters = (uint32_t*)L.d_ctr.p;
  a.cp = cost_params(S->costs, S->cap);
  a.out = (PairOut*)out_dev;
  a.runs = (uint32_t*)runs_dev;
  WaveGeom G{};
  uint32_t slots = n;
  if (L.tier == TIER_LDS) {
    G = lds_geometry(lds_tier_ring(S->engine, S->costs) ? ring_frames(S->engine, S->costs) : 0,
                     max_seq_words, S->wave_lds_kib);
    slots = std::min<uint32_t>(n, std::max<uint32_t>(kWaveWaves, S->wave_slots));
  }
  hipError_t e = hipMemcpyAsync(L.d_seq.p, L.h_seq.p, words * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(L.d_pairs.p, L.h_pairs.p, n * sizeof(PairDesc), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(L.d_work.p, L.h_work.p, n * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_zero((uint32_t*)L.d_ctr.p, 2, st);
  if (e == hipSuccess) e = launch_tier(L.tier, S->mode, a, G, slots, st);
  if (e == hipSuccess) e = hipEventRecord(L.done, st);
  if (e != hipSuccess) {
    (void)hipStreamSynchronize(st);
    for (Req& r : g) {
      if (L.tier != TIER_LDS) S->arena.release(r.res_off, r.res_bytes);
      fin.push_back(make_error(r, PHX_E_DEVICE, t));
    }
    return true;
  }
  L.reqs = std::move(g);
  L.busy = true;
  S->launches += 1;
  S->n_flight += n;
  S->arena_used = S->arena.used();
  S->arena_peak = S->arena.peak();
  return true;
}

// Collect a finished lane: deliver results, re-queue overflowed requests.
void finish(phx_server* S, Lane& L, bool device_ok, std::vector<Done>& fin) {
  const double t = now_ms();
  const PairOut* outs = (const PairOut*)L.h_out.p;
  const uint32_t* runs = (const uint32_t*)L.h_runs.p;
  std::vector<Req> again;
  const uint32_t n = (uint32_t)L.reqs.size();
  for (uint32_t i = 0; i < n; ++i) {
    Req& r = L.reqs[i];
    if (L.tier != TIER_LDS) S->arena.release(r.res_off, r.res_bytes);
    if (!device_ok) {
      fin.push_back(make_error(r, PHX_E_DEVICE, t));
      continue;
    }
    const PairOut& o = outs[i];
    if (o.status >= ST_OVERFLOW_HIST && o.status <= ST_OVERFLOW_RUNS && r.attempts < 6) {
      ++r.attempts;
      S->reruns += 1;
      if (L.tier == TIER_LDS) {
        r.tier = hbm_tier_for((uint64_t)r.qn + r.tn);
        const uint8_t* tp = r.ref ? S->ref_host.data() + r.ref_pos : r.t.data();
        r.est = estimate_steps_seq(S->engine, S->costs, r.q.data(), r.q.size(), tp, (size_t)r.tn);
      } else {
        // History overflow: the kernel kept going and o.steps is exact; a run
        // overflow also reports the exact count.
        r.steps = std::max(o.steps, r.steps + 1);
      }
      again.push_back(std::move(r));
      continue;
    }
    Done d;
    d.tag = r.tag;
    d.latency_ms = t - r.t_submit;
    phx_alignment& a = d.aln;
    if (o.status != ST_OK && o.status != ST_ALIGNMENT_FAILED) {
      a.status = PHX_E_INTERNAL;
    } else {
      a.status = o.status;
      if (o.status == ST_OK) {
        a.score = o.score;
        a.edit_distance = o.edit_distance;
        a.query_start = (int64_t)r.seed_q;
        a.query_end = (int64_t)(r.seed_q + (uint64_t)r.qn);
        a.target_start = (int64_t)r.seed_t;
        a.target_end = (int64_t)r.seed_t + o.t_end;
        if (o.nruns > 0) {
          const uint32_t* src = runs + L.run_off[i] + (uint64_t)o.run_start;
          d.runs.assign(src, src + o.nruns);
        }
      }
    }
    fin.push_back(std::move(d));
  }
  S->n_flight -= n;
  for (auto it = again.rbegin(); it != again.rend(); ++it) S->waiting[it->tier].push_front(std::move(*it));
  L.reqs.clear();
  L.busy = false;
  S->arena_used = S->arena.used();
}

void scheduler(phx_server* S) {
  (void)hipSetDevice(S->device);
  std::vector<Done> fin;
  for (;;) {
    bool progress = false;
    bool stopping = false;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      while (!S->incoming.empty()) {
        Req r = std::move(S->incoming.front());
        S->incoming.pop_front();
        S->waiting[r.tier].push_back(std::move(r));
        progress = true;
      }
      stopping = S->stop;
    }
    for (Lane& L : S->lanes) {
      if (!L.busy) continue;
      const hipError_t q = hipEventQuery(L.done);
      if (q == hipErrorNotReady) continue;
      finish(S, L, q == hipSuccess, fin);
      progress = true;
    }
    for (Lane& L : S->lanes)
      if (!L.busy && !S->waiting[L.tier].empty() && launch(S, L, fin)) progress = true;
    deliver(S, fin);
    uint64_t nw = 0;
    for (auto& w : S->waiting) nw += w.size();
    S->n_waiting = nw;
    bool busy = false;
    for (Lane& L : S->lanes) busy |= L.busy;
    if (stopping && !busy && nw == 0) {
      std::lock_guard<std::mutex> lk(S->mu);
      if (S->incoming.empty()) break;
      continue;
    }
    if (!progress) {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_sched.wait_for(lk, std::chrono::microseconds(busy ? 20 : 2000),
                           [&] { return !S->incoming.empty() || S->stop; });
    }
  }
}

void free_server(phx_server* S) {
  for (Lane& L : S->lanes) {
    for (DBuf* b : {&L.d_seq, &L.d_pairs, &L.d_work, &L.d_ctr})
      if (b->p) (void)hipFree(b->p);
    for (HBuf* b : {&L.h_seq, &L.h_pairs, &L.h_out, &L.h_runs, &L.h_work})
      if (b->p) (void)hipHostFree(b->p);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
  }
  for (DBuf* b : {&S->d_ref2, &S->d_ref8, &S->d_arena})
    if (b->p) (void)hipFree(b->p);
  delete S;
}

}  // namespace

extern "C" {

phx_server* phx_serve_start(phx_ctx* ctx, const phx_serve_config* cfg) {
  if (!ctx || !cfg) return nullptr;
  if (cfg->engine != PHX_ENGINE_AFFINE && cfg->engine != PHX_ENGINE_LINEAR) {
    set_error(ctx, "the server runs PHX_ENGINE_AFFINE or PHX_ENGINE_LINEAR");
    return nullptr;
  }
  if (cfg->engine == PHX_ENGINE_AFFINE) {
    const phx_costs& c = cfg->costs;
    if (c.mismatch < 1 || c.gap_extend < 1 || c.gap_open < 0) {
      set_error(ctx, "affine costs need mismatch >= 1, gap_extend >= 1, gap_open >= 0");
      return nullptr;
    }
    if (ring_slots(cfg->engine, c) > kRingSlotsMax) {
      set_error(ctx, "affine costs span more score frames than the device ring holds");
      return nullptr;
    }
  }
  if (hipSetDevice(ctx->device) != hipSuccess) {
    set_error(ctx, "hipSetDevice failed", PHX_E_DEVICE);
    return nullptr;
  }
  phx_server* S = new phx_server;
  S->ctx = ctx;
  S->device = ctx->device;
  S->engine = cfg->engine;
  S->mode = cfg->engine == PHX_ENGINE_AFFINE ? ENG_AFFINE : ENG_LINEAR;
  S->costs = cfg->engine == PHX_ENGINE_AFFINE ? cfg->costs : phx_costs{0, 0, 0};
  S->cap = cfg->cap < 0 ? -1 : (int32_t)std::min<int64_t>(cfg->cap, INT32_MAX);
  S->R = ring_slots(cfg->engine, S->costs);
  S->max_group = cfg->max_group ? cfg->max_group : 8192;
  S->wave_lds_kib = ctx->tun.wave_lds_kib;
  S->wave_slots = ctx->tun.wave_slots;
  auto fail = [&](const std::string& m) -> phx_server* {
    set_error(ctx, m);
    free_server(S);
    return nullptr;
  };
  // Reference: 2-bit copy (non-ACGT bytes pack as A; windows holding them use
  // the 8-bit copy) and 8-bit copy, each with two pad words.
  if (cfg->reference && cfg->reference_len) {
    const uint64_t n = cfg->reference_len;
    S->ref_len = n;
    S->ref_host.assign(cfg->reference, cfg->reference + n);
    const uint8_t* code = acgt_table();
    for (uint64_t i = 0; i < n; ++i)
      if (code[cfg->reference[i]] == 0xFF) S->ref_bad.push_back(i);
    std::vector<uint64_t> w2((n + 31) / 32 + 2, 0), w8((n + 7) / 8 + 2, 0);
    for (uint64_t i = 0; i < n; ++i) {
      const uint8_t c = code[cfg->reference[i]];
      w2[i >> 5] |= (uint64_t)(c == 0xFF ? 0 : c) << ((i & 31) * 2);
    }
    std::memcpy(w8.data(), cfg->reference, n);
    if (!grow_dev(S->d_ref2, w2.size() * 8) || !grow_dev(S->d_ref8, w8.size() * 8) ||
        hipMemcpy(S->d_ref2.p, w2.data(), w2.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(S->d_ref8.p, w8.data(), w8.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
      return fail("reference upload failed");
  }
  // Arena for the wavefront state of HBM-tier requests.
  uint64_t budget = cfg->hbm_budget_mib << 20;
  if (!budget) {
    size_t free_b = 0, total_b = 0;
    budget = 16ull << 30;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b)
      budget = std::max<uint64_t>(1ull << 30, (uint64_t)(0.5 * (double)free_b));
  }
  if (hipMalloc(&S->d_arena.p, budget) != hipSuccess) return fail("arena allocation failed");
  S->d_arena.bytes = budget;
  S->arena.init(budget);
  const uint32_t per_tier = cfg->lanes_per_tier ? cfg->lanes_per_tier : 1;
  for (int t = 0; t < 4; ++t)
    for (uint32_t j = 0; j < per_tier; ++j) {
      Lane L;
      L.tier = t;
      const bool made = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking) == hipSuccess &&
                        hipEventCreateWithFlags(&L.done, hipEventDisableTiming) == hipSuccess;
      // Staging sized for a full group of short reads up front, so buffers
      // rarely grow (a free may wait for the device) once serving.
      const size_t g = S->max_group;
      const bool sized = made && grow_dev(L.d_seq, g * 128) && grow_dev(L.d_pairs, g * sizeof(PairDesc)) &&
                         grow_dev(L.d_work, g * 4) && grow_dev(L.d_ctr, 64) &&
                         grow_host(L.h_seq, g * 128) && grow_host(L.h_pairs, g * sizeof(PairDesc)) &&
                         grow_host(L.h_out, g * sizeof(PairOut)) && grow_host(L.h_runs, g * 1100) &&
                         grow_host(L.h_work, g * 4);
      S->lanes.push_back(L);
      if (!sized) return fail("stream or staging allocation failed");
    }
  S->thr = std::thread(scheduler, S);
  return S;
}

int phx_serve_submit(phx_server* S, const phx_request* reqs, uint32_t n) {
  if (!S || (n && !reqs)) return PHX_E_ARG;
  std::vector<Req> batch;
  batch.reserve(n);
  std::vector<Done> fin;
  const double t = now_ms();
  for (uint32_t i = 0; i < n; ++i) {
    const phx_request& q = reqs[i];
    if (!q.query && q.query_len) return PHX_E_ARG;
    uint64_t tfull;
    if (q.target) {
      tfull = q.target_len;
    } else {
      if (q.ref_start > S->ref_len || q.ref_len > S->ref_len - q.ref_start) return PHX_E_ARG;
      tfull = q.ref_len;
    }
    Req r;
    r.tag = q.tag;
    r.t_submit = t;
    r.seed_q = q.seed_q;
    r.seed_t = q.seed_t;
    if (q.seed_q > q.query_len || q.seed_t > tfull) {  // lib.rs:107-111
      fin.push_back(make_error(r, PHX_INVALID_INPUT, t));
      continue;
    }
    const uint64_t qn = q.query_len - q.seed_q, tn = tfull - q.seed_t;
    if (qn > (uint64_t)kMaxLen || tn > (uint64_t)kMaxLen) return PHX_E_ARG;
    r.qn = (int32_t)qn;
    r.tn = (int32_t)tn;
    r.q.assign(q.query + q.seed_q, q.query + q.query_len);
    bool acgt = is_acgt(r.q.data(), r.q.size());
    if (q.target) {
      r.t.assign(q.target + q.seed_t, q.target + tfull);
      acgt = acgt && is_acgt(r.t.data(), r.t.size());
    } else {
      r.ref = true;
      r.ref_pos = q.ref_start + q.seed_t;
      acgt = acgt && ref_window_acgt(S, r.ref_pos, tn);
    }
    r.b8 = !acgt;
    r.tier = initial_tier(S, r.qn, r.tn);
    if (r.tier != TIER_LDS) {
      const uint8_t* tp = r.ref ? S->ref_host.data() + r.ref_pos : r.t.data();
      r.est = estimate_steps_seq(S->engine, S->costs, r.q.data(), r.q.size(), tp, (size_t)r.tn);
    }
    batch.push_back(std::move(r));
  }
  S->submitted += n;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (Req& r : batch) S->incoming.push_back(std::move(r));
    for (Done& d : fin) S->done.push_back(std::move(d));
  }
  S->completed += fin.size();
  S->cv_sched.notify_one();
  if (!fin.empty()) S->cv_done.notify_all();
  return PHX_OK;
}

int phx_serve_poll(phx_server* S, phx_response* out, uint32_t max_out, uint32_t* runs,
                   uint64_t runs_cap, uint64_t* run_off, uint32_t* n_out, int32_t timeout_us) {
  if (!S || !n_out || (max_out && (!out || !run_off))) return PHX_E_ARG;
  *n_out = 0;
  std::unique_lock<std::mutex> lk(S->mu);
  if (S->done.empty() && timeout_us > 0)
    S->cv_done.wait_for(lk, std::chrono::microseconds(timeout_us), [&] { return !S->done.empty(); });
  if (max_out) run_off[0] = 0;
  uint32_t k = 0;
  uint64_t pos = 0;
  while (k < max_out && !S->done.empty()) {
    Done& d = S->done.front();
    const uint64_t nr = d.runs.size();
    if (nr && (!runs || pos + nr > runs_cap)) {
      if (k == 0) {
        run_off[0] = nr;
        return PHX_E_CAPACITY;
      }
      break;
    }
    if (nr) std::memcpy(runs + pos, d.runs.data(), nr * 4);
    out[k].tag = d.tag;
    out[k].aln = d.aln;
    out[k].latency_ms = d.latency_ms;
    pos += nr;
    run_off[k + 1] = pos;
    ++k;
    S->done.pop_front();
  }
  *n_out = k;
  return PHX_OK;
}

int phx_serve_stats_get(phx_server* S, phx_serve_stats* st) {
  if (!S || !st) return PHX_E_ARG;
  st->submitted = S->submitted;
  st->completed = S->completed;
  st->waiting = S->n_waiting;
  st->in_flight = S->n_flight;
  st->launches = S->launches;
  st->reruns = S->reruns;
  st->arena_bytes = S->arena.cap();
  st->arena_used_bytes = S->arena_used;
  st->arena_peak_bytes = S->arena_peak;
  return PHX_OK;
}

int phx_serve_stop(phx_server* S) {
  if (!S) return PHX_E_ARG;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_sched.notify_all();
  if (S->thr.joinable()) S->thr.join();
  free_server(S);
  return PHX_OK;
}

}  // extern "C"
