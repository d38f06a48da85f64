// phx_host.h -- host-side helpers shared by the batch runtime (phx_api.hip)
// and the request server (phx_serve.hip): sequence packing, the memory model
// of a WFA pair (score steps and history cells), the LDS layout of the
// wave-resident tier, and tier launches.
#ifndef PHX_HOST_H
#define PHX_HOST_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "phraya_hip.h"
#include "phx_device.h"

namespace phx {

hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t st);
// pairs[iv[2 k]].bound = iv[2 k + 1] for k < n.
hipError_t launch_set_bounds(PairDesc* pairs, const uint32_t* iv, uint32_t n, hipStream_t st);
hipError_t launch_myers(int nb_template, const MyersArgs& a, hipStream_t st);
hipError_t launch_wfa_wave(int mode, bool lds, const WfaArgs& a, const WaveGeom& G,
                           uint32_t grid, hipStream_t st);
hipError_t launch_wfa_block(int mode, int nt, int storage, const WfaArgs& a, const WaveGeom& G,
                            uint32_t grid, hipStream_t st);
hipError_t launch_wfa_rows(bool b8, const WfaArgs& a, const WaveGeom& G, uint32_t grid, hipStream_t st);
hipError_t launch_stage_rows(const PairDesc* pairs, const uint32_t* idx, const RowDesc* rd, uint64_t* stage,
                             uint32_t n, hipStream_t st);
hipError_t launch_gather_runs(const uint32_t* runs, const uint64_t* src, const uint64_t* off,
                              uint32_t* dst, uint32_t n, hipStream_t st);

namespace host {

constexpr int kRingSlotsMax = 128;          // score frames a ring may span
constexpr int32_t kMaxLen = (1 << 27) - 2;  // history cells hold offset + 1 in 28 bits
constexpr uint64_t kWaveMaxLen = 2048;      // qn + tn limit of the LDS tier
constexpr int kWaveRingMax = 64;            // ring slots the LDS tier allows
constexpr uint32_t kWaveSteps = 128;        // LDS tier step-table entries per wave
constexpr int kWaveWaves = 4;               // waves per workgroup of the wave kernels
constexpr uint32_t kWaveWidth = 128;        // LDS tier ring width (diagonals)
constexpr uint32_t kMaxGroups = 64;         // launch groups of a WFA plan (counters and flags per group)
constexpr int32_t kRowsMaxKend = 48;        // banded rows: |qn - tn| limit (history steps per row)
// LDS-tier launch groups run as banded rows when their storage code is
// kRowsStorage (2-bit pairs) or kRowsStorage + 1 (bytes).
constexpr int kRowsStorage = 3;

// Tiers of the WFA kernels.
enum Tier : int {
  TIER_LDS = 0,      // wave per pair, everything in LDS
  TIER_WAVE = 1,     // wave per pair, ring and history in an HBM slot
  TIER_BLOCK256 = 2, // 256-thread workgroup per pair, HBM slot
  TIER_BLOCK1024 = 3,// 1024-thread workgroup per pair, HBM slot
  TIER_BLOCK512 = 4, // 512-thread workgroup per pair, register wavefront only (two per CU)
  TIER_BLOCK64 = 5   // one-wave workgroup per pair, register wavefront only (16 per CU)
};

// Threads per pair of a workgroup tier.
inline int tier_threads(int t) {
  return t == TIER_BLOCK64 ? 64 : t == TIER_BLOCK256 ? 256 : t == TIER_BLOCK512 ? 512 : 1024;
}
// Workgroups of a workgroup tier a CU holds at once (register tiers: 121
// VGPRs per lane, so four waves per SIMD).
inline uint64_t tier_per_cu(int t) {
  return t == TIER_BLOCK64 ? 16 : t == TIER_BLOCK256 ? 4 : t == TIER_BLOCK512 ? 2 : 1;
}

inline const uint8_t* acgt_table() {
  static const std::array<uint8_t, 256> t = [] {
    std::array<uint8_t, 256> a;
    a.fill(0xFF);
    a['A'] = 0;
    a['C'] = 1;
    a['G'] = 2;
    a['T'] = 3;
    return a;
  }();
  return t.data();
}

// Whether every byte of the 8-byte word x is one of A, C, G, T: a byte is
// zero after XOR with the letter exactly when it equals it; the classic
// has-zero-byte test per letter, ORed, must flag all eight bytes.
inline bool word_acgt(uint64_t x) {
  constexpr uint64_t L = 0x0101010101010101ull, H = 0x8080808080808080ull;
  auto zero_bytes = [&](uint64_t v) {  // 0x80 in each byte of v that is zero, exactly
    return ~(((v & ~H) + ~H) | v) & H;
  };
  const uint64_t hit = zero_bytes(x ^ (L * 'A')) | zero_bytes(x ^ (L * 'C')) | zero_bytes(x ^ (L * 'G')) |
                       zero_bytes(x ^ (L * 'T'));
  return hit == H;
}

inline bool is_acgt(const uint8_t* s, size_t n) {
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t x;
    std::memcpy(&x, s + i, 8);
    if (!word_acgt(x)) return false;
  }
  const uint8_t* code = acgt_table();
  for (; i < n; ++i)
    if (code[s[i]] == 0xFF) return false;
  return true;
}

// 2-bit codes of 32 ACGT bytes at s as one word (symbol i in bits 2i..2i+1):
// for A, C, G, T (0x41, 0x43, 0x47, 0x54), ((c >> 1) ^ (c >> 2)) & 3 is 0, 1,
// 2, 3; eight codes per 64-bit lane are then gathered to 16 bits.
inline uint64_t pack32_acgt(const uint8_t* s) {
  uint64_t out = 0;
  for (int k = 0; k < 4; ++k) {
    uint64_t x;
    std::memcpy(&x, s + 8 * k, 8);
    uint64_t y = ((x >> 1) ^ (x >> 2)) & 0x0303030303030303ull;  // a code per byte
    y = (y | (y >> 6)) & 0x000F000F000F000Full;                    // 4 bits per 16
    y = (y | (y >> 12)) & 0x000000FF000000FFull;                   // 8 bits per 32
    y = (y | (y >> 24)) & 0xFFFFull;                               // 16 bits
    out |= y << (16 * k);
  }
  return out;
}

// Words of an n-symbol sequence plus the zero pad word.
inline uint64_t words_for(size_t n, bool b8) { return (b8 ? (n + 7) / 8 : (n + 31) / 32) + 1; }

inline void pack_into(uint64_t* w, const uint8_t* s, size_t n, bool b8) {
  const uint64_t nw = words_for(n, b8);
  std::memset(w, 0, nw * 8);
  if (b8) {
    std::memcpy(w, s, n);
    return;
  }
  size_t i = 0;
  for (; i + 32 <= n; i += 32) w[i >> 5] = pack32_acgt(s + i);
  const uint8_t* code = acgt_table();
  for (; i < n; ++i) w[i >> 5] |= (uint64_t)code[s[i]] << ((i & 31) * 2);
}

inline uint32_t next_pow2(uint64_t v) {
  uint32_t p = 1;
  while (p < v && p < (1u << 30)) p <<= 1;
  return p;
}

inline uint32_t log2u(uint32_t v) {
  uint32_t l = 0;
  while ((1u << l) < v) ++l;
  return l;
}

inline int32_t gcd3(int32_t a, int32_t b, int32_t c) {
  auto g2 = [](int32_t x, int32_t y) {
    while (y) {
      const int32_t r = x % y;
      x = y;
      y = r;
    }
    return x;
  };
  return g2(g2(a, b), c);
}

// Kernel cost arguments: the costs, the score cap and their score-step form
// (valid affine costs have g >= 1; other costs get g = 1, unused).
inline phx::CostParams cost_params(const phx_costs& c, int32_t cap) {
  int32_t g = gcd3(c.mismatch, c.gap_open, c.gap_extend);
  if (g < 1) g = 1;
  return phx::CostParams{c.mismatch, c.gap_open, c.gap_extend, cap, g, c.mismatch / g,
                         (c.gap_open + c.gap_extend) / g, c.gap_extend / g};
}

// Score frames the affine ring spans: max(x, o+e, e)/g + 1 (linear: 2).
inline int32_t ring_slots(int engine, const phx_costs& c) {
  if (engine != PHX_ENGINE_AFFINE) return 2;
  const int32_t g = gcd3(c.mismatch, c.gap_open, c.gap_extend);
  return std::max(c.mismatch, std::max(c.gap_open + c.gap_extend, c.gap_extend)) / g + 1;
}

// Frames of score ring a pair keeps: affine max(x, o+e)/g + 1 for the m layer
// and e/g + 1 for each gap layer; linear 2.
inline int32_t ring_frames(int engine, const phx_costs& c) {
  if (engine != PHX_ENGINE_AFFINE) return 2;
  const int32_t g = gcd3(c.mismatch, c.gap_open, c.gap_extend);
  return (std::max(c.mismatch, c.gap_open + c.gap_extend) / g + 1) + 2 * (c.gap_extend / g + 1);
}

// Whether the LDS tier's layout needs an LDS ring: only for affine cost
// shapes without a register wavefront (reg_wavefront_shape).
inline bool lds_tier_ring(int engine, const phx_costs& c) {
  if (engine != PHX_ENGINE_AFFINE) return false;
  const int32_t g = gcd3(c.mismatch, c.gap_open, c.gap_extend);
  return !reg_wavefront_shape(c.mismatch / g, (c.gap_open + c.gap_extend) / g, c.gap_extend / g);
}

// sum_{u=1..S} min(u, a)
inline uint64_t sum_min(uint64_t S, uint64_t a) {
  return S <= a ? S * (S + 1) / 2 : a * (a + 1) / 2 + (S - a) * a;
}

// History words a pair needs for `steps` score steps: the cells of every step
// (width min(u, qn) + min(u, tn) + 1) plus one offset word per step.
inline uint64_t exact_hist_words(int32_t steps, int32_t qn, int32_t tn) {
  const uint64_t S = (uint64_t)std::max(steps, 0);
  return sum_min(S, (uint64_t)qn) + sum_min(S, (uint64_t)tn) + 2 * S + 16;
}

// History words of a pair computed within a bound of B score steps (the
// band of the register-wavefront workgroup tiers): at step s only diagonals
// within (B - s) / en of kend = qn - tn (fitting: not below kend - (B - s) /
// en) are stored, plus one offset word per step.
inline uint64_t band_hist_words(int32_t bound, int32_t qn, int32_t tn, int32_t en, bool fitting) {
  const int64_t kend = (int64_t)qn - tn;
  uint64_t sum = 0;
  for (int64_t s = 1; s <= bound; ++s) {
    const int64_t r = (bound - s) / en;
    const int64_t lo = std::max(std::max(-s, -(int64_t)tn), kend - r);
    const int64_t hi = fitting ? std::min<int64_t>(s, qn) : std::min(std::min<int64_t>(s, qn), kend + r);
    sum += (uint64_t)std::max<int64_t>(hi - lo + 1, 0) + 1;
  }
  return sum + 16;
}

// Generous estimate of the score steps of a pair, used to size HBM history
// slots before a run: an end gap for the length difference plus 9% events per
// base (40% mismatches, 60% indels of mean length 2), +15%.  A pair that needs
// more reports its exact step count and is re-run in a slot of that size.
inline int32_t estimate_steps(int engine, const phx_costs& c, int32_t qn, int32_t tn) {
  const double m = std::max(qn, tn), d = std::abs(tn - qn);
  double steps;
  if (engine == PHX_ENGINE_AFFINE) {
    const int32_t g = gcd3(c.mismatch, c.gap_open, c.gap_extend);
    const double score = c.gap_open + c.gap_extend * d +
                         0.09 * m * (0.4 * c.mismatch + 0.6 * (c.gap_open + 2.0 * c.gap_extend));
    steps = score / g;
  } else {
    steps = d + 0.09 * m;
  }
  return (int32_t)std::min(1.15 * steps + 16.0, (double)(qn + tn + 16));
}

// Per-base event rate of a pair, from how many of the query's 12-mers occur
// in the target: a k-mer survives p events per base with probability about
// (1 - p)^k, so p = 1 - f^(1/k) for a surviving fraction f.  Target 12-mers go
// into an open-addressing table; up to 1024 evenly spaced query 12-mers are
// looked up.  Returns -1 when either sequence is too short to say.
inline double kmer_event_rate(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn) {
  constexpr int K = 12;
  constexpr uint32_t MASK = (1u << (2 * K)) - 1;
  if (qn < 4 * K || tn < K) return -1.0;
  const uint8_t* code = acgt_table();
  const uint32_t cap = next_pow2(std::max<uint64_t>(64, 2 * (uint64_t)tn));
  static thread_local std::vector<uint32_t> table;
  table.assign(cap, 0);
  auto slot = [&](uint32_t key) { return (key * 0x9E3779B1u) >> (32 - log2u(cap)); };
  uint32_t h = 0;
  int valid = 0;
  for (size_t j = 0; j < tn; ++j) {
    const uint8_t c = code[t[j]];
    if (c == 0xFF) { valid = 0; continue; }
    h = ((h << 2) | c) & MASK;
    if (++valid < K) continue;
    const uint32_t key = h + 1;
    for (uint32_t i = slot(key);; i = (i + 1) & (cap - 1)) {
      if (table[i] == key) break;
      if (table[i] == 0) { table[i] = key; break; }
    }
  }
  const size_t positions = qn - K + 1;
  const size_t samples = std::min<size_t>(1024, positions);
  size_t hits = 0, tried = 0;
  for (size_t s = 0; s < samples; ++s) {
    const size_t at = s * positions / samples;
    uint32_t key = 0;
    bool ok = true;
    for (int k = 0; k < K; ++k) {
      const uint8_t c = code[q[at + k]];
      if (c == 0xFF) { ok = false; break; }
      key = (key << 2) | c;
    }
    if (!ok) continue;
    ++tried;
    key += 1;
    for (uint32_t i = slot(key);; i = (i + 1) & (cap - 1)) {
      if (table[i] == key) { ++hits; break; }
      if (table[i] == 0) break;
    }
  }
  if (tried < 16) return -1.0;
  const double f = (double)hits / (double)tried;
  if (f <= 0.0) return 0.35;
  return std::min(0.35, std::max(0.0, 1.0 - std::pow(f, 1.0 / K)));
}

// Score steps of a pair estimated from its sequences: the end gap for the
// length difference plus the k-mer event rate, each event costed as the mean
// of a mismatch and a 2-base gap, with 25% margin.  Pairs the k-mer probe
// cannot judge fall back to estimate_steps.  A pair that needs more steps
// than estimated still completes: it reports its exact count and is re-run.
inline int32_t estimate_steps_seq(int engine, const phx_costs& c, const uint8_t* q, size_t qn,
                                  const uint8_t* t, size_t tn) {
  const double p = kmer_event_rate(q, qn, t, tn);
  if (p < 0) return estimate_steps(engine, c, (int32_t)qn, (int32_t)tn);
  const double m = (double)std::max(qn, tn), d = std::fabs((double)tn - (double)qn);
  const double rate = std::max(p, 0.002) * 1.25;
  double steps;
  if (engine == PHX_ENGINE_AFFINE) {
    const int32_t g = gcd3(c.mismatch, c.gap_open, c.gap_extend);
    const double per_event = 0.5 * (c.mismatch + c.gap_open + 2.0 * c.gap_extend);
    steps = (c.gap_open + c.gap_extend * d + m * rate * per_event) / g;
  } else {
    steps = d + m * rate * 2.0;  // a mismatch costs 1, an indel event its length
  }
  return (int32_t)std::min(1.1 * steps + 32.0, (double)(qn + tn + 16));
}

// Persistent host worker threads (up to 15, plus the calling thread) for
// parallel_for: spawning threads per call would cost as much as packing a
// 10,000-pair batch.  One job at a time; a call that finds the pool busy
// (another host thread's job, or a nested call) runs on its own thread.
class HostPool {
 public:
  static HostPool& get() {
    static HostPool* p = new HostPool();  // lives until the process ends
    return *p;
  }
  unsigned workers() const { return (unsigned)threads_.size(); }
  // Runs f(i) for i in [0, n) on the workers and the caller; false (nothing
  // run) when the pool is busy.
  bool run(size_t n, const std::function<void(size_t)>& f) {
    std::unique_lock<std::mutex> busy(busy_, std::try_to_lock);
    if (!busy.owns_lock()) return false;
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &f;
      n_ = n;
      next_.store(0);
      pending_ = (unsigned)threads_.size();
      ++gen_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return pending_ == 0; });
    job_ = nullptr;
    return true;
  }

 private:
  HostPool() {
    const unsigned T = std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
    for (unsigned k = 1; k < T; ++k) threads_.emplace_back([this] { loop(); });
    for (auto& t : threads_) t.detach();
  }
  // Indices are claimed in blocks (about eight per thread), so the threads
  // do not contend on the counter once per index.
  void drain() {
    const size_t block = std::max<size_t>(1, n_ / (8 * (threads_.size() + 1)));
    for (size_t i; (i = next_.fetch_add(block)) < n_;)
      for (size_t j = i; j < std::min(n_, i + block); ++j) (*job_)(j);
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
      }
      drain();
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_all();
    }
  }
  std::mutex busy_, mu_;
  std::condition_variable cv_, done_;
  std::vector<std::thread> threads_;
  const std::function<void(size_t)>* job_ = nullptr;
  size_t n_ = 0;
  std::atomic<size_t> next_{0};
  unsigned pending_ = 0;
  uint64_t gen_ = 0;
};

// f(i) for i in [0, n) on the host pool (up to 16 threads).
template <class F>
void parallel_for(size_t n, F f) {
  if (n >= 32) {
    HostPool& pool = HostPool::get();
    if (pool.workers() > 0 && pool.run(n, std::function<void(size_t)>(f))) return;
  }
  for (size_t i = 0; i < n; ++i) f(i);
}

// Ring words of one HBM slot (`frames` from ring_frames) for pairs with
// qn + tn <= max_sum.
inline uint64_t ring_words_for(int32_t frames, uint64_t max_sum, uint32_t* wcap_out) {
  // Every diagonal of the pair plus -1 margins of up to `frames` entries on
  // both sides and one clamped read past the top: the kernels index an HBM
  // ring without wrap-around.
  const uint32_t wcap = next_pow2(max_sum + 1 + 2 * (uint64_t)frames + 2);
  if (wcap_out) *wcap_out = wcap;
  return (uint64_t)frames * wcap;
}

// Bytes of an HBM history slot holding `steps` steps of a qn x tn pair.
inline uint64_t hist_slot_bytes(int32_t steps, int32_t qn, int32_t tn) {
  return (exact_hist_words(steps, qn, tn) * 4 + 255) & ~(uint64_t)255;
}

// LDS layout of one wave of the LDS tier: int16 ring [frames][128], the
// packed query and target (max_seq_words words), the u16 step table and u16
// history cells filling the wave's share of lds_kib.
inline WaveGeom lds_geometry(int32_t frames, uint32_t max_seq_words, uint32_t lds_kib) {
  auto al = [](uint32_t v) { return (v + 15u) & ~15u; };
  WaveGeom G{};
  G.ring_off = 0;
  const uint32_t ring_bytes = (uint32_t)frames * kWaveWidth * 2u;
  G.seq_off = al(ring_bytes);
  G.seq_cap_words = std::max<uint32_t>(max_seq_words, 2);
  G.tab_off = G.seq_off + al(G.seq_cap_words * 8);
  G.steps_cap = kWaveSteps;
  G.cell_off = G.tab_off + al(2 * G.steps_cap);
  const uint32_t budget = std::max<uint32_t>(lds_kib * 1024u, std::min<uint32_t>(G.cell_off + 4096u, 15u * 1024u));
  const uint32_t cell_bytes = budget > G.cell_off + 256 ? (budget - G.cell_off) & ~15u : 256u;
  G.cells_cap = std::min<uint32_t>(65535, cell_bytes / 2);
  G.per_wave_bytes = G.cell_off + al(2 * G.cells_cap);
  return G;
}

// Waves of an LDS-tier launch over n pairs with `wave_slots` waves wanted:
// whole workgroups of kWaveWaves, at most one wave per pair (rounded up).
inline uint32_t lds_tier_waves(uint64_t n, uint32_t wave_slots) {
  const uint32_t slots = (uint32_t)std::min<uint64_t>(n, std::max<uint32_t>(kWaveWaves, wave_slots));
  return std::max<uint32_t>(1, (slots + kWaveWaves - 1) / kWaveWaves) * kWaveWaves;
}

// Stable sort of idx by key(i) descending (a 32-bit key): LSD radix passes of
// 8 bits, skipping the digits every key shares (lengths and step estimates
// usually need two).
template <class K>
void radix_sort_desc(std::vector<uint32_t>& idx, K key) {
  const size_t n = idx.size();
  if (n < 2) return;
  std::vector<uint32_t> keys(n), tmp(n), tk(n);
  uint32_t all_or = 0, all_and = ~0u;
  for (size_t i = 0; i < n; ++i) {
    keys[i] = ~(uint32_t)key(idx[i]);  // ascending on the complement = descending
    all_or |= keys[i];
    all_and &= keys[i];
  }
  for (int shift = 0; shift < 32; shift += 8) {
    if ((((all_or ^ all_and) >> shift) & 0xFF) == 0) continue;  // one digit value only
    size_t count[257] = {0};
    for (size_t i = 0; i < n; ++i) ++count[((keys[i] >> shift) & 0xFF) + 1];
    for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
    for (size_t i = 0; i < n; ++i) {
      const size_t at = count[(keys[i] >> shift) & 0xFF]++;
      tmp[at] = idx[i];
      tk[at] = keys[i];
    }
    idx.swap(tmp);
    keys.swap(tk);
  }
}

// Order of an LDS-tier work list for static striding over `waves` waves
// (wave g takes positions g, g + waves, ...): pairs by estimated cost,
// largest first, dealt to the waves in alternating direction each round.
inline void stride_balance(std::vector<uint32_t>& list, const std::vector<int32_t>& est, uint32_t waves) {
  radix_sort_desc(list, [&](uint32_t i) { return (uint32_t)std::max(est[i], 0); });
  std::vector<uint32_t> out(list.size());
  const size_t S = waves;
  for (size_t r0 = 0; r0 < list.size(); r0 += S) {
    const size_t m = std::min(S, list.size() - r0);
    const bool rev = (r0 / S) & 1;
    for (size_t g = 0; g < m; ++g) out[r0 + g] = list[r0 + (rev ? m - 1 - g : g)];
  }
  list.swap(out);
}

// Layout of a banded-row launch: four 16-lane or two 32-lane row regions
// of `row_seq_words` packed words and row_steps (+ 16) history steps each
// (rows_region_bytes), or the single-pair LDS layout G in the same region
// (for the pairs a 32-lane row cannot finish).
inline WaveGeom rows_geometry(WaveGeom G, uint32_t row_seq_words, uint32_t row_steps, bool wide_rows) {
  G.rows_hist_off = (row_seq_words * 8 + 15u) & ~15u;
  G.rows_steps = row_steps;
  const uint32_t quads = 4u * rows_region_bytes(G.rows_hist_off, row_steps, 16);
  G.per_wave_bytes = wide_rows ? std::max(std::max(G.per_wave_bytes, quads),
                                          2u * rows_region_bytes(G.rows_hist_off, row_steps, 32))
                               : quads;
  return G;
}

// Mismatching bytes of a[0, n) against b[0, n), eight at a time.
inline uint32_t mismatched_bytes(const uint8_t* a, const uint8_t* b, size_t n) {
  uint32_t d = 0;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t x, y;
    std::memcpy(&x, a + i, 8);
    std::memcpy(&y, b + i, 8);
    uint64_t z = x ^ y;
    z |= z >> 4;
    z |= z >> 2;
    z |= z >> 1;
    d += (uint32_t)__builtin_popcountll(z & 0x0101010101010101ull);
  }
  for (; i < n; ++i) d += a[i] != b[i];
  return d;
}

// Upper bound (score steps) on the global gap-affine optimum of q against t
// for a cost shape with extend = 1 step: the cost of the better of two
// alignments that both exist, one diagonal of min(qn, tn) bases (xn steps a
// mismatch) with the |qn - tn| gap (oen + |qn - tn| - 1 steps) at the end,
// or with the gap at the start.  A pair whose bound is at most |qn - tn| +
// 15 has its optimum inside a 16-lane row's band.
inline int64_t rows_upper_bound(const uint8_t* q, size_t qn, const uint8_t* t, size_t tn, int32_t xn,
                                int32_t oen) {
  const size_t m = std::min(qn, tn), d = std::max(qn, tn) - m;
  const int64_t gap = d ? (int64_t)oen + (int64_t)d - 1 : 0;
  const int64_t end = (int64_t)mismatched_bytes(q, t, m) * xn + gap;
  if (d == 0) return end;
  const int64_t start = (int64_t)mismatched_bytes(qn > tn ? q + d : q, qn > tn ? t : t + d, m) * xn + gap;
  return std::min(end, start);
}

// Order of a banded-row work list for `waves` waves (wave g takes sets g,
// g + waves, ...; a set is `per` consecutive entries, one per row): pairs by
// estimated cost, largest first, so each set holds pairs of similar cost;
// sets dealt to the waves in alternating direction each round.
inline void quad_balance(std::vector<uint32_t>& list, const std::vector<int32_t>& est, uint32_t waves,
                         uint32_t per = 4) {
  radix_sort_desc(list, [&](uint32_t i) { return (uint32_t)std::max(est[i], 0); });
  const size_t nq = (list.size() + per - 1) / per, S = waves;
  std::vector<uint32_t> out;
  out.reserve(list.size());
  for (size_t r0 = 0; r0 < nq; r0 += S) {
    const size_t m = std::min(S, nq - r0);
    const bool rev = (r0 / S) & 1;
    for (size_t g = 0; g < m; ++g) {
      const size_t q = r0 + (rev ? m - 1 - g : g);
      for (size_t j = per * q; j < std::min(list.size(), per * q + per); ++j) out.push_back(list[j]);
    }
  }
  list.swap(out);
}

// Words the LDS tier stages for one pair (query then target, with shifts).
inline uint32_t lds_seq_words(const PairDesc& pd) {
  const bool b8 = pd.flags & PF_BYTES;
  return (uint32_t)(words_for((size_t)pd.q_shift + (size_t)pd.qn, b8) +
                    words_for((size_t)pd.t_shift + (size_t)pd.tn, b8));
}

// Launch one tier over a.work_list: `slots` pairs in flight (waves for the
// wave tiers, workgroups for the block tiers); storage of the block tiers:
// 0 ring in HBM, 1 ring and sequences in LDS, 2 register wavefront (layout
// in G); of the LDS tier: 0 a pair per wave, kRowsStorage.. banded rows.
inline hipError_t launch_tier(int tier, int mode, const WfaArgs& a, const WaveGeom& G,
                              uint32_t slots, hipStream_t st, int storage = 0) {
  if (tier == TIER_LDS && storage >= kRowsStorage) {
    const uint32_t wgs = std::max<uint32_t>(1, (slots + kWaveWaves - 1) / kWaveWaves);
    return launch_wfa_rows(storage != kRowsStorage, a, G, wgs, st);
  }
  if (tier <= TIER_WAVE) {
    const uint32_t wgs = std::max<uint32_t>(1, (slots + kWaveWaves - 1) / kWaveWaves);
    return launch_wfa_wave(mode, tier == TIER_LDS, a, G, wgs, st);
  }
  return launch_wfa_block(mode, tier_threads(tier), storage, a, G, std::max<uint32_t>(1, slots), st);
}

// LDS layout of a block-tier workgroup with the ring in LDS: int32 ring
// [frames][width] then the packed sequences (seq_words).
inline WaveGeom block_lds_geometry(int32_t frames, uint32_t width, uint32_t seq_words) {
  WaveGeom G{};
  G.ring_off = 0;
  G.seq_off = ((uint32_t)frames * width * 4u + 15u) & ~15u;
  G.seq_cap_words = seq_words;
  G.per_wave_bytes = G.seq_off + seq_words * 8u;
  return G;
}
constexpr uint32_t kBlockLdsMax = 64u * 1024u;  // dynamic LDS a block-tier workgroup may take

// Register wavefront of the workgroup tiers (storage 2): 16 chunks of 64
// diagonals per wave, so the 1024-thread tier covers 16,384 diagonals and the
// 256-thread tier 4,096; chunk-end tables of kRegEdgeSlots steps after the
// staged sequences.
constexpr uint32_t kRegEdgeSlots = 6;
inline uint32_t reg_chunks(int tier_threads) { return (uint32_t)tier_threads / 64u * 16u; }
// Longest query the register wavefront takes: its frames pack offset + 1
// (at most qn + 1) into 16-bit halves.
constexpr uint64_t kRegMaxQuery = 65534;
inline uint32_t reg_window(int tier_threads) { return 64u * reg_chunks(tier_threads); }
inline WaveGeom block_regs_geometry(int tier_threads, uint32_t seq_words) {
  WaveGeom G{};
  G.ring_off = 0;
  G.seq_off = 0;
  G.seq_cap_words = seq_words;
  G.tab_off = (seq_words * 8u + 15u) & ~15u;
  // One 16-byte entry per (slot, chunk), plus a zero entry at each end.
  G.per_wave_bytes = G.tab_off + kRegEdgeSlots * (reg_chunks(tier_threads) + 2u) * 16u;
  return G;
}
// Whether a register-wavefront cost shape (affine) is one the workgroup
// tiers have (4/6/2 and 2/4/1 in score steps).
inline bool block_regs_shape(int engine, const phx_costs& c) {
  if (engine != PHX_ENGINE_AFFINE) return false;
  const int32_t g = gcd3(c.mismatch, c.gap_open, c.gap_extend);
  const int32_t xn = c.mismatch / g, oen = (c.gap_open + c.gap_extend) / g, en = c.gap_extend / g;
  return xn == 2 && en == 1 && (oen == 4 || oen == 5);
}

// A batch plan over windows of resident sequences: the descriptors carry
// device pointers, shifts, lengths and flags; qh / th are host copies of each
// window (memory estimates of WFA pairs beyond the LDS tier).  Any engine;
// Myers pairs need word-aligned queries (q_shift 0).  oneshot: the plan
// borrows the context's cached buffers (destroy it before the next one).
// arena: host words the plan uploads as its own sequence buffer; pair i's
// query / target then starts at word q_word[i] / t_word[i] of it (kNoWord:
// the descriptor's own pointer).
// Run and read it with phx_plan_run / phx_plan_results.  Defined in
// phx_api.hip.
constexpr uint64_t kNoWord = ~0ull;  // a descriptor that keeps its own device pointer
phx_plan* plan_create_windows(phx_ctx* ctx, int engine, phx_costs costs, int64_t cap,
                              std::vector<PairDesc> pairs, const std::vector<const uint8_t*>& qh,
                              const std::vector<const uint8_t*>& th, bool oneshot = false,
                              const std::vector<uint64_t>* arena = nullptr,
                              const std::vector<uint64_t>* q_word = nullptr,
                              const std::vector<uint64_t>* t_word = nullptr);

// Tier of an HBM pair by qn + tn: a wave up to 8192, then workgroups.
inline int hbm_tier_for(uint64_t L) {
  return L <= 8192 ? TIER_WAVE : L <= 16384 ? TIER_BLOCK256 : TIER_BLOCK1024;
}

}  // namespace host
}  // namespace phx

constexpr int kStreams = 4;  // launch groups of one run overlap on these

struct PlanCache;  // device buffers kept between one-shot calls (phx_api.hip)

struct phx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;         // = streams[0]
  hipStream_t streams[kStreams] = {};
  hipEvent_t join[32] = {};             // cross-stream joins
  std::mutex mu;
  phx_tuning tun{};
  int cu_count = 256;
  char impl_name[64] = "hip";           // "hip-<arch>" (phx_last_impl)
  uint8_t* noalloc_scratch = nullptr;   // device address of this device's allocation-free scratch
  PlanCache* plan_cache = nullptr;      // buffers of one-shot plans, reused by the next one
  std::mutex cache_mu;
};

namespace phx {
namespace host {

// Error text and code of the last failed call the calling thread made, and
// the context it was made on.  Per thread, so threads sharing a context never
// race on it and the text stays valid until that thread's next failure.
struct LastError {
  const phx_ctx* ctx = nullptr;
  std::string msg;
  int code = PHX_OK;
};
inline LastError& last_error() {
  static thread_local LastError e;
  return e;
}
inline void set_error(const phx_ctx* ctx, const std::string& msg, int code = PHX_E_ARG) {
  LastError& e = last_error();
  e.ctx = ctx;
  e.msg = msg;
  e.code = code;
}
// Code of the calling thread's last failure on ctx (PHX_E_ARG when it made
// none there): what a one-shot entry point returns when plan creation fails.
inline int last_error_code(const phx_ctx* ctx) {
  const LastError& e = last_error();
  return e.ctx == ctx && e.code != PHX_OK ? e.code : PHX_E_ARG;
}

}  // namespace host
}  // namespace phx

#endif
