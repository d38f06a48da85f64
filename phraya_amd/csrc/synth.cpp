// synth.cpp -- seeded synthetic genomes and read/window pairs for the
// benchmark workloads (host code, no GPU).
//
// PRNG: SplitMix64 (the generator phraya-align's tests use,
// phraya-align/src/wfa_simd.rs:1536-1548).  The mutation walk follows the
// upstream benchmark generator (scripts/benchmark/local/gen_synthetic.py:44-98):
// at each site of the source fragment at most one event fires, an indel
// (checked first) or a substitution to a different base; an insertion emits
// random bases before the site without consuming it.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

struct Gen {
  std::vector<uint8_t> q, t;
  std::vector<uint64_t> qo{0}, to{0};
  std::vector<uint8_t> cls;
  std::vector<uint64_t> pos;  // window start of each pair in the genome
};

}  // namespace

extern "C" {

typedef struct synth_params {
  uint64_t genome_len;
  uint64_t genome_seed;
  uint64_t read_seed;
  uint32_t n;
  uint32_t len_lo, len_hi;   // read length range
  int32_t log_uniform;       // 1: log-uniform lengths, 0: uniform
  double window_scale;       // window = ceil(L * scale) + window_add
  int32_t window_add;
  double class_a_frac;       // probability a read is of class A
  double a_err_lo, a_err_hi; // per-read event rate ~ U[lo, hi] (events per source site)
  double a_indel_frac;       // fraction of events that are indels
  double b_err_lo, b_err_hi;
  double b_indel_frac;
  double indel_mean;         // geometric indel length mean (p = 1 / mean)
  int32_t indel_max;         // indel length truncation
  int32_t indel_uniform;     // 1: indel lengths uniform in [1, indel_max] instead
} synth_params;

void synth_genome(uint64_t len, uint64_t seed, uint8_t* out) {
  SplitMix64 r{seed};
  uint64_t i = 0;
  while (i < len) {
    uint64_t w = r.next();
    for (int k = 0; k < 32 && i < len; ++k, ++i) out[i] = (uint8_t)kBases[(w >> (2 * k)) & 3];
  }
}

// Mutate src[0..n) into out; stops once out holds `want` bases.
static void mutate(const uint8_t* src, uint64_t n, double sub, double indel, double p_len,
                   int32_t max_len, bool uniform_len, SplitMix64& r, std::vector<uint8_t>& out,
                   uint64_t want) {
  uint64_t i = 0;
  while (i < n && out.size() < want) {
    const uint8_t b = src[i];
    if (indel > 0 && r.uniform() < indel) {
      int32_t len = 1;
      if (uniform_len) len = 1 + (int32_t)r.below((uint64_t)std::max(max_len, 1));
      else
        while (len < max_len && r.uniform() >= p_len) ++len;
      if (r.uniform() < 0.5) {
        for (int32_t k = 0; k < len; ++k) out.push_back((uint8_t)kBases[r.below(4)]);
        continue;
      }
      i += (uint64_t)len;
      continue;
    }
    if (sub > 0 && r.uniform() < sub) {
      uint8_t alt = b;
      while (alt == b) alt = (uint8_t)kBases[r.below(4)];
      out.push_back(alt);
    } else {
      out.push_back(b);
    }
    ++i;
  }
  if (out.size() > want) out.resize(want);
}

void* synth_pairs(const synth_params* p, const uint8_t* genome) {
  Gen* g = new Gen;
  SplitMix64 r{p->read_seed};
  const double p_len = p->indel_mean > 1.0 ? 1.0 / p->indel_mean : 1.0;
  std::vector<uint8_t> buf;
  g->qo.reserve(p->n + 1);
  g->to.reserve(p->n + 1);
  g->cls.reserve(p->n);
  for (uint32_t i = 0; i < p->n; ++i) {
    uint64_t L;
    if (p->log_uniform) {
      const double u = r.uniform();
      L = (uint64_t)std::floor(std::exp(std::log((double)p->len_lo) +
                                        u * (std::log((double)p->len_hi) - std::log((double)p->len_lo))));
    } else {
      L = p->len_lo + r.below((uint64_t)(p->len_hi - p->len_lo) + 1);
    }
    if (L < 1) L = 1;
    const uint64_t W = (uint64_t)std::ceil((double)L * p->window_scale) + (uint64_t)p->window_add;
    const uint64_t src_len = L + L / 5 + 64;
    const uint64_t span = std::max<uint64_t>(W, src_len);
    const uint64_t pos = r.below(p->genome_len - span);
    const bool a = r.uniform() < p->class_a_frac;
    const double lo = a ? p->a_err_lo : p->b_err_lo, hi = a ? p->a_err_hi : p->b_err_hi;
    const double ifrac = a ? p->a_indel_frac : p->b_indel_frac;
    const double e = lo + (hi - lo) * r.uniform();
    buf.clear();
    mutate(genome + pos, src_len, e * (1.0 - ifrac), e * ifrac, p_len, p->indel_max,
           p->indel_uniform != 0, r, buf, L);
    g->q.insert(g->q.end(), buf.begin(), buf.end());
    g->qo.push_back(g->q.size());
    g->t.insert(g->t.end(), genome + pos, genome + pos + W);
    g->to.push_back(g->t.size());
    g->cls.push_back(a ? 0 : 1);
    g->pos.push_back(pos);
  }
  return g;
}

uint64_t synth_q_bytes(const void* h) { return ((const Gen*)h)->q.size(); }
uint64_t synth_t_bytes(const void* h) { return ((const Gen*)h)->t.size(); }

void synth_copy(const void* h, uint8_t* q, uint64_t* qo, uint8_t* t, uint64_t* to, uint8_t* cls) {
  const Gen* g = (const Gen*)h;
  std::memcpy(q, g->q.data(), g->q.size());
  std::memcpy(qo, g->qo.data(), g->qo.size() * 8);
  std::memcpy(t, g->t.data(), g->t.size());
  std::memcpy(to, g->to.data(), g->to.size() * 8);
  if (cls) std::memcpy(cls, g->cls.data(), g->cls.size());
}

void synth_positions(const void* h, uint64_t* pos) {
  const Gen* g = (const Gen*)h;
  std::memcpy(pos, g->pos.data(), g->pos.size() * 8);
}

void synth_free(void* h) { delete (Gen*)h; }

}  // extern "C"
