// phx_sketch.hip -- minimizer sketches of a batch of sequences on the device.
//
// sketch_alphabet (phraya-core/src/types.rs:1846-1867): canonical (k, w)
// minimizers for DNA, forward ones for protein, both over the 2-bit ASCII
// code (b >> 1) & 3 (A 0, C 1, T 2, G 3), as (value, position) lists.  Each
// k-mer's value packs its codes first base most significant; a canonical
// sketch keeps the smaller of that and the reverse complement's value.  A
// window of w consecutive k-mers selects the k-mer of smallest 64-bit hash
// (the SplitMix64 finalizer of the value, a bijection, so equal hashes mean
// equal values): the leftmost of equal ones for forward sketches; for
// canonical ones the leftmost when more than half of the window's
// l = k + w - 1 bases are T or G and the rightmost otherwise, which makes a
// sequence and its reverse complement select mirrored positions.
// Consecutive windows selecting the same k-mer report it once.
//
// Work unit: a wave takes T = 253 - w consecutive windows of one sequence
// (so that its k-mers, T + w - 1 plus the four before, fill 256 slots, four
// per lane).  Its bytes are staged in LDS as 2-bit codes, twice: first base
// most significant (a k-mer's forward value is then one funnel shift) and
// first base least significant (its reverse complement's value is a funnel
// shift and an XOR), both built from two ballots per 64 bytes, the second
// of which is also the T/G mask.  Each lane computes the values and hashes
// of 4 consecutive k-mers into LDS; lane 0 then takes the four windows
// before the task's first (only the last is used: the predecessor that
// decides whether the first window repeats it) and lanes 1.. four windows
// each, sharing the minimum of the w - 3 k-mers common to its windows.
// Reported k-mers go to the task's slot of a staging area (T entries), in
// order; a scan of the per-task counts places them, and a copy launch moves
// them to their final offsets and writes each sequence's offset.  Reading a
// byte per base and writing 12 bytes per reported k-mer (about 2 / (w + 1)
// per base) is the algorithmic traffic; the hashing and window work bound
// the compute launch well before HBM does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "phraya_hip.h"
#include "phx_host.h"

namespace {

constexpr uint32_t SK_PER = 4;                   // windows (and k-mers) per lane
constexpr uint32_t SK_SPAN = 64 * SK_PER;        // k-mer slots per wave
constexpr uint32_t SK_WAVES = 4;                 // tasks per workgroup
constexpr uint32_t SK_MAX_K = 32;
constexpr uint32_t SK_MAX_W = 128;
constexpr uint32_t SCAN_THREADS = 1024;
constexpr uint32_t SCAN_PER = 16;                // counts per thread and pass of the scan
constexpr uint32_t SK_ROUNDS = (SK_SPAN + SK_MAX_K - 1 + 63) / 64;  // 64-base rounds staged at most

// Windows per task: the four predecessor slots, T windows and their w - 1
// trailing k-mers fill the 256 k-mer slots.
__host__ __device__ constexpr uint32_t sk_task_windows(uint32_t w) { return SK_SPAN - SK_PER - w + 1; }

struct SkTask {
  uint64_t start;   // the sequence's first byte in the batch buffer
  uint32_t len;     // its length
  uint32_t x0;      // first window (= first k-mer position) of the task
  uint32_t nx;      // windows in the task, 1..T
  uint32_t seq;     // sequence index
  uint32_t seq_lo;  // sequences seq_lo..seq take this task's offset as theirs (empty when seq_lo > seq)
  uint32_t pad;
};

struct SkArgs {
  const uint8_t* seq;
  const SkTask* tasks;
  uint32_t ntasks, k, w, canonical, tw;   // tw: windows (staging slots) per task
  uint32_t* counts;          // per task
  uint64_t* stage_v;         // per task, tw slots
  uint32_t* stage_p;
};

// LDS of one wave: the two 2-bit code streams (and a zero pad word), the
// T/G masks, then the k-mers' hashes and values.
struct SkLayout {
  uint32_t bytes;   // staged bases, a multiple of 64
  uint32_t words;   // 64-bit words of one code stream, plus the pad word
  uint32_t stride;  // bytes per wave
  __host__ __device__ SkLayout(uint32_t k) {
    bytes = (SK_SPAN + k - 1 + 63) & ~63u;
    words = bytes / 32 + 1;
    stride = 2 * words * 8 + (bytes / 64) * 8 + SK_SPAN * 16;
  }
};

__device__ __forceinline__ void sk_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t sk_mix64(uint64_t z) {  // SplitMix64 finalizer
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Bits of x at the even positions of the result.
__device__ __forceinline__ uint64_t sk_spread(uint64_t x) {
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  return (x | (x << 1)) & 0x5555555555555555ull;
}

// Order of the 32 2-bit groups reversed.
__device__ __forceinline__ uint64_t sk_rev2(uint64_t x) {
  x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
  x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
  return __builtin_bswap64(x);
}

// T/G bases among staged bases [a, a + l) from the per-64-base masks.
__device__ __forceinline__ uint32_t sk_tg(const uint64_t* tg, uint32_t a, uint32_t l) {
  uint32_t n = 0;
  uint32_t i = a;
  const uint32_t e = a + l;
  while (i < e) {
    const uint32_t wd = i >> 6, sh = i & 63;
    const uint32_t take = min(64u - sh, e - i);
    uint64_t m = tg[wd] >> sh;
    if (take < 64) m &= (1ull << take) - 1;
    n += __popcll(m);
    i += take;
  }
  return n;
}

__device__ __forceinline__ uint32_t sk_bit(const uint64_t* tg, uint32_t i) { return (tg[i >> 6] >> (i & 63)) & 1; }

__global__ __launch_bounds__(64 * SK_WAVES) void sketch_kernel(SkArgs a) {
  extern __shared__ __align__(16) unsigned char sk_lds[];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * SK_WAVES + wave;
  if (t >= a.ntasks) return;
  const SkLayout L(a.k);
  uint64_t* be = (uint64_t*)(sk_lds + (size_t)wave * L.stride);  // first base most significant
  uint64_t* le = be + L.words;                                     // first base least significant
  uint64_t* tg = le + L.words;
  uint64_t* hs = tg + L.bytes / 64;
  uint64_t* vs = hs + SK_SPAN;
  const SkTask T = a.tasks[t];
  const uint32_t k = a.k, w = a.w, l = k + w - 1;

  // Stage the bases of positions x0 - 4 + i (an 'A' outside the sequence:
  // such k-mers belong to no window that is reported or consulted): all
  // loads first, then per 64 bases one ballot per code bit (the second is
  // the T/G mask); lane q then interleaves the two ballots of its 32-base
  // half into the little-endian word q and reverses it.
  const int64_t p0 = (int64_t)T.x0 - (int64_t)SK_PER;
  uint32_t bv[SK_ROUNDS];
#pragma unroll
  for (uint32_t r = 0; r < SK_ROUNDS; ++r) {
    const uint32_t i = 64 * r + lane;
    const int64_t p = p0 + i;
    bv[r] = (i < L.bytes && p >= 0 && p < (int64_t)T.len) ? a.seq[T.start + (uint64_t)p] : (uint32_t)'A';
  }
  uint64_t lo = 0, hi = 0;
  const uint32_t half = lane >> 1;
#pragma unroll
  for (uint32_t r = 0; r < SK_ROUNDS; ++r) {
    const uint64_t bl = __ballot((bv[r] >> 1) & 1), bh = __ballot((bv[r] >> 2) & 1);
    if (half == r) {
      lo = bl;
      hi = bh;
    }
    if (lane == 0 && 64 * r < L.bytes) tg[r] = bh;
  }
  if (lane + 1 < L.words) {
    const uint32_t sh = 32 * (lane & 1);
    const uint64_t word = sk_spread((lo >> sh) & 0xFFFFFFFFull) | (sk_spread((hi >> sh) & 0xFFFFFFFFull) << 1);
    le[lane] = word;
    be[lane] = sk_rev2(word);
  } else if (lane + 1 == L.words) {
    le[lane] = 0;
    be[lane] = 0;
  }
  sk_wave_sync();

  // k-mers m0 .. m0 + 3: forward value F (first base most significant),
  // reverse complement's value R (complemented codes, first base least
  // significant), canonical min, hash.
  const uint64_t kmask = k == 32 ? ~0ull : ((1ull << (2 * k)) - 1);
  const uint64_t comp = 0xAAAAAAAAAAAAAAAAull & kmask;
  const uint32_t m0 = SK_PER * lane;
#pragma unroll
  for (uint32_t j = 0; j < SK_PER; ++j) {
    const uint32_t m = m0 + j;
    const uint32_t q = m >> 5, o = 2 * (m & 31);
    const uint64_t b0 = be[q], b1 = be[q + 1], l0 = le[q], l1 = le[q + 1];
    const uint64_t W = o ? (b0 << o) | (b1 >> (64 - o)) : b0;
    const uint64_t Z = o ? (l0 >> o) | (l1 << (64 - o)) : l0;
    const uint64_t F = W >> (64 - 2 * k);
    const uint64_t R = (Z & kmask) ^ comp;
    const uint64_t V = (a.canonical && R < F) ? R : F;
    vs[m] = V;
    hs[m] = sk_mix64(V);
  }
  sk_wave_sync();

  // Windows r = 4 lane - 4 + j (relative to x0), first k-mer m0 + j; a
  // window is looked at only when reported or consulted (lane 0: the last,
  // as predecessor).  For w >= 4 the k-mers m0 + 3 .. m0 + w - 1 are common
  // to the lane's windows: their minimum, with its first and last index,
  // is found once.
  const bool lane_need = lane == 0 ? T.x0 > 0 : (int64_t)m0 - SK_PER < (int64_t)T.nx;
  const bool shared = w >= SK_PER;
  uint64_t shm = ~0ull;
  uint32_t shl = 0, shr = 0;
  if (shared && lane_need) {
    shm = hs[m0 + 3];
    shl = shr = m0 + 3;
    for (uint32_t i = m0 + 4; i < m0 + w; ++i) {
      const uint64_t h = hs[i];
      if (h < shm) {
        shm = h;
        shl = shr = i;
      } else if (h == shm) {
        shr = i;
      }
    }
  }
  uint32_t best[SK_PER];
  bool emit[SK_PER];
  uint32_t ntg = 0;
  bool rolled = false;  // ntg holds the previous window's count
#pragma unroll
  for (uint32_t j = 0; j < SK_PER; ++j) {
    const uint32_t m = m0 + j;
    const int64_t r = (int64_t)m - SK_PER;
    const bool need = lane == 0 ? (j == SK_PER - 1 && T.x0 > 0) : r < (int64_t)T.nx;
    uint32_t bi = m;
    if (need) {
      bool left = true;
      if (a.canonical) {
        ntg = rolled ? ntg + sk_bit(tg, m + l - 1) - sk_bit(tg, m - 1) : sk_tg(tg, m, l);
        rolled = true;
        left = 2 * ntg > l;
      }
      uint64_t hb;
      if (shared && m == m0 + 3) {
        hb = shm;  // the window starts at the common block
        bi = left ? shl : shr;
      } else {
        hb = hs[m];
        const uint32_t e = shared ? m0 + 3 : m + w;
        for (uint32_t i = m + 1; i < e; ++i) {
          const uint64_t h = hs[i];
          if (h < hb || (!left && h == hb)) {
            hb = h;
            bi = i;
          }
        }
        if (shared && (shm < hb || (!left && shm == hb))) {
          hb = shm;
          bi = left ? shl : shr;
        }
      }
      if (shared) {
        for (uint32_t i = m0 + w; i < m + w; ++i) {
          const uint64_t h = hs[i];
          if (h < hb || (!left && h == hb)) {
            hb = h;
            bi = i;
          }
        }
      }
    }
    best[j] = bi;
    emit[j] = lane > 0 && r < (int64_t)T.nx;
  }
  const uint32_t prev_lane = __shfl_up(best[SK_PER - 1], 1);
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t j = 0; j < SK_PER; ++j) {
    const uint32_t prev = j == 0 ? prev_lane : best[j - 1];
    const bool first = T.x0 == 0 && lane == 1 && j == 0;  // the sequence's first window
    emit[j] = emit[j] && (first || best[j] != prev);
    cnt += emit[j];
  }
  // Rank of the lane's first entry: exclusive prefix of cnt (0..4) over the
  // lanes, from one ballot per bit.
  uint32_t excl = 0, total = 0;
#pragma unroll
  for (uint32_t q = 0; q < 3; ++q) {
    const uint64_t bq = __ballot((cnt >> q) & 1);
    const uint64_t below = lane ? (bq & ((1ull << lane) - 1)) : 0ull;
    excl += (uint32_t)__popcll(below) << q;
    total += (uint32_t)__popcll(bq) << q;
  }
  uint64_t o = (uint64_t)t * a.tw + excl;
#pragma unroll
  for (uint32_t j = 0; j < SK_PER; ++j) {
    if (emit[j]) {
      a.stage_v[o] = vs[best[j]];
      a.stage_p[o] = (uint32_t)(p0 + best[j]);
      ++o;
    }
  }
  if (lane == 0) a.counts[t] = total;
}

// Exclusive scan of the per-task counts into base[0 .. n] (base[n] =
// total), 16384 counts per pass (16 per thread, four 16-byte loads), and
// the offsets of the sequences after the last task's (= total).
__global__ __launch_bounds__(SCAN_THREADS) void sketch_scan_kernel(const uint32_t* counts, uint64_t n,
                                                                   uint64_t* base, uint64_t* out_off,
                                                                   uint64_t tail_from, uint64_t nseq) {
  __shared__ uint64_t wsum[SCAN_THREADS / 64];
  __shared__ uint64_t carry_s;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint64_t carry = 0;
  for (uint64_t t0 = 0; t0 < n; t0 += (uint64_t)SCAN_PER * SCAN_THREADS) {
    const uint64_t i = t0 + (uint64_t)SCAN_PER * tid;
    uint32_t c[SCAN_PER];
    if (i + SCAN_PER <= n) {
#pragma unroll
      for (uint32_t j = 0; j < SCAN_PER; j += 4) {
        const uint4 v = *(const uint4*)(counts + i + j);
        c[j] = v.x;
        c[j + 1] = v.y;
        c[j + 2] = v.z;
        c[j + 3] = v.w;
      }
    } else {
#pragma unroll
      for (uint32_t j = 0; j < SCAN_PER; ++j) c[j] = i + j < n ? counts[i + j] : 0;
    }
    uint64_t mine = 0;
#pragma unroll
    for (uint32_t j = 0; j < SCAN_PER; ++j) mine += c[j];
    uint64_t inc = mine;  // inclusive scan over the wave
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint64_t y = __shfl_up(inc, d);
      if (lane >= d) inc += y;
    }
    if (lane == 63) wsum[wv] = inc;
    __syncthreads();
    uint64_t run = carry + inc - mine;
    for (uint32_t u = 0; u < wv; ++u) run += wsum[u];
    if (i + SCAN_PER <= n) {
#pragma unroll
      for (uint32_t j = 0; j < SCAN_PER; ++j) {
        base[i + j] = run;
        run += c[j];
      }
    } else {
      for (uint32_t j = 0; j < SCAN_PER; ++j) {
        if (i + j < n) base[i + j] = run;
        run += c[j];
      }
    }
    if (tid == SCAN_THREADS - 1) carry_s = run;
    __syncthreads();
    carry = carry_s;
    __syncthreads();
  }
  if (tid == 0) base[n] = carry;
  for (uint64_t s = tail_from + tid; s <= nseq; s += SCAN_THREADS) out_off[s] = carry;
}

// Each task's staged entries to their final offsets, and the offsets of
// the sequences that start at this task.
__global__ __launch_bounds__(64 * SK_WAVES) void sketch_copy_kernel(SkArgs a, const uint64_t* base, uint64_t* vals,
                                                                    uint32_t* pos, uint64_t* out_off) {
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * SK_WAVES + wave;
  if (t >= a.ntasks) return;
  const uint32_t n = a.counts[t];
  const uint64_t b = base[t];
  const uint64_t src = (uint64_t)t * a.tw;
  for (uint32_t i = lane; i < n; i += 64) {
    vals[b + i] = a.stage_v[src + i];
    pos[b + i] = a.stage_p[src + i];
  }
  const SkTask T = a.tasks[t];
  for (uint32_t s = T.seq_lo + lane; s <= T.seq; s += 64) out_off[s] = b;
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

int sfail(phx_ctx* ctx, int code, const std::string& msg) {
  phx::host::set_error(ctx, msg, code);
  return code;
}

}  // namespace

struct phx_sketch_plan {
  phx_ctx* ctx = nullptr;
  uint32_t k = 0, w = 0, canonical = 0, nseq = 0, tw = 0;
  uint64_t ntasks = 0, cap = 0, tail_from = 0;
  Buf seq, tasks, counts, base, out_off, vals, pos, stage_v, stage_p;
  bool ran = false;
  uint64_t total = 0;
  bool total_known = false;
  hipEvent_t ev[4] = {};  // around the compute, scan and copy launches
};

extern "C" {

phx_sketch_plan* phx_sketch_plan_create(phx_ctx* ctx, const uint8_t* seqs, const uint64_t* seq_off, uint32_t nseq,
                                        uint32_t k, uint32_t w, int32_t canonical) {
  if (!ctx) return nullptr;
  if (!seq_off) {
    phx::host::set_error(ctx, "phx_sketch_plan_create: null seq_off");
    return nullptr;
  }
  if (k == 0 || k > SK_MAX_K || w == 0 || w > SK_MAX_W) {
    phx::host::set_error(ctx, "phx_sketch_plan_create: k must lie in [1, 32] and w in [1, 128]");
    return nullptr;
  }
  const uint64_t l = (uint64_t)k + w - 1;
  if (canonical && l % 2 == 0) {
    phx::host::set_error(ctx, "phx_sketch_plan_create: canonical minimizers need k + w - 1 odd");
    return nullptr;
  }
  const uint64_t nbytes = seq_off[nseq];
  if (nbytes && !seqs) {
    phx::host::set_error(ctx, "phx_sketch_plan_create: null sequence bytes");
    return nullptr;
  }
  const uint32_t tw = sk_task_windows(w);
  std::vector<SkTask> tasks;
  uint64_t cap = 0;
  uint32_t next_seq = 0;  // first sequence whose offset no task has taken yet
  for (uint32_t s = 0; s < nseq; ++s) {
    const uint64_t st = seq_off[s], en = seq_off[s + 1];
    if (en < st || en > nbytes || en - st >= (1ull << 32)) {
      phx::host::set_error(ctx, "phx_sketch_plan_create: seq_off must be non-decreasing, sequences under 2^32 bytes");
      return nullptr;
    }
    const uint64_t n = en - st;
    if (n < l) continue;
    const uint64_t nw = n - l + 1;
    cap += nw;
    for (uint64_t x = 0; x < nw; x += tw) {
      SkTask T{st, (uint32_t)n, (uint32_t)x, (uint32_t)std::min<uint64_t>(tw, nw - x), s, x ? s + 1 : next_seq, 0};
      tasks.push_back(T);
    }
    next_seq = s + 1;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) {
    phx::host::set_error(ctx, "phx_sketch_plan_create: hipSetDevice", PHX_E_DEVICE);
    return nullptr;
  }
  phx_sketch_plan* P = new phx_sketch_plan();
  P->ctx = ctx;
  P->k = k;
  P->w = w;
  P->tw = tw;
  P->canonical = canonical ? 1 : 0;
  P->nseq = nseq;
  P->ntasks = tasks.size();
  P->cap = cap;
  P->tail_from = next_seq;
  const size_t slots = tasks.size() * (size_t)tw;
  bool ok = alloc_copy(P->seq, seqs, nbytes) && alloc_copy(P->tasks, tasks.data(), tasks.size() * sizeof(SkTask)) &&
            alloc_copy(P->counts, nullptr, tasks.size() * 4) && alloc_copy(P->base, nullptr, (tasks.size() + 1) * 8) &&
            alloc_copy(P->out_off, nullptr, ((size_t)nseq + 1) * 8) && alloc_copy(P->vals, nullptr, cap * 8) &&
            alloc_copy(P->pos, nullptr, cap * 4) && alloc_copy(P->stage_v, nullptr, slots * 8) &&
            alloc_copy(P->stage_p, nullptr, slots * 4);
  for (hipEvent_t& e : P->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  if (!ok) {
    phx::host::set_error(ctx, "phx_sketch_plan_create: device allocation or upload failed", PHX_E_NOMEM);
    phx_sketch_plan_destroy(P);
    return nullptr;
  }
  return P;
}

int phx_sketch_plan_run(phx_sketch_plan* P) {
  if (!P) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "hipSetDevice");
  hipStream_t s = ctx->stream;
  P->ran = false;
  P->total_known = false;
  if (hipEventRecord(P->ev[0], s) != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "event");
  SkArgs a{};
  a.seq = (const uint8_t*)P->seq.p;
  a.tasks = (const SkTask*)P->tasks.p;
  a.ntasks = (uint32_t)P->ntasks;
  a.k = P->k;
  a.w = P->w;
  a.canonical = P->canonical;
  a.tw = P->tw;
  a.counts = (uint32_t*)P->counts.p;
  a.stage_v = (uint64_t*)P->stage_v.p;
  a.stage_p = (uint32_t*)P->stage_p.p;
  const SkLayout L(P->k);
  const size_t lds = (size_t)L.stride * SK_WAVES;
  const uint32_t grid = (uint32_t)((P->ntasks + SK_WAVES - 1) / SK_WAVES);
  if (grid) {
    hipLaunchKernelGGL(sketch_kernel, dim3(grid), dim3(64 * SK_WAVES), lds, s, a);
    if (hipGetLastError() != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "sketch launch");
  }
  if (hipEventRecord(P->ev[1], s) != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "event");
  hipLaunchKernelGGL(sketch_scan_kernel, dim3(1), dim3(SCAN_THREADS), 0, s, (const uint32_t*)P->counts.p,
                     P->ntasks, (uint64_t*)P->base.p, (uint64_t*)P->out_off.p, P->tail_from, (uint64_t)P->nseq);
  if (hipGetLastError() != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "sketch scan launch");
  if (hipEventRecord(P->ev[2], s) != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "event");
  if (grid) {
    hipLaunchKernelGGL(sketch_copy_kernel, dim3(grid), dim3(64 * SK_WAVES), 0, s, a, (const uint64_t*)P->base.p,
                       (uint64_t*)P->vals.p, (uint32_t*)P->pos.p, (uint64_t*)P->out_off.p);
    if (hipGetLastError() != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "sketch copy launch");
  }
  if (hipEventRecord(P->ev[3], s) != hipSuccess || hipEventSynchronize(P->ev[3]) != hipSuccess)
    return sfail(ctx, PHX_E_DEVICE, "sketch run");
  P->ran = true;
  return PHX_OK;
}

uint64_t phx_sketch_plan_total(phx_sketch_plan* P) {
  if (!P || !P->ran) return 0;
  if (!P->total_known) {
    std::lock_guard<std::mutex> lk(P->ctx->mu);
    uint64_t t = 0;
    if (hipSetDevice(P->ctx->device) != hipSuccess ||
        hipMemcpy(&t, (const uint64_t*)P->out_off.p + P->nseq, 8, hipMemcpyDeviceToHost) != hipSuccess)
      return 0;
    P->total = t;
    P->total_known = true;
  }
  return P->total;
}

int phx_sketch_plan_results(phx_sketch_plan* P, uint64_t* values, uint32_t* positions, uint64_t cap,
                            uint64_t* off, uint64_t* total) {
  if (!P || !off) return PHX_E_ARG;
  phx_ctx* ctx = P->ctx;
  if (!P->ran) return sfail(ctx, PHX_E_ARG, "phx_sketch_plan_results before a run");
  std::lock_guard<std::mutex> lk(ctx->mu);
  const size_t no = (size_t)P->nseq + 1;
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipMemcpy(off, P->out_off.p, no * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return sfail(ctx, PHX_E_DEVICE, "sketch offsets download");
  const uint64_t n = off[P->nseq];
  P->total = n;
  P->total_known = true;
  if (total) *total = n;
  if (cap < n || (n && (!values || !positions))) return PHX_E_CAPACITY;
  if (n && (hipMemcpy(values, P->vals.p, n * 8, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(positions, P->pos.p, n * 4, hipMemcpyDeviceToHost) != hipSuccess))
    return sfail(ctx, PHX_E_DEVICE, "sketch download");
  return PHX_OK;
}

uint64_t phx_sketch_plan_capacity(const phx_sketch_plan* P) { return P ? P->cap : 0; }

double phx_sketch_plan_last_kernel_ms(const phx_sketch_plan* P) {
  if (!P || !P->ran) return 0.0;
  float ms = 0;
  if (hipEventElapsedTime(&ms, P->ev[0], P->ev[3]) != hipSuccess) return 0.0;
  return ms;
}

double phx_sketch_plan_last_launch_ms(const phx_sketch_plan* P, uint32_t i) {
  if (!P || !P->ran || i > 2) return 0.0;
  float ms = 0;
  if (hipEventElapsedTime(&ms, P->ev[i], P->ev[i + 1]) != hipSuccess) return 0.0;
  return ms;
}

void phx_sketch_plan_destroy(phx_sketch_plan* P) {
  if (!P) return;
  (void)hipStreamSynchronize(P->ctx->stream);
  for (Buf* b : {&P->seq, &P->tasks, &P->counts, &P->base, &P->out_off, &P->vals, &P->pos, &P->stage_v,
                 &P->stage_p})
    free_buf(*b);
  for (hipEvent_t e : P->ev)
    if (e) (void)hipEventDestroy(e);
  delete P;
}

int phx_sketch_batch(phx_ctx* ctx, const uint8_t* seqs, const uint64_t* seq_off, uint32_t nseq, uint32_t k,
                     uint32_t w, int32_t canonical, uint64_t* values, uint32_t* positions, uint64_t cap,
                     uint64_t* off, uint64_t* total) {
  if (!ctx || !seq_off || !off) return PHX_E_ARG;
  phx_sketch_plan* P = phx_sketch_plan_create(ctx, seqs, seq_off, nseq, k, w, canonical);
  if (!P) return phx::host::last_error_code(ctx);
  int rc = phx_sketch_plan_run(P);
  if (rc == PHX_OK) rc = phx_sketch_plan_results(P, values, positions, cap, off, total);
  phx_sketch_plan_destroy(P);
  return rc;
}

}  // extern "C"
