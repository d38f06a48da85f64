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
// Work unit: a wave takes up to 252 consecutive windows of one sequence.
// Its bytes (and a T/G bit mask from one ballot per 64 bytes) are staged in
// LDS; each lane packs 4 consecutive k-mers by rolling shifts and stores
// their values and hashes in LDS; lane 0 then handles the four windows
// before the task's first (only the last one is used: the predecessor that
// decides whether the first window repeats it) and lanes 1..63 four windows
// each.  Three launches per run: count the reported k-mers per task, scan
// the counts (task and sequence offsets), then recompute and write each
// task's entries at its offset.  Reading a byte per base and writing 12
// bytes per reported k-mer (about 2 / (w + 1) per base) is the algorithmic
// traffic; the selection work per window bounds it long before HBM does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "phraya_hip.h"
#include "phx_host.h"

namespace {

constexpr uint32_t SK_PER = 4;                   // windows (and k-mers) per lane
constexpr uint32_t SK_SPAN = 64 * SK_PER;        // window slots per wave
constexpr uint32_t SK_TASK = SK_SPAN - SK_PER;   // windows per task; lane 0 holds the predecessors
constexpr uint32_t SK_WAVES = 4;                 // tasks per workgroup
constexpr uint32_t SK_MAX_K = 32;
constexpr uint32_t SK_MAX_W = 128;
constexpr uint32_t SCAN_THREADS = 1024;

struct SkTask {
  uint64_t start;   // the sequence's first byte in the batch buffer
  uint32_t len;     // its length
  uint32_t x0;      // first window (= first k-mer position) of the task
  uint32_t nx;      // windows in the task, 1..SK_TASK
  uint32_t seq;     // sequence index
};

struct SkArgs {
  const uint8_t* seq;
  const SkTask* tasks;
  uint32_t ntasks, k, w, canonical;
  uint32_t* counts;          // per task (count pass)
  const uint64_t* base;      // per task (write pass)
  uint64_t* vals;
  uint32_t* pos;
};

// LDS of one wave: staged bytes, T/G masks, then k-mer hashes and values.
struct SkLayout {
  uint32_t bytes;   // staged bytes, a multiple of 64
  uint32_t kmers;   // k-mers of the span
  uint32_t words;   // T/G mask words
  uint32_t stride;  // bytes per wave
  __host__ __device__ SkLayout(uint32_t k, uint32_t w) {
    kmers = SK_SPAN + w - 1;
    bytes = (kmers + k - 1 + 63) & ~63u;
    words = bytes / 64;
    stride = bytes + words * 8 + kmers * 16;
    stride = (stride + 15) & ~15u;
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

// T/G bases among staged bytes [a, a + l) from the per-64-byte masks.
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

template <bool WRITE>
__global__ __launch_bounds__(64 * SK_WAVES) void sketch_kernel(SkArgs a) {
  extern __shared__ __align__(16) unsigned char sk_lds[];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t t = blockIdx.x * SK_WAVES + wave;
  if (t >= a.ntasks) return;
  const SkLayout L(a.k, a.w);
  unsigned char* by = sk_lds + (size_t)wave * L.stride;
  uint64_t* tg = (uint64_t*)(by + L.bytes);
  uint64_t* hs = tg + L.words;
  uint64_t* vs = hs + L.kmers;
  const SkTask T = a.tasks[t];
  const uint32_t k = a.k, w = a.w, l = k + w - 1;

  // Stage bytes of positions x0 - 4 + i; outside the sequence an 'A' (such
  // k-mers belong to no window that is reported or consulted).
  const int64_t p0 = (int64_t)T.x0 - (int64_t)SK_PER;
  for (uint32_t i = lane; i < L.bytes; i += 64) {
    const int64_t p = p0 + i;
    const uint8_t b = (p >= 0 && p < (int64_t)T.len) ? a.seq[T.start + (uint64_t)p] : (uint8_t)'A';
    by[i] = b;
    const uint64_t m = __ballot((b >> 2) & 1);  // code T 2 / G 3
    if (lane == 0) tg[i >> 6] = m;
  }
  sk_wave_sync();

  // k-mers m .. m + 3 per lane: forward value F (first base most
  // significant), reverse complement's value R, canonical min.
  const uint64_t kmask = k == 32 ? ~0ull : ((1ull << (2 * k)) - 1);
  const uint32_t rsh = 2 * (k - 1);
  for (uint32_t m0 = SK_PER * lane; m0 < L.kmers; m0 += SK_SPAN) {
    uint64_t F = 0, R = 0;
    for (uint32_t j = 0; j + 1 < k; ++j) {
      const uint64_t c = (by[m0 + j] >> 1) & 3;
      F = (F << 2) | c;
      R = (R >> 2) | ((c ^ 2) << rsh);
    }
#pragma unroll
    for (uint32_t j = 0; j < SK_PER; ++j) {
      const uint32_t m = m0 + j;
      if (m < L.kmers) {
        const uint64_t c = (by[m + k - 1] >> 1) & 3;
        F = ((F << 2) | c) & kmask;
        R = (R >> 2) | ((c ^ 2) << rsh);
        const uint64_t V = (a.canonical && R < F) ? R : F;
        vs[m] = V;
        hs[m] = sk_mix64(V);
      }
    }
  }
  sk_wave_sync();

  // Windows r = 4 lane - 4 + j (relative to x0), first k-mer m = 4 lane + j.
  uint32_t best[SK_PER];
  bool emit[SK_PER];
#pragma unroll
  for (uint32_t j = 0; j < SK_PER; ++j) {
    const uint32_t m = SK_PER * lane + j;
    const int64_t r = (int64_t)m - SK_PER;
    const bool need = lane == 0 ? (j == SK_PER - 1 && T.x0 > 0) : r < (int64_t)T.nx;
    uint32_t bi = m;
    if (need) {
      const bool left = !a.canonical || 2 * sk_tg(tg, m, l) > l;
      uint64_t hb = hs[m];
      for (uint32_t i = 1; i < w; ++i) {
        const uint64_t h = hs[m + i];
        if (h < hb || (!left && h == hb)) {
          hb = h;
          bi = m + i;
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
  if (!WRITE) {
    uint32_t s = cnt;
    for (uint32_t o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) a.counts[t] = s;
    return;
  }
  // Rank of the lane's first entry: exclusive prefix of cnt (0..4) over the
  // lanes, from one ballot per bit.
  uint32_t excl = 0;
#pragma unroll
  for (uint32_t q = 0; q < 3; ++q) {
    const uint64_t bq = __ballot((cnt >> q) & 1);
    const uint64_t below = lane ? (bq & ((1ull << lane) - 1)) : 0ull;
    excl += (uint32_t)__popcll(below) << q;
  }
  uint64_t o = a.base[t] + excl;
#pragma unroll
  for (uint32_t j = 0; j < SK_PER; ++j) {
    if (emit[j]) {
      a.vals[o] = vs[best[j]];
      a.pos[o] = (uint32_t)(p0 + best[j]);
      ++o;
    }
  }
}

// Exclusive scan of the per-task counts into base[0 .. n] (base[n] = total),
// then each sequence's offset: out_off[s] = base[first_task[s]].
__global__ __launch_bounds__(SCAN_THREADS) void sketch_scan_kernel(const uint32_t* counts, uint64_t n,
                                                                   uint64_t* base, const uint64_t* first_task,
                                                                   uint64_t nseq, uint64_t* out_off) {
  __shared__ uint64_t part[SCAN_THREADS];
  const uint32_t tid = threadIdx.x;
  const uint64_t chunk = (n + SCAN_THREADS - 1) / SCAN_THREADS;
  const uint64_t lo0 = (uint64_t)tid * chunk;
  const uint64_t lo = lo0 < n ? lo0 : n, hi = lo + chunk < n ? lo + chunk : n;
  uint64_t s = 0;
  for (uint64_t i = lo; i < hi; ++i) s += counts[i];
  part[tid] = s;
  __syncthreads();
  for (uint32_t d = 1; d < SCAN_THREADS; d <<= 1) {
    const uint64_t v = tid >= d ? part[tid - d] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint64_t run = part[tid] - s;  // exclusive
  for (uint64_t i = lo; i < hi; ++i) {
    base[i] = run;
    run += counts[i];
  }
  if (tid == SCAN_THREADS - 1) base[n] = part[tid];
  __syncthreads();
  __threadfence_block();
  for (uint64_t i = tid; i <= nseq; i += SCAN_THREADS) out_off[i] = base[first_task[i]];
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
  uint32_t k = 0, w = 0, canonical = 0, nseq = 0;
  uint64_t ntasks = 0, cap = 0;
  Buf seq, tasks, first_task, counts, base, out_off, vals, pos;
  bool ran = false;
  uint64_t total = 0;
  bool total_known = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
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
  std::vector<SkTask> tasks;
  std::vector<uint64_t> first_task((size_t)nseq + 1);
  uint64_t cap = 0;
  for (uint32_t s = 0; s < nseq; ++s) {
    const uint64_t st = seq_off[s], en = seq_off[s + 1];
    if (en < st || en > nbytes || en - st >= (1ull << 32)) {
      phx::host::set_error(ctx, "phx_sketch_plan_create: seq_off must be non-decreasing, sequences under 2^32 bytes");
      return nullptr;
    }
    first_task[s] = tasks.size();
    const uint64_t n = en - st;
    if (n < l) continue;
    const uint64_t nw = n - l + 1;
    cap += nw;
    for (uint64_t x = 0; x < nw; x += SK_TASK)
      tasks.push_back(SkTask{st, (uint32_t)n, (uint32_t)x, (uint32_t)std::min<uint64_t>(SK_TASK, nw - x), s});
  }
  first_task[nseq] = tasks.size();
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) {
    phx::host::set_error(ctx, "phx_sketch_plan_create: hipSetDevice", PHX_E_DEVICE);
    return nullptr;
  }
  phx_sketch_plan* P = new phx_sketch_plan();
  P->ctx = ctx;
  P->k = k;
  P->w = w;
  P->canonical = canonical ? 1 : 0;
  P->nseq = nseq;
  P->ntasks = tasks.size();
  P->cap = cap;
  const bool ok = alloc_copy(P->seq, seqs, nbytes) &&
                  alloc_copy(P->tasks, tasks.data(), tasks.size() * sizeof(SkTask)) &&
                  alloc_copy(P->first_task, first_task.data(), first_task.size() * 8) &&
                  alloc_copy(P->counts, nullptr, tasks.size() * 4) &&
                  alloc_copy(P->base, nullptr, (tasks.size() + 1) * 8) &&
                  alloc_copy(P->out_off, nullptr, first_task.size() * 8) && alloc_copy(P->vals, nullptr, cap * 8) &&
                  alloc_copy(P->pos, nullptr, cap * 4) && hipEventCreate(&P->ev0) == hipSuccess &&
                  hipEventCreate(&P->ev1) == hipSuccess;
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
  if (hipEventRecord(P->ev0, s) != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "event");
  SkArgs a{};
  a.seq = (const uint8_t*)P->seq.p;
  a.tasks = (const SkTask*)P->tasks.p;
  a.ntasks = (uint32_t)P->ntasks;
  a.k = P->k;
  a.w = P->w;
  a.canonical = P->canonical;
  a.counts = (uint32_t*)P->counts.p;
  a.base = (const uint64_t*)P->base.p;
  a.vals = (uint64_t*)P->vals.p;
  a.pos = (uint32_t*)P->pos.p;
  const SkLayout L(P->k, P->w);
  const size_t lds = (size_t)L.stride * SK_WAVES;
  const uint32_t grid = (uint32_t)((P->ntasks + SK_WAVES - 1) / SK_WAVES);
  if (grid) {
    hipLaunchKernelGGL(sketch_kernel<false>, dim3(grid), dim3(64 * SK_WAVES), lds, s, a);
    if (hipGetLastError() != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "sketch count launch");
  }
  hipLaunchKernelGGL(sketch_scan_kernel, dim3(1), dim3(SCAN_THREADS), 0, s, (const uint32_t*)P->counts.p,
                     P->ntasks, (uint64_t*)P->base.p, (const uint64_t*)P->first_task.p, (uint64_t)P->nseq,
                     (uint64_t*)P->out_off.p);
  if (hipGetLastError() != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "sketch scan launch");
  if (grid) {
    hipLaunchKernelGGL(sketch_kernel<true>, dim3(grid), dim3(64 * SK_WAVES), lds, s, a);
    if (hipGetLastError() != hipSuccess) return sfail(ctx, PHX_E_DEVICE, "sketch write launch");
  }
  if (hipEventRecord(P->ev1, s) != hipSuccess || hipEventSynchronize(P->ev1) != hipSuccess)
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
  if (hipEventElapsedTime(&ms, P->ev0, P->ev1) != hipSuccess) return 0.0;
  return ms;
}

void phx_sketch_plan_destroy(phx_sketch_plan* P) {
  if (!P) return;
  (void)hipStreamSynchronize(P->ctx->stream);
  for (Buf* b : {&P->seq, &P->tasks, &P->first_task, &P->counts, &P->base, &P->out_off, &P->vals, &P->pos})
    free_buf(*b);
  if (P->ev0) (void)hipEventDestroy(P->ev0);
  if (P->ev1) (void)hipEventDestroy(P->ev1);
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
