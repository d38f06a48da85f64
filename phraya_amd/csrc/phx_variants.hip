// phx_variants.hip -- post-extension step: variant sites from CIGAR runs.
//
// Batched form of extract_variants_from_cigar (phraya-align/src/executor.rs:
// 814-1044) with the coverage windows of compute_windowed_coverage
// (:1071-1090, summed as get_abs_multi, :218-220) and the repeat / hotspot
// lookups (is_in_repeat_region / is_in_hotspot, :795-811).  Upstream walks
// one read's CIGAR string per call on a rayon thread; here one thread walks
// one alignment's runs, twice: a counting pass, an exclusive scan of the
// counts, then a writing pass at the scanned offsets, so the records come out
// in alignment order and, per alignment, in CIGAR order, exactly as
// upstream's Vec would be concatenated; a last pass gives each variant a wave
// for its coverage window and interval lookups.
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
};

// Membership of [start, end) intervals sorted by start, non-overlapping:
// the last interval starting at or before pos must end after it.
template <class T>
__device__ bool in_sorted(uint64_t pos, const T* iv, uint64_t n) {
  uint64_t lo = 0, hi = n;  // first interval with start > pos
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if ((uint64_t)iv[2 * mid] <= pos) lo = mid + 1;
    else hi = mid;
  }
  return lo > 0 && pos < (uint64_t)iv[2 * lo - 1];
}

// One alignment's variants.  WRITE = false: counts only (*nv variants, *nc
// coverage words); WRITE = true: the records at out[0 ..], their windows at
// coverage offsets from cov_base (the values come from the coverage pass).
template <bool WRITE>
__device__ void walk(const VarDev& a, uint32_t i, phx_variant* out, uint64_t cov_base,
                     uint64_t* nv, uint64_t* nc) {
  const uint8_t* q = a.q_seq + a.q_off[i];
  const uint64_t qn = a.q_off[i + 1] - a.q_off[i], tn = a.target_len;
  uint64_t q_pos = 0, t_pos = a.target_start[i], v = 0, c = 0;
  auto emit = [&](uint64_t p, uint8_t type, uint8_t ref, uint8_t alt, uint64_t iq, uint64_t il) {
    const uint64_t ws = p >= a.radius ? p - a.radius : 0;
    const uint64_t we = min(p + a.radius + 1, tn);
    if (WRITE) {
      phx_variant r;
      r.alignment = i;
      r.position = (uint32_t)p;
      r.type = type;
      r.ref_base = ref;
      r.alt_base = alt;
      r.flags = 0;  // set by the coverage pass (its thread for the window's first word)
      r.ins_query_pos = (uint32_t)iq;
      r.ins_len = (uint32_t)il;
      r.window_start = (uint32_t)ws;
      r.window_len = (uint32_t)(we - ws);
      r.window_offset = (uint32_t)(p - ws);
      r.coverage_off = cov_base + c;
      out[v] = r;
    }
    ++v;
    c += we - ws;
  };
  for (uint64_t r = a.run_off[i]; r < a.run_off[i + 1]; ++r) {
    const uint32_t run = a.runs[r];
    const uint64_t len = run >> 4;
    switch (run & 15u) {
      case phx::OP_M:
        q_pos += len;
        t_pos += len;
        break;
      case phx::OP_X:
        for (uint64_t k = 0; k < len; ++k)
          if (q_pos + k < qn && t_pos + k < tn)
            emit(t_pos + k, PHX_VAR_SNV, a.target[t_pos + k], q[q_pos + k], 0, 0);
        q_pos += len;
        t_pos += len;
        break;
      case phx::OP_I:
        if (t_pos < tn && q_pos < qn) emit(t_pos, PHX_VAR_DELETION, a.target[t_pos], '.', 0, 0);
        t_pos += len;
        break;
      case phx::OP_D:
        if (q_pos < qn && t_pos < tn) emit(t_pos, PHX_VAR_INSERTION, '.', 0, q_pos, min(q_pos + len, qn) - q_pos);
        q_pos += len;
        break;
      default:
        break;
    }
  }
  *nv = v;
  *nc = c;
}

__global__ __launch_bounds__(256) void variant_count_kernel(VarDev a, uint64_t* nvar, uint64_t* ncov) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > a.n) return;
  uint64_t v = 0, c = 0;
  if (i < a.n) walk<false>(a, i, nullptr, 0, &v, &c);
  nvar[i] = v;  // entry n stays 0, so the exclusive scan's last entry is the total
  ncov[i] = c;
}

// Exclusive prefix sums of two arrays of m entries in one workgroup: each
// thread sums a contiguous chunk, the workgroup scans the chunk sums in LDS,
// then each thread writes its chunk's prefixes.  (Batches are thousands to
// hundreds of thousands of alignments: a few loads per thread.)
__global__ __launch_bounds__(1024) void scan2_kernel(const uint64_t* a, const uint64_t* b, uint64_t* ao,
                                                     uint64_t* bo, uint32_t m, uint64_t* totals) {
  __shared__ uint64_t sa[1024], sb[1024];
  const uint32_t t = threadIdx.x, T = blockDim.x;
  const uint32_t chunk = (m + T - 1) / T;
  const uint32_t lo = min(m, t * chunk), hi = min(m, lo + chunk);
  uint64_t xa = 0, xb = 0;
  for (uint32_t i = lo; i < hi; ++i) {
    xa += a[i];
    xb += b[i];
  }
  sa[t] = xa;
  sb[t] = xb;
  __syncthreads();
  for (uint32_t d = 1; d < T; d <<= 1) {  // inclusive scan of the chunk sums
    const uint64_t va = t >= d ? sa[t - d] : 0, vb = t >= d ? sb[t - d] : 0;
    __syncthreads();
    sa[t] += va;
    sb[t] += vb;
    __syncthreads();
  }
  uint64_t pa = sa[t] - xa, pb = sb[t] - xb;
  for (uint32_t i = lo; i < hi; ++i) {
    const uint64_t va = a[i], vb = b[i];
    ao[i] = pa;
    bo[i] = pb;
    pa += va;
    pb += vb;
  }
  // The totals straight to pinned host memory (system-scope stores), so the
  // host reads them after the stream sync with no copy.
  if (t == T - 1) {
    __hip_atomic_store(&totals[0], sa[T - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&totals[1], sb[T - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

__global__ __launch_bounds__(256) void variant_write_kernel(VarDev a, const uint64_t* var_off,
                                                            const uint64_t* cov_off, phx_variant* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  uint64_t v = 0, c = 0;
  walk<true>(a, i, out + var_off[i], cov_off[i], &v, &c);
}

// One wave per variant: lane 0 looks the variant up in the repeat and
// hotspot intervals (two binary searches), and the lanes fill its coverage
// window, consecutive lanes consecutive words: for each target position, how
// many of the alignment's windows [min(start, |target|), min(end, |target|))
// hold it.  The grid comes from the host's bound on the variant count; waves
// past the total (read on the device) return.
__global__ __launch_bounds__(256) void variant_coverage_kernel(VarDev a, phx_variant* out,
                                                               const uint64_t* var_off, uint32_t* cov) {
  const uint64_t v = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t lane = threadIdx.x & 63;
  if (v >= var_off[a.n]) return;
  phx_variant& r = out[v];
  const uint32_t ws = r.window_start, wl = r.window_len, al = r.alignment;
  const uint64_t co = r.coverage_off, tn = a.target_len;
  if (lane == 0)
    r.flags = (uint8_t)((in_sorted(r.position, a.repeats, a.n_repeats) ? PHX_VAR_IN_REPEAT : 0) |
                        (in_sorted(r.position, a.hotspots, a.n_hotspots) ? PHX_VAR_IN_HOTSPOT : 0));
  const uint64_t s0 = a.span_off[al], s1 = a.span_off[al + 1];
  for (uint32_t j = lane; j < wl; j += 64) {
    const uint64_t pos = (uint64_t)ws + j;
    uint32_t k = 0;
    for (uint64_t s = s0; s < s1; ++s) {
      const uint64_t lo = min(a.spans[2 * s], tn), hi = min(a.spans[2 * s + 1], tn);
      k += (lo < hi && pos >= lo && pos < hi) ? 1u : 0u;
    }
    cov[co + j] = k;
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
  Buf nvar, ncov, var_off, cov_off, out, cov;
  uint64_t cap_v = 0, cap_c = 0;  // device capacities (upper bounds from the runs)
  uint64_t n_var = 0, n_cov = 0;  // totals of the last run
  bool ran = false;
  uint64_t* tot_host = nullptr;  // pinned, mapped: the scan kernel's totals
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
  if (ctx) ctx->err = msg;
  return code;
}

}  // namespace

extern "C" {

phx_variant_plan* phx_variant_plan_create(phx_ctx* ctx, const phx_variant_batch* b) {
  if (!ctx || !b) return nullptr;
  const uint32_t n = b->n;
  if (n && (!b->run_off || !b->target_start || !b->q_off || !b->span_off)) {
    ctx->err = "phx_variant_plan_create: missing per-alignment arrays";
    return nullptr;
  }
  const uint64_t nruns = n ? b->run_off[n] : 0, nq = n ? b->q_off[n] : 0,
                 nspans = n ? b->span_off[n] : 0;
  if ((nruns && !b->runs) || (nq && !b->q_seq) || (b->target_len && !b->target) ||
      (nspans && !b->spans) || (b->n_repeats && !b->repeats) || (b->n_hotspots && !b->hotspots)) {
    ctx->err = "phx_variant_plan_create: null array with a non-zero length";
    return nullptr;
  }
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
            alloc_copy(P->nvar, nullptr, (size_t)(n + 1) * 8) &&
            alloc_copy(P->ncov, nullptr, (size_t)(n + 1) * 8) &&
            alloc_copy(P->var_off, nullptr, (size_t)(n + 1) * 8) &&
            alloc_copy(P->cov_off, nullptr, (size_t)(n + 1) * 8) &&
            alloc_copy(P->out, nullptr, P->cap_v * sizeof(phx_variant)) &&
            alloc_copy(P->cov, nullptr, P->cap_c * 4);
  ok = ok && hipEventCreate(&P->ev0) == hipSuccess && hipEventCreate(&P->ev1) == hipSuccess &&
       hipHostMalloc((void**)&P->tot_host, 16, hipHostMallocMapped) == hipSuccess &&
       hipHostGetDevicePointer((void**)&P->tot_dev, P->tot_host, 0) == hipSuccess;
  if (!ok) {
    ctx->err = "phx_variant_plan_create: device allocation or upload failed";
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
  const dim3 block(256), grid((n + 1 + 255) / 256);
  uint64_t* nv = (uint64_t*)P->nvar.p;
  uint64_t* nc = (uint64_t*)P->ncov.p;
  uint64_t* vo = (uint64_t*)P->var_off.p;
  uint64_t* co = (uint64_t*)P->cov_off.p;
  if (hipEventRecord(P->ev0, st) != hipSuccess) return vfail(ctx, PHX_E_DEVICE, "event");
  hipLaunchKernelGGL(variant_count_kernel, grid, block, 0, st, P->a, nv, nc);
  P->tot_host[0] = P->tot_host[1] = ~0ull;
  hipLaunchKernelGGL(scan2_kernel, dim3(1), dim3(1024), 0, st, (const uint64_t*)nv, (const uint64_t*)nc,
                     vo, co, n + 1, P->tot_dev);
  if (hipGetLastError() != hipSuccess) return vfail(ctx, PHX_E_DEVICE, "variant count / scan");
  if (n) {
    hipLaunchKernelGGL(variant_write_kernel, dim3((n + 255) / 256), block, 0, st, P->a,
                       (const uint64_t*)vo, (const uint64_t*)co, (phx_variant*)P->out.p);
    hipLaunchKernelGGL(variant_coverage_kernel, dim3((unsigned)((P->cap_v + 3) / 4)), block, 0, st,
                       P->a, (phx_variant*)P->out.p, (const uint64_t*)vo, (uint32_t*)P->cov.p);
    if (hipGetLastError() != hipSuccess) return vfail(ctx, PHX_E_DEVICE, "variant write");
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
      hipMemcpy(out->var_off, P->var_off.p, ((size_t)P->a.n + 1) * 8, hipMemcpyDeviceToHost) != hipSuccess)
    return vfail(ctx, PHX_E_DEVICE, "variant download");
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
                 &P->span_off, &P->repeats, &P->hotspots, &P->nvar, &P->ncov, &P->var_off, &P->cov_off,
                 &P->out, &P->cov})
    free_buf(*b);
  if (P->ev0) (void)hipEventDestroy(P->ev0);
  if (P->ev1) (void)hipEventDestroy(P->ev1);
  if (P->tot_host) (void)hipHostFree(P->tot_host);
  delete P;
}

int phx_extract_variants(phx_ctx* ctx, const phx_variant_batch* b, phx_variant_results* out) {
  if (!ctx || !b || !out) return PHX_E_ARG;
  phx_variant_plan* P = phx_variant_plan_create(ctx, b);
  if (!P) return PHX_E_NOMEM;
  int rc = phx_variant_plan_run(P);
  if (rc == PHX_OK) rc = phx_variant_plan_results(P, out);
  phx_variant_plan_destroy(P);
  return rc;
}

}  // extern "C"
