// phx_io.cpp -- the output files of an alignment run: .phraya and
// .phraya.queries (host code; the records come from the device).
//
// write_phraya (phraya-io/src/phraya.rs:67-80) writes rmp_serde::to_vec of a
// PhrayaFile (:33-38, built by PhrayaFile::new, :40-62) compressed by zstd at
// level 3; write_queries (phraya-io/src/queries.rs:62-103) and
// write_cross_space_queries (:132-167) do the same for the query indexes
// after their filter and sort.  rmp_serde::to_vec writes a struct as a
// MessagePack array of its fields in declaration order (the reason
// VariantObservation keeps mate_info last, types.rs:281-286), integers in
// their shortest form (non-negative signed ones as unsigned), f64 always as
// float 64, strings as str, maps as map, unit enum variants as their
// (renamed) names.  Each observation is built as extract_variants_from_cigar
// (phraya-align/src/executor.rs:814-1044) builds it from a device variant
// record, its coverage window and the per-read fields the caller passes.
//
// zstd is the system's libzstd.so.1, loaded at first use; its documented
// one-shot API is declared here (no header on this system).  The compressed
// bytes differ from the Rust encoder's stream framing; every zstd decoder
// reads both to the same MessagePack payload.
#include <dlfcn.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "phraya_hip.h"

namespace {

// MessagePack writer with rmp's shortest encodings.
struct Mp {
  std::vector<uint8_t> b;
  void byte(uint8_t x) { b.push_back(x); }
  void be(uint64_t v, int n) {
    for (int i = n - 1; i >= 0; --i) b.push_back((uint8_t)(v >> (8 * i)));
  }
  void uint(uint64_t v) {
    if (v < 128) byte((uint8_t)v);
    else if (v < 256) { byte(0xcc); be(v, 1); }
    else if (v < 65536) { byte(0xcd); be(v, 2); }
    else if (v < (1ull << 32)) { byte(0xce); be(v, 4); }
    else { byte(0xcf); be(v, 8); }
  }
  void sint(int64_t v) {
    if (v >= 0) uint((uint64_t)v);
    else if (v >= -32) byte((uint8_t)(int8_t)v);
    else if (v >= -128) { byte(0xd0); be((uint64_t)v, 1); }
    else if (v >= -32768) { byte(0xd1); be((uint64_t)v, 2); }
    else if (v >= -2147483648ll) { byte(0xd2); be((uint64_t)v, 4); }
    else { byte(0xd3); be((uint64_t)v, 8); }
  }
  void f64(double d) {
    uint64_t u;
    std::memcpy(&u, &d, 8);
    byte(0xcb);
    be(u, 8);
  }
  void boolean(bool x) { byte(x ? 0xc3 : 0xc2); }
  void str(const char* s, size_t n) {
    if (n < 32) byte((uint8_t)(0xa0 | n));
    else if (n < 256) { byte(0xd9); be(n, 1); }
    else if (n < 65536) { byte(0xda); be(n, 2); }
    else { byte(0xdb); be(n, 4); }
    b.insert(b.end(), (const uint8_t*)s, (const uint8_t*)s + n);
  }
  void str(const char* s) { str(s, std::strlen(s)); }
  void arr(size_t n) {
    if (n < 16) byte((uint8_t)(0x90 | n));
    else if (n < 65536) { byte(0xdc); be(n, 2); }
    else { byte(0xdd); be(n, 4); }
  }
  void map(size_t n) {
    if (n < 16) byte((uint8_t)(0x80 | n));
    else if (n < 65536) { byte(0xde); be(n, 2); }
    else { byte(0xdf); be(n, 4); }
  }
};

constexpr uint32_t kPhrayaVersion = 1;      // PHRAYA_VERSION (phraya.rs:6)
constexpr double kScoreThreshold = 0.95;    // queries.rs:63, :137

// CoverageTrack::new (phraya-core/src/types.rs:626-651): each value
// quantized to the nearest multiple of 5 (capped at 255), then run-length
// encoded as (value, length) pairs.
void coverage_track(Mp& m, const uint32_t* cov, uint64_t n) {
  std::vector<std::pair<uint8_t, uint32_t>> runs;
  uint8_t cur = 0;
  uint32_t len = 0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t r = ((uint64_t)cov[i] + 2) / 5 * 5;
    const uint8_t q = (uint8_t)std::min<uint64_t>(r, 255);
    if (q == cur) {
      ++len;
    } else {
      if (len) runs.emplace_back(cur, len);
      cur = q;
      len = 1;
    }
  }
  if (len) runs.emplace_back(cur, len);
  m.arr(2);
  m.arr(runs.size());
  for (const auto& r : runs) {
    m.arr(2);
    m.uint(r.first);
    m.uint(r.second);
  }
  m.uint((uint32_t)n);
}

int encode_phraya(Mp& m, const phx_phraya_header* h, const phx_variant* v, uint64_t nv, const uint32_t* wcov,
                  uint64_t n_wcov, const phx_obs_read* reads, uint32_t nreads, const uint32_t* cov,
                  uint64_t cov_len) {
  if (!h || !h->sample_id || !h->timestamp || (nv && (!v || !reads)) || (cov_len && !cov)) return PHX_E_ARG;
  if (cov_len >= (1ull << 32)) return PHX_E_ARG;
  for (uint64_t i = 0; i < nv; ++i) {
    const phx_variant& x = v[i];
    if (x.alignment >= nreads || x.type > PHX_VAR_INSERTION) return PHX_E_ARG;
    if (x.window_len && (!wcov || x.coverage_off + x.window_len > n_wcov)) return PHX_E_ARG;
    const phx_obs_read& r = reads[x.alignment];
    if (!r.cigar || !r.provenance || (r.mate && !r.mate->mate_id) || r.strand > 1) return PHX_E_ARG;
    if (x.type == PHX_VAR_INSERTION && (uint64_t)x.ins_query_pos + x.ins_len > r.query_len) return PHX_E_ARG;
    if (x.type == PHX_VAR_INSERTION && x.ins_len && !r.query) return PHX_E_ARG;
  }
  // PhrayaFile { header, observations, coverage_track } (phraya.rs:33-38).
  m.arr(3);
  // PhrayaHeader (phraya.rs:22-30).
  m.arr(5);
  m.uint(kPhrayaVersion);
  m.uint(h->reference_length);
  m.str(h->sample_id);
  m.str(h->timestamp);
  m.uint(nv);
  m.arr(nv);
  for (uint64_t i = 0; i < nv; ++i) {
    const phx_variant& x = v[i];
    const phx_obs_read& r = reads[x.alignment];
    const bool hotspot = x.flags & PHX_VAR_IN_HOTSPOT;
    const double ku = hotspot ? 0.0 : 1.0;
    // VariantObservation (types.rs:217-287): 19 fields, 20 with mate_info.
    m.arr(r.mate ? 20 : 19);
    m.uint(x.position);
    m.uint(x.ref_base);
    // all_alleles, ascending keys (serialize_alleles_sorted).
    if (x.type == PHX_VAR_INSERTION) {
      uint32_t cnt[256] = {0};
      for (uint32_t j = 0; j < x.ins_len; ++j) ++cnt[r.query[x.ins_query_pos + j]];
      size_t keys = 0;
      for (int c = 0; c < 256; ++c) keys += cnt[c] != 0;
      m.map(keys);
      for (int c = 0; c < 256; ++c)
        if (cnt[c]) {
          m.uint((uint64_t)c);
          m.uint(cnt[c]);
        }
    } else {
      m.map(1);
      m.uint(x.type == PHX_VAR_SNV ? x.alt_base : (uint8_t)'.');
      m.uint(1);
    }
    m.f64(ku * r.confidence);
    m.str(r.cigar);
    m.uint(r.mapq);
    m.uint(r.edit_distance);
    m.arr(x.window_len);
    for (uint32_t j = 0; j < x.window_len; ++j) m.uint(wcov[x.coverage_off + j]);
    m.f64(r.avg_base_quality);
    m.str(r.provenance);
    m.boolean(x.flags & PHX_VAR_IN_REPEAT);
    m.str(x.type == PHX_VAR_SNV ? "snp" : x.type == PHX_VAR_DELETION ? "deletion" : "insertion");
    m.f64(ku);
    m.str(r.strand ? "-" : "+");
    // with_pair_counts(1, proper) and with_insert_stats only for reads with mate info.
    m.uint(r.mate ? 1 : 0);
    m.uint(r.mate && r.mate->proper_pair ? 1 : 0);
    m.sint(r.mate && r.has_insert_stats ? r.insert_size_sum : 0);
    m.uint(r.mate && r.has_insert_stats ? r.insert_size_count : 0);
    m.uint(x.window_offset);
    if (r.mate) {
      // MateInfo (types.rs:2031-2050).
      m.arr(6);
      m.str(r.mate->mate_id);
      m.boolean(r.mate->proper_pair);
      m.sint(r.mate->insert_size);
      m.boolean(r.mate->is_first_in_pair);
      m.boolean(r.mate->is_second_in_pair);
      m.boolean(r.mate->mate_mapped);
    }
  }
  coverage_track(m, cov, cov_len);
  return PHX_OK;
}

// QueryIndex after write_queries' filter (score >= 0.95, reads left with
// none dropped) and order (BTreeMap by id; per read by position, then score
// descending, stable).
int encode_queries(Mp& m, const phx_query_hits* q, uint32_t nq) {
  if (nq && !q) return PHX_E_ARG;
  std::map<std::string, std::vector<std::pair<uint32_t, double>>> kept;
  std::vector<std::string> seen;
  for (uint32_t i = 0; i < nq; ++i) {
    if (!q[i].query_id || (q[i].n && (!q[i].pos || !q[i].score))) return PHX_E_ARG;
    seen.emplace_back(q[i].query_id);
    std::vector<std::pair<uint32_t, double>> a;
    for (uint32_t j = 0; j < q[i].n; ++j)
      if (q[i].score[j] >= kScoreThreshold) a.emplace_back(q[i].pos[j], q[i].score[j]);
    if (a.empty()) continue;
    std::stable_sort(a.begin(), a.end(), [](const auto& x, const auto& y) {
      return x.first != y.first ? x.first < y.first : x.second > y.second;
    });
    kept.emplace(q[i].query_id, std::move(a));
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) return PHX_E_ARG;  // a map has unique keys
  m.map(kept.size());
  for (const auto& e : kept) {
    m.str(e.first.data(), e.first.size());
    m.arr(e.second.size());
    for (const auto& p : e.second) {
      m.arr(2);
      m.uint(p.first);
      m.f64(p.second);
    }
  }
  return PHX_OK;
}

// CrossSpaceQueryIndex after write_cross_space_queries' filter (identity
// >= 0.95) and order (by id; per read by space, position, then identity
// descending, stable); CrossSpacePlacement is [space, pos, identity].
int encode_cross_space(Mp& m, const phx_cross_space_hits* q, uint32_t nq) {
  if (nq && !q) return PHX_E_ARG;
  struct P {
    std::string space;
    uint32_t pos;
    double id;
  };
  std::map<std::string, std::vector<P>> kept;
  std::vector<std::string> seen;
  for (uint32_t i = 0; i < nq; ++i) {
    if (!q[i].query_id || (q[i].n && !q[i].placements)) return PHX_E_ARG;
    seen.emplace_back(q[i].query_id);
    std::vector<P> a;
    for (uint32_t j = 0; j < q[i].n; ++j) {
      const phx_placement& p = q[i].placements[j];
      if (!p.space) return PHX_E_ARG;
      if (p.identity >= kScoreThreshold) a.push_back(P{p.space, p.pos, p.identity});
    }
    if (a.empty()) continue;
    std::stable_sort(a.begin(), a.end(), [](const P& x, const P& y) {
      if (x.space != y.space) return x.space < y.space;
      if (x.pos != y.pos) return x.pos < y.pos;
      return x.id > y.id;
    });
    kept.emplace(q[i].query_id, std::move(a));
  }
  std::sort(seen.begin(), seen.end());
  if (std::adjacent_find(seen.begin(), seen.end()) != seen.end()) return PHX_E_ARG;
  m.map(kept.size());
  for (const auto& e : kept) {
    m.str(e.first.data(), e.first.size());
    m.arr(e.second.size());
    for (const auto& p : e.second) {
      m.arr(3);
      m.str(p.space.data(), p.space.size());
      m.uint(p.pos);
      m.f64(p.id);
    }
  }
  return PHX_OK;
}

// libzstd's one-shot API (zstd.h, stable since 1.0).
struct Zstd {
  size_t (*bound)(size_t) = nullptr;
  size_t (*compress)(void*, size_t, const void*, size_t, int) = nullptr;
  unsigned (*is_error)(size_t) = nullptr;
};

const Zstd* zstd() {
  static Zstd z;
  static bool ok = false;
  static std::once_flag once;
  std::call_once(once, [] {
    void* h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    z.bound = (size_t(*)(size_t))dlsym(h, "ZSTD_compressBound");
    z.compress = (size_t(*)(void*, size_t, const void*, size_t, int))dlsym(h, "ZSTD_compress");
    z.is_error = (unsigned (*)(size_t))dlsym(h, "ZSTD_isError");
    ok = z.bound && z.compress && z.is_error;
  });
  return ok ? &z : nullptr;
}

int copy_out(const Mp& m, uint8_t* out, uint64_t cap, uint64_t* size) {
  if (size) *size = m.b.size();
  if (cap < m.b.size() || (!out && !m.b.empty())) return PHX_E_CAPACITY;
  if (!m.b.empty()) std::memcpy(out, m.b.data(), m.b.size());
  return PHX_OK;
}

int compress_write(const char* path, const Mp& m, int level) {
  if (!path) return PHX_E_ARG;
  const Zstd* z = zstd();
  if (!z) return PHX_E_DEVICE;  // no libzstd on this system
  std::vector<uint8_t> c(z->bound(m.b.size()));
  const size_t n = z->compress(c.data(), c.size(), m.b.data(), m.b.size(), level);
  if (z->is_error(n)) return PHX_E_INTERNAL;
  FILE* f = std::fopen(path, "wb");
  if (!f) return PHX_E_ARG;
  const bool ok = std::fwrite(c.data(), 1, n, f) == n;
  return (std::fclose(f) == 0 && ok) ? PHX_OK : PHX_E_ARG;
}

}  // namespace

extern "C" {

int phx_zstd_available(void) { return zstd() != nullptr; }

int phx_phraya_encode(const phx_phraya_header* header, const phx_variant* variants, uint64_t n_variants,
                      const uint32_t* window_coverage, uint64_t n_window_coverage, const phx_obs_read* reads,
                      uint32_t n_reads, const uint32_t* coverage, uint64_t coverage_len, uint8_t* out,
                      uint64_t cap, uint64_t* size) {
  Mp m;
  const int rc = encode_phraya(m, header, variants, n_variants, window_coverage, n_window_coverage, reads, n_reads,
                               coverage, coverage_len);
  return rc != PHX_OK ? rc : copy_out(m, out, cap, size);
}

int phx_phraya_write(const char* path, const phx_phraya_header* header, const phx_variant* variants,
                     uint64_t n_variants, const uint32_t* window_coverage, uint64_t n_window_coverage,
                     const phx_obs_read* reads, uint32_t n_reads, const uint32_t* coverage, uint64_t coverage_len,
                     int level) {
  Mp m;
  const int rc = encode_phraya(m, header, variants, n_variants, window_coverage, n_window_coverage, reads, n_reads,
                               coverage, coverage_len);
  return rc != PHX_OK ? rc : compress_write(path, m, level);
}

int phx_queries_encode(const phx_query_hits* queries, uint32_t n, uint8_t* out, uint64_t cap, uint64_t* size) {
  Mp m;
  const int rc = encode_queries(m, queries, n);
  return rc != PHX_OK ? rc : copy_out(m, out, cap, size);
}

int phx_queries_write(const char* path, const phx_query_hits* queries, uint32_t n, int level) {
  Mp m;
  const int rc = encode_queries(m, queries, n);
  return rc != PHX_OK ? rc : compress_write(path, m, level);
}

int phx_cross_space_queries_encode(const phx_cross_space_hits* queries, uint32_t n, uint8_t* out, uint64_t cap,
                                   uint64_t* size) {
  Mp m;
  const int rc = encode_cross_space(m, queries, n);
  return rc != PHX_OK ? rc : copy_out(m, out, cap, size);
}

int phx_cross_space_queries_write(const char* path, const phx_cross_space_hits* queries, uint32_t n, int level) {
  Mp m;
  const int rc = encode_cross_space(m, queries, n);
  return rc != PHX_OK ? rc : compress_write(path, m, level);
}

}  // extern "C"
