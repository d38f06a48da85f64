// wfa_kernels.hip -- gap-linear and gap-affine WFA on gfx950.
//
// One workgroup owns one alignment at a time and pulls the next alignment from
// a device work queue (an atomic counter over a cost-sorted work list), so a
// launch of G workgroups drains any number of pairs of uneven length.  Within
// a workgroup, thread t computes diagonals k = lo + t, lo + t + NT, ... of the
// current score step.  The last R score frames (R = largest predecessor
// distance + 1) sit in a ring indexed by diagonal modulo a power-of-two width,
// in LDS for pairs narrow enough, otherwise in the workgroup's HBM slot.  For
// traceback each step also stores one 32-bit history cell per diagonal:
//   bits 31..4  pre-extension offset + 1 (0 = cell unset)
//   bits  3..0  op codes (affine: m_op | ins_op << 2 | del_op << 3;
//                linear: op 1 X, 2 query-only, 3 target-only)
// which is everything the upstream traceback reads: the M run before an edit
// is always (current offset - pre-extension offset of the cell).
//
// Semantics restated from phraya-align/src/wfa_simd.rs: linear recurrence and
// tie rule :283-330 (last maximum wins, bounds check after the choice),
// fitting end :940-955, linear traceback :959-1130, affine recurrence and
// endpoint :581-779 (strict-greater first-candidate rule), affine traceback
// :785-886, fitting/global switch :374 and :930.
#include <hip/hip_runtime.h>

#include "phx_device.h"
#include "wfa_common.h"

namespace phx {

constexpr int RMAX = 128;  // ring slots the LDS metadata can describe

// --- shared (LDS) control block -------------------------------------------------

struct WfaShared {
  int32_t lo[3][RMAX];  // per ring slot and layer (m, ins, del): diagonal range
  int32_t hi[3][RMAX];
  int32_t end_key[3];   // triple-buffered endpoint reduction
  uint32_t work;
};

// History region of one workgroup: cells grow up from the start, the step
// table ({lo, hi, cell offset, 0} per step) grows down from the end.
struct Hist {
  uint32_t* cells;
  int4* table_end;  // one past the last table entry
  uint64_t cap_cells;

  __device__ int4* entry(int32_t step) const { return table_end - 1 - step; }
  // Whether `w` more cells plus the table entry of `step` still fit.
  __device__ bool fits(uint64_t used, int32_t w, int32_t step) const {
    const uint64_t table_cells = (uint64_t)(step + 1) * 4;
    return used + (uint64_t)w + table_cells <= cap_cells;
  }
};

// --- affine ------------------------------------------------------------------------

template <int NT, bool RING_LDS, bool B8>
__device__ void affine_pair(const WfaArgs& a, WfaShared& sh, int32_t* ring, const PairDesc& pd,
                            PairOut* po, Hist H) {
  const int tid = threadIdx.x;
  const uint64_t* qw = a.seq + pd.q_word;
  const uint64_t* tw = a.seq + pd.t_word;
  const int32_t qn = pd.qn, tn = pd.tn;
  const int32_t x = a.cp.x, o = a.cp.o, e = a.cp.e;

  if (qn == 0) {  // fill_wfa_affine_fitting_impl: empty query (wfa_simd.rs:925-927)
    if (tid == 0) *po = PairOut{ST_OK, 0, 0, 0, 0, 0, 0, 0};
    return;
  }
  const bool fitting = tn > qn + qn / 2 + 10;
  // Scores that are not multiples of g = gcd(x, o, e) have empty frames, so
  // the loop steps in units of g.
  int32_t g = x;
  for (int32_t b = o; b != 0;) { int32_t r = g % b; g = b; b = r; }
  for (int32_t b = e; b != 0;) { int32_t r = g % b; g = b; b = r; }
  const int32_t xn = x / g, oen = (o + e) / g, en = e / g;
  const int32_t R = max(xn, max(oen, en)) + 1;
  const int64_t hard_max = (int64_t)(qn + tn) * max(x, 1) + o + e + 4;
  const int64_t explore_to = a.cp.cap >= 0 ? min((int64_t)a.cp.cap, hard_max) : hard_max;
  const int32_t max_step = (int32_t)(explore_to / g);
  const uint32_t wmask = (1u << a.wcap_log2) - 1;
  const int32_t wcap = (int32_t)(wmask + 1);
  int32_t* Rm = ring;
  int32_t* Ri = ring + (size_t)R * wcap;
  int32_t* Rd = ring + (size_t)2 * R * wcap;

  // Step 0: m[0][0] = extend(0, 0).
  int32_t m0 = 0;
  if (tid == 0) {
    m0 = extend_diag<B8>(qw, tw, qn, tn, 0, 0);
    Rm[0] = m0;
    sh.lo[0][0] = 0; sh.hi[0][0] = 0;
    sh.lo[1][0] = 1; sh.hi[1][0] = 0;
    sh.lo[2][0] = 1; sh.hi[2][0] = 0;
    sh.end_key[0] = INT32_MAX;
    sh.end_key[1] = INT32_MAX;
    sh.end_key[2] = INT32_MAX;
    // Endpoint at s = 0 (global: diagonal qn - tn; fitting: diagonal 0 gives t_end = qn).
    const int32_t kend = qn - tn;
    bool hit = fitting ? (0 >= kend && 0 <= qn && m0 == qn) : (kend == 0 && m0 == qn);
    if (hit) sh.end_key[0] = (tn << 2) | 0;
  }
  __syncthreads();

  int32_t step = 0, found_key = sh.end_key[0];
  uint64_t used = 0;
  int32_t status = ST_OK;
  int32_t slot = 0;
  while (found_key == INT32_MAX) {
    ++step;
    if (step > max_step) { status = ST_ALIGNMENT_FAILED; break; }
    slot = (slot + 1 == R) ? 0 : slot + 1;
    const int32_t sx = step - xn, so = step - oen, se = step - en;
    const int32_t rx = sx >= 0 ? (slot - xn + R) % R : 0;
    const int32_t ro = so >= 0 ? (slot - oen + R) % R : 0;
    const int32_t re = se >= 0 ? (slot - en + R) % R : 0;
    // Predecessor ranges (empty when the frame does not exist).
    const int32_t mx_lo = sx >= 0 ? sh.lo[0][rx] : 1, mx_hi = sx >= 0 ? sh.hi[0][rx] : 0;
    const int32_t mo_lo = so >= 0 ? sh.lo[0][ro] : 1, mo_hi = so >= 0 ? sh.hi[0][ro] : 0;
    const int32_t ie_lo = se >= 0 ? sh.lo[1][re] : 1, ie_hi = se >= 0 ? sh.hi[1][re] : 0;
    const int32_t de_lo = se >= 0 ? sh.lo[2][re] : 1, de_hi = se >= 0 ? sh.hi[2][re] : 0;
    int32_t i_lo = INT32_MAX, i_hi = INT32_MIN, d_lo = INT32_MAX, d_hi = INT32_MIN;
    if (mo_lo <= mo_hi) { i_lo = mo_lo + 1; i_hi = mo_hi + 1; d_lo = mo_lo - 1; d_hi = mo_hi - 1; }
    if (ie_lo <= ie_hi) { i_lo = min(i_lo, ie_lo + 1); i_hi = max(i_hi, ie_hi + 1); }
    if (de_lo <= de_hi) { d_lo = min(d_lo, de_lo - 1); d_hi = max(d_hi, de_hi - 1); }
    i_hi = min(i_hi, qn);
    d_lo = max(d_lo, -tn);
    int32_t m_lo = i_lo, m_hi = i_hi;
    if (d_lo <= d_hi) { m_lo = min(m_lo, d_lo); m_hi = max(m_hi, d_hi); }
    if (mx_lo <= mx_hi) { m_lo = min(m_lo, mx_lo); m_hi = max(m_hi, mx_hi); }
    if (i_lo > i_hi) { i_lo = 1; i_hi = 0; }
    if (d_lo > d_hi) { d_lo = 1; d_hi = 0; }
    const bool empty = m_lo > m_hi;
    const int32_t width = empty ? 0 : m_hi - m_lo + 1;
    if (width > wcap) { status = ST_OVERFLOW_WIDTH; break; }
    if (!H.fits(used, width, step)) { status = ST_OVERFLOW_HIST; break; }
    const int32_t kb = step % 3;
    if (tid == 0) {
      sh.lo[0][slot] = empty ? 1 : m_lo; sh.hi[0][slot] = empty ? 0 : m_hi;
      sh.lo[1][slot] = i_lo; sh.hi[1][slot] = i_hi;
      sh.lo[2][slot] = d_lo; sh.hi[2][slot] = d_hi;
      *H.entry(step) = make_int4(empty ? 1 : m_lo, empty ? 0 : m_hi, (int32_t)used, 0);
      sh.end_key[(step + 1) % 3] = INT32_MAX;
    }
    const int32_t kend_lo = fitting ? max(qn - tn, m_lo) : qn - tn;
    const int32_t kend_hi = fitting ? min(qn, m_hi) : qn - tn;
    uint32_t* cells = H.cells + used;
    const int32_t* Rmx = Rm + (size_t)rx * wcap;
    const int32_t* Rmo = Rm + (size_t)ro * wcap;
    const int32_t* Rie = Ri + (size_t)re * wcap;
    const int32_t* Rde = Rd + (size_t)re * wcap;
    int32_t* Wm = Rm + (size_t)slot * wcap;
    int32_t* Wi = Ri + (size_t)slot * wcap;
    int32_t* Wd = Rd + (size_t)slot * wcap;
    int32_t my_key = INT32_MAX;
    for (int32_t k = m_lo + tid; k <= m_hi; k += NT) {
      // ins (query-only, 'D'): open m[s-o-e][k-1]+1, extend ins[s-e][k-1]+1
      int32_t ins = UNSET; uint32_t iop = 0;
      if (in_range(k - 1, mo_lo, mo_hi)) {
        const int32_t p = Rmo[(k - 1) & wmask];
        if (p != UNSET && pos_valid(p + 1, k, qn, tn)) { ins = p + 1; iop = 0; }
      }
      if (in_range(k - 1, ie_lo, ie_hi)) {
        const int32_t p = Rie[(k - 1) & wmask];
        if (p != UNSET && pos_valid(p + 1, k, qn, tn) && p + 1 > ins) { ins = p + 1; iop = 1; }
      }
      // del (target-only, 'I'): open m[s-o-e][k+1], extend del[s-e][k+1]
      int32_t del = UNSET; uint32_t dop = 0;
      if (in_range(k + 1, mo_lo, mo_hi)) {
        const int32_t p = Rmo[(k + 1) & wmask];
        if (p != UNSET && pos_valid(p, k, qn, tn)) { del = p; dop = 0; }
      }
      if (in_range(k + 1, de_lo, de_hi)) {
        const int32_t p = Rde[(k + 1) & wmask];
        if (p != UNSET && pos_valid(p, k, qn, tn) && p > del) { del = p; dop = 1; }
      }
      // m: mismatch m[s-x][k]+1, close ins, close del
      int32_t best = UNSET; uint32_t mop = 0;
      if (in_range(k, mx_lo, mx_hi)) {
        const int32_t p = Rmx[k & wmask];
        if (p != UNSET && pos_valid(p + 1, k, qn, tn)) { best = p + 1; mop = 0; }
      }
      if (ins != UNSET && ins > best) { best = ins; mop = 1; }
      if (del != UNSET && del > best) { best = del; mop = 2; }
      const int32_t mval = best == UNSET ? UNSET : extend_diag<B8>(qw, tw, qn, tn, best, k);
      Wm[k & wmask] = mval;
      Wi[k & wmask] = ins;
      Wd[k & wmask] = del;
      cells[k - m_lo] = ((uint32_t)(best == UNSET ? 0 : best + 1) << 4) | mop | (iop << 2) | (dop << 3);
      if (k >= kend_lo && k <= kend_hi) {
        int32_t layer = -1;
        if (mval == qn) layer = 0;
        else if (ins == qn) layer = 1;
        else if (del == qn) layer = 2;
        if (layer >= 0) my_key = min(my_key, ((k + tn) << 2) | layer);
      }
    }
    if (my_key != INT32_MAX) atomicMin(&sh.end_key[kb], my_key);
    used += (uint64_t)width;
    __syncthreads();
    found_key = sh.end_key[kb];
  }

  // Traceback (affine_traceback, wfa_simd.rs:785-886) on thread 0.
  if (tid == 0) {
    PairOut r{status, 0, 0, 0, 0, 0, step, (int32_t)used};
    if (status == ST_OK) {
      const int32_t kend = (found_key >> 2) - tn;
      int32_t layer = found_key & 3;
      const int32_t t_end = qn - kend;
      RunWriter w;
      w.init(a.runs + pd.run_off, pd.run_cap);
      int32_t qi = qn, ti = t_end, k = kend, s = step;
      while (qi > 0 || ti > 0) {
        if (s == 0) {
          const int32_t n = min(qi, ti);
          w.emit(OP_M, n);
          qi -= n; ti -= n;
          break;
        }
        const int4 ent = *H.entry(s);
        if (k < ent.x || k > ent.y) { r.status = ST_INTERNAL; break; }
        const uint32_t cell = H.cells[(uint32_t)ent.z + (uint32_t)(k - ent.x)];
        const int32_t pre = (int32_t)(cell >> 4) - 1;
        if (layer == 0) {
          if (pre < 0) { r.status = ST_INTERNAL; break; }
          const uint32_t mop = cell & 3;
          const int32_t run = max(qi - pre, 0);
          w.emit(OP_M, run);
          qi -= run; ti -= run;
          if (mop == 0) {
            w.emit(OP_X, 1);
            --qi; --ti;
            s -= xn;
          } else {
            layer = (int32_t)mop;
          }
        } else if (layer == 1) {
          w.emit(OP_D, 1);
          --qi;
          if (((cell >> 2) & 1) == 0) { s -= oen; layer = 0; } else { s -= en; }
          --k;
        } else {
          w.emit(OP_I, 1);
          --ti;
          if (((cell >> 3) & 1) == 0) { s -= oen; layer = 0; } else { s -= en; }
          ++k;
        }
        if (s < 0) { r.status = ST_INTERNAL; break; }
      }
      w.flush();
      if (w.overflow && r.status == ST_OK) r.status = ST_OVERFLOW_RUNS;
      r.score = step * g;
      r.edit_distance = w.non_m;
      r.t_end = t_end;
      r.nruns = (int32_t)w.n;
      r.run_start = (int32_t)(pd.run_cap - w.n);
    }
    *po = r;
  }
}

// --- linear ------------------------------------------------------------------------

template <int NT, bool RING_LDS, bool B8>
__device__ void linear_pair(const WfaArgs& a, WfaShared& sh, int32_t* ring, const PairDesc& pd,
                            PairOut* po, Hist H) {
  const int tid = threadIdx.x;
  const uint64_t* qw = a.seq + pd.q_word;
  const uint64_t* tw = a.seq + pd.t_word;
  const int32_t qn = pd.qn, tn = pd.tn;
  if (qn == 0) {  // both-empty and empty-query cases (wfa_simd.rs:194-203, :365-367)
    if (tid == 0) *po = PairOut{ST_OK, 0, 0, 0, 0, 0, 0, 0};
    return;
  }
  const bool fitting = tn > qn + qn / 2 + 10;
  const int32_t max_s = qn + tn;
  const int32_t explore = (fitting && a.cp.cap >= 0) ? min(a.cp.cap, max_s) : max_s;
  const uint32_t wmask = (1u << a.wcap_log2) - 1;
  const int32_t wcap = (int32_t)(wmask + 1);
  const int32_t kend_global = qn - tn;

  if (tid == 0) {
    const int32_t w0 = extend_diag<B8>(qw, tw, qn, tn, 0, 0);
    ring[0] = w0;
    sh.lo[0][0] = 0; sh.hi[0][0] = 0;
    sh.end_key[0] = INT32_MAX; sh.end_key[1] = INT32_MAX; sh.end_key[2] = INT32_MAX;
    const bool hit = fitting ? (0 >= qn - tn && w0 >= qn) : (kend_global == 0 && w0 == qn);
    if (hit) sh.end_key[0] = tn;  // key = k + tn
  }
  __syncthreads();
  int32_t step = 0, found_key = sh.end_key[0];
  uint64_t used = 0;
  int32_t status = ST_OK;
  bool exhausted = false;
  while (found_key == INT32_MAX) {
    ++step;
    if (step > explore) { exhausted = true; break; }
    const int32_t slot = step & 1, ps = slot ^ 1;
    const int32_t p_lo = sh.lo[0][ps], p_hi = sh.hi[0][ps];
    int32_t lo = 1, hi = 0;
    if (p_lo <= p_hi) {
      lo = max(p_lo - 1, -min(step, tn));
      hi = min(p_hi + 1, min(step, qn));
    }
    const bool empty = lo > hi;
    const int32_t width = empty ? 0 : hi - lo + 1;
    if (width > wcap) { status = ST_OVERFLOW_WIDTH; break; }
    if (!H.fits(used, width, step)) { status = ST_OVERFLOW_HIST; break; }
    const int32_t kb = step % 3;
    if (tid == 0) {
      sh.lo[0][slot] = empty ? 1 : lo; sh.hi[0][slot] = empty ? 0 : hi;
      *H.entry(step) = make_int4(empty ? 1 : lo, empty ? 0 : hi, (int32_t)used, 0);
      sh.end_key[(step + 1) % 3] = INT32_MAX;
    }
    const int32_t* P = ring + (size_t)ps * wcap;
    int32_t* Wn = ring + (size_t)slot * wcap;
    uint32_t* cells = H.cells + used;
    const int32_t ke_lo = fitting ? max(qn - tn, lo) : kend_global;
    const int32_t ke_hi = fitting ? min(qn, hi) : kend_global;
    int32_t my_key = INT32_MAX;
    for (int32_t k = lo + tid; k <= hi; k += NT) {
      // Candidates [mismatch, query-only, target-only]; the last maximum wins.
      int32_t best = UNSET; uint32_t op = 0; bool have = false;
      if (in_range(k, p_lo, p_hi)) {
        const int32_t p = P[k & wmask];
        if (p != UNSET) { best = p + 1; op = 1; have = true; }
      }
      if (k > -tn && in_range(k - 1, p_lo, p_hi)) {
        const int32_t p = P[(k - 1) & wmask];
        if (p != UNSET && (!have || p + 1 >= best)) { best = p + 1; op = 2; have = true; }
      }
      if (k < qn && in_range(k + 1, p_lo, p_hi)) {
        const int32_t p = P[(k + 1) & wmask];
        if (p != UNSET && (!have || p >= best)) { best = p; op = 3; have = true; }
      }
      int32_t val = UNSET;
      uint32_t cell = 0;
      if (have) {
        const int32_t j = best - k;
        if (!(best < 0 || best > qn || j < 0 || j > tn)) {
          val = extend_diag<B8>(qw, tw, qn, tn, best, k);
          cell = ((uint32_t)(best + 1) << 4) | op;
        }
      }
      Wn[k & wmask] = val;
      cells[k - lo] = cell;
      if (val != UNSET && k >= ke_lo && k <= ke_hi && val >= qn) my_key = min(my_key, k + tn);
    }
    if (my_key != INT32_MAX) atomicMin(&sh.end_key[kb], my_key);
    used += (uint64_t)width;
    __syncthreads();
    found_key = sh.end_key[kb];
  }

  if (tid == 0) {
    PairOut r{status, 0, 0, 0, 0, 0, step, (int32_t)used};
    if (status == ST_OK && exhausted) {
      if (fitting) {
        r.status = ST_ALIGNMENT_FAILED;
      } else {
        // fill_wfa's fallthrough result ("", 0) (wfa_simd.rs:343-344).
        r.t_end = tn;
        r.run_start = (int32_t)pd.run_cap;
      }
    } else if (status == ST_OK) {
      const int32_t kend = found_key - tn;
      const int32_t t_end = qn - kend;
      const int32_t ed = step;
      if (!fitting && a.cp.cap >= 0 && ed > a.cp.cap) {
        r.status = ST_ALIGNMENT_FAILED;  // global branch cap check (wfa_simd.rs:380-384)
      } else {
        RunWriter w;
        w.init(a.runs + pd.run_off, pd.run_cap);
        int32_t qi = qn, ti = t_end, k = qi - ti, s = ed;
        while (qi > 0 || ti > 0) {
          if (s == 0) {
            const int32_t n = min(qi, ti);
            w.emit(OP_M, n);
            break;
          }
          const int4 ent = *H.entry(s);
          if (k < ent.x || k > ent.y) { r.status = ST_INTERNAL; break; }
          const uint32_t cell = H.cells[(uint32_t)ent.z + (uint32_t)(k - ent.x)];
          const uint32_t op = cell & 15;
          if (op == 0) { r.status = ST_INTERNAL; break; }
          const int32_t pre = (int32_t)(cell >> 4) - 1;
          const int32_t run = max(qi - pre, 0);
          w.emit(OP_M, run);
          qi -= run; ti -= run;
          if (op == 1) { w.emit(OP_X, 1); --qi; --ti; }
          else if (op == 2) { w.emit(OP_D, 1); --qi; --k; }
          else { w.emit(OP_I, 1); --ti; ++k; }
          --s;
        }
        w.flush();
        if (w.overflow && r.status == ST_OK) r.status = ST_OVERFLOW_RUNS;
        r.score = ed;
        r.edit_distance = ed;
        r.t_end = t_end;
        r.nruns = (int32_t)w.n;
        r.run_start = (int32_t)(pd.run_cap - w.n);
      }
    }
    *po = r;
  }
}

// --- kernel entry ------------------------------------------------------------------

template <int MODE, int NT, bool RING_LDS>
__global__ __launch_bounds__(NT) void wfa_kernel(WfaArgs a) {
  __shared__ WfaShared sh;
  extern __shared__ int32_t lds_ring[];
  int32_t* ring = RING_LDS ? lds_ring : a.ring_g + (size_t)blockIdx.x * a.ring_stride;
  uint8_t* hbase = a.hist + (size_t)blockIdx.x * a.hist_stride;
  Hist H;
  H.cells = reinterpret_cast<uint32_t*>(hbase);
  H.table_end = reinterpret_cast<int4*>(hbase + a.hist_stride);
  H.cap_cells = a.hist_stride / 4;
  for (;;) {
    if (threadIdx.x == 0) sh.work = atomicAdd(&a.counters[0], 1u);
    __syncthreads();
    const uint32_t w = sh.work;
    if (w >= a.nwork) break;
    const uint32_t pi = a.work_list[w];
    const PairDesc pd = a.pairs[pi];
    PairOut* po = a.out + pi;
    if (MODE == ENG_AFFINE) {
      if (pd.flags & PF_BYTES) affine_pair<NT, RING_LDS, true>(a, sh, ring, pd, po, H);
      else affine_pair<NT, RING_LDS, false>(a, sh, ring, pd, po, H);
    } else {
      if (pd.flags & PF_BYTES) linear_pair<NT, RING_LDS, true>(a, sh, ring, pd, po, H);
      else linear_pair<NT, RING_LDS, false>(a, sh, ring, pd, po, H);
    }
    if (threadIdx.x == 0 && po->status >= ST_OVERFLOW_HIST) atomicAdd(&a.counters[1], 1u);
    __syncthreads();  // history and ring are reused by the next pair
  }
}

__global__ void zero_u32_kernel(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace phx

// Host-visible launch helpers (called from phx_api.hip).
namespace phx {

hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t st) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

template <int MODE, int NT, bool RING_LDS>
static hipError_t launch_wfa_t(const WfaArgs& a, uint32_t grid, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((wfa_kernel<MODE, NT, RING_LDS>), dim3(grid), dim3(NT), lds, st, a);
  return hipGetLastError();
}

// variant: 0 = one wave per pair with LDS ring, 1 = 256 threads with LDS ring,
// 2 = 256 threads with HBM ring, 3 = 1024 threads with HBM ring.
hipError_t launch_wfa(int mode, int variant, const WfaArgs& a, uint32_t grid, size_t lds,
                      hipStream_t st) {
  if (mode == ENG_AFFINE) {
    switch (variant) {
      case 0: return launch_wfa_t<ENG_AFFINE, 64, true>(a, grid, lds, st);
      case 1: return launch_wfa_t<ENG_AFFINE, 256, true>(a, grid, lds, st);
      case 2: return launch_wfa_t<ENG_AFFINE, 256, false>(a, grid, 0, st);
      default: return launch_wfa_t<ENG_AFFINE, 1024, false>(a, grid, 0, st);
    }
  }
  switch (variant) {
    case 0: return launch_wfa_t<ENG_LINEAR, 64, true>(a, grid, lds, st);
    case 1: return launch_wfa_t<ENG_LINEAR, 256, true>(a, grid, lds, st);
    case 2: return launch_wfa_t<ENG_LINEAR, 256, false>(a, grid, 0, st);
    default: return launch_wfa_t<ENG_LINEAR, 1024, false>(a, grid, 0, st);
  }
}

int wfa_variant_threads(int variant) {
  switch (variant) {
    case 0: return 64;
    case 1: return 256;
    case 2: return 256;
    default: return 1024;
  }
}

}  // namespace phx
