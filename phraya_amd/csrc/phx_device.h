// phx_device.h -- device-side data layout shared by the HIP kernels and the
// host runtime of libphraya_hip.so.
//
// Sequences live in u64 word arenas in HBM.  A pair's query and target are
// each either 2-bit packed (A=0 C=1 G=2 T=3, 32 bases per word, base i in
// bits [2i, 2i+2) of word i/32) or, when either sequence holds a byte outside
// uppercase ACGT, 8-bit raw (8 bytes per word, byte i in bits [8i, 8i+8)).
// Both encodings compare exactly what upstream compares (raw bytes), because
// the 2-bit code is a bijection on {A,C,G,T}.  Every sequence is followed by
// one zero pad word so a two-word funnel read never leaves its allocation.
#ifndef PHX_DEVICE_H
#define PHX_DEVICE_H

#include <stdint.h>

namespace phx {

enum Engine : int32_t {
  ENG_LINEAR = 0,
  ENG_AFFINE = 1,
  ENG_MYERS = 2,         // myers_extend: fitting/global + CIGAR
  ENG_MYERS_GLOBAL = 3,  // myers_edit_distance: global + CIGAR
  ENG_MYERS_DIST = 4     // myers_extend distance + end column only
};

// Device status codes (the C-ABI reports the same numbers).
enum Status : int32_t {
  ST_OK = 0,
  ST_INVALID_INPUT = 1,
  ST_ALIGNMENT_FAILED = 2,
  ST_PENDING = 15,          // not yet run
  ST_OVERFLOW_HIST = 16,    // history region too small: rerun with more
  ST_OVERFLOW_WIDTH = 17,   // wavefront wider than the kernel's ring
  ST_OVERFLOW_RUNS = 18,    // CIGAR run capacity exceeded
  ST_INTERNAL = 19          // traceback reached a cell that was never set
};

enum PairFlags : uint32_t {
  PF_BYTES = 1u,   // 8-bit encoding
  PF_GLOBAL = 2u,  // always the global branch (both ends pinned, e.g. a gap between two anchors):
                   // fill_wfa / fill_wfa_affine_generic(fitting = false), wfa_simd.rs:231-345, :581-779
};

// A sequence is addressed by the device address of the word holding its
// first symbol plus that symbol's index inside the word (shift: 0..31 for
// 2-bit, 0..7 for 8-bit), so windows of a resident reference or contig arena
// need no copy.  The word after the last symbol's word must be readable.
struct PairDesc {
  const uint64_t* q;  // query words
  const uint64_t* t;  // target words
  int32_t qn, tn;     // suffix lengths (the host applies the seed)
  uint32_t flags;
  uint32_t run_cap;   // CIGAR run capacity
  uint64_t run_off;   // offset (in runs) of this pair's run slots
  uint32_t q_shift;   // index of the first query symbol in q[0]
  uint32_t t_shift;   // index of the first target symbol in t[0]
  // Optional storage of its own for the HBM tiers (request server): history
  // region of hist_bytes and a ring of (layers x R) << ring_log2 ints.  When
  // hist is null the pair uses the slot of the wave / workgroup running it.
  uint8_t* hist;
  uint64_t hist_bytes;
  int32_t* ring;
  uint32_t ring_log2;
  // Upper bound on the pair's optimum in score steps (0 = none).  Launches
  // with WfaArgs::band set compute only the cells from which the end is
  // still reachable within it; a pair that needs more stops and is re-run
  // without it.
  uint32_t bound;
};

struct PairOut {
  int32_t status;
  int32_t score;          // linear s / affine s / Myers distance
  int32_t edit_distance;  // reported distance (affine: non-M bases of the CIGAR)
  int32_t t_end;          // target bases consumed
  int32_t nruns;          // CIGAR runs
  int32_t run_start;      // first run slot (runs occupy [run_start, run_start + nruns))
  int32_t steps;          // score steps explored (diagnostic / memory model)
  int32_t cells;          // history cells written (diagnostic / memory model)
};

struct CostParams {
  int32_t x, o, e;  // mismatch, gap open, gap extend (affine); unused for linear
  int32_t cap;      // max score explored; -1 = none
  // Affine: g = gcd(x, o, e) and the costs in score steps of g (mismatch,
  // open + extend, extend), computed once by the host (cost_params).
  int32_t g, xn, oen, en;
};

// CIGAR run encoding: (length << 4) | op.
enum RunOp : uint32_t { OP_M = 0, OP_X = 1, OP_I = 2, OP_D = 3 };

// Cost shapes (mismatch, gap open + extend and gap extend in score steps of
// g = gcd(x, o, e)) whose LDS-tier kernel keeps the wavefront in registers;
// their LDS-tier layout has no ring.  Linear WFA always does.
__host__ __device__ inline bool reg_wavefront_shape(int32_t xn, int32_t oen, int32_t en) {
  return (xn == 2 && oen == 4 && en == 1) || (xn == 2 && oen == 5 && en == 1) ||
         (xn == 1 && oen == 1 && en == 1) || (xn == 3 && oen == 7 && en == 2);
}

// Cost shapes of the banded-row kernel (a register shape with en = 1: one
// diagonal per gap-extend step, so a row's window slides by one diagonal per
// score step).
__host__ __device__ inline bool rows_shape(int32_t xn, int32_t oen, int32_t en) {
  return en == 1 && ((xn == 2 && oen == 4) || (xn == 2 && oen == 5) || (xn == 1 && oen == 1));
}

// Kernel arguments of one WFA tier launch.
// A banded-row pair as its launch reads it: the plan copies each row pair's
// query and target, shifted to bit 0, into a staging arena in the order of
// the launch list (so the pairs of one wave sit together), and keeps this
// 16-byte record per list position (queries and targets of at most 2,048
// bases: banded rows hold qn + tn <= 2048).
struct RowDesc {
  uint32_t q_word;  // first staged word of the query; the target's follows its words
  uint32_t run_off, run_cap;
  uint16_t qn, tn;
};
static_assert(sizeof(RowDesc) == 16, "RowDesc is one 16-byte load");

struct WfaArgs {
  const PairDesc* pairs;
  const uint32_t* work_list;
  uint32_t nwork;
  uint32_t* counters;    // [0] work queue head (zero at launch unless static_order)
  uint32_t* overflow;    // host-visible flag set to 1 when a pair overflows; may be null
  CostParams cp;
  PairOut* out;
  uint32_t* runs;
  uint8_t* hist;         // one slot of hist_stride bytes per workgroup
  uint64_t hist_stride;
  int32_t* ring_g;       // one slot of ring_stride ints per workgroup (HBM ring tiers)
  uint64_t ring_stride;
  uint32_t wcap_log2;    // ring width in diagonals = 1 << wcap_log2
  uint32_t static_order; // wave kernels: wave g takes items g, g + waves, ... (no queue)
  uint32_t band;         // register-wavefront workgroup launches: prune to PairDesc::bound
  // Banded-row launches: the list is four segments of these many entries,
  // in this order: 32-lane rows with qn <= tn, 32-lane rows with qn > tn,
  // 16-lane rows with qn <= tn, 16-lane rows with qn > tn.
  uint32_t rows_seg[4];
  // Banded-row launches of a plan with a row stage: the group's records,
  // parallel to work_list, and the staging arena; null: the launch reads
  // PairDesc (pairs[work_list[pos]]).
  const RowDesc* rows_desc;
  const uint64_t* rows_seq;
};

// LDS layout of one wave of the wave-resident WFA tier (byte offsets inside
// the wave's region; the workgroup's dynamic LDS holds WAVES regions).
struct WaveGeom {
  uint32_t per_wave_bytes;
  uint32_t ring_off;       // int16 ring [layers][R][128]
  uint32_t seq_off;        // u64 packed query then target
  uint32_t tab_off;        // u16 first-cell offset of each step [steps_cap]
  uint32_t cell_off;       // u16 history cells [cells_cap]
  uint32_t seq_cap_words;
  uint32_t steps_cap;
  uint32_t cells_cap;
  // Banded-row launches (four pairs per wave in 16-lane rows, or two in
  // 32-lane rows): offset of a row's history inside its region (after the
  // packed query and target) and history steps of a 16-lane row
  // (rows_region_bytes gives the rest).  A wave's region holds the largest
  // of four 16-lane row regions, two 32-lane ones and the single-pair layout
  // above, which serves the pairs a 32-lane row cannot finish.
  uint32_t rows_hist_off;
  uint32_t rows_steps;
};

// Bytes of one row region of `rw` lanes (16 or 32): sequences, then
// (rows_steps + rw - 16) steps of rw u16 history cells (steps 0 .. the last
// step a row can reach).
__host__ __device__ inline uint32_t rows_region_bytes(uint32_t hist_off, uint32_t rows_steps, uint32_t rw) {
  return hist_off + (((rows_steps + rw - 16u) * rw * 2u + 15u) & ~15u);
}

// Kernel arguments of one Myers launch.
struct MyersArgs {
  const PairDesc* pairs;  // Myers queries are word aligned (q_shift 0); targets may be shifted
  const uint32_t* work_list;
  uint32_t nwork;
  PairOut* out;
  uint32_t* runs;
  uint8_t* scratch;        // one slot of scratch_stride bytes per work item
  uint64_t scratch_stride;
  int32_t engine;          // ENG_MYERS, ENG_MYERS_GLOBAL or ENG_MYERS_DIST
};

}  // namespace phx

#endif
