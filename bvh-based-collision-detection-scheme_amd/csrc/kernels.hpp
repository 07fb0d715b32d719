// Host-side launchers of the gfx950 kernels (stream-ordered, no host synchronisation).
#pragma once
#include <string>

#include "bvhcd_internal.hpp"

namespace bvhcd {

// lbvh.hip
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
void fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
void launch_scene_box(DevState& d, int32_t T, hipStream_t s);
void launch_codes(DevState& d, int32_t T, uint32_t* codes, int32_t* vals, hipStream_t s);
void launch_exscan(DevState& d, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, hipStream_t s);
void radix_sort_u32(DevState& d, uint32_t* keys, int32_t* vals, uint32_t* tkeys, int32_t* tvals, int64_t n, int nbits,
                    hipStream_t s);
// n_max bounds the count; the actual count is read on the device from *dn. Returns true when
// the result is in (tkeys, tvals).
bool radix_sort_u32_dev(DevState& d, uint32_t* keys, int32_t* vals, uint32_t* tkeys, int32_t* tvals, int64_t n_max,
                        const uint32_t* dn, int nbits, hipStream_t s);
void radix_sort_u64(DevState& d, uint64_t* keys, int32_t* vals, uint64_t* tkeys, int32_t* tvals, int64_t n, int nbits,
                    hipStream_t s);
void launch_emit_full(DevState& d, int32_t n, hipStream_t s);
// Launch-shape tuning: grid sizes (workgroups) of the grid-stride / persistent kernels per size
// class ceil(log2 n), process-wide; 0 entries fall back to the defaults.
enum TunedKernel { TK_KEEP = 0, TK_NARROW = 1, TK_BATCH_CD = 2, TK_COUNT = 3 };
// A launch shape: grid size (low 24 bits) and items per thread per loop iteration (1, 2, 4).
// launch shape: grid (24 bits), entries per thread (4 bits), waves per workgroup (4 bits, 0 = 4)
inline unsigned make_shape(unsigned grid, unsigned unroll, unsigned waves = 4) {
    return grid | (unroll << 24) | ((waves & 15u) << 28);
}
inline unsigned shape_grid(unsigned sh) { return sh & 0xFFFFFFu; }
inline unsigned shape_unroll(unsigned sh) { return ((sh >> 24) & 15u) ? ((sh >> 24) & 15u) : 1u; }
inline unsigned shape_waves(unsigned sh) { return (sh >> 28) ? (sh >> 28) : 4u; }
unsigned tuned_shape(int kernel, int64_t n);
void set_tuned_shape(int kernel, int64_t n, unsigned sh);
void record_tuning(int kernel, int64_t n, unsigned sh, float ms);
std::string tuning_report();
// Emission from sorted 64-bit keys (batched scenes: scene id above the Morton bits).
void launch_emit_full64(DevState& d, const uint64_t* keys, int32_t n, hipStream_t s);
// Stable LSD sort of 64-bit keys over bits [0, nbits); returns true when the result sits in
// (tkeys, tvals).
bool radix_sort_u64_dev(DevState& d, uint64_t* keys, int32_t* vals, uint64_t* tkeys, int32_t* tvals, int64_t n,
                        int nbits, hipStream_t s);
// Batched self-CD (persistent, work queue) over per-scene subtrees; pairs into d.bp.
void launch_batch_cd(DevState& d, int32_t n, const uint64_t* keys, const int32_t* scene_last, uint32_t* queue,
                     unsigned grid, hipStream_t s);
void launch_emit_runs(DevState& d, int32_t n, const int32_t* rt_of_leaf, int32_t* leaf_lo, int32_t* leaf_hi,
                      int32_t* split_lo, int32_t* split_hi, int32_t* base, int32_t* keep_par, uint32_t* keep_mark,
                      hipStream_t s);
// zero_step_counters: the top-tree kernel also zeroes the step's output counters, for a KEEP
// that follows right after (launch_keep(..., true)) without a kernel of its own.
void launch_refit(DevState& d, int32_t n, hipStream_t s, bool zero_step_counters = false);
void launch_refit_prep(DevState& d, int32_t n, hipStream_t s);  // after every topology change
int32_t refit_chunks(int32_t n);  // refit workgroups (treelet chunks) for n leaves
// Top trees of up to this many nodes are merged in LDS by one workgroup; larger ones are split
// into at most kRefitTopGrid segments of up to this size (one workgroup each) and the upper
// nodes above them. The top list holds kRefitTopLds entries past its end (read ahead).
constexpr int32_t kRefitTopLds = 4096;
constexpr int32_t kRefitTopGrid = 64;
// Capacity (boxes) of the hand-over array between the chunk refit and the top-tree merge.
inline int64_t refit_tcb_boxes(int64_t T) { return 2 * (int64_t)kRefitTopLds + T / 128; }
void launch_prim_map(DevState& d, int32_t n, hipStream_t s);

// front.hip
void launch_pure_cd(DevState& d, int32_t n, hipStream_t s);
// order = false keeps the emission order (entries of neighbouring queries together) and skips
// the node-id sort and bucket offsets that only the front-quality heuristic reads.
void launch_generate(DevState& d, int32_t n, hipStream_t s, bool order = true);  // traversal + ordering
void launch_update(DevState& d, int32_t n, int quality, hipStream_t s, bool order = true);
void launch_keep(DevState& d, int32_t n, hipStream_t s, bool counters_zeroed = false);
void launch_reorder(DevState& d, int32_t n, hipStream_t s);
// Incremental front restructure: classify the current fronts against the runs in rt_of_leaf
// (old topology; kept entries to the other buffers, walks to d.rtasks with their count in
// ctr[C_TASKS]), then, after the rebuild, pairs of the kept entries plus the walks.
void launch_front_restr_prepare(DevState& d, int32_t n, hipStream_t s);
void launch_front_restr_apply(DevState& d, int32_t n, uint32_t ntasks, hipStream_t s);
void launch_order_fronts(DevState& d, int32_t n, hipStream_t s, bool generated);  // by node + bucket starts
void launch_query_order(DevState& d, int32_t n, hipStream_t s);       // stable radix sort by query leaf
void launch_snapshot(DevState& d, int32_t n, hipStream_t s);
void launch_narrow(DevState& d, hipStream_t s, bool counters_zeroed = false);  // true: right after a step's reset
// Segmented pair lists: totals (and largest segments) into ctr[C_BP / C_NP / *_SEGMAX]; clearing
// the narrow-phase list; a contiguous copy of list `kind` (0 broad phase, mapped from leaf to
// triangle ids; 1 narrow phase) of at most cap pairs into dst.
void launch_pair_totals(DevState& d, hipStream_t s, bool narrow);
void clear_np(DevState& d, hipStream_t s);
void launch_pair_gather(DevState& d, int kind, int2* dst, int64_t first, int64_t count, hipStream_t s);
void launch_digest(DevState& d, int filtered, hipStream_t s);
void launch_restr_decision(DevState& d, int32_t n, hipStream_t s);    // counts into ctr[C_RT_*]

// restructure.hip
int restructure_subtrees(DevState& d, int32_t n, hipStream_t s, bool prepare_front);      // uses rt marks; returns status

}  // namespace bvhcd
