// Host-side launchers of the gfx950 kernels (stream-ordered, no host synchronisation).
#pragma once
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
void launch_emit_runs(DevState& d, int32_t n, const int32_t* rt_of_leaf, int32_t* leaf_lo, int32_t* leaf_hi,
                      int32_t* split_lo, int32_t* split_hi, int32_t* base, int32_t* keep_par, uint32_t* keep_mark,
                      hipStream_t s);
void launch_refit(DevState& d, int32_t n, hipStream_t s);
void launch_refit_prep(DevState& d, int32_t n, hipStream_t s);  // after every topology change
void launch_prim_map(DevState& d, int32_t n, hipStream_t s);

// front.hip
void launch_pure_cd(DevState& d, int32_t n, hipStream_t s);
void launch_generate(DevState& d, int32_t n, hipStream_t s);          // traversal + ordering
void launch_update(DevState& d, int32_t n, int quality, hipStream_t s);// prune/sprout, quality, ordering
void launch_keep(DevState& d, int32_t n, hipStream_t s);
void launch_reorder(DevState& d, int32_t n, hipStream_t s);
// Incremental front restructure: classify the current fronts against the runs in rt_of_leaf
// (old topology; kept entries to the other buffers, walks to d.rtasks with their count in
// ctr[C_TASKS]), then, after the rebuild, pairs of the kept entries plus the walks.
void launch_front_restr_prepare(DevState& d, int32_t n, hipStream_t s);
void launch_front_restr_apply(DevState& d, int32_t n, uint32_t ntasks, hipStream_t s);
void launch_order_fronts(DevState& d, int32_t n, hipStream_t s);      // radix sort by node + bucket starts
void launch_snapshot(DevState& d, int32_t n, hipStream_t s);
void launch_narrow(DevState& d, hipStream_t s);
void launch_digest(DevState& d, int filtered, hipStream_t s);
void launch_restr_decision(DevState& d, int32_t n, hipStream_t s);    // counts into ctr[C_RT_*]

// restructure.hip
int restructure_subtrees(DevState& d, int32_t n, hipStream_t s, bool prepare_front);      // uses rt marks; returns status

}  // namespace bvhcd
