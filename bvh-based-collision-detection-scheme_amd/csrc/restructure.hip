// Subtree (treelet) restructure for gfx950.
//
// Intent of the reference (collision/lbvh/LBvhFixedDeformable.cu:210-352,
// collision/lbvh/LBvhRestructureKernels.cu:37-215, Source/Project/Mine/ of
// littlemine/BVH-based-Collision-Detection-Scheme): every outermost subtree flagged by the
// front-quality check is rebuilt from fresh Morton codes of its own leaves (current scene box),
// bounded by its leaf range, and spliced back under the same root id so that the rest of the
// tree and its DFS numbering stay untouched. The shipped version writes through unsized
// buffers; here every buffer is sized to the tree.
//
// All flagged subtrees are handled by one stable radix sort over 64-bit keys
// (run start << 30 | code): leaves outside the runs keep key (slot << 30 | old code) and
// therefore stay in place, leaves inside a run are ordered by their new code.
#include "bvhcd_internal.hpp"
#include "kernels.hpp"

namespace bvhcd {

__global__ void k_rt_of_leaf(int32_t n, const int32_t* __restrict__ diff_excl, const int32_t* __restrict__ diff,
                             int32_t* __restrict__ rt_of_leaf) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    rt_of_leaf[x] = diff_excl[x] + diff[x] - 1;  // inclusive sum of (root + 1) marks, -1 = none
}

__global__ void __launch_bounds__(kBlock) k_restr_keys(int32_t n, DevState d, uint64_t* __restrict__ keys,
                                                       int32_t* __restrict__ vals) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int32_t r = d.rt_of_leaf[x];
    const int32_t tri = d.prim_idx[x];
    uint64_t key;
    if (r >= 0) key = ((uint64_t)(uint32_t)d.int_rx[r] << 30) | box_code(prim_box(d, tri), d.scene_box);
    else key = ((uint64_t)(uint32_t)x << 30) | d.codes[x];
    keys[x] = key;
    vals[x] = tri;
}

__global__ void __launch_bounds__(kBlock) k_restr_unpack(int32_t n, const uint64_t* __restrict__ keys,
                                                         uint32_t* __restrict__ codes) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    codes[x] = (uint32_t)(keys[x] & 0x3FFFFFFFull);
}

int restructure_subtrees(DevState& d, int32_t n, hipStream_t s, bool prepare_front) {
    if (n < 3) return BVHCD_OK;
    // per-leaf subtree root from the difference marks left by launch_restr_decision
    launch_exscan(d, reinterpret_cast<const uint32_t*>(d.r_leaf_lo), reinterpret_cast<uint32_t*>(d.r_leaf_hi), n,
                  nullptr, s);
    hipLaunchKernelGGL(k_rt_of_leaf, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d.r_leaf_hi, d.r_leaf_lo, d.rt_of_leaf);
    // the front is classified while the old topology is still in place
    if (prepare_front) launch_front_restr_prepare(d, n, s);
    d.front_version++;  // the topology under the front changes
    // fresh scene box and codes, stable sort inside each run
    launch_scene_box(d, n, s);
    hipLaunchKernelGGL(k_restr_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d, d.r_keys, d.vals_tmp);
    int bits = 30;
    while (((int64_t)1 << (bits - 30)) < (int64_t)n) ++bits;
    radix_sort_u64(d, d.r_keys, d.vals_tmp, d.r_keys2, d.prim_map, n, bits, s);
    hipLaunchKernelGGL(k_restr_unpack, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d.r_keys, d.codes);
    hipMemcpyAsync(d.prim_idx, d.vals_tmp, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s);
    launch_prim_map(d, n, s);
    launch_emit_runs(d, n, d.rt_of_leaf, d.r_leaf_lo, d.r_leaf_hi, d.r_split_lo, d.r_split_hi, d.r_base, d.r_keep_par,
                     d.r_keep_mark, s);
    launch_refit_prep(d, n, s);
    launch_refit(d, n, s);
    return BVHCD_OK;
}

}  // namespace bvhcd
