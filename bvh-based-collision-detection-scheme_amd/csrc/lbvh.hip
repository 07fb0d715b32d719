// LBVH construction and refit kernels for gfx950.
//
// Reference behaviour (Source/Project/Mine/ of littlemine/BVH-based-Collision-Detection-Scheme):
//   scene box + Morton          collision/lbvh/LBvhBuildKernels.cu:55-91, utility/CudaDeviceUtils.cu:82-130
//   sort_by_key + prim map      collision/lbvh/LBvhFixedDeformable.cu:355-363
//   split metric                collision/lbvh/BvhExtNode.cu:128-134
//   topology (Apetrei)          collision/lbvh/LBvhBuildKernels.cu:135-172
//   DFS renumbering             collision/lbvh/LBvhFixedDeformable.cu:365-380, LBvhBuildKernels.cu:174-217
//   refit                       collision/lbvh/LBvhUpdateKernels.cu:35-74
//
// Design here:
//  * The topology rule (a node [a,b] is the right child of split a-1 if delta(a-1) < delta(b),
//    else the left child of split b) makes the tree the Cartesian tree of delta with the
//    leftmost maximum on top. Each split node's range therefore follows from two prefix
//    searches over the sorted codes (nearest delta >= d on the left, nearest delta > d on
//    the right), so emission is one independent thread per node: no atomics, no climbing,
//    no memory-ordering hazards.
//  * Sorting is an LSD radix sort with 8-bit digits over an LDS-staged tile of 4096 keys:
//    per-tile digit counts, a device-wide exclusive scan of the digit-major count table,
//    then a stable wave64 multisplit ranking in LDS and coalesced scatter.
//  * Refit is level-synchronous over treelet-aligned chunks: each workgroup merges its
//    treelets' nodes one depth level per barrier in LDS (no atomics), then the top tree above
//    the treelets is merged in LDS by one workgroup per segment of at most 4096 nodes.
#include <type_traits>

#include "bvhcd_internal.hpp"
#include "kernels.hpp"

namespace bvhcd {

// ======================================================================= scene box + codes
// Order-preserving float <-> uint map for atomic min/max (boxes here never hold NaN:
// fminf/fmaxf against the FLT_MAX start value discard NaN vertex coordinates).
__device__ inline uint32_t f2o(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float o2f(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __builtin_bit_cast(float, u);
}

__global__ void k_scene_box_init(uint32_t* box_u) {
    if (threadIdx.x < 3) box_u[threadIdx.x] = 0xFFFFFFFFu;       // min slots
    else if (threadIdx.x < 6) box_u[threadIdx.x] = 0u;           // max slots
}

// pbox: per-triangle boxes given by the caller (box mode), else null (boxes from the vertices).
__global__ void __launch_bounds__(kBlock) k_scene_box(int32_t T, const int32_t* __restrict__ faces,
                                                      const float* __restrict__ verts, const float* __restrict__ pbox,
                                                      uint32_t* box_u) {
    Box acc = empty_box();
    for (int32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x)
        acc = box_merge(acc, pbox ? given_box(pbox, t) : tri_box(faces, verts, t));
    // wave reduction
    float v[6] = {acc.mnx, acc.mny, acc.mnz, acc.mxx, acc.mxy, acc.mxz};
    for (int off = 32; off > 0; off >>= 1)
        for (int k = 0; k < 6; ++k) {
            float o = __shfl_down(v[k], off);
            v[k] = k < 3 ? fminf(v[k], o) : fmaxf(v[k], o);
        }
    __shared__ float red[kBlock / 64][6];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0)
        for (int k = 0; k < 6; ++k) red[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        float r = red[0][k];
        for (int i = 1; i < kBlock / 64; ++i) r = k < 3 ? fminf(r, red[i][k]) : fmaxf(r, red[i][k]);
        if (k < 3) atomicMin(box_u + k, f2o(r));
        else atomicMax(box_u + k, f2o(r));
    }
}

__global__ void k_scene_box_final(const uint32_t* box_u, float* box) {
    if (threadIdx.x < 6) box[threadIdx.x] = o2f(box_u[threadIdx.x]);
}

__global__ void __launch_bounds__(kBlock) k_codes(int32_t T, const int32_t* __restrict__ faces,
                                                  const float* __restrict__ verts, const float* __restrict__ pbox,
                                                  const float* __restrict__ scene, uint32_t* __restrict__ codes,
                                                  int32_t* __restrict__ vals) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    codes[t] = box_code(pbox ? given_box(pbox, t) : tri_box(faces, verts, t), scene);
    vals[t] = t;
}

void launch_scene_box(DevState& d, int32_t T, hipStream_t s) {
    hipLaunchKernelGGL(k_scene_box_init, dim3(1), dim3(64), 0, s, d.scene_box_u);
    unsigned g = grid_for(T);
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(k_scene_box, dim3(g), dim3(kBlock), 0, s, T, d.faces, d.verts, d.pbox, d.scene_box_u);
    hipLaunchKernelGGL(k_scene_box_final, dim3(1), dim3(64), 0, s, d.scene_box_u, d.scene_box);
}

void launch_codes(DevState& d, int32_t T, uint32_t* codes, int32_t* vals, hipStream_t s) {
    hipLaunchKernelGGL(k_codes, dim3(grid_for(T)), dim3(kBlock), 0, s, T, d.faces, d.verts, d.pbox, d.scene_box, codes,
                       vals);
}

// ======================================================================= exclusive scan (u32)
constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;  // 4096

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ inline uint32_t block_exscan(uint32_t v, uint32_t& excl) {
    __shared__ uint32_t wsum[kBlock / 64];
    const int lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t wpre = 0, total = 0;
    for (int i = 0; i < kBlock / 64; ++i) {
        uint32_t s = wsum[i];
        if (i < w) wpre += s;
        total += s;
    }
    __syncthreads();
    excl = wpre + x - v;
    return total;
}

__global__ void __launch_bounds__(kBlock) k_scan_sums(const uint32_t* __restrict__ in, int64_t n, uint32_t* sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
    for (int i = 0; i < kScanItems; ++i) {
        int64_t j = base + (int64_t)i * kBlock + threadIdx.x;
        if (j < n) s += in[j];
    }
    uint32_t ex;
    uint32_t tot = block_exscan(s, ex);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Single-block exclusive scan of m block sums (in place); optional total.
__global__ void __launch_bounds__(kBlock) k_scan_single(uint32_t* sums, int64_t m, uint32_t* total_out) {
    uint32_t carry = 0;
    for (int64_t base = 0; base < m; base += kBlock) {
        int64_t j = base + threadIdx.x;
        uint32_t v = j < m ? sums[j] : 0u;
        uint32_t ex;
        uint32_t tot = block_exscan(v, ex);
        if (j < m) sums[j] = carry + ex;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total_out) *total_out = carry;
}

// `in` may equal `out` (in-place scan): every thread reads its items before the barrier in
// block_exscan and writes them after it.
__global__ void __launch_bounds__(kBlock) k_scan_apply(const uint32_t* in, int64_t n, const uint32_t* __restrict__ sums,
                                                       uint32_t* out) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
    for (int i = 0; i < kScanItems; ++i) {
        int64_t j = base + i;
        v[i] = j < n ? in[j] : 0u;
        s += v[i];
    }
    uint32_t ex;
    block_exscan(s, ex);
    uint32_t run = sums[blockIdx.x] + ex;
    for (int i = 0; i < kScanItems; ++i) {
        int64_t j = base + i;
        if (j < n) out[j] = run;
        run += v[i];
    }
}

// out[0..n-1] = exclusive scan of in[0..n-1]; if total != nullptr, *total = sum (device).
void launch_exscan(DevState& d, const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, hipStream_t s) {
    if (n <= 0) return;
    const int64_t blocks = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_sums, dim3((unsigned)blocks), dim3(kBlock), 0, s, in, n, d.scan_tmp);
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(kBlock), 0, s, d.scan_tmp, blocks, total);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)blocks), dim3(kBlock), 0, s, in, n, d.scan_tmp, out);
}

// ======================================================================= radix sort
constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;  // 4096 keys per tile
constexpr int kRadix = 256;

// Digit counts of one 4096-key tile. Each wave takes 64 keys at a time and groups equal digits
// with a wave64 multisplit (8 ballots), so one LDS atomic per distinct digit per wave replaces
// one per key: partially sorted input (later passes, front node ids) has long runs of equal
// digits that would otherwise serialise on the same LDS word.
template <class K>
__global__ void __launch_bounds__(kBlock) k_sort_hist(const K* __restrict__ keys, int64_t n, const uint32_t* dn,
                                                      int shift, uint32_t* __restrict__ hist, int64_t tiles) {
    __shared__ uint32_t h[kRadix];
    if (dn) n = min(n, (int64_t)*dn);
    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    if (base >= n) {  // block-uniform: an empty tile contributes zero counts
        hist[(int64_t)tid * tiles + blockIdx.x] = 0;
        return;
    }
    h[tid] = 0;
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    constexpr int kPerWave = kSortTile / (kBlock / 64);
    for (int r = 0; r < kPerWave / 64; ++r) {
        const int64_t j = base + w * kPerWave + r * 64 + lane;
        const bool ok = j < n;
        const uint32_t dig = ok ? (uint32_t)(keys[j] >> shift) & 0xFFu : 0u;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot(ok && ((dig >> b) & 1u));
            peers &= ((dig >> b) & 1u) ? m : ~m;
        }
        if (ok && (peers & below) == 0) atomicAdd(&h[dig], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    hist[(int64_t)tid * tiles + blockIdx.x] = h[tid];
}

// Stable scatter of one 4096-key tile. Wave w owns the tile's items [1024 w, 1024 w + 1024),
// taken 64 at a time in index order; a wave64 multisplit (8 ballots) gives each item its rank
// among the wave's earlier items with the same digit, using only the wave's own LDS counters
// (no block barrier inside the loop). One block-level pass then turns the per-wave digit counts
// into offsets, so an item's tile position is
//   tile_start[digit] + (same-digit items of earlier waves) + (its rank inside the wave),
// i.e. the stable order by index. Keys stay in registers between the two steps.
template <class K>
__global__ void __launch_bounds__(kBlock) k_sort_scatter(const K* __restrict__ keys, const int32_t* __restrict__ vals,
                                                         int64_t n, const uint32_t* dn, int shift,
                                                         const uint32_t* __restrict__ hist, int64_t tiles,
                                                         K* __restrict__ okeys, int32_t* __restrict__ ovals) {
    constexpr int kWavesS = kBlock / 64;
    constexpr int kPerWave = kSortTile / kWavesS;   // 1024
    constexpr int kRounds = kPerWave / 64;          // 16
    __shared__ uint32_t wcnt[kWavesS][kRadix];      // per-wave digit counts, then offsets
    __shared__ uint32_t tile_start[kRadix];
    __shared__ uint32_t gofs[kRadix];
    __shared__ K skeys[kSortTile];
    __shared__ int32_t svals[kSortTile];
    __shared__ uint8_t sdig[kSortTile];

    const int tid = threadIdx.x, lane = lane_id(), w = tid >> 6;
    if (dn) n = min(n, (int64_t)*dn);
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    if (base >= n) return;  // whole block uniform: before any barrier
    const int valid = (int)((n - base) < kSortTile ? (n - base) : kSortTile);

    for (int k = lane; k < kRadix; k += 64) wcnt[w][k] = 0;
    K key[kRounds];
    uint32_t rk[kRounds];
    const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int j = w * kPerWave + r * 64 + lane;
        const bool ok = j < valid;
        key[r] = ok ? keys[base + j] : (K)0;
        const uint32_t dig = (uint32_t)(key[r] >> shift) & 0xFFu;
        uint64_t peers = __ballot(ok);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t m = __ballot(ok && ((dig >> b) & 1u));
            peers &= ((dig >> b) & 1u) ? m : ~m;
        }
        const uint32_t before = ok ? wcnt[w][dig] : 0u;
        rk[r] = before + (uint32_t)__popcll(peers & below);
        // the lowest lane of each peer group advances the wave's count (its read above was
        // issued before this write, and LDS operations of one wave complete in order)
        if (ok && (peers & below) == 0) wcnt[w][dig] = before + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();  // keep the next round's reads after this write
    }
    __syncthreads();
    {   // digit tid: offsets of each wave's group, then the tile-wide exclusive scan
        uint32_t acc = 0;
#pragma unroll
        for (int ww = 0; ww < kWavesS; ++ww) {
            const uint32_t c = wcnt[ww][tid];
            wcnt[ww][tid] = acc;
            acc += c;
        }
        uint32_t ex;
        block_exscan(acc, ex);   // contains a barrier; every thread reaches it
        tile_start[tid] = ex;
        gofs[tid] = hist[(int64_t)tid * tiles + blockIdx.x];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int j = w * kPerWave + r * 64 + lane;
        if (j < valid) {
            const uint32_t dig = (uint32_t)(key[r] >> shift) & 0xFFu;
            const uint32_t pos = tile_start[dig] + wcnt[w][dig] + rk[r];
            skeys[pos] = key[r];
            svals[pos] = vals[base + j];
            sdig[pos] = (uint8_t)dig;
        }
    }
    __syncthreads();
    // coalesced write-out in tile-sorted order
    for (int i = 0; i < kSortItems; ++i) {
        const int j = i * kBlock + tid;
        if (j < valid) {
            const uint32_t dg = sdig[j];
            const int64_t o = (int64_t)gofs[dg] + (j - (int64_t)tile_start[dg]);
            okeys[o] = skeys[j];
            ovals[o] = svals[j];
        }
    }
}

// Stable LSD sort of (keys, vals) over bits [0, nbits), ping-ponging between the given
// buffers. n is the element count, or (dn != nullptr) its upper bound with the actual count
// read on the device. Returns true when the result sits in (tkeys, tvals) (odd pass count).
template <class K>
bool radix_sort_passes(DevState& d, K* keys, int32_t* vals, K* tkeys, int32_t* tvals, int64_t n, const uint32_t* dn,
                       int nbits, hipStream_t s) {
    if (n <= 1 && !dn) return false;
    if (n < 1) return false;
    const int64_t tiles = (n + kSortTile - 1) / kSortTile;
    const int passes = (nbits + 7) / 8;
    K* ik = keys; int32_t* iv = vals;
    K* ok = tkeys; int32_t* ov = tvals;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        hipLaunchKernelGGL(k_sort_hist<K>, dim3((unsigned)tiles), dim3(kBlock), 0, s, ik, n, dn, shift, d.sort_hist,
                           tiles);
        launch_exscan(d, d.sort_hist, d.sort_hist, (int64_t)kRadix * tiles, nullptr, s);
        hipLaunchKernelGGL(k_sort_scatter<K>, dim3((unsigned)tiles), dim3(kBlock), 0, s, ik, iv, n, dn, shift,
                           d.sort_hist, tiles, ok, ov);
        std::swap(ik, ok);
        std::swap(iv, ov);
    }
    return passes & 1;
}

// Result always ends in (keys, vals).
template <class K>
void radix_sort_impl(DevState& d, K* keys, int32_t* vals, K* tkeys, int32_t* tvals, int64_t n, int nbits,
                     hipStream_t s) {
    if (n <= 1) return;
    if (radix_sort_passes<K>(d, keys, vals, tkeys, tvals, n, nullptr, nbits, s)) {
        hipMemcpyAsync(keys, tkeys, sizeof(K) * n, hipMemcpyDeviceToDevice, s);
        hipMemcpyAsync(vals, tvals, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s);
    }
}

bool radix_sort_u32_dev(DevState& d, uint32_t* keys, int32_t* vals, uint32_t* tkeys, int32_t* tvals, int64_t n_max,
                        const uint32_t* dn, int nbits, hipStream_t s) {
    return radix_sort_passes<uint32_t>(d, keys, vals, tkeys, tvals, n_max, dn, nbits, s);
}

void radix_sort_u32(DevState& d, uint32_t* keys, int32_t* vals, uint32_t* tkeys, int32_t* tvals, int64_t n, int nbits,
                    hipStream_t s) {
    radix_sort_impl<uint32_t>(d, keys, vals, tkeys, tvals, n, nbits, s);
}
bool radix_sort_u64_dev(DevState& d, uint64_t* keys, int32_t* vals, uint64_t* tkeys, int32_t* tvals, int64_t n,
                        int nbits, hipStream_t s) {
    return radix_sort_passes<uint64_t>(d, keys, vals, tkeys, tvals, n, nullptr, nbits, s);
}
void radix_sort_u64(DevState& d, uint64_t* keys, int32_t* vals, uint64_t* tkeys, int32_t* tvals, int64_t n, int nbits,
                    hipStream_t s) {
    radix_sort_impl<uint64_t>(d, keys, vals, tkeys, tvals, n, nbits, s);
}

// ======================================================================= emission
// delta(i) = bit length of c[i] ^ c[i+1]; the last split position holds a sentinel above every
// real value (33 for 32-bit codes, 65 for 64-bit keys).
__device__ inline int delta_at(const uint32_t* __restrict__ c, int32_t i, int32_t n) {
    if (i >= n - 1) return 33;
    const uint32_t x = c[i] ^ c[i + 1];
    return x ? 32 - __clz(x) : 0;
}
__device__ inline int delta_at(const uint64_t* __restrict__ c, int32_t i, int32_t n) {
    if (i >= n - 1) return 65;
    const uint64_t x = c[i] ^ c[i + 1];
    return x ? 64 - __clzll((long long)x) : 0;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    if (n <= 0) return;
    unsigned g = grid_for(n);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3(g), dim3(kBlock), 0, s, p, n, v);
}
void fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return;
    unsigned g = grid_for(n);
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(kBlock), 0, s, p, n, v);
}

// Split node k of the sorted run [lo, hi] (leaves lo..hi; splits lo..hi-1). Boundaries of the
// run act as delta = +inf, so a run is built exactly as a standalone tree over its leaves.
// Range [a, b]: a-1 = nearest split left of k with delta >= d (a = lo if none); b = nearest
// split right of k with delta > d (b = hi if none).
template <class C>
__device__ inline void split_range(const C* __restrict__ c, int32_t k, int32_t lo, int32_t hi, int d,
                                   int32_t& a, int32_t& b) {
    constexpr int kBits = 8 * (int)sizeof(C);
    // a: smallest index in [lo, k] whose code agrees with c[k] on bits >= d-1 (d >= 1).
    if (d == 0) {
        a = k;
    } else {
        const C ck = c[k];
        const int sh = d - 1;
        int32_t L = lo, R = k;  // answer in [L, R]
        while (L < R) {
            const int32_t m = L + ((R - L) >> 1);
            if (((c[m] ^ ck) >> sh) == (C)0) R = m;
            else L = m + 1;
        }
        a = L;
    }
    // b: largest index in [k+1, hi] whose code agrees with c[k+1] on bits >= d.
    {
        const C ck1 = c[k + 1];
        int32_t L = k + 1, R = hi;
        if (d >= kBits) {
            L = hi;
        } else {
            while (L < R) {
                const int32_t m = L + ((R - L + 1) >> 1);
                if (((c[m] ^ ck1) >> d) == (C)0) L = m;
                else R = m - 1;
            }
        }
        b = L;
    }
}

// One thread per split k of the sorted codes; segments: runs [seg_lo, seg_hi] given per split
// through seg_of (nullptr = one run [0, n-1]).
template <class C>
__global__ void __launch_bounds__(kBlock) k_emit_internal(int32_t n, const C* __restrict__ c,
                                                          const int32_t* __restrict__ run_lo,
                                                          const int32_t* __restrict__ run_hi,
                                                          int32_t* __restrict__ s_a, int32_t* __restrict__ s_b,
                                                          int32_t* __restrict__ s_par, int32_t* __restrict__ s_lc,
                                                          int32_t* __restrict__ s_rc, int32_t* __restrict__ s_top,
                                                          uint32_t* __restrict__ depths) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n - 1) return;
    const int32_t lo = run_lo ? run_lo[k] : 0;
    const int32_t hi = run_hi ? run_hi[k] : n - 1;
    if (lo < 0) return;  // split not inside a run being (re)built
    const int d = delta_at(c, k, n);
    int32_t a, b;
    split_range(c, k, lo, hi, d, a, b);
    s_a[k] = a;
    s_b[k] = b;
    atomicAdd(&depths[a], 1u);
    if (a == lo && b == hi) {  // run root
        s_par[k] = -1;
        s_top[a] = k;
        return;
    }
    const int dl = a > lo ? delta_at(c, a - 1, n) : 1 << 20;
    const int dr = b < hi ? delta_at(c, b, n) : 1 << 20;
    if (a > lo && dl < dr) {  // right child of split a-1
        s_par[k] = (a - 1) | (1 << 30);
        s_rc[a - 1] = k << 1;
        s_top[a] = k;
    } else {                  // left child of split b
        s_par[k] = b;
        s_lc[b] = k << 1;
    }
}

template <class C>
__global__ void __launch_bounds__(kBlock) k_emit_leaves(int32_t n, const C* __restrict__ c,
                                                        const int32_t* __restrict__ leaf_lo,
                                                        const int32_t* __restrict__ leaf_hi,
                                                        int32_t* __restrict__ s_lpar, int32_t* __restrict__ s_lc,
                                                        int32_t* __restrict__ s_rc) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t lo = leaf_lo ? leaf_lo[i] : 0;
    const int32_t hi = leaf_hi ? leaf_hi[i] : n - 1;
    if (lo < 0) return;
    const int dl = i > lo ? delta_at(c, i - 1, n) : 1 << 20;
    const int dr = i < hi ? delta_at(c, i, n) : 1 << 20;
    if (i > lo && dl < dr) {
        s_lpar[i] = (i - 1) | (1 << 30);
        s_rc[i - 1] = (i << 1) | 1;
    } else {
        s_lpar[i] = i;
        s_lc[i] = (i << 1) | 1;
    }
}

// Preorder ids: the internal nodes whose leftmost leaf is x form a left spine starting at
// s_top[x]; they get ids base + offsets[x] .. in top-down order.
__global__ void __launch_bounds__(kBlock) k_assign_preorder(int32_t n, const int32_t* __restrict__ s_top,
                                                            const uint32_t* __restrict__ depths,
                                                            const uint32_t* __restrict__ offsets,
                                                            const int32_t* __restrict__ s_lc,
                                                            const int32_t* __restrict__ leaf_base,
                                                            int32_t* __restrict__ tkmap) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    int32_t node = s_top[x];
    if (node < 0) return;
    int32_t id = (int32_t)offsets[x] + (leaf_base ? leaf_base[x] : 0);
    for (uint32_t dd = depths[x]; dd > 0; --dd) {
        tkmap[node] = id++;
        node = s_lc[node] >> 1;
    }
}

__device__ inline int32_t enc_lca(int32_t x, int32_t n, const int32_t* __restrict__ s_top,
                                  const int32_t* __restrict__ tkmap) {
    if (x >= n) return -1;
    const int32_t t = s_top[x];
    return t >= 0 ? tkmap[t] << 1 : (x << 1) | 1;
}

// Writes preorder SoA topology and the w words of the node records.
__global__ void __launch_bounds__(kBlock) k_write_internal(int32_t n, const int32_t* __restrict__ run_lo,
                                                           const int32_t* __restrict__ s_a, const int32_t* __restrict__ s_b,
                                                           const int32_t* __restrict__ s_par,
                                                           const int32_t* __restrict__ s_lc, const int32_t* __restrict__ s_rc,
                                                           const int32_t* __restrict__ s_top, const int32_t* __restrict__ tkmap,
                                                           const int32_t* __restrict__ keep_par,
                                                           const uint32_t* __restrict__ keep_mark, DevState d) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n - 1) return;
    if (run_lo && run_lo[k] < 0) return;
    const int32_t id = tkmap[k];
    const int32_t lce = s_lc[k], rce = s_rc[k];
    const bool lleaf = lce & 1, rleaf = rce & 1;
    const int32_t lc = lleaf ? (lce >> 1) : tkmap[lce >> 1];
    const int32_t rc = rleaf ? (rce >> 1) : tkmap[rce >> 1];
    const int32_t sp = s_par[k];
    uint32_t mark = (lleaf ? 1u : 0u) | (rleaf ? 2u : 0u);
    int32_t par;
    if (sp == -1) {  // run root: keeps the parent link and right-child bit it had
        par = keep_par ? keep_par[id] : -1;
        mark |= keep_mark ? (keep_mark[id] & 4u) : 0u;
    } else {
        par = tkmap[sp & ~(1 << 30)];
        if (sp & (1 << 30)) mark |= 4u;
    }
    d.int_lc[id] = lc;
    d.int_rc[id] = rc;
    d.int_par[id] = par;
    d.int_mark[id] = mark;
    d.itopo[id] = make_int4(lc, rc, par, (int32_t)mark);
    d.int_rx[id] = s_a[k];
    d.int_ry[id] = s_b[k];
    d.inode[id].lo.w = i2f(s_a[k]);
    d.inode[id].hi.w = i2f(enc_lca(s_b[k] + 1, n, s_top, tkmap));
}

__global__ void __launch_bounds__(kBlock) k_write_leaves(int32_t n, const int32_t* __restrict__ leaf_lo,
                                                         const int32_t* __restrict__ s_lpar,
                                                         const int32_t* __restrict__ s_top,
                                                         const int32_t* __restrict__ tkmap, DevState d) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    if (leaf_lo && leaf_lo[x] < 0) return;
    const int32_t sp = s_lpar[x];
    const int32_t par = tkmap[sp & ~(1 << 30)];
    d.leaf_par[x] = par;
    d.leaf_mark[x] = (sp & (1 << 30)) ? 7u : 3u;
    // the run's first leaf keeps its lca (its topmost spine node may lie above the run)
    if (!leaf_lo || leaf_lo[x] != x) d.leaf_lca[x] = enc_lca(x, n, s_top, tkmap);
    d.leaf[x].lo.w = i2f(par);
    if (!leaf_lo) {
        d.leaf[x].hi.w = i2f(enc_lca(x + 1, n, s_top, tkmap));
        if (x == n - 1) d.leaf_lca[n] = -1;
    }
}

// escape words of leaves whose successor's lca changed (restructure runs)
__global__ void k_fix_leaf_escape(int32_t n, const int32_t* __restrict__ leaf_lca, NodeRec* leaf) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    leaf[x].hi.w = i2f(leaf_lca[x + 1]);
}
__global__ void k_fix_int_escape(int32_t I, const int32_t* __restrict__ int_ry, const int32_t* __restrict__ leaf_lca,
                                 NodeRec* inode) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= I) return;
    inode[k].hi.w = i2f(leaf_lca[int_ry[k] + 1]);
}

// Full-tree emission from sorted codes d.codes[0..n-1] (n >= 2).
template <class C>
static void emit_full(DevState& d, const C* codes, int32_t n, hipStream_t s) {
    fill_u32(d.depths, n, 0u, s);
    fill_i32(d.s_top, n, -1, s);
    hipLaunchKernelGGL(k_emit_internal<C>, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, codes, nullptr, nullptr,
                       d.s_a, d.s_b, d.s_par, d.s_lc, d.s_rc, d.s_top, d.depths);
    hipLaunchKernelGGL(k_emit_leaves<C>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, codes, nullptr, nullptr, d.s_lpar,
                       d.s_lc, d.s_rc);
    launch_exscan(d, d.depths, d.offsets, n, nullptr, s);
    hipLaunchKernelGGL(k_assign_preorder, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d.s_top, d.depths, d.offsets,
                       d.s_lc, nullptr, d.tkmap);
    hipLaunchKernelGGL(k_write_internal, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, nullptr, d.s_a, d.s_b, d.s_par,
                       d.s_lc, d.s_rc, d.s_top, d.tkmap, nullptr, nullptr, d);
    hipLaunchKernelGGL(k_write_leaves, dim3(grid_for(n)), dim3(kBlock), 0, s, n, nullptr, d.s_lpar, d.s_top, d.tkmap,
                       d);
}

void launch_emit_full(DevState& d, int32_t n, hipStream_t s) { emit_full<uint32_t>(d, d.codes, n, s); }
void launch_emit_full64(DevState& d, const uint64_t* keys, int32_t n, hipStream_t s) {
    emit_full<uint64_t>(d, keys, n, s);
}

// ======================================================================= restructure emission
// Runs: leaves of each restructured subtree [a, b]; run_lo/run_hi per leaf (and per split,
// with -1 outside runs). Preorder ids of run r start at its root id R: base[x] = R - offsets[a].
__global__ void k_run_bounds(int32_t n, const int32_t* __restrict__ rt_of_leaf, const int32_t* __restrict__ int_rx,
                             const int32_t* __restrict__ int_ry, int32_t* __restrict__ leaf_lo,
                             int32_t* __restrict__ leaf_hi, int32_t* __restrict__ split_lo,
                             int32_t* __restrict__ split_hi) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int32_t r = rt_of_leaf[x];
    if (r < 0) {
        leaf_lo[x] = -1; leaf_hi[x] = -1;
        if (x < n - 1) { split_lo[x] = -1; split_hi[x] = -1; }
        return;
    }
    const int32_t a = int_rx[r], b = int_ry[r];
    leaf_lo[x] = a; leaf_hi[x] = b;
    if (x < n - 1) {
        if (x < b) { split_lo[x] = a; split_hi[x] = b; }
        else { split_lo[x] = -1; split_hi[x] = -1; }
    }
}

__global__ void k_run_base(int32_t n, const int32_t* __restrict__ rt_of_leaf, const int32_t* __restrict__ int_rx,
                           const uint32_t* __restrict__ offsets, int32_t* __restrict__ base) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int32_t r = rt_of_leaf[x];
    base[x] = r < 0 ? 0 : r - (int32_t)offsets[int_rx[r]];
}

// Rebuild the topology inside the runs (codes in d.codes already re-sorted per run).
// Scratch: s_a/s_b/... as in the full build; run arrays passed in.
void launch_emit_runs(DevState& d, int32_t n, const int32_t* rt_of_leaf, int32_t* leaf_lo, int32_t* leaf_hi,
                      int32_t* split_lo, int32_t* split_hi, int32_t* base, int32_t* keep_par, uint32_t* keep_mark,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_run_bounds, dim3(grid_for(n)), dim3(kBlock), 0, s, n, rt_of_leaf, d.int_rx, d.int_ry, leaf_lo,
                       leaf_hi, split_lo, split_hi);
    fill_u32(d.depths, n, 0u, s);
    fill_i32(d.s_top, n, -1, s);
    hipLaunchKernelGGL(k_emit_internal<uint32_t>, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, d.codes, split_lo, split_hi,
                       d.s_a, d.s_b, d.s_par, d.s_lc, d.s_rc, d.s_top, d.depths);
    hipLaunchKernelGGL(k_emit_leaves<uint32_t>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d.codes, leaf_lo, leaf_hi, d.s_lpar,
                       d.s_lc, d.s_rc);
    launch_exscan(d, d.depths, d.offsets, n, nullptr, s);
    hipLaunchKernelGGL(k_run_base, dim3(grid_for(n)), dim3(kBlock), 0, s, n, rt_of_leaf, d.int_rx, d.offsets, base);
    // keep_par / keep_mark: copies of int_par / int_mark before rewriting
    hipMemcpyAsync(keep_par, d.int_par, sizeof(int32_t) * (n - 1), hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(keep_mark, d.int_mark, sizeof(uint32_t) * (n - 1), hipMemcpyDeviceToDevice, s);
    hipLaunchKernelGGL(k_assign_preorder, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d.s_top, d.depths, d.offsets,
                       d.s_lc, base, d.tkmap);
    hipLaunchKernelGGL(k_write_internal, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, split_lo, d.s_a, d.s_b, d.s_par,
                       d.s_lc, d.s_rc, d.s_top, d.tkmap, keep_par, keep_mark, d);
    hipLaunchKernelGGL(k_write_leaves, dim3(grid_for(n)), dim3(kBlock), 0, s, n, leaf_lo, d.s_lpar, d.s_top, d.tkmap,
                       d);
    hipLaunchKernelGGL(k_fix_leaf_escape, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d.leaf_lca, d.leaf);
    hipLaunchKernelGGL(k_fix_int_escape, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n - 1, d.int_ry, d.leaf_lca,
                       d.inode);
}

// ======================================================================= refit
// Treelet-aligned, level-synchronous refit: no atomics, no cross-workgroup hand-off.
//
// Topology preparation (after every topology change):
//  * a treelet is a maximal subtree with at most kTreelet leaves (a leaf whose parent holds
//    more is a treelet of its own); the internal nodes holding more form the top tree;
//  * chunk c is the run of treelets whose first leaf lies in [c*kTreelet, (c+1)*kTreelet).
//    The treelet holding leaf c*kTreelet either starts there or ends before (c+1)*kTreelet - 1,
//    so every window but the tail of the last treelet starts one, the chunks tile the leaves
//    and each spans fewer than kChunkCap = 2*kTreelet leaves;
//  * preorder numbering gives the internal nodes of a treelet with root r and leaves [a, b]
//    the ids [r, r + b - a): node p's entry sits at p - r + a, so a chunk's entries fill its
//    own leaf range;
//  * LDS slots are leaf positions in the chunk: leaf x holds slot x - chunk start and every
//    internal node the slot of its leftmost leaf, i.e. it overwrites its left child's box.
//    Nodes merged at one depth are disjoint subtrees, so each slot is written by at most one
//    of them, and a right child's slot (its own leftmost leaf) is written only by deeper nodes;
//  * each chunk lists its internal nodes deepest first (depth below the treelet root) as
//    8-byte entries (node, own slot | right child's slot << 10 | depth << 20), followed by
//    its treelet roots' hand-over offsets; the top tree is listed deepest first (depth below
//    the root) with its children's kinds, split into LDS-sized segments when it is large.
// Per refit: one workgroup per chunk writes the leaf boxes into LDS, merges one depth level
// per barrier in LDS and writes its node boxes at the end; then k_refit_top merges the top
// tree: in LDS (one workgroup per segment), the upper nodes above the segments through the
// node records, or, for top trees deeper than the kDepthCap level bins (long chains of equal
// Morton codes), by climbing. Within a chunk the depth is below the treelet size.
constexpr int kTreelet = 512;
constexpr int kChunkCap = 2 * kTreelet;
constexpr int kRefitPer = kChunkCap / kBlock;  // entries (leaves, internal nodes) per thread
constexpr int kDepthCap = 1024;
constexpr int kTopBlock = 1024;
constexpr int kRefitBlock = 512;                       // per-frame chunk refit: 8 waves per chunk
constexpr int kRefitPerT = kChunkCap / kRefitBlock;     // its leaves / nodes per thread
constexpr int kTopLds = kRefitTopLds;  // top trees up to this size are merged in LDS
constexpr int kTopGrid = kRefitTopGrid;  // at most this many LDS segments of larger top trees
constexpr int kTopPer = kTopLds / kTopBlock;
static_assert(kChunkCap % kBlock == 0, "chunk capacity is a whole number of block passes");
static_assert(kChunkCap <= 65536, "slots are packed in 16 bits");

__device__ inline bool rp_is_top(const DevState& d, int32_t p) { return d.int_ry[p] - d.int_rx[p] >= kTreelet; }
// Chunk entry (node, own slot | right child's slot << 10 | depth << 20); the left child's slot
// is the node's own. Slots are below kChunkCap, depths below kTreelet.
static_assert(kChunkCap <= 1024 && kTreelet <= 1024, "chunk entries pack slots and depths in 10 bits");
__host__ __device__ inline int32_t rp_pack(int32_t own, int32_t rs, int32_t depth) { return own | (rs << 10) | (depth << 20); }
__device__ inline int32_t rp_depth(const int2& v) { return v.y >> 20; }
__device__ inline int32_t rp_own(const int2& v) { return v.y & 0x3FF; }
__device__ inline int32_t rp_right(const int2& v) { return (v.y >> 10) & 0x3FF; }

// Treelet starts (atomic min per window) and the unordered top-tree list.
__global__ void __launch_bounds__(kBlock) k_rp_starts(int32_t n, DevState d) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int32_t lp = d.leaf_par[x];
    if (lp < 0 || rp_is_top(d, lp)) atomicMin(&d.rp_cstart[x / kTreelet], x);
    bool top = false;
    int4 e = make_int4(0, 0, 0, 0);
    if (x < n - 1) {
        const int32_t rx = d.int_rx[x];
        top = d.int_ry[x] - rx >= kTreelet;
        const int32_t par = d.int_par[x];
        if (!top && (par < 0 || rp_is_top(d, par))) atomicMin(&d.rp_cstart[rx / kTreelet], rx);
        if (top) {
            // depth below the root; a top tree deeper than the level bins (chains of equal codes)
            // is merged by climbing instead
            int32_t depth = 0, q = par;
            for (; q >= 0 && depth < kDepthCap - 1; q = d.int_par[q]) ++depth;
            if (q >= 0) atomicOr(d.rp_tdeep, 1u);
            const int4 tp = d.itopo[x];
            e = make_int4(x, tp.x, tp.y, depth | ((tp.w & 1) << 16) | ((tp.w & 2) << 17));
        }
    }
    const uint32_t slot = wave_append(d.rp_tcount, top);
    if (top) d.rp_tent[slot] = e;
}

// Per leaf: the treelet root and depth, then the entries of the internal nodes whose leftmost
// leaf it is (each node written once; slot b of every treelet written empty by leaf b).
__global__ void __launch_bounds__(kBlock) k_rp_walk(int32_t n, DevState d) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int2 empty = make_int2(-1, -1);
    const int32_t p = d.leaf_par[x];
    int32_t len = 0, r = -1;
    for (int32_t q = p; q >= 0 && len < kTreelet && !rp_is_top(d, q); q = d.int_par[q]) {
        r = q;
        ++len;
    }
    if (len == 0) {  // a treelet of its own
        d.rp_ent[x] = empty;
        return;
    }
    const int32_t a = d.int_rx[r];
    if (x == d.int_ry[r]) d.rp_ent[x] = empty;
    const int32_t cs = d.rp_cstart[a / kTreelet];
    const int32_t own = x - cs;  // slot of every node whose leftmost leaf is x (and of x)
    int32_t q = p;
    for (int32_t i = 0; i < len && d.int_rx[q] == x; ++i) {
        const int4 tp = d.itopo[q];
        const int32_t rs = ((tp.w & 2) ? tp.y : d.int_rx[tp.y]) - cs;
        d.rp_ent[q - r + a] = make_int2(q, rp_pack(own, rs, len - 1 - i));
        q = tp.z;
    }
}

// Per chunk: entries sorted in place by depth, deepest first; (count, deepest depth | treelet
// roots << 16). The roots (depth 0) end the list; each one's hand-over offset for the top kernel
// (or -1) follows the list at the root's position + roots: a chunk's treelets of L >= 2 leaves
// hold L - 1 entries each, so the chunk's leaf range has room for one more per treelet.
__global__ void __launch_bounds__(kBlock) k_rp_chunk_sort(DevState d) {
    __shared__ uint32_t hist[kDepthCap];
    __shared__ uint32_t dmax, roots;
    const int32_t c = blockIdx.x;
    const int32_t s = d.rp_cstart[c], e = d.rp_cstart[c + 1];
    for (int i = threadIdx.x; i < kDepthCap; i += kBlock) hist[i] = 0;
    if (threadIdx.x == 0) dmax = 0;
    __syncthreads();
    int2 v[kRefitPer];
#pragma unroll
    for (int k = 0; k < kRefitPer; ++k) {
        const int32_t i = s + threadIdx.x + k * kBlock;
        v[k] = i < e ? d.rp_ent[i] : make_int2(-1, -1);
        if (v[k].x >= 0) {
            atomicAdd(&hist[rp_depth(v[k])], 1u);
            atomicMax(&dmax, (uint32_t)rp_depth(v[k]));
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        roots = hist[0];
        uint32_t acc = 0;
        for (int t = (int)dmax; t >= 0; --t) {
            const uint32_t h = hist[t];
            hist[t] = acc;
            acc += h;
        }
        d.rp_cinfo[c] = make_int2((int32_t)acc, (int32_t)(dmax | (roots << 16)));
    }
    __syncthreads();
    // hand-over offsets exist for children of segment nodes (list positions below the upper
    // nodes), unless the top tree goes through the records as a whole
    const int32_t seg_end = d.rp_tinfo[2] > 0 ? d.rp_tinfo[3] : 0;
#pragma unroll
    for (int k = 0; k < kRefitPer; ++k) {
        if (v[k].x < 0) continue;
        const uint32_t pos = s + atomicAdd(&hist[rp_depth(v[k])], 1u);
        d.rp_ent[pos] = v[k];
        if (rp_depth(v[k]) == 0) {
            // the treelet root hands its box to the top kernel: dense offset from its parent's
            // entry, side by mark bit 2 (right child)
            const int4 tp = d.itopo[v[k].x];
            int32_t dest = -1;
            // rp_tpos holds positions only when the top tree has a list (seg_end > 0); deep
            // top trees merged by climbing never write it
            if (seg_end > 0 && tp.z >= 0 && d.rp_tpos[tp.z] < seg_end) {
                const int4 pe = d.rp_tlist[d.rp_tpos[tp.z]];
                dest = (tp.w & 4) ? pe.z : pe.y;
            }
            d.rp_ent[pos + roots] = make_int2(dest, 0);
        }
    }
}

// Top tree sorted deepest first; level t occupies [rp_tlev[t], t ? rp_tlev[t-1] : count).
// Entry (node, left, right, depth | left kind << 16 | right kind << 18), kind 0: internal node
// record, 1: leaf record, 2: list position of a top node (set here once positions are known,
// for top trees merged in LDS only; larger ones keep the node ids).
__device__ inline int32_t top_depth(const int4& e) { return e.w & 0xFFFF; }
__device__ inline int top_kind(const int4& e, int side) { return (e.w >> (16 + 2 * side)) & 3; }

__global__ void __launch_bounds__(kTopBlock) k_rp_top_sort(DevState d) {
    __shared__ uint32_t hist[kDepthCap];
    __shared__ uint32_t dmax;
    const int32_t m = (int32_t)*d.rp_tcount;
    if (*d.rp_tdeep) {  // deeper than the level bins: no list, the top kernel climbs (rp_tent)
        if (threadIdx.x == 0) {
            d.rp_tinfo[0] = m;
            d.rp_tinfo[1] = -1;
            d.rp_tinfo[2] = -2;
            d.rp_tinfo[3] = 0;
            d.rp_tinfo[4] = 0;
            d.rp_tinfo[5] = -1;
        }
        return;
    }
    for (int i = threadIdx.x; i < kDepthCap; i += kTopBlock) hist[i] = 0;
    if (threadIdx.x == 0) dmax = 0;
    __syncthreads();
    for (int32_t i = threadIdx.x; i < m; i += kTopBlock) {
        const uint32_t t = (uint32_t)top_depth(d.rp_tent[i]);
        atomicAdd(&hist[t], 1u);
        atomicMax(&dmax, t);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int t = (int)dmax; t >= 0; --t) {
            const uint32_t h = hist[t];
            hist[t] = acc;
            d.rp_tlev[t] = (int32_t)acc;
            acc += h;
        }
        d.rp_tinfo[0] = m;
        d.rp_tinfo[1] = m > 0 ? (int32_t)dmax : -1;
        // the shallowest levels holding at most one wave's worth of nodes (positions from p1 on)
        int32_t l1 = -1, p1 = m;
        for (int t = 0; t <= (int)dmax && m > 0; ++t) {
            if (m - (int32_t)hist[t] > 64) break;
            l1 = t;
            p1 = (int32_t)hist[t];
        }
        // one segment holding the whole top tree, no upper nodes; larger trees are split by
        // k_rp_top_segments, which overwrites these
        d.rp_tinfo[2] = 1;
        d.rp_tinfo[3] = m;
        d.rp_tinfo[4] = 0;
        d.rp_tinfo[5] = -1;
        d.rp_seg[0] = make_int4(0, m, m > 0 ? (int32_t)dmax : -1, (l1 + 1) | (p1 << 16));
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < m; i += kTopBlock) {
        const int4 e = d.rp_tent[i];
        const uint32_t pos = atomicAdd(&hist[top_depth(e)], 1u);
        d.rp_tlist[pos] = e;
        d.rp_tpos[e.x] = (int32_t)pos;
    }
    __syncthreads();
    // LDS-resident merge only (larger trees keep node ids): top children by list position, and
    // treelet-root children by dense hand-over offsets in list order, so that k_refit writes
    // exactly the boxes the top kernel loads. Thread t renumbers positions [t*kTopPer, ...).
    if (m > kTopLds) return;
    __shared__ uint32_t part[kTopBlock];
    int4 ev[kTopPer];
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < kTopPer; ++k) {
        const int32_t i = (int32_t)threadIdx.x * kTopPer + k;
        if (i >= m) continue;
        int4 e = d.rp_tlist[i];
        if (top_kind(e, 0) == 0 && rp_is_top(d, e.y)) {
            e.y = d.rp_tpos[e.y];
            e.w |= 2 << 16;
        }
        if (top_kind(e, 1) == 0 && rp_is_top(d, e.z)) {
            e.z = d.rp_tpos[e.z];
            e.w |= 2 << 18;
        }
        cnt += (top_kind(e, 0) == 0 ? 1u : 0u) + (top_kind(e, 1) == 0 ? 1u : 0u);
        ev[k] = e;
    }
    part[threadIdx.x] = cnt;
    __syncthreads();
    for (int off = 1; off < kTopBlock; off <<= 1) {  // inclusive scan (prep only)
        const uint32_t add = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    int32_t o = (int32_t)(part[threadIdx.x] - cnt);
#pragma unroll
    for (int k = 0; k < kTopPer; ++k) {
        const int32_t i = (int32_t)threadIdx.x * kTopPer + k;
        if (i >= m) continue;
        int4 e = ev[k];
        if (top_kind(e, 0) == 0) e.y = o++;
        if (top_kind(e, 1) == 0) e.z = o++;
        d.rp_tlist[i] = e;
    }
}

// Top trees of more than kTopLds nodes: segments (maximal top subtrees of at most kTopLds
// nodes), merged in LDS by one k_refit_top workgroup each, and the upper nodes above them,
// merged through the records by the segment workgroup that finishes last. One workgroup, on the
// list k_rp_top_sort left sorted deepest first with node ids as children:
//  1. top-tree subtree sizes bottom-up, one level per barrier;
//  2. segment ids top-down: a node whose size fits and whose parent's does not starts one;
//  3. the list regrouped by segment (segment 0 first, upper nodes last), top children of
//     segment nodes as positions within their segment, node positions renumbered, and dense
//     hand-over offsets for the segment nodes' treelet-root children in list order.
// More than kTopGrid segments, or more hand-over boxes than the buffer holds, leave the list as
// it is and send the whole top tree through the records (rp_tinfo[2] = -1).
__global__ void __launch_bounds__(kTopBlock) k_rp_top_segments(DevState d, int32_t tcb_cap) {
    const int32_t m = d.rp_tinfo[0];
    if (m <= kTopLds || d.rp_tinfo[2] == -2) return;
    const int32_t top = d.rp_tinfo[1];
    __shared__ int32_t ls[kDepthCap];
    __shared__ uint32_t nseg, nhand, cnt[kTopGrid + 1], fill[kTopGrid + 1], deep[kTopGrid + 1], mind[kTopGrid + 1];
    __shared__ uint32_t part[kTopBlock];
    __shared__ int32_t sbase[kTopGrid + 1];
    for (int32_t t = threadIdx.x; t <= top; t += kTopBlock) ls[t] = d.rp_tlev[t];
    if (threadIdx.x == 0) { nseg = 0; nhand = 0; }
    __syncthreads();
    int32_t* ts = d.rp_tsz;   // per list position: top nodes in the subtree, later the new position
    int32_t* sg = d.rp_tseg;  // per list position: segment, -1 for upper nodes
    for (int32_t lev = top; lev >= 0; --lev) {
        const int32_t b0 = ls[lev], b1 = lev > 0 ? ls[lev - 1] : m;
        for (int32_t i = b0 + (int32_t)threadIdx.x; i < b1; i += kTopBlock) {
            const int4 e = d.rp_tlist[i];
            int32_t sz = 1;
            if (top_kind(e, 0) == 0 && rp_is_top(d, e.y)) sz += ts[d.rp_tpos[e.y]];
            if (top_kind(e, 1) == 0 && rp_is_top(d, e.z)) sz += ts[d.rp_tpos[e.z]];
            ts[i] = sz;
        }
        __syncthreads();
    }
    for (int32_t lev = 0; lev <= top; ++lev) {
        const int32_t b0 = ls[lev], b1 = lev > 0 ? ls[lev - 1] : m;
        for (int32_t i = b0 + (int32_t)threadIdx.x; i < b1; i += kTopBlock) {
            const int32_t par = d.itopo[d.rp_tlist[i].x].z;  // the parent of a top node is a top node
            const int32_t pi = par >= 0 ? d.rp_tpos[par] : -1;
            int32_t sid = -1;
            if (ts[i] <= kTopLds) sid = (pi < 0 || ts[pi] > kTopLds) ? (int32_t)atomicAdd(&nseg, 1u) : sg[pi];
            sg[i] = sid;
        }
        __syncthreads();
    }
    const uint32_t ns = nseg;
    for (uint32_t g = threadIdx.x; g <= ns && g <= kTopGrid; g += kTopBlock) {
        cnt[g] = 0; fill[g] = 0; deep[g] = 0; mind[g] = 0xFFFF;
    }
    __syncthreads();
    if (ns > kTopGrid) {
        if (threadIdx.x == 0) d.rp_tinfo[2] = -1;
        return;
    }
    for (int32_t i = threadIdx.x; i < m; i += kTopBlock) {
        const int4 e = d.rp_tlist[i];
        const int32_t gi = sg[i] < 0 ? (int32_t)ns : sg[i];
        atomicAdd(&cnt[gi], 1u);
        atomicMax(&deep[gi], (uint32_t)top_depth(e));
        atomicMin(&mind[gi], (uint32_t)top_depth(e));
        if (sg[i] >= 0) {
            const uint32_t h = (top_kind(e, 0) == 0 && !rp_is_top(d, e.y) ? 1u : 0u) +
                               (top_kind(e, 1) == 0 && !rp_is_top(d, e.z) ? 1u : 0u);
            if (h) atomicAdd(&nhand, h);
        }
    }
    __syncthreads();
    if (nhand > (uint32_t)tcb_cap) {
        if (threadIdx.x == 0) d.rp_tinfo[2] = -1;
        return;
    }
    if (threadIdx.x == 0) {
        int32_t acc = 0;
        for (uint32_t g = 0; g <= ns; ++g) {
            sbase[g] = acc;
            acc += (int32_t)cnt[g];
        }
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < m; i += kTopBlock) {
        const int32_t gi = sg[i] < 0 ? (int32_t)ns : sg[i];
        ts[i] = sbase[gi] + (int32_t)atomicAdd(&fill[gi], 1u);
    }
    __syncthreads();
    for (int32_t i = threadIdx.x; i < m; i += kTopBlock) {
        int4 e = d.rp_tlist[i];
        const int32_t gi = sg[i];
        if (gi >= 0) {
            if (top_kind(e, 0) == 0 && rp_is_top(d, e.y)) {
                e.y = ts[d.rp_tpos[e.y]] - sbase[gi];
                e.w |= 2 << 16;
            }
            if (top_kind(e, 1) == 0 && rp_is_top(d, e.z)) {
                e.z = ts[d.rp_tpos[e.z]] - sbase[gi];
                e.w |= 2 << 18;
            }
        }
        d.rp_tent[ts[i]] = e;
    }
    __syncthreads();
    for (int32_t j = threadIdx.x; j < m; j += kTopBlock) d.rp_tpos[d.rp_tent[j].x] = j;
    // dense hand-over offsets over the segment positions, thread t a contiguous run of them
    const int32_t nsp = sbase[ns];
    const int32_t per = (nsp + kTopBlock - 1) / kTopBlock;
    const int32_t j0 = min((int32_t)threadIdx.x * per, nsp), j1 = min(j0 + per, nsp);
    uint32_t c = 0;
    for (int32_t j = j0; j < j1; ++j) {
        const int4 e = d.rp_tent[j];
        c += (top_kind(e, 0) == 0 ? 1u : 0u) + (top_kind(e, 1) == 0 ? 1u : 0u);
    }
    part[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < kTopBlock; off <<= 1) {  // inclusive scan (prep only)
        const uint32_t add = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    int32_t o = (int32_t)(part[threadIdx.x] - c);
    for (int32_t j = j0; j < j1; ++j) {
        int4 e = d.rp_tent[j];
        if (top_kind(e, 0) == 0) e.y = o++;
        if (top_kind(e, 1) == 0) e.z = o++;
        d.rp_tlist[j] = e;
    }
    for (int32_t j = nsp + (int32_t)threadIdx.x; j < m; j += kTopBlock) d.rp_tlist[j] = d.rp_tent[j];
    // segments: merged from their deepest level down to their root's (no one-wave levels)
    for (uint32_t g = threadIdx.x; g < ns; g += kTopBlock)
        d.rp_seg[g] = make_int4(sbase[g], (int32_t)cnt[g], (int32_t)deep[g], (int32_t)mind[g] | ((int32_t)cnt[g] << 16));
    if (threadIdx.x == 0) {
        d.rp_tinfo[2] = (int32_t)ns;
        d.rp_tinfo[3] = nsp;
        d.rp_tinfo[4] = (int32_t)cnt[ns];
        d.rp_tinfo[5] = (int32_t)deep[ns];
    }
}

int32_t refit_chunks(int32_t n) { return (n + kTreelet - 1) / kTreelet; }

void launch_refit_prep(DevState& d, int32_t n, hipStream_t s) {
    d.topo_version++;  // every topology change passes here
    if (n < 1) return;
    const int32_t nc = refit_chunks(n);
    fill_i32(d.rp_cstart, nc + 1, n, s);
    fill_u32(d.rp_tcount, 1, 0u, s);
    hipLaunchKernelGGL(k_rp_starts, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d);
    fill_u32(d.rp_tdone, 1, 0u, s);
    fill_u32(d.rp_tdeep, 1, 0u, s);
    fill_i32(d.rp_tsz, n, 0, s);  // climb flags of deep top trees (prep scratch otherwise)
    hipLaunchKernelGGL(k_rp_top_sort, dim3(1), dim3(kTopBlock), 0, s, d);
    hipLaunchKernelGGL(k_rp_top_segments, dim3(1), dim3(kTopBlock), 0, s, d, refit_tcb_boxes(n));
    hipLaunchKernelGGL(k_rp_walk, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d);
    hipLaunchKernelGGL(k_rp_chunk_sort, dim3(nc), dim3(kBlock), 0, s, d);
}

__device__ inline void store_box(NodeRec* rec, const Box& b) {
    float4* p = reinterpret_cast<float4*>(rec);
    p[0].x = b.mnx; p[0].y = b.mny; p[0].z = b.mnz;
    p[1].x = b.mxx; p[1].y = b.mxy; p[1].z = b.mxz;
}

__global__ void __launch_bounds__(kRefitBlock) k_refit(int32_t n, DevState d) {
    __shared__ Box sb[kChunkCap];
    const int32_t c = blockIdx.x;
    const int32_t s = d.rp_cstart[c], e = d.rp_cstart[c + 1];
    const int2 info = d.rp_cinfo[c];  // (internal nodes, deepest level | treelet roots << 16)
    // leaf boxes: LDS + records; per-triangle finiteness for the narrow phase. A leaf's face
    // indices arrive in one load and its three vertices together, the leaves of a thread one
    // after another: measured fastest on the 2M-triangle cloth (k_refit 46.8 us, against 50.4
    // with every index and vertex loaded on its own and 63-69 with all of a thread's loads in
    // flight at once, which congests the gathers).
    int bad = 0;
#pragma unroll
    for (int k = 0; k < kRefitPerT; ++k) {
        const int32_t x = s + threadIdx.x + k * kRefitBlock;
        if (x < e) {
            bool fin = true;
            Box b = empty_box();
            if (d.pbox) {  // box mode: the caller's box; a shared vertex need not lie in both boxes
                b = given_box(d.pbox, d.prim_idx[x]);
                fin = false;
            } else {
                const int32_t* fp = d.lfaces + 3 * (int64_t)x;
                const int32_t f[3] = {fp[0], fp[1], fp[2]};
                float pv[3][3];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const float* v = d.verts + 3 * (int64_t)f[j];
                    pv[j][0] = v[0]; pv[j][1] = v[1]; pv[j][2] = v[2];
                }
                int ok = 1;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    // |x| <= 1e8 keeps every SAT product finite (dots of edge x normal crosses
                    // stay below 1536 * 1e32): no NaN can drop the shared vertex's projection.
                    ok &= (int)(fabsf(pv[j][0]) <= 1e8f) & (int)(fabsf(pv[j][1]) <= 1e8f) &
                          (int)(fabsf(pv[j][2]) <= 1e8f);
                    box_add(b, pv[j][0], pv[j][1], pv[j][2]);
                }
                fin = ok != 0;
            }
            d.tri_finite[x] = fin ? 1 : 0;
            bad |= fin ? 0 : 1;
            sb[x - s] = b;
            store_box(&d.leaf[x], b);
        }
    }
    int2 v[kRefitPerT];
#pragma unroll
    for (int k = 0; k < kRefitPerT; ++k) {
        const int32_t i = threadIdx.x + k * kRefitBlock;
        v[k] = i < info.x ? d.rp_ent[s + i] : make_int2(-1, -1);
    }
    Box out[kRefitPerT];
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) d.rp_cbad[c] = (uint32_t)bad;
    // treelet roots: their hand-over offsets, loaded now and used at the end
    const int32_t roots = info.y >> 16;
    int32_t dest[kRefitPerT];
#pragma unroll
    for (int k = 0; k < kRefitPerT; ++k) {
        const int32_t i = threadIdx.x + k * kRefitBlock;
        dest[k] = v[k].x >= 0 && rp_depth(v[k]) == 0 ? d.rp_ent[s + i + roots].x : -1;
    }
    for (int32_t lev = info.y & 0xFFFF; lev >= 0; --lev) {
#pragma unroll
        for (int k = 0; k < kRefitPerT; ++k) {
            if (rp_depth(v[k]) == lev) {
                out[k] = box_merge(sb[rp_own(v[k])], sb[rp_right(v[k])]);
                sb[rp_own(v[k])] = out[k];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kRefitPerT; ++k) {
        if (v[k].x < 0) continue;
        store_box(&d.inode[v[k].x], out[k]);
        // a treelet root under an LDS-sized top tree also hands its box to the top kernel, which
        // then loads it without a dependent load
        if (dest[k] >= 0) {
            float* dst = d.rp_tcb + 6 * (int64_t)dest[k];
            dst[0] = out[k].mnx; dst[1] = out[k].mny; dst[2] = out[k].mnz;
            dst[3] = out[k].mxx; dst[4] = out[k].mxy; dst[5] = out[k].mxz;
        }
    }
}

__device__ inline Box record_box(const DevState& d, int32_t v, int kind) {
    const NodeRec* r = kind == 1 ? &d.leaf[v] : &d.inode[v];
    return rec_box(r->lo, r->hi);
}

// Box of a top node's child below the top tree: kind 0 a treelet root (dense hand-over box),
// 1 a leaf (record), 2 none (a valid address is read and the value ignored). Branch-free, so
// that every load of a thread is in flight before the first one is used.
__device__ inline Box below_top_box(const DevState& d, int kind, int32_t ref) {
    const float* p = kind == 1 ? reinterpret_cast<const float*>(&d.leaf[ref]) : d.rp_tcb + 6 * (int64_t)(kind == 0 ? ref : 0);
    const float* q = p + (kind == 1 ? 4 : 3);
    return {p[0], p[1], p[2], q[0], q[1], q[2]};
}

// Record box of node v (kind 1: leaf, otherwise internal), without a branch.
__device__ inline Box below_top_record(const DevState& d, int kind, int32_t v) {
    const NodeRec* r = kind == 1 ? &d.leaf[v] : &d.inode[v];
    const float* p = reinterpret_cast<const float*>(r);
    return {p[0], p[1], p[2], p[4], p[5], p[6]};
}

// Top-tree nodes merged level by level through the node records (children are node ids,
// kinds 0 and 1; records of this workgroup are ordered by its barriers). whole: the list sorted
// deepest first with level starts in rp_tlev (m nodes, deepest level top); otherwise the m upper
// nodes after the segments, deepest level top, in no order (each level scans them). A level goes
// in batches of up to four entries per thread, a batch's entry loads and then its child-record
// loads all in flight together.
__device__ inline void refit_top_records(const DevState& d, Box* lds, int32_t m, int32_t top, bool whole) {
    int32_t* ls = reinterpret_cast<int32_t*>(lds);  // ls[t]: first position of level t
    const int32_t ub = d.rp_tinfo[3];               // first upper position
    if (whole)
        for (int32_t t = threadIdx.x; t <= top; t += kTopBlock) ls[t] = d.rp_tlev[t];
    __syncthreads();
    for (int32_t lev = top; lev >= 0; --lev) {
        const int32_t b0 = whole ? ls[lev] : ub, b1 = whole ? (lev > 0 ? ls[lev - 1] : m) : ub + m;
        // batches of K rounds of kTopBlock entries, K = 4 while more than two rounds remain, then
        // 2, then 1; indices clamped (a repeated entry is merged twice into the same box)
        auto batch = [&](int32_t at, auto kc) {
            constexpr int K = decltype(kc)::value;
            int4 e[K];
            Box l[K], r[K];
#pragma unroll
            for (int j = 0; j < K; ++j) e[j] = d.rp_tlist[min(at + (int32_t)threadIdx.x + j * kTopBlock, b1 - 1)];
#pragma unroll
            for (int j = 0; j < K; ++j) {
                l[j] = below_top_record(d, top_kind(e[j], 0), e[j].y);
                r[j] = below_top_record(d, top_kind(e[j], 1), e[j].z);
            }
#pragma unroll
            for (int j = 0; j < K; ++j)
                if (at + (int32_t)threadIdx.x + j * kTopBlock < b1 && (whole || top_depth(e[j]) == lev))
                    store_box(&d.inode[e[j].x], box_merge(l[j], r[j]));
        };
        int32_t at = b0;
        for (; b1 - at > 2 * kTopBlock; at += 4 * kTopBlock) batch(at, std::integral_constant<int, 4>{});
        for (; b1 - at > kTopBlock; at += 2 * kTopBlock) batch(at, std::integral_constant<int, 2>{});
        for (; at < b1; at += kTopBlock) batch(at, std::integral_constant<int, 1>{});
        __syncthreads();
    }
}

// Top trees deeper than the level bins (long chains of equal Morton codes): one workgroup
// climbs from every top node whose children are both below the top tree; at a node with two
// top children the second arrival merges (flag per node, workgroup-scope acquire/release,
// reset by the second arrival), at a node with one it merges at once. Each step reads the
// sibling's record, so the time is the longest chain; only such inputs take this path.
__device__ inline void refit_top_climb(const DevState& d, int32_t m) {
    uint32_t* flag = reinterpret_cast<uint32_t*>(d.rp_tsz);
    for (int32_t i = threadIdx.x; i < m; i += kTopBlock) {
        const int4 e = d.rp_tent[i];
        const int lk = top_kind(e, 0), rk = top_kind(e, 1);
        if ((lk == 0 && rp_is_top(d, e.y)) || (rk == 0 && rp_is_top(d, e.z))) continue;
        Box b = box_merge(below_top_record(d, lk, e.y), below_top_record(d, rk, e.z));
        store_box(&d.inode[e.x], b);
        for (int32_t q = e.x;;) {
            const int4 tq = d.itopo[q];
            const int32_t u = tq.z;
            if (u < 0) break;
            const int4 tu = d.itopo[u];
            const bool right = (tq.w & 4) != 0;  // q is u's right child
            const int32_t sib = right ? tu.x : tu.y;
            const int sk = (right ? (tu.w & 1) : (tu.w & 2)) ? 1 : 0;
            if (sk == 0 && rp_is_top(d, sib)) {
                if (__hip_atomic_fetch_add(&flag[u], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) break;
                flag[u] = 0;
            }
            b = box_merge(b, below_top_record(d, sk, sib));
            store_box(&d.inode[u], b);
            q = u;
        }
    }
}

// One workgroup per segment (the whole top tree when it has at most kTopLds nodes): the
// segment stays in LDS, slot = position in the segment; a node whose children are both below
// the top tree (treelet roots or leaves) is merged at once, a node with one such child starts
// its slot with that child's box and merges the other at its level. The workgroup that
// finishes last merges the upper nodes above the segments.

__global__ void __launch_bounds__(kTopBlock) k_refit_top(DevState d, int32_t chunks, int zero_step) {
    __shared__ Box tb[kTopLds];
    __shared__ uint32_t last;
    const int32_t g = blockIdx.x;
    if (zero_step && g == 0 && threadIdx.x < 64) zero_step_counters(d, threadIdx.x);
    const int32_t nseg = d.rp_tinfo[2];
    // frame-wide (workgroup 0): may KEEP count shared-vertex pairs without a test (every leaf box
    // built from finite vertices); OR over the chunks, four loads per thread in flight at once
    // (indices clamped: a repeated flag does not change the OR)
    auto chunks_bad = [&]() {
        int bad = 0;
        for (int32_t c0 = 0; c0 < chunks; c0 += 4 * kTopBlock) {
            uint32_t a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = d.rp_cbad[min(c0 + (int32_t)threadIdx.x + u * kTopBlock, chunks - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u) bad |= (int)a[u];
        }
        return bad;
    };
    if (nseg < 0) {  // too many segments: the whole top tree through the records, deepest first
        if (g != 0) return;
        const int bad = __syncthreads_or(chunks_bad());
        if (threadIdx.x == 0) d.ctr[C_NO_IMPLICIT] = (uint32_t)bad;
        if (nseg == -2) refit_top_climb(d, d.rp_tinfo[0]);
        else refit_top_records(d, tb, d.rp_tinfo[0], d.rp_tinfo[1], true);
        return;
    }
    if (g >= nseg) return;
    // segment g: positions [base, base + m) of the list, all of its nodes' children in the
    // segment (kind 2, position relative to base) or below the top tree. Segment 0 starts at
    // position 0, so its list loads go out before the segment record arrives (the list holds
    // kTopLds entries past its end); positions past the segment are masked afterwards.
    int4 e[kTopPer];
    int32_t base = 0;
    if (g != 0) base = d.rp_seg[g].x;
#pragma unroll
    for (int k = 0; k < kTopPer; ++k) e[k] = d.rp_tlist[base + threadIdx.x + k * kTopBlock];
    const int4 sinfo = d.rp_seg[g];  // (base, nodes, deepest level, l1 + 1 | p1 << 16)
    const int32_t m = sinfo.y, top = sinfo.z, l1 = (sinfo.w & 0xFFFF) - 1, p1 = sinfo.w >> 16;
#pragma unroll
    for (int k = 0; k < kTopPer; ++k)
        if ((int32_t)threadIdx.x + k * kTopBlock >= m) e[k] = make_int4(0, 0, 0, 0xFFFF | (2 << 16) | (2 << 18));
    // children below the top tree, all loads of the thread in flight together: treelet roots
    // from the dense hand-over array k_refit wrote, leaves from their records
    Box cl[kTopPer], cr[kTopPer];
#pragma unroll
    for (int k = 0; k < kTopPer; ++k) {
        cl[k] = below_top_box(d, top_kind(e[k], 0), e[k].y);
        cr[k] = below_top_box(d, top_kind(e[k], 1), e[k].z);
    }
    // the shallowest levels (at most 64 nodes, positions p1..m-1) are merged by wave 0 alone,
    // one level after another without workgroup barriers; lane l holds position p1 + l
    int4 ew = make_int4(0, 0, 0, 0xFFFF | (2 << 16) | (2 << 18));
    if (threadIdx.x < 64 && p1 + (int32_t)threadIdx.x < m) ew = d.rp_tlist[base + p1 + threadIdx.x];
    int bad = g == 0 ? chunks_bad() : 0;
#pragma unroll
    for (int k = 0; k < kTopPer; ++k) {
        const int32_t i = threadIdx.x + k * kTopBlock;
        const int lk = top_kind(e[k], 0), rk = top_kind(e[k], 1);
        if (lk != 2 && rk != 2) {
            tb[i] = box_merge(cl[k], cr[k]);
            e[k].w |= 0xFFFF;  // done
        } else if (lk != 2) {
            tb[i] = cl[k];
        } else if (rk != 2) {
            tb[i] = cr[k];
        }
    }
    bad = __syncthreads_or(bad);
    if (g == 0 && threadIdx.x == 0) d.ctr[C_NO_IMPLICIT] = (uint32_t)bad;
    for (int32_t lev = top; lev > l1; --lev) {
#pragma unroll
        for (int k = 0; k < kTopPer; ++k) {
            if (top_depth(e[k]) == lev) {
                const int32_t i = threadIdx.x + k * kTopBlock;
                const int lk = top_kind(e[k], 0), rk = top_kind(e[k], 1);
                const Box l = tb[lk == 2 ? e[k].y : i];
                const Box r = tb[rk == 2 ? e[k].z : i];
                tb[i] = box_merge(l, r);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kTopPer; ++k) {
        const int32_t i = threadIdx.x + k * kTopBlock;
        if (i < p1) store_box(&d.inode[e[k].x], tb[i]);
    }
    if (threadIdx.x < 64) {
        // wave 0: LDS operations of one wave complete in issue order, so a level's writes are
        // seen by the next level's reads; the fences keep the compiler from reordering them
        const int32_t i = p1 + (int32_t)threadIdx.x;
        const int lk = top_kind(ew, 0), rk = top_kind(ew, 1);
        for (int32_t lev = l1; lev >= 0; --lev) {
            if (top_depth(ew) == lev && (lk == 2 || rk == 2)) {
                const Box l = tb[lk == 2 ? ew.y : i];
                const Box r = tb[rk == 2 ? ew.z : i];
                tb[i] = box_merge(l, r);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (i < m) store_box(&d.inode[ew.x], tb[i]);
    }
    // the upper nodes above the segments: merged through the node records by the segment
    // workgroup that finishes last (release of this workgroup's records, then a counter; the
    // last one acquires every segment's records and resets the counter for the next refit)
    if (d.rp_tinfo[4] == 0) return;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(d.rp_tdone, 1u) == (uint32_t)(nseg - 1) ? 1u : 0u;
    __syncthreads();
    if (!last) return;
    __threadfence();
    refit_top_records(d, tb, d.rp_tinfo[4], d.rp_tinfo[5], false);
    if (threadIdx.x == 0) *d.rp_tdone = 0;
}

void launch_refit(DevState& d, int32_t n, hipStream_t s, bool zero_step_counters) {
    if (n < 1) return;
    hipLaunchKernelGGL(k_refit, dim3(refit_chunks(n)), dim3(kRefitBlock), 0, s, n, d);
    hipLaunchKernelGGL(k_refit_top, dim3(kTopGrid), dim3(kTopBlock), 0, s, d, refit_chunks(n),
                       zero_step_counters ? 1 : 0);
}

// ======================================================================= prim map / misc
__global__ void k_prim_map(int32_t n, const int32_t* __restrict__ prim_idx, int32_t* __restrict__ prim_map) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    prim_map[prim_idx[s]] = s;
}
// Inverse permutation, and the faces gathered into leaf order once per sort so that refit and
// the narrow phase read them sequentially instead of gathering by triangle id every frame.
__global__ void k_leaf_faces(int32_t n, const int32_t* __restrict__ prim_idx, const int32_t* __restrict__ faces,
                             int32_t* __restrict__ lfaces) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int64_t t = prim_idx[x];
    lfaces[3 * (int64_t)x] = faces[3 * t];
    lfaces[3 * (int64_t)x + 1] = faces[3 * t + 1];
    lfaces[3 * (int64_t)x + 2] = faces[3 * t + 2];
}
void launch_prim_map(DevState& d, int32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_prim_map, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d.prim_idx, d.prim_map);
    hipLaunchKernelGGL(k_leaf_faces, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d.prim_idx, d.faces, d.lfaces);
}

}  // namespace bvhcd
