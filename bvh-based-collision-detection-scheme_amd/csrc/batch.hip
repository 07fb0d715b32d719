// Batched scenes for gfx950: many independent triangle soups of uneven sizes per launch.
//
// The reference handles one scene per LBvh / BvttFront pair (collision/lbvh/LBvhFixedDeformable,
// collision/bvtt_front/BvttFrontLooseIntra of littlemine/BVH-based-Collision-Detection-Scheme,
// Source/Project/Mine/), sized by compile-time constants and launched kernel by kernel with a
// per-kernel occupancy block size (system/CudaDevice/CudaDevice.cu:54-94). Here a batch is one
// tree over all scenes:
//  * sort keys are (scene id << 30 | 30-bit Morton code in the scene's own box), so the LSD
//    radix sort groups the scenes and the Karras/Apetrei emission turns every scene into one
//    subtree (a scene boundary differs in a bit above the Morton bits, i.e. a larger delta than
//    anything inside a scene); each scene's subtree is the LBVH the scene alone would get;
//  * refit and the narrow phase run once over all triangles;
//  * the self-CD traversal is one persistent launch: waves take 64-query chunks from a global
//    counter and each query walks only its own scene's leaves to its right;
//  * the narrow-phase hits are grouped per scene on the device (counts, exclusive scan,
//    scatter) into scene-local triangle ids;
//  * every launch of a frame reads its sizes from device memory and clears its counters with
//    kernels, so the whole frame is captured once in a hipGraph and replayed afterwards.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "bvhcd_internal.hpp"
#include "kernels.hpp"

using namespace bvhcd;

struct bvhcd_batch {
    bvhcd_config cfg{};
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    int32_t S = 0;
    int64_t T = 0, V = 0;
    std::vector<int64_t> toff, voff;  // per-scene triangle / vertex offsets (S + 1)
    DevState d;
    std::vector<void*> allocs;
    int32_t* tri_scene = nullptr;     // T
    int32_t* scene_last = nullptr;    // S: last sorted leaf of each scene
    int64_t* voff_dev = nullptr;      // S + 1
    uint32_t* sbox_u = nullptr;       // 8 S ordered-uint scene boxes
    float* sbox = nullptr;            // 8 S float scene boxes
    uint64_t* keys = nullptr;         // T
    uint64_t* keys_tmp = nullptr;     // T
    uint32_t* scnt = nullptr;         // S + 1 hits per scene, then exclusive offsets
    uint32_t* scur = nullptr;         // S + 1 scatter cursors
    uint32_t* queue = nullptr;        // traversal work counter
    int2* grouped = nullptr;          // pair_cap, hits grouped by scene (scene-local ids)
    bool topo = false, built = false, keys_in_tmp = false;
    hipGraph_t graph[2] = {nullptr, nullptr};          // [0] build frame, [1] refit frame
    hipGraphExec_t exec[2] = {nullptr, nullptr};
};

namespace {

int bfail(bvhcd_batch* b, int code, const std::string& m) {
    if (b) b->err = m;
    return code;
}

template <class T>
bool balloc(bvhcd_batch* b, T*& p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&q, n * sizeof(T)) != hipSuccess) return false;
    b->allocs.push_back(q);
    p = static_cast<T*>(q);
    return true;
}

void drop_graphs(bvhcd_batch* b) {
    for (int k = 0; k < 2; ++k) {
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->exec[k] = nullptr;
        b->graph[k] = nullptr;
    }
}

__device__ inline uint32_t ford(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float funord(uint32_t u) {
    u = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
    return __builtin_bit_cast(float, u);
}

__global__ void k_tri_scene(int64_t T, int32_t S, const int64_t* __restrict__ toff, int32_t* __restrict__ tri_scene) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = 0, hi = S - 1;  // last scene with toff[s] <= t
        while (lo < hi) {
            const int32_t m = (lo + hi + 1) >> 1;
            if (toff[m] <= t) lo = m;
            else hi = m - 1;
        }
        tri_scene[t] = lo;
    }
}

// Scene-local vertex ids to global ones. An id outside its scene's vertex range sets *bad and
// is replaced by the scene's first vertex, so no later kernel gathers out of bounds.
__global__ void k_globalize(int64_t T, const int32_t* __restrict__ tri_scene, const int64_t* __restrict__ voff,
                            int32_t* __restrict__ faces, uint32_t* bad) {
    bool any = false;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = tri_scene[t];
        const int32_t o = (int32_t)voff[s], nv = (int32_t)(voff[s + 1] - voff[s]);
        for (int k = 0; k < 3; ++k) {
            const int32_t v = faces[3 * t + k];
            const bool ok = v >= 0 && v < nv;
            any |= !ok;
            faces[3 * t + k] = (ok ? v : 0) + o;
        }
    }
    if (any) atomicOr(bad, 1u);
}

__global__ void k_sbox_init(int32_t S, uint32_t* __restrict__ box_u) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 8 * S; i += gridDim.x * blockDim.x)
        box_u[i] = (i & 7) < 3 ? 0xFFFFFFFFu : 0u;
}

// Per-scene boxes. Triangles of a scene are contiguous, so a wave whose 64 triangles share one
// scene (almost every wave) reduces in registers and issues 6 atomics; mixed waves fall back
// to per-lane atomics.
__global__ void __launch_bounds__(kBlock) k_sbox(int64_t T, const int32_t* __restrict__ faces,
                                                 const float* __restrict__ verts, const int32_t* __restrict__ tri_scene,
                                                 uint32_t* __restrict__ box_u) {
    for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) & ~63ll; base < T;
         base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = base + lane_id();
        const bool ok = t < T;
        Box b = ok ? tri_box(faces, verts, (int32_t)t) : empty_box();
        const int32_t sc = ok ? tri_scene[t] : -1;
        const int32_t s0 = __shfl(sc, 0);
        const bool uniform = __ballot(ok && sc != s0) == 0;
        if (uniform) {
            float v[6] = {b.mnx, b.mny, b.mnz, b.mxx, b.mxy, b.mxz};
            for (int off = 32; off > 0; off >>= 1)
                for (int k = 0; k < 6; ++k) {
                    const float o = __shfl_xor(v[k], off);
                    v[k] = k < 3 ? fminf(v[k], o) : fmaxf(v[k], o);
                }
            if (lane_id() < 6) {
                const int k = lane_id();
                if (k < 3) atomicMin(box_u + 8 * s0 + k, ford(v[k]));
                else atomicMax(box_u + 8 * s0 + k, ford(v[k]));
            }
        } else if (ok) {
            atomicMin(box_u + 8 * sc + 0, ford(b.mnx));
            atomicMin(box_u + 8 * sc + 1, ford(b.mny));
            atomicMin(box_u + 8 * sc + 2, ford(b.mnz));
            atomicMax(box_u + 8 * sc + 3, ford(b.mxx));
            atomicMax(box_u + 8 * sc + 4, ford(b.mxy));
            atomicMax(box_u + 8 * sc + 5, ford(b.mxz));
        }
    }
}

__global__ void k_sbox_final(int32_t S, const uint32_t* __restrict__ box_u, float* __restrict__ box) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 8 * S; i += gridDim.x * blockDim.x)
        box[i] = (i & 7) < 6 ? funord(box_u[i]) : 0.f;
}

__global__ void __launch_bounds__(kBlock) k_batch_keys(int64_t T, const int32_t* __restrict__ faces,
                                                       const float* __restrict__ verts,
                                                       const int32_t* __restrict__ tri_scene,
                                                       const float* __restrict__ sbox, uint64_t* __restrict__ keys,
                                                       int32_t* __restrict__ vals) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = tri_scene[t];
        const uint32_t code = box_code(tri_box(faces, verts, (int32_t)t), sbox + 8 * s);
        keys[t] = ((uint64_t)(uint32_t)s << 30) | code;
        vals[t] = (int32_t)t;
    }
}

__global__ void k_copy_sorted(int64_t n, const uint64_t* __restrict__ k_in, const int32_t* __restrict__ v_in,
                              uint64_t* __restrict__ k_out, int32_t* __restrict__ v_out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        k_out[i] = k_in[i];
        v_out[i] = v_in[i];
    }
}

// Hits per scene with one atomic per group of equal scenes in a wave.
__device__ inline void wave_add(uint32_t* counter, int32_t key, bool pred, uint32_t* base_out) {
    uint64_t todo = __ballot(pred);
    const uint64_t below = (1ull << lane_id()) - 1ull;
    uint32_t mine = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int32_t k = __shfl(key, leader);
        const uint64_t grp = __ballot(pred && key == k);
        uint32_t b = 0;
        if (lane_id() == leader) b = atomicAdd(&counter[k], (uint32_t)__popcll(grp));
        b = __shfl(b, leader);
        if (pred && key == k) mine = b + (uint32_t)__popcll(grp & below);
        todo &= ~grp;
    }
    if (base_out) *base_out = mine;
}

__global__ void __launch_bounds__(kBlock) k_pair_scene_count(DevState d, const int32_t* __restrict__ tri_scene,
                                                             uint32_t* __restrict__ scnt) {
    __shared__ uint32_t pre[kPairSegs + 1];
    const uint32_t m = seg_prefix(d.pcnt + kPairSegs * kCntStride, (uint32_t)d.pair_seg_cap, pre);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m + 63u; k += gridDim.x * blockDim.x) {
        if ((k & ~63u) >= m) break;  // wave-uniform
        const bool ok = k < m;
        const int32_t s = ok ? tri_scene[d.np[seg_slot(pre, k, d.pair_seg_cap)].x] : 0;
        wave_add(scnt, s, ok, nullptr);
    }
}

__global__ void __launch_bounds__(kBlock) k_pair_scene_scatter(DevState d, const int32_t* __restrict__ tri_scene,
                                                               const int64_t* __restrict__ toff,
                                                               const uint32_t* __restrict__ soff,
                                                               uint32_t* __restrict__ scur, int2* __restrict__ out) {
    __shared__ uint32_t pre[kPairSegs + 1];
    const uint32_t m = seg_prefix(d.pcnt + kPairSegs * kCntStride, (uint32_t)d.pair_seg_cap, pre);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m + 63u; k += gridDim.x * blockDim.x) {
        if ((k & ~63u) >= m) break;
        const bool ok = k < m;
        const int2 p = ok ? d.np[seg_slot(pre, k, d.pair_seg_cap)] : make_int2(0, 0);
        const int32_t s = ok ? tri_scene[p.x] : 0;
        uint32_t slot = 0;
        wave_add(scur, s, ok, &slot);
        if (ok) {
            const int32_t o = (int32_t)toff[s];
            out[soff[s] + slot] = make_int2(p.x - o, p.y - o);
        }
    }
}

int64_t* toff_dev_of(bvhcd_batch* b) { return reinterpret_cast<int64_t*>(b->d.intbeg); }

// One frame's launches (stream-ordered, capturable): bvh_op 0 builds, 1 refits.
void record_frame(bvhcd_batch* b, int bvh_op, unsigned grid) {
    DevState& d = b->d;
    hipStream_t s = b->stream;
    const int32_t n = (int32_t)b->T;
    if (bvh_op == BVHCD_BVH_BUILD) {
        hipLaunchKernelGGL(k_sbox_init, dim3(grid_for(8 * b->S)), dim3(kBlock), 0, s, b->S, b->sbox_u);
        hipLaunchKernelGGL(k_sbox, dim3(std::min<unsigned>(grid_for(n), 4096)), dim3(kBlock), 0, s, (int64_t)n, d.faces,
                           d.verts, b->tri_scene, b->sbox_u);
        hipLaunchKernelGGL(k_sbox_final, dim3(grid_for(8 * b->S)), dim3(kBlock), 0, s, b->S, b->sbox_u, b->sbox);
        hipLaunchKernelGGL(k_batch_keys, dim3(std::min<unsigned>(grid_for(n), 8192)), dim3(kBlock), 0, s, (int64_t)n,
                           d.faces, d.verts, b->tri_scene, b->sbox, b->keys, d.prim_idx);
        int bits = 30;
        while (((int64_t)1 << (bits - 30)) < (int64_t)b->S) ++bits;
        if (radix_sort_u64_dev(d, b->keys, d.prim_idx, b->keys_tmp, d.vals_tmp, n, bits, s))
            hipLaunchKernelGGL(k_copy_sorted, dim3(std::min<unsigned>(grid_for(n), 8192)), dim3(kBlock), 0, s,
                               (int64_t)n, b->keys_tmp, d.vals_tmp, b->keys, d.prim_idx);
        launch_prim_map(d, n, s);
        launch_emit_full64(d, b->keys, n, s);
        launch_refit_prep(d, n, s);
    }
    launch_refit(d, n, s);
    launch_batch_cd(d, n, b->keys, b->scene_last, b->queue, grid, s);
    launch_narrow(d, s, true);  // launch_batch_cd reset the counters
    launch_pair_totals(d, s, true);
    fill_u32(b->scnt, b->S + 1, 0u, s);
    fill_u32(b->scur, b->S + 1, 0u, s);
    hipLaunchKernelGGL(k_pair_scene_count, dim3(1024), dim3(kBlock), 0, s, d, b->tri_scene, b->scnt);
    launch_exscan(d, b->scnt, b->scnt, b->S + 1, nullptr, s);
    hipLaunchKernelGGL(k_pair_scene_scatter, dim3(1024), dim3(kBlock), 0, s, d, b->tri_scene, toff_dev_of(b), b->scnt,
                       b->scur, b->grouped);
}

unsigned batch_grid(bvhcd_batch* b) { return shape_grid(tuned_shape(TK_BATCH_CD, b->T)); }

int check_batch_overflow(bvhcd_batch* b) {
    uint32_t ovf = 0;
    if (hipMemcpyAsync(&ovf, b->d.ctr + C_OVF, sizeof(ovf), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "device error");
    if (ovf & 16u) return bfail(b, BVHCD_E_DEVICE, "traversal step bound exceeded");
    if (ovf) return bfail(b, BVHCD_E_CAPACITY, "pair capacity exceeded");
    return BVHCD_OK;
}

}  // namespace

extern "C" {

int bvhcd_batch_create(const bvhcd_config* cfg, int32_t S, const int64_t* tps, const int64_t* vps, bvhcd_batch** out) {
    if (!cfg || !out || S < 1 || !tps || !vps || S > (1 << 20)) return BVHCD_E_INVALID;
    *out = nullptr;
    int64_t T = 0, V = 0;
    for (int32_t s = 0; s < S; ++s) {
        if (tps[s] < 1 || vps[s] < 1) return BVHCD_E_INVALID;
        T += tps[s];
        V += vps[s];
    }
    if (T < 2 || T > (1ll << 30) || V > (1ll << 30)) return BVHCD_E_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || cfg->device < 0 || cfg->device >= ndev)
        return BVHCD_E_DEVICE;
    if (hipSetDevice(cfg->device) != hipSuccess) return BVHCD_E_DEVICE;
    bvhcd_batch* b = new (std::nothrow) bvhcd_batch();
    if (!b) return BVHCD_E_OOM;
    b->cfg = *cfg;
    b->S = S;
    b->T = T;
    b->V = V;
    b->toff.assign(S + 1, 0);
    b->voff.assign(S + 1, 0);
    for (int32_t s = 0; s < S; ++s) {
        b->toff[s + 1] = b->toff[s] + tps[s];
        b->voff[s + 1] = b->voff[s] + vps[s];
    }
    if (cfg->stream) {
        b->stream = static_cast<hipStream_t>(cfg->stream);
    } else {
        if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return BVHCD_E_DEVICE;
        }
        b->own_stream = true;
    }
    DevState& d = b->d;
    d.maxT = T;
    d.maxV = V;
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    int64_t cp = cfg->pair_cap > 0 ? cfg->pair_cap : 24 * T + (1 << 20);
    cp = std::min<int64_t>(cp, 0x7FFFFFFFll);
    if (cfg->pair_cap <= 0) cp = std::min<int64_t>(cp, (int64_t)(0.5 * (double)free_b / 24.0));
    d.pair_seg_cap = std::max<int64_t>((std::max<int64_t>(cp, 1) + kPairSegs - 1) / kPairSegs, 1);
    d.pair_cap = d.pair_seg_cap * kPairSegs;
    const int64_t tiles = (T + 4095) / 4096;
    d.sort_hist_n = (size_t)(256 * tiles);
    d.scan_tmp_n = (size_t)((std::max<int64_t>((int64_t)d.sort_hist_n, T + 2) + 4095) / 4096 + 1);
    bool ok = true;
    ok &= balloc(b, d.faces, 3 * T);
    ok &= balloc(b, d.verts, 3 * V);
    ok &= balloc(b, d.prim_idx, T);
    ok &= balloc(b, d.vals_tmp, T);
    ok &= balloc(b, d.prim_map, T);
    ok &= balloc(b, d.inode, T);
    ok &= balloc(b, d.leaf, T);
    ok &= balloc(b, d.int_lc, T);
    ok &= balloc(b, d.int_rc, T);
    ok &= balloc(b, d.int_par, T);
    ok &= balloc(b, d.int_mark, T);
    ok &= balloc(b, d.int_rx, T);
    ok &= balloc(b, d.int_ry, T);
    ok &= balloc(b, d.itopo, T);
    ok &= balloc(b, d.tri_finite, T);
    ok &= balloc(b, d.lfaces, 3 * T);
    ok &= balloc(b, d.rp_cstart, (size_t)refit_chunks((int32_t)T) + 1);
    ok &= balloc(b, d.rp_cinfo, (size_t)refit_chunks((int32_t)T));
    ok &= balloc(b, d.rp_ent, T);
    ok &= balloc(b, d.rp_tent, T);
    ok &= balloc(b, d.rp_tlist, (size_t)T + kRefitTopLds);
    ok &= balloc(b, d.rp_tpos, T);
    ok &= balloc(b, d.rp_tlev, 1024);
    ok &= balloc(b, d.rp_tinfo, 6);
    ok &= balloc(b, d.rp_seg, kRefitTopGrid);
    ok &= balloc(b, d.rp_tsz, T);
    ok &= balloc(b, d.rp_tseg, T);
    ok &= balloc(b, d.rp_tdone, 1);
    ok &= balloc(b, d.rp_tdeep, 1);
    ok &= balloc(b, d.rp_tcount, 1);
    ok &= balloc(b, d.rp_tcb, 6 * (size_t)refit_tcb_boxes((int64_t)T));
    ok &= balloc(b, d.rp_cbad, (size_t)refit_chunks((int32_t)T));
    ok &= balloc(b, d.leaf_par, T);
    ok &= balloc(b, d.leaf_mark, T);
    ok &= balloc(b, d.leaf_lca, T + 1);
    ok &= balloc(b, d.s_a, T);
    ok &= balloc(b, d.s_b, T);
    ok &= balloc(b, d.s_par, T);
    ok &= balloc(b, d.s_lc, T);
    ok &= balloc(b, d.s_rc, T);
    ok &= balloc(b, d.s_lpar, T);
    ok &= balloc(b, d.s_top, T);
    ok &= balloc(b, d.depths, T + 1);
    ok &= balloc(b, d.offsets, T + 1);
    ok &= balloc(b, d.tkmap, T);
    ok &= balloc(b, d.sort_hist, d.sort_hist_n);
    ok &= balloc(b, d.scan_tmp, d.scan_tmp_n);
    ok &= balloc(b, d.ctr, 32);
    ok &= balloc(b, d.bp, d.pair_cap);
    ok &= balloc(b, d.np, d.pair_cap);
    ok &= balloc(b, b->grouped, d.pair_cap);
    ok &= balloc(b, d.pcnt, 2 * kPairSegs * kCntStride);
    ok &= balloc(b, d.intbeg, 2 * (S + 1));  // holds the device copy of the triangle offsets
    ok &= balloc(b, b->tri_scene, T);
    ok &= balloc(b, b->scene_last, S);
    ok &= balloc(b, b->voff_dev, S + 1);
    ok &= balloc(b, b->sbox_u, 8 * S);
    ok &= balloc(b, b->sbox, 8 * S);
    ok &= balloc(b, b->keys, T);
    ok &= balloc(b, b->keys_tmp, T);
    ok &= balloc(b, b->scnt, S + 1);
    ok &= balloc(b, b->scur, S + 1);
    ok &= balloc(b, b->queue, 4);
    if (!ok) {
        for (void* p : b->allocs) (void)hipFree(p);
        if (b->own_stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return BVHCD_E_OOM;
    }
    d.qleaf = d.leaf;  // self-CD queries: the tree's own leaves
    d.qprim = d.prim_idx;
    // static per-scene tables: sorted leaves of scene s are [toff[s], toff[s + 1])
    std::vector<int32_t> last(S);
    for (int32_t s = 0; s < S; ++s) last[s] = (int32_t)(b->toff[s + 1] - 1);
    bool okc = hipMemcpyAsync(b->scene_last, last.data(), sizeof(int32_t) * S, hipMemcpyHostToDevice, b->stream) ==
                   hipSuccess &&
               hipMemcpyAsync(toff_dev_of(b), b->toff.data(), sizeof(int64_t) * (S + 1), hipMemcpyHostToDevice,
                              b->stream) == hipSuccess &&
               hipMemcpyAsync(b->voff_dev, b->voff.data(), sizeof(int64_t) * (S + 1), hipMemcpyHostToDevice,
                              b->stream) == hipSuccess &&
               hipMemsetAsync(d.ctr, 0, 32 * sizeof(uint32_t), b->stream) == hipSuccess &&
               hipMemsetAsync(d.pcnt, 0, 2 * kPairSegs * kCntStride * sizeof(uint32_t), b->stream) == hipSuccess;
    if (okc) {
        hipLaunchKernelGGL(k_tri_scene, dim3(std::min<unsigned>(grid_for(T), 4096)), dim3(kBlock), 0, b->stream, T, S,
                           toff_dev_of(b), b->tri_scene);
        okc = hipStreamSynchronize(b->stream) == hipSuccess;
    }
    if (!okc) {
        for (void* p : b->allocs) (void)hipFree(p);
        if (b->own_stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return BVHCD_E_DEVICE;
    }
    *out = b;
    return BVHCD_OK;
}

int bvhcd_batch_destroy(bvhcd_batch* b) {
    if (!b) return BVHCD_E_INVALID;
    (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return BVHCD_OK;
}

int bvhcd_batch_set_topology(bvhcd_batch* b, const int32_t* faces, int where) {
    if (!b || !faces) return BVHCD_E_INVALID;
    if (where != BVHCD_MEM_HOST && where != BVHCD_MEM_DEVICE) return bfail(b, BVHCD_E_INVALID, "bad memory kind");
    if (where == BVHCD_MEM_HOST) {  // vertex ids must lie inside their scene
        for (int32_t s = 0; s < b->S; ++s) {
            const int64_t nv = b->voff[s + 1] - b->voff[s];
            for (int64_t k = 3 * b->toff[s]; k < 3 * b->toff[s + 1]; ++k)
                if (faces[k] < 0 || faces[k] >= nv) return bfail(b, BVHCD_E_INVALID, "face vertex id out of range");
        }
    }
    const hipMemcpyKind kind = where == BVHCD_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (hipMemcpyAsync(b->d.faces, faces, sizeof(int32_t) * 3 * b->T, kind, b->stream) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "faces copy failed");
    uint32_t bad = 0;
    b->topo = false;
    if (hipMemsetAsync(b->d.ctr + C_BADFACE, 0, sizeof(uint32_t), b->stream) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "set_topology failed");
    hipLaunchKernelGGL(k_globalize, dim3(std::min<unsigned>(grid_for(b->T), 4096)), dim3(kBlock), 0, b->stream, b->T,
                       b->tri_scene, b->voff_dev, b->d.faces, b->d.ctr + C_BADFACE);
    if (hipMemcpyAsync(&bad, b->d.ctr + C_BADFACE, sizeof(uint32_t), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "set_topology failed");
    if (bad) return bfail(b, BVHCD_E_INVALID, "face vertex id out of range");
    b->topo = true;
    b->built = false;
    drop_graphs(b);
    return BVHCD_OK;
}

int bvhcd_batch_update_vertices(bvhcd_batch* b, const float* xyz, int where) {
    if (!b || !xyz) return BVHCD_E_INVALID;
    if (where != BVHCD_MEM_HOST && where != BVHCD_MEM_DEVICE) return bfail(b, BVHCD_E_INVALID, "bad memory kind");
    const hipMemcpyKind kind = where == BVHCD_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (hipMemcpyAsync(b->d.verts, xyz, sizeof(float) * 3 * b->V, kind, b->stream) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "vertex copy failed");
    // host memory (pinned or not) may be reused by the caller as soon as this returns
    if (where != BVHCD_MEM_DEVICE && hipStreamSynchronize(b->stream) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "vertex copy failed");
    return BVHCD_OK;
}

int bvhcd_batch_step(bvhcd_batch* b, int bvh_op, int use_graph) {
    if (!b) return BVHCD_E_INVALID;
    if (!b->topo) return bfail(b, BVHCD_E_INVALID, "set_topology first");
    if (bvh_op != BVHCD_BVH_BUILD && bvh_op != BVHCD_BVH_REFIT) return bfail(b, BVHCD_E_INVALID, "bvh_op");
    if (bvh_op == BVHCD_BVH_REFIT && !b->built) return bfail(b, BVHCD_E_NOT_BUILT, "first step must BUILD");
    const int g = bvh_op == BVHCD_BVH_BUILD ? 0 : 1;
    if (!use_graph) {
        record_frame(b, bvh_op, batch_grid(b));
    } else {
        if (!b->exec[g]) {
            if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return bfail(b, BVHCD_E_DEVICE, "stream capture failed");
            record_frame(b, bvh_op, batch_grid(b));
            if (hipStreamEndCapture(b->stream, &b->graph[g]) != hipSuccess ||
                hipGraphInstantiate(&b->exec[g], b->graph[g], nullptr, nullptr, 0) != hipSuccess) {
                drop_graphs(b);
                return bfail(b, BVHCD_E_DEVICE, "graph capture failed");
            }
        }
        if (hipGraphLaunch(b->exec[g], b->stream) != hipSuccess) return bfail(b, BVHCD_E_DEVICE, "graph launch failed");
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, BVHCD_E_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    if (bvh_op == BVHCD_BVH_BUILD) b->built = true;
    return BVHCD_OK;
}

int bvhcd_batch_pair_counts(bvhcd_batch* b, int64_t* counts) {
    if (!b || !counts) return BVHCD_E_INVALID;
    if (!b->built) return bfail(b, BVHCD_E_NOT_BUILT, "no step yet");
    int st = check_batch_overflow(b);
    if (st) return st;
    std::vector<uint32_t> off(b->S + 1);
    if (hipMemcpyAsync(off.data(), b->scnt, sizeof(uint32_t) * (b->S + 1), hipMemcpyDeviceToHost, b->stream) !=
            hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "device error");
    uint32_t total = 0;
    if (hipMemcpy(&total, b->d.ctr + C_NP, sizeof(total), hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "device error");
    for (int32_t s = 0; s < b->S; ++s) counts[s] = (int64_t)((s + 1 < b->S ? off[s + 1] : total) - off[s]);
    return BVHCD_OK;
}

int bvhcd_batch_get_pairs(bvhcd_batch* b, int32_t scene, int32_t* out, int64_t cap, int64_t* n) {
    if (!b || !n || scene < 0 || scene >= b->S) return BVHCD_E_INVALID;
    std::vector<int64_t> counts(b->S);
    int st = bvhcd_batch_pair_counts(b, counts.data());
    if (st) return st;
    *n = counts[scene];
    if (!out) return BVHCD_OK;
    if (cap < *n) return bfail(b, BVHCD_E_CAPACITY, "output buffer too small");
    uint32_t off = 0;
    if (hipMemcpy(&off, b->scnt + scene, sizeof(off), hipMemcpyDeviceToHost) != hipSuccess ||
        (*n > 0 && hipMemcpy(out, b->grouped + off, sizeof(int2) * (size_t)*n, hipMemcpyDeviceToHost) != hipSuccess))
        return bfail(b, BVHCD_E_DEVICE, "pair copy failed");
    return BVHCD_OK;
}

const char* bvhcd_batch_last_error(const bvhcd_batch* b) { return b ? b->err.c_str() : "null batch"; }

// Autotune of the persistent traversal's grid (workgroups resident on the device): every
// candidate replays the refit frame a few times on the batch's current data, timed with
// events; the fastest is kept for this size class (process-wide table) and the frame graphs are
// re-captured on their next use.
int bvhcd_batch_autotune(bvhcd_batch* b) {
    if (!b) return BVHCD_E_INVALID;
    if (!b->built) return bfail(b, BVHCD_E_NOT_BUILT, "autotune needs a built batch");
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return bfail(b, BVHCD_E_DEVICE, "event create failed");
    float best_ms = 1e30f;
    unsigned best = batch_grid(b);
    for (unsigned g : {256u, 512u, 1024u, 2048u, 4096u, 8192u}) {
        launch_batch_cd(b->d, (int32_t)b->T, b->keys, b->scene_last, b->queue, g, b->stream);  // warm-up
        (void)hipEventRecord(e0, b->stream);
        for (int r = 0; r < 3; ++r) launch_batch_cd(b->d, (int32_t)b->T, b->keys, b->scene_last, b->queue, g, b->stream);
        (void)hipEventRecord(e1, b->stream);
        if (hipEventSynchronize(e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return bfail(b, BVHCD_E_DEVICE, "autotune run failed");
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        record_tuning(TK_BATCH_CD, b->T, g, ms / 3.0f);
        if (ms < best_ms) {
            best_ms = ms;
            best = g;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    set_tuned_shape(TK_BATCH_CD, b->T, best);
    drop_graphs(b);
    // leave the batch with a complete frame result (pairs of the current vertices)
    record_frame(b, BVHCD_BVH_REFIT, best);
    return hipStreamSynchronize(b->stream) == hipSuccess ? BVHCD_OK : bfail(b, BVHCD_E_DEVICE, "device error");
}

}  // extern "C"
