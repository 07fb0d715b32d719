// Internal declarations of the gfx950 collision-detection pipeline (HIP + C++ host).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bvhcd.h"

namespace bvhcd {

constexpr int kPairSegs = 64;   // pair-list segments (one append counter each)
constexpr int kCntStride = 32;  // words between two segment counters (128 bytes)
constexpr int kDroppedWord = 2; // word of a list's counter block: entries no segment had room for

// ------------------------------------------------------------------ node records (HBM layout)
// One 32-byte record per node, read with two 16-byte loads by the traversal kernels:
//   internal node n (preorder): lo = (minx, miny, minz, rangex), hi = (maxx, maxy, maxz, escape)
//   leaf j (sorted slot):       lo = (minx, miny, minz, parent), hi = (maxx, maxy, maxz, escape)
// escape = encoded lca of the first leaf after the node's range (lca[rangey + 1], -1 at the end):
// (internal id << 1) or (leaf << 1 | 1). The w words hold int bits.
struct alignas(16) NodeRec {
    float4 lo, hi;
};

__host__ __device__ inline int32_t f2i(float f) { return __builtin_bit_cast(int32_t, f); }
__host__ __device__ inline float i2f(int32_t i) { return __builtin_bit_cast(float, i); }

struct Box {
    float mnx, mny, mnz, mxx, mxy, mxz;
};

__device__ inline Box empty_box() { return {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX}; }
__device__ inline void box_add(Box& b, float x, float y, float z) {
    b.mnx = fminf(b.mnx, x); b.mny = fminf(b.mny, y); b.mnz = fminf(b.mnz, z);
    b.mxx = fmaxf(b.mxx, x); b.mxy = fmaxf(b.mxy, y); b.mxz = fmaxf(b.mxz, z);
}
__device__ inline Box box_merge(const Box& a, const Box& b) {
    return {fminf(a.mnx, b.mnx), fminf(a.mny, b.mny), fminf(a.mnz, b.mnz),
            fmaxf(a.mxx, b.mxx), fmaxf(a.mxy, b.mxy), fmaxf(a.mxz, b.mxz)};
}
// Closed-interval overlap; false as soon as one axis separates (base/Bv.h:40-45).
__device__ inline bool box_overlap(const Box& a, float4 lo, float4 hi) {
    return !(lo.x > a.mxx || hi.x < a.mnx || lo.y > a.mxy || hi.y < a.mny || lo.z > a.mxz || hi.z < a.mnz);
}
__device__ inline bool box_overlap(const Box& a, const Box& b) {
    return !(b.mnx > a.mxx || b.mxx < a.mnx || b.mny > a.mxy || b.mxy < a.mny || b.mnz > a.mxz || b.mxz < a.mnz);
}
__device__ inline Box rec_box(const float4& lo, const float4& hi) { return {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z}; }

__device__ inline Box tri_box(const int32_t* __restrict__ faces, const float* __restrict__ verts, int32_t t) {
    const int32_t a = faces[3 * (int64_t)t], b = faces[3 * (int64_t)t + 1], c = faces[3 * (int64_t)t + 2];
    Box bx = empty_box();
    box_add(bx, verts[3 * (int64_t)a], verts[3 * (int64_t)a + 1], verts[3 * (int64_t)a + 2]);
    box_add(bx, verts[3 * (int64_t)b], verts[3 * (int64_t)b + 1], verts[3 * (int64_t)b + 2]);
    box_add(bx, verts[3 * (int64_t)c], verts[3 * (int64_t)c + 1], verts[3 * (int64_t)c + 2]);
    return bx;
}

// ------------------------------------------------------------------ Morton (30-bit)
__host__ __device__ inline uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
__device__ inline uint32_t morton3d(float x, float y, float z) {
    x = fminf(fmaxf(x * 1024.0f, 0.0f), 1023.0f);
    y = fminf(fmaxf(y * 1024.0f, 0.0f), 1023.0f);
    z = fminf(fmaxf(z * 1024.0f, 0.0f), 1023.0f);
    return expand_bits((uint32_t)x) * 4u + expand_bits((uint32_t)y) * 2u + expand_bits((uint32_t)z);
}
// Centre = float sum halved exactly; offset / extent with IEEE division (base/Bv.h:65-67).
__device__ inline uint32_t box_code(const Box& b, const float* __restrict__ scene) {
    const float cx = (float)((double)(b.mnx + b.mxx) * 0.5);
    const float cy = (float)((double)(b.mny + b.mxy) * 0.5);
    const float cz = (float)((double)(b.mnz + b.mxz) * 0.5);
    const float qx = (cx - scene[0]) / (scene[3] - scene[0]);
    const float qy = (cy - scene[1]) / (scene[4] - scene[1]);
    const float qz = (cz - scene[2]) / (scene[5] - scene[2]);
    return morton3d(qx, qy, qz);
}

// ------------------------------------------------------------------ wave64 helpers
__device__ inline int lane_id() { return __lane_id(); }
// Wave-aggregated append: returns this lane's slot; one atomic per wave.
__device__ inline uint32_t wave_append(uint32_t* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (!pred) return 0;
    const int lane = lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}

// Segmented pair lists read by a virtual index: seg_prefix builds, per workgroup, the exclusive
// prefix of the kPairSegs segment counts (clamped to the segment capacity) in LDS `pre`
// (kPairSegs + 1 words; all threads call it, it holds a barrier) and returns the total;
// seg_slot maps virtual index k < total to its array slot.
static_assert(kPairSegs == 64, "one wave scans the segment counts");
__device__ inline uint32_t seg_prefix(const uint32_t* __restrict__ cnt, uint32_t seg_cap, uint32_t* pre) {
    if (threadIdx.x < 64) {
        const uint32_t v = min(cnt[threadIdx.x * kCntStride], seg_cap);
        uint32_t x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if ((int)threadIdx.x >= off) x += y;
        }
        pre[threadIdx.x + 1] = x;
        if (threadIdx.x == 0) pre[0] = 0;
    }
    __syncthreads();
    return pre[kPairSegs];
}
__device__ inline int64_t seg_slot(const uint32_t* pre, uint32_t k, int64_t seg_cap) {
    int lo = 0, hi = kPairSegs - 1;  // last segment g with pre[g] <= k
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (pre[m] <= k) lo = m;
        else hi = m - 1;
    }
    return (int64_t)lo * seg_cap + (int64_t)(k - pre[lo]);
}

// ------------------------------------------------------------------ device buffers
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
};

// Per-context device state. All arrays are owned by the context.
struct DevState {
    int64_t maxT = 0, maxV = 0;
    // inputs
    int32_t* faces = nullptr;   // 3T
    float* verts = nullptr;     // 3V
    // BVH
    uint32_t* codes = nullptr;      // T sorted codes
    uint32_t* keys_tmp = nullptr;   // T
    uint32_t* keys_tmp2 = nullptr;  // T
    int32_t* prim_idx = nullptr;    // T slot -> triangle
    int32_t* vals_tmp = nullptr;    // T
    int32_t* prim_map = nullptr;    // T triangle -> slot
    NodeRec* inode = nullptr;       // T-1 (preorder)
    NodeRec* leaf = nullptr;        // T
    int32_t* int_lc = nullptr;      // T-1
    int32_t* int_rc = nullptr;
    int32_t* int_par = nullptr;
    uint32_t* int_mark = nullptr;
    int32_t* int_rx = nullptr;
    int32_t* int_ry = nullptr;
    int4* itopo = nullptr;          // T-1 packed (lc, rc, par, mark)
    // refit schedule, rebuilt after every topology change (treelet chunks + top tree)
    int32_t* rp_cstart = nullptr;   // chunks + 1: first leaf of each chunk, then T
    int2* rp_cinfo = nullptr;       // chunks: (internal nodes, deepest level | treelet roots << 16)
    int2* rp_ent = nullptr;         // T: per chunk (node, slots | depth), deepest first, then the
                                    //    roots' hand-over offsets
    int4* rp_tent = nullptr;        // T: top-tree (node, child, child, depth), unordered
    int4* rp_tlist = nullptr;       // T + 4096: the same, deepest first (or by segment), top
                                    //    children as positions, hand-over offsets
    int32_t* rp_tpos = nullptr;     // T-1: list position of each top node
    int32_t* rp_tlev = nullptr;     // 1024: first entry of each top-tree depth
    int32_t* rp_tinfo = nullptr;    // 6: (top nodes, deepest level or -1, segments or -1 (all
                                    //     through the records) or -2 (climb), first upper
                                    //     position, upper nodes, deepest upper level)
    int4* rp_seg = nullptr;         // 64: per segment (first position, nodes, deepest level,
                                    //     (last level merged per barrier + 1) | one-wave start << 16)
    int32_t* rp_tsz = nullptr;      // T: top-tree prep scratch (subtree sizes, new positions),
                                    //    climb flags per node when the top tree is too deep
    int32_t* rp_tseg = nullptr;     // T: top-tree prep scratch (segment per position)
    uint32_t* rp_tdone = nullptr;   // 1: segment workgroups finished (reset by the last one)
    uint32_t* rp_tdeep = nullptr;   // 1: the top tree is deeper than the level bins
    uint32_t* rp_tcount = nullptr;  // 1: append counter
    uint32_t* rp_cbad = nullptr;    // chunks: some leaf of the chunk has a non-finite coordinate
    float* rp_tcb = nullptr;        // refit_tcb_boxes(T) boxes: treelet-root children of segment
                                    // nodes, dense in list order, written by the chunk refit
    float* sa_ref = nullptr;        // T-1 internal-node surface area at its last rebuild
    float* sa_ref_leaf = nullptr;   // T leaf-box surface area at its last rebuild
    int32_t* built_frame = nullptr; // T frame at which each leaf's subtree was last rebuilt
    uint8_t* tri_finite = nullptr;  // T (by leaf): all 9 |coordinates| <= 1e8 (set by refit)
    int32_t* lfaces = nullptr;      // 3T: the faces in leaf (sorted) order, rebuilt with prim_idx
    const float* pbox = nullptr;    // box mode: 6T per-triangle boxes (min xyz, max xyz) given by the
                                    // caller, used instead of the vertices; null otherwise
    int skip_adj = 0;               // never report pairs of triangles sharing a vertex
    int32_t* leaf_par = nullptr;    // T
    uint32_t* leaf_mark = nullptr;  // T
    int32_t* leaf_lca = nullptr;    // T+1
    // emission scratch (split-id space)
    int32_t* s_a = nullptr;         // T range start
    int32_t* s_b = nullptr;         // T range end
    int32_t* s_par = nullptr;       // T parent split id (-1 root); leaves use s_lpar
    int32_t* s_lc = nullptr;        // T child (id << 1 | is_leaf)
    int32_t* s_rc = nullptr;
    int32_t* s_lpar = nullptr;      // T leaf parent split id (sign bit clear), right-child flag in bit 30
    int32_t* s_top = nullptr;       // T topmost split node with leftmost leaf x (-1 none)
    uint32_t* depths = nullptr;     // T
    uint32_t* offsets = nullptr;    // T+1
    int32_t* tkmap = nullptr;       // T split id -> preorder id
    // scan / sort scratch
    uint32_t* scan_tmp = nullptr;   // block sums
    size_t scan_tmp_n = 0;
    uint32_t* sort_hist = nullptr;  // digit histograms
    size_t sort_hist_n = 0;
    float* scene_box = nullptr;     // 6 floats (+ pad)
    uint32_t* scene_box_u = nullptr;// ordered-int reduction scratch, 6 words
    // fronts: SoA (node id, query leaf), double buffered per kind (0 internal, 1 external)
    uint32_t* fnode[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [kind][buf]
    int32_t* fquery[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int64_t front_cap[2] = {0, 0};  // entries per buffer, per kind
    int fcur[2] = {0, 0};           // current buffer per kind
    int64_t front_version = 0;      // host-side: bumped whenever a step changes the front's contents
    int64_t adj_version = -1;       // host-side: front version of the last shared-vertex partition
    uint32_t* adj_cnt = nullptr;    // partition scratch: shared-vertex entries per tile, then offsets
    uint32_t* adj_off = nullptr;
    uint32_t* fsize = nullptr;      // [kind*2 + buf] sizes (4 words)
    uint32_t* fcnt = nullptr;       // [kind][kPairSegs * kCntStride] segment counters of a front
                                    // being produced (front_cap is kPairSegs equal segments)
    // Wave-ordered production (self-CD GENERATE without node ordering): every staging flush of a
    // front list also records one chunk (offset in its segment, producing wave, the wave's
    // entries before it, count) at [segment][chunk seq], the seq being the high word of the
    // segment's 64-bit counter; each wave leaves its total in wave_tot. A scan of the totals
    // then places every chunk: the front comes out in wave order with no sort.
    uint4* fchunk[2] = {nullptr, nullptr};      // [kind][kPairSegs * chunk_seg_cap]
    int64_t chunk_seg_cap = 0;
    uint32_t* wave_tot[2] = {nullptr, nullptr};  // [kind][wave_cap]
    uint32_t* wave_base = nullptr;              // [wave_cap] exclusive scan of one kind's totals
    uint32_t* wave_nch = nullptr;               // [wave_cap] chunk records per wave (list cursors)
    uint32_t* wave_cbase = nullptr;             // [wave_cap] exclusive scan of wave_nch
    uint4* clist = nullptr;                     // [kPairSegs * chunk_seg_cap] chunks grouped by wave:
                                                // (offset in the segments: low, high word; c.z; count)
    int64_t wave_cap = 0;
    uint32_t* intcnt = nullptr;     // T
    uint32_t* extcnt = nullptr;     // T+1
    uint32_t* intbeg = nullptr;     // T+1
    uint32_t* extbeg = nullptr;     // T+2
    uint32_t* intback = nullptr;    // T
    uint32_t* cursor = nullptr;     // T (REORDER front cursor)
    float* snapshot = nullptr;      // T
    int32_t* intrange = nullptr;    // T+1
    int32_t* rtroot = nullptr;      // T
    int32_t* intmark = nullptr;     // T
    // restructure scratch
    int32_t* rt_of_leaf = nullptr;  // T+1
    int32_t* r_leaf_lo = nullptr;   // T
    int32_t* r_leaf_hi = nullptr;
    int32_t* r_split_lo = nullptr;
    int32_t* r_split_hi = nullptr;
    int32_t* r_base = nullptr;
    int32_t* r_keep_par = nullptr;
    uint32_t* r_keep_mark = nullptr;
    uint64_t* r_keys = nullptr;     // T
    uint64_t* r_keys2 = nullptr;    // T
    // incremental front restructure: (query, rebuilt subtree root) walks
    int2* rtasks = nullptr;
    int64_t rtask_cap = 0;
    // pairs
    // Pair lists are split into kPairSegs segments with one counter each (128 bytes apart), a
    // workgroup appending to segment blockIdx % kPairSegs: one shared counter serialises the
    // appends of the whole device. Segment g holds [g * pair_seg_cap, g * pair_seg_cap + count).
    int2* bp = nullptr;             // broad-phase pairs (leaf ids, query leaf first)
    int2* np = nullptr;             // narrow-phase hits (triangle ids, lower first)
    int2* pair_tmp = nullptr;       // gathered contiguous copy for read-out
    int64_t pair_cap = 0;           // kPairSegs * pair_seg_cap
    int64_t pair_seg_cap = 0;
    uint32_t* pcnt = nullptr;       // [2][kPairSegs * kCntStride]: bp, np segment counts
    // counters block: see Counter enum
    uint32_t* ctr = nullptr;        // 32 words
    // Queries of the front kernels: the tree's own leaves for self-CD (qleaf = leaf, qprim =
    // prim_idx, each query walking only the leaves after it). An inter-object view (inter = 1)
    // points them at another object's sorted leaves instead, its nq queries each walking the
    // whole tree; qlfaces / qverts are that object's leaf-ordered faces and vertices.
    const NodeRec* qleaf = nullptr;
    const int32_t* qprim = nullptr;
    const int32_t* qlfaces = nullptr;
    const float* qverts = nullptr;
    int32_t nq = 0;
    int inter = 0;
    int64_t topo_version = 0;       // host-side: bumped by every topology change (refit preparation)
    int front_by_query = 0;         // host-side: the last produced fronts are ordered by query leaf
};

enum Counter : int {
    C_BP = 0,          // broad-phase pair count (listed + implicit)
    C_NP = 1,          // narrow-phase hit count (listed + implicit)
    C_OVF = 2,         // overflow bits: 1 int front, 2 ext front, 4 bp pairs, 8 np pairs
    C_IMPLICIT = 3,    // broad-phase pairs of the last step held implicitly: the head of the
                       // external front (shared-vertex entries, every coordinate finite)
    C_IMPL_NP = 4,     // of those, counted as narrow-phase hits (0 without a narrow phase)
    C_RT_SUB = 5,      // restructure: subtree count
    C_RT_INT = 6,      // restructure: internal node count
    C_RT_LEAF = 7,     // restructure: leaf count
    C_TASKS = 8,       // incremental front restructure: (query, subtree) walks
    C_KEPT_I = 9,      // incremental front restructure: kept internal / external entries
    C_KEPT_E = 10,
    C_BP_SEGMAX = 11,  // largest broad-phase segment count requested (capacity reporting)
    C_NP_SEGMAX = 12,
    C_FREQ_I = 13,     // capacity the last produced internal / external front needed
    C_FREQ_E = 14,
    // words 16..23 hold 64-bit counters (see Counter64)
    C_ADJ = 24,        // shared-vertex entries heading the current external front (last partition)
    C_NO_IMPLICIT = 25,// the last refit cannot vouch for shared-vertex pairs: a vertex coordinate
                       // above 1e8 in magnitude (or NaN / inf), or boxes given by the caller
    C_BADFACE = 26,    // set_topology: a face vertex id outside its vertex range
    C_WAVE_TOTAL = 27, // wave-ordered production: entries of the list being gathered
    C_WAVE_CHUNKS = 28,// wave-ordered production: chunk records of the list being gathered
    C_NUM = 32
};
enum Counter64 : int {
    C64_VISITS = 0,    // node tests of the last step
    C64_DIGEST = 1,    // pair-set digest
    C64_NUM = 4
};
__host__ __device__ inline uint64_t* ctr64(uint32_t* ctr) { return reinterpret_cast<uint64_t*>(ctr + 16); }
// Zeroes a step's output counters (pair-list segment counters, pair totals, overflow flags,
// implicit pairs, node tests); lane g of one wave of 64 threads.
__device__ inline void zero_step_counters(const DevState& d, int g) {
    d.pcnt[g * kCntStride] = 0u;                // broad-phase segment counters
    d.pcnt[(kPairSegs + g) * kCntStride] = 0u;  // narrow-phase segment counters
    if (g == 0) {
        d.pcnt[kDroppedWord] = 0u;
        d.pcnt[kPairSegs * kCntStride + kDroppedWord] = 0u;
    }
    if (g <= C_IMPL_NP) d.ctr[g] = 0u;           // C_BP, C_NP, C_OVF, C_IMPLICIT, C_IMPL_NP
    if (g == 8) ctr64(d.ctr)[C64_VISITS] = 0ull;
}

__device__ inline Box given_box(const float* __restrict__ b, int64_t t) {
    const float* p = b + 6 * t;
    return {p[0], p[1], p[2], p[3], p[4], p[5]};
}
// Box of triangle t: the caller's (box mode) or the vertices'.
__device__ inline Box prim_box(const DevState& d, int32_t t) {
    return d.pbox ? given_box(d.pbox, t) : tri_box(d.faces, d.verts, t);
}
// Do leaves a and b (faces in leaf order) share a vertex (the reference's covertex,
// utility/CudaDeviceUtils.cu:153-157).
__device__ inline bool faces_share_vertex(const int32_t* __restrict__ lfaces, int32_t a, int32_t b) {
    const int32_t* fa = lfaces + 3 * (int64_t)a;
    const int32_t* fb = lfaces + 3 * (int64_t)b;
    const int32_t a0 = fa[0], a1 = fa[1], a2 = fa[2], b0 = fb[0], b1 = fb[1], b2 = fb[2];
    return a0 == b0 || a0 == b1 || a0 == b2 || a1 == b0 || a1 == b1 || a1 == b2 || a2 == b0 || a2 == b1 ||
           a2 == b2;
}
__device__ inline bool leaves_share_vertex(const DevState& d, int32_t a, int32_t b) {
    return faces_share_vertex(d.lfaces, a, b);
}

// Launch helpers
constexpr int kBlock = 256;
constexpr uint32_t kAdjTile = 16 * kBlock;  // external-front entries per shared-vertex partition block
inline unsigned grid_for(int64_t n, int block = kBlock) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2. xcd_loop gives
// workgroup b the XCD b % 8 and the rank b / 8 inside it, so that every XCD sweeps one
// contiguous eighth of [0, total) with a grid-stride loop over its own workgroups: on a
// spatially ordered list, an XCD's L2 then holds the nodes of one region instead of every L2
// fetching every node. Grids that are not a multiple of 8 fall back to the plain loop.
constexpr uint32_t kXcds = 8;
struct XcdLoop {
    uint32_t begin, end, stride;
};
__device__ inline XcdLoop xcd_loop(uint32_t total) {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    if (g % kXcds != 0) return {b * blockDim.x + threadIdx.x, total, g * blockDim.x};
    const uint32_t per = total / kXcds + 1;
    const uint32_t x = b % kXcds, lo = min(x * per, total), hi = min(lo + per, total);
    return {lo + (b / kXcds) * blockDim.x + threadIdx.x, hi, (g / kXcds) * blockDim.x};
}

}  // namespace bvhcd
