// BVTT-front kernels for gfx950: self-CD traversal, front generate / UPDATE / KEEP / REORDER,
// front ordering, quality snapshot and check, narrow phase.
//
// Reference behaviour (Source/Project/Mine/ of littlemine/BVH-based-Collision-Detection-Scheme):
//   pure traversal       collision/bvtt_front/BvhCDKernels.cu:17-49
//   generate             collision/bvtt_front/BvttFrontGenerateKernels.cu:51-90
//   UPDATE prune/sprout  collision/bvtt_front/BvttFrontPruneSproutKernels.cu:12-93
//   KEEP                 collision/bvtt_front/BvttFrontKeepKernels.cu:11-58
//   REORDER (balance)    collision/bvtt_front/BvttFrontReorderKernels.cu:27-63, BvttFrontLooseIntra.cu:285-312
//   ordering             collision/bvtt_front/BvttFrontReorderKernels.cu:12-24, collision/auxiliary/FlOrderLog.cu:7-12
//   snapshot / quality   collision/bvtt_front/BvttFrontGenerateKernels.cu:26-49
//   restructure marks    collision/lbvh/LBvhFixedDeformable.cu:216-241, LBvhRestructureKernels.cu:9-35
//   narrow phase         collision/narrow_phase/narrow_phase.cu:8-24 (17-axis SAT written here from the
//                        separating-axis theorem; canonical argument order: lower triangle id first)
//
// Every node visit is one 32-byte record (two 16-byte loads): box + rangex/escape (internal) or
// box + parent/escape (leaf). Appends to fronts and pair lists are wave-aggregated (one atomic
// per wave64) and bounds-checked against capacity, with an overflow bit instead of a write.
// Front kernels fetch work dynamically per wave so that uneven traversals balance.
#include "bvhcd_internal.hpp"
#include "kernels.hpp"

namespace bvhcd {

struct FrontIO {
    const int2* fi;       // current internal front
    const int2* fe;       // current external front
    const uint32_t* fsz;  // current sizes [int, ext]
    int2* oi;             // next internal front
    int2* oe;             // next external front
    uint32_t* osz;        // next sizes [int, ext]
    uint32_t cap_i, cap_e;
};

__device__ inline void push_pair(const DevState& d, int32_t qi, int32_t j) {
    const int32_t a = d.prim_idx[qi], b = d.prim_idx[j];
    const uint32_t slot = wave_append(&d.ctr[C_BP], true);
    if (slot < (uint32_t)d.pair_cap) d.bp[slot] = a < b ? make_int2(a, b) : make_int2(b, a);
    else atomicOr(&d.ctr[C_OVF], 4u);
}
__device__ inline void push_int(const DevState& d, const FrontIO& io, int32_t qi, int32_t node) {
    const uint32_t slot = wave_append(&io.osz[0], true);
    if (slot < io.cap_i) io.oi[slot] = make_int2(qi, node);
    else atomicOr(&d.ctr[C_OVF], 1u);
    atomicAdd(&d.intcnt[node], 1u);
}
__device__ inline void push_ext(const DevState& d, const FrontIO& io, int32_t qi, int32_t leaf) {
    const uint32_t slot = wave_append(&io.osz[1], true);
    if (slot < io.cap_e) io.oe[slot] = make_int2(qi, leaf);
    else atomicOr(&d.ctr[C_OVF], 2u);
    atomicAdd(&d.extcnt[leaf], 1u);
}

__device__ inline bool node_overlap(const DevState& d, int32_t nd, const Box& q) {
    const float4 lo = d.inode[nd].lo, hi = d.inode[nd].hi;
    return box_overlap(q, lo, hi);
}

// Stackless DFS from encoded node `st` until `stop` (lca-escape encoding). With FRONT, the
// non-overlapping internal nodes and all reached leaves are appended to the next front.
template <bool FRONT>
__device__ uint32_t stackless(const DevState& d, const FrontIO& io, int32_t qi, const Box& q, int32_t st,
                              int32_t stop) {
    uint32_t visits = 0;
    // A correct traversal visits every node at most once; the bound only guards against a
    // corrupted tree (flag 16 reported as a device error instead of a hang).
    const int64_t limit = 2 * d.maxT + 16;
    int64_t steps = 0;
    while (st != stop) {
        if (++steps > limit) {
            atomicOr(&d.ctr[C_OVF], 16u);
            break;
        }
        int32_t j;
        if (!(st & 1)) {
            int32_t nd = st >> 1;
            float4 lo = d.inode[nd].lo, hi = d.inode[nd].hi;
            j = f2i(lo.w);                  // leftmost leaf of the whole left spine
            const int32_t t = d.leaf_par[j];  // bottom of the spine
            bool reached = false;
            for (;;) {
                ++visits;
                if (!box_overlap(q, lo, hi)) break;
                if (nd == t) { reached = true; break; }
                ++nd;
                lo = d.inode[nd].lo;
                hi = d.inode[nd].hi;
            }
            if (!reached) {
                if (FRONT) push_int(d, io, qi, nd);
                st = f2i(hi.w);
                continue;
            }
        } else {
            j = st >> 1;
        }
        const float4 llo = d.leaf[j].lo, lhi = d.leaf[j].hi;
        ++visits;
        if (box_overlap(q, llo, lhi)) push_pair(d, qi, j);
        if (FRONT) push_ext(d, io, qi, j);
        st = f2i(lhi.w);
    }
    return visits;
}

__device__ inline void add_visits(const DevState& d, uint32_t v) {
    // wave reduction, one 64-bit atomic per wave
    uint64_t x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (lane_id() == 0 && x) atomicAdd((unsigned long long*)&ctr64(d.ctr)[C64_VISITS], (unsigned long long)x);
}

// ---------------------------------------------------------------- pure traversal / generate
template <bool FRONT>
__global__ void __launch_bounds__(kBlock) k_traverse(int32_t n, DevState d, FrontIO io) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t v = 0;
    if (i < n - 1) {
        const float4 lo = d.leaf[i].lo, hi = d.leaf[i].hi;
        v = stackless<FRONT>(d, io, i, rec_box(lo, hi), f2i(hi.w), -1);
    }
    add_visits(d, v);
}

static FrontIO make_io(DevState& d) {
    FrontIO io;
    const int c = d.cur, nx = d.cur ^ 1;
    io.fi = d.front[c][0];
    io.fe = d.front[c][1];
    io.fsz = d.fsize + 2 * c;
    io.oi = d.front[nx][0];
    io.oe = d.front[nx][1];
    io.osz = d.fsize + 2 * nx;
    io.cap_i = (uint32_t)std::min<int64_t>(d.front_cap[0], 0xFFFFFFFFll);
    io.cap_e = (uint32_t)std::min<int64_t>(d.front_cap[1], 0xFFFFFFFFll);
    return io;
}

static void reset_step_counters(DevState& d, hipStream_t s) {
    // pair counts, overflow, visits
    hipMemsetAsync(d.ctr, 0, sizeof(uint32_t) * 3, s);
    hipMemsetAsync(ctr64(d.ctr) + C64_VISITS, 0, sizeof(uint64_t), s);
}

void launch_pure_cd(DevState& d, int32_t n, hipStream_t s) {
    reset_step_counters(d, s);
    FrontIO io = make_io(d);
    if (n >= 2) hipLaunchKernelGGL(k_traverse<false>, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, d, io);
}

// ---------------------------------------------------------------- ordering (counting sort by node)
__global__ void __launch_bounds__(kBlock) k_order_scatter(DevState d, FrontIO io) {
    const uint32_t ni = io.fsz[0], ne = io.fsz[1];
    const uint32_t ci = min(ni, io.cap_i), ce = min(ne, io.cap_e);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ci + ce; k += gridDim.x * blockDim.x) {
        if (k < ci) {
            const int2 e = io.fi[k];
            const uint32_t pos = d.intbeg[e.y] + atomicAdd(&d.intcnt[e.y], 1u);
            if (pos < io.cap_i) io.oi[pos] = e;
        } else {
            const int2 e = io.fe[k - ci];
            const uint32_t pos = d.extbeg[e.y] + atomicAdd(&d.extcnt[e.y], 1u);
            if (pos < io.cap_e) io.oe[pos] = e;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        io.osz[0] = ni;
        io.osz[1] = ne;
    }
}

static unsigned persistent_grid() { return 2048; }

// prepare (exclusive scans of the per-node counts), clear, scatter into the other buffer, flip.
void launch_order_fronts(DevState& d, int32_t n, hipStream_t s) {
    launch_exscan(d, d.intcnt, d.intbeg, n, nullptr, s);
    launch_exscan(d, d.extcnt, d.extbeg, (int64_t)n + 1, nullptr, s);
    hipMemsetAsync(d.intcnt, 0, sizeof(uint32_t) * n, s);
    hipMemsetAsync(d.extcnt, 0, sizeof(uint32_t) * ((size_t)n + 1), s);
    FrontIO io = make_io(d);
    hipLaunchKernelGGL(k_order_scatter, dim3(persistent_grid()), dim3(kBlock), 0, s, d, io);
    d.cur ^= 1;
}

void launch_generate(DevState& d, int32_t n, hipStream_t s) {
    reset_step_counters(d, s);
    hipMemsetAsync(d.intcnt, 0, sizeof(uint32_t) * n, s);
    hipMemsetAsync(d.extcnt, 0, sizeof(uint32_t) * ((size_t)n + 1), s);
    hipMemsetAsync(d.fsize + 2 * (d.cur ^ 1), 0, sizeof(uint32_t) * 2, s);
    FrontIO io = make_io(d);
    if (n >= 2) hipLaunchKernelGGL(k_traverse<true>, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, d, io);
    d.cur ^= 1;
    launch_order_fronts(d, n, s);
}

// ---------------------------------------------------------------- dynamic work fetch
// Each wave takes 64 consecutive items at a time from a device cursor.
__device__ inline uint32_t wave_fetch(uint32_t* cursor) {
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(cursor, 64u);
    return __shfl(base, 0);
}

// ---------------------------------------------------------------- UPDATE (prune / sprout)
__global__ void __launch_bounds__(kBlock) k_update(DevState d, FrontIO io) {
    const uint32_t ni = min(io.fsz[0], io.cap_i), ne = min(io.fsz[1], io.cap_e), total = ni + ne;
    uint32_t visits = 0;
    for (;;) {
        const uint32_t base = wave_fetch(&d.ctr[C_WORK]);
        if (base >= total) break;
        const uint32_t k = base + lane_id();
        if (k >= total) continue;
        if (k < ni) {
            const int2 e = io.fi[k];
            int32_t st = e.y;
            const float4 qlo = d.leaf[e.x].lo, qhi = d.leaf[e.x].hi;
            const Box q = rec_box(qlo, qhi);
            const float4 lo = d.inode[st].lo, hi = d.inode[st].hi;
            ++visits;
            if (!box_overlap(q, lo, hi)) {  // prune: climb the left spine while not overlapping
                const int32_t t = max(d.leaf_lca[f2i(lo.w)] >> 1, 0);
                for (st--; st >= t; st--) {
                    ++visits;
                    if (node_overlap(d, st, q)) break;
                }
                if (st < t && st + 1 > 0) {
                    ++visits;
                    if (!node_overlap(d, d.int_par[st + 1], q)) continue;  // covered by a left sibling
                }
                push_int(d, io, e.x, st + 1);
                continue;
            }
            const int32_t start = (d.int_lc[st] << 1) | (int32_t)(d.int_mark[st] & 1u);
            visits += stackless<true>(d, io, e.x, q, start, f2i(hi.w));  // sprout
        } else {
            const int2 e = io.fe[k - ni];
            const int32_t j = e.y;
            const float4 qlo = d.leaf[e.x].lo, qhi = d.leaf[e.x].hi;
            const Box q = rec_box(qlo, qhi);
            const float4 llo = d.leaf[j].lo, lhi = d.leaf[j].hi;
            ++visits;
            if (box_overlap(q, llo, lhi)) {
                push_ext(d, io, e.x, j);
                push_pair(d, e.x, j);
                continue;
            }
            const int32_t gfa = f2i(llo.w);
            ++visits;
            if (node_overlap(d, gfa, q)) {
                push_ext(d, io, e.x, j);
                continue;
            }
            if (d.leaf_mark[j] & 4u) continue;  // right child: the left sibling's entry climbs
            const int32_t top = max(d.leaf_lca[j] >> 1, 0);
            int32_t st;
            for (st = gfa - 1; st >= top; st--) {
                ++visits;
                if (node_overlap(d, st, q)) break;
            }
            if (st < top && st + 1 > 0) {
                ++visits;
                if (!node_overlap(d, d.int_par[st + 1], q)) continue;
            }
            push_int(d, io, e.x, st + 1);
        }
    }
    add_visits(d, visits);
}

// checkFrontQuality: uses the offsets of the previous ordering (intbeg/extbeg) and the snapshot.
__global__ void __launch_bounds__(kBlock) k_quality(int32_t n, DevState d) {
    const int32_t I = n - 1;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= I) return;
    const int32_t x = d.int_rx[i], y = d.int_ry[i];
    const int intCnt = (int)(d.intbeg[i + y - x] - d.intbeg[i]);
    const int extCnt = (int)(d.extbeg[y + 1] - d.extbeg[x]);
    const float snap1 = d.snapshot[i] + 1.0f;
    if (1.0 * intCnt / extCnt > (double)snap1 - 0.7 * intCnt / (double)d.intbeg[I]) {
        atomicAdd(&d.intrange[i], 1);
        atomicAdd(&d.intrange[i + y - x], -1);
        atomicMin(&d.rtroot[x], i);
    }
}

__global__ void __launch_bounds__(kBlock) k_snapshot(int32_t n, DevState d) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n - 1) return;
    const int32_t x = d.int_rx[i], y = d.int_ry[i];
    const uint32_t ic = d.intbeg[i + y - x] - d.intbeg[i];
    const uint32_t ec = d.extbeg[y + 1] - d.extbeg[x];
    d.snapshot[i] = (float)(1.0 * ic / (double)(ec + 1u));
}

void launch_snapshot(DevState& d, int32_t n, hipStream_t s) {
    if (n >= 2) hipLaunchKernelGGL(k_snapshot, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, d);
}

void launch_update(DevState& d, int32_t n, int quality, hipStream_t s) {
    reset_step_counters(d, s);
    hipMemsetAsync(d.intcnt, 0, sizeof(uint32_t) * n, s);
    hipMemsetAsync(d.extcnt, 0, sizeof(uint32_t) * ((size_t)n + 1), s);
    hipMemsetAsync(d.fsize + 2 * (d.cur ^ 1), 0, sizeof(uint32_t) * 2, s);
    hipMemsetAsync(d.ctr + C_WORK, 0, sizeof(uint32_t), s);
    FrontIO io = make_io(d);
    hipLaunchKernelGGL(k_update, dim3(persistent_grid()), dim3(kBlock), 0, s, d, io);
    d.cur ^= 1;
    if (quality && n >= 2) {
        hipMemsetAsync(d.intrange, 0, sizeof(int32_t) * n, s);
        fill_i32(d.rtroot, n, INT32_MAX, s);
        hipLaunchKernelGGL(k_quality, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, d);
    }
    launch_order_fronts(d, n, s);
}

// ---------------------------------------------------------------- KEEP
__global__ void __launch_bounds__(kBlock) k_keep(DevState d, FrontIO io) {
    const uint32_t ni = min(io.fsz[0], io.cap_i), ne = min(io.fsz[1], io.cap_e), total = ni + ne;
    uint32_t visits = 0;
    for (;;) {
        const uint32_t base = wave_fetch(&d.ctr[C_WORK]);
        if (base >= total) break;
        const uint32_t k = base + lane_id();
        if (k >= total) continue;
        if (k < ni) {
            const int2 e = io.fi[k];
            const float4 qlo = d.leaf[e.x].lo, qhi = d.leaf[e.x].hi;
            const Box q = rec_box(qlo, qhi);
            const float4 hi = d.inode[e.y].hi, lo = d.inode[e.y].lo;
            ++visits;
            if (!box_overlap(q, lo, hi)) continue;
            const int32_t start = (d.int_lc[e.y] << 1) | (int32_t)(d.int_mark[e.y] & 1u);
            visits += stackless<false>(d, io, e.x, q, start, f2i(hi.w));
        } else {
            const int2 e = io.fe[k - ni];
            const float4 alo = d.leaf[e.x].lo, ahi = d.leaf[e.x].hi;
            const float4 blo = d.leaf[e.y].lo, bhi = d.leaf[e.y].hi;
            ++visits;
            if (box_overlap(rec_box(alo, ahi), blo, bhi)) push_pair(d, e.x, e.y);
        }
    }
    add_visits(d, visits);
}

void launch_keep(DevState& d, int32_t n, hipStream_t s) {
    reset_step_counters(d, s);
    hipMemsetAsync(d.ctr + C_WORK, 0, sizeof(uint32_t), s);
    FrontIO io = make_io(d);
    hipLaunchKernelGGL(k_keep, dim3(persistent_grid()), dim3(kBlock), 0, s, d, io);
}

// ---------------------------------------------------------------- REORDER (balance)
// Internal entries re-bucketed by node: non-overlapping ones from the bucket start, overlapping
// ones from the bucket end; the external front is copied. Pairs as in KEEP.
__global__ void __launch_bounds__(kBlock) k_reorder_int(DevState d, FrontIO io) {
    const uint32_t ni = min(io.fsz[0], io.cap_i);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ni; k += gridDim.x * blockDim.x) {
        const int2 e = io.fi[k];
        const float4 qlo = d.leaf[e.x].lo, qhi = d.leaf[e.x].hi;
        const bool ov = node_overlap(d, e.y, rec_box(qlo, qhi));
        uint32_t pos;
        if (!ov) pos = d.intbeg[e.y] + atomicAdd(&d.cursor[e.y], 1u);
        else pos = d.intbeg[e.y + 1] - 1u - atomicAdd(&d.intback[e.y], 1u);
        if (pos < io.cap_i) io.oi[pos] = e;
    }
}
__global__ void __launch_bounds__(kBlock) k_copy_ext(FrontIO io) {
    const uint32_t ne = min(io.fsz[1], io.cap_e);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < ne; k += gridDim.x * blockDim.x) io.oe[k] = io.fe[k];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        io.osz[0] = io.fsz[0];
        io.osz[1] = io.fsz[1];
    }
}

void launch_reorder(DevState& d, int32_t n, hipStream_t s) {
    launch_keep(d, n, s);  // pairs from the unchanged cut
    launch_exscan(d, d.intcnt, d.intbeg, n, nullptr, s);
    hipMemsetAsync(d.cursor, 0, sizeof(uint32_t) * n, s);
    hipMemsetAsync(d.intback, 0, sizeof(uint32_t) * n, s);
    FrontIO io = make_io(d);
    hipLaunchKernelGGL(k_reorder_int, dim3(persistent_grid()), dim3(kBlock), 0, s, d, io);
    hipLaunchKernelGGL(k_copy_ext, dim3(persistent_grid()), dim3(kBlock), 0, s, io);
    d.cur ^= 1;
}

// ---------------------------------------------------------------- restructure decision
// intmark = inclusive scan of intrange; outermost flagged roots are rtroot[x] with intmark 1.
__global__ void k_intmark_fix(int32_t I, const int32_t* __restrict__ intrange, int32_t* __restrict__ intmark) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < I) intmark[i] += intrange[i];
}
__global__ void k_select_roots(int32_t n, DevState d) {
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int32_t r = d.rtroot[x];
    if (r != INT32_MAX && d.intmark[r] == 1) {
        const int32_t a = d.int_rx[r], b = d.int_ry[r];
        atomicAdd(&d.ctr[C_RT_SUB], 1u);
        atomicAdd(&d.ctr[C_RT_INT], (uint32_t)(b - a));
        atomicAdd(&d.ctr[C_RT_LEAF], (uint32_t)(b - a + 1));
        // difference array of (root + 1) over the leaf range
        atomicAdd(&d.r_leaf_lo[a], r + 1);
        atomicAdd(&d.r_leaf_lo[b + 1], -(r + 1));
    }
}

void launch_restr_decision(DevState& d, int32_t n, hipStream_t s) {
    const int32_t I = n - 1;
    hipMemsetAsync(d.ctr + C_RT_SUB, 0, sizeof(uint32_t) * 3, s);
    launch_exscan(d, reinterpret_cast<const uint32_t*>(d.intrange), reinterpret_cast<uint32_t*>(d.intmark), I, nullptr,
                  s);
    hipLaunchKernelGGL(k_intmark_fix, dim3(grid_for(I)), dim3(kBlock), 0, s, I, d.intrange, d.intmark);
    hipMemsetAsync(d.r_leaf_lo, 0, sizeof(int32_t) * ((size_t)n + 1), s);
    hipLaunchKernelGGL(k_select_roots, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d);
}

// ---------------------------------------------------------------- narrow phase
__device__ inline float3 ld3(const float* __restrict__ v, int32_t i) {
    return make_float3(v[3 * (int64_t)i], v[3 * (int64_t)i + 1], v[3 * (int64_t)i + 2]);
}
__device__ inline float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Projections of both triangles on `ax` overlap (closed intervals).
__device__ inline bool axis_overlap(float3 ax, const float3* p, const float3* q) {
    const float P0 = dot3(ax, p[0]), P1 = dot3(ax, p[1]), P2 = dot3(ax, p[2]);
    const float Q0 = dot3(ax, q[0]), Q1 = dot3(ax, q[1]), Q2 = dot3(ax, q[2]);
    const float mx1 = fmaxf(fmaxf(P0, P1), P2), mn1 = fminf(fminf(P0, P1), P2);
    const float mx2 = fmaxf(fmaxf(Q0, Q1), Q2), mn2 = fminf(fminf(Q0, Q1), Q2);
    return !(mn1 > mx2) && !(mn2 > mx1);
}

// 17 candidate axes: two face normals, 9 edge-edge crosses, 6 in-plane edge normals.
// Coordinates are translated so that the first vertex of P is the origin.
__device__ bool sat_tri_tri(float3 P1, float3 P2, float3 P3, float3 Q1, float3 Q2, float3 Q3) {
    float3 p[3] = {make_float3(0.f, 0.f, 0.f), sub3(P2, P1), sub3(P3, P1)};
    float3 q[3] = {sub3(Q1, P1), sub3(Q2, P1), sub3(Q3, P1)};
    const float3 e[3] = {sub3(p[1], p[0]), sub3(p[2], p[1]), sub3(p[0], p[2])};
    const float3 f[3] = {sub3(q[1], q[0]), sub3(q[2], q[1]), sub3(q[0], q[2])};
    const float3 nrm = cross3(e[0], e[1]), mrm = cross3(f[0], f[1]);
    if (!axis_overlap(nrm, p, q)) return false;
    if (!axis_overlap(mrm, p, q)) return false;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            if (!axis_overlap(cross3(e[i], f[j]), p, q)) return false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (!axis_overlap(cross3(e[k], nrm), p, q)) return false;
#pragma unroll
    for (int k = 0; k < 3; ++k)
        if (!axis_overlap(cross3(f[k], mrm), p, q)) return false;
    return true;
}

__global__ void __launch_bounds__(kBlock) k_narrow(DevState d) {
    const uint32_t np_ = min(d.ctr[C_BP], (uint32_t)d.pair_cap);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < np_; k += gridDim.x * blockDim.x) {
        const int2 pr = d.bp[k];  // pr.x < pr.y
        const int32_t* fa = d.faces + 3 * (int64_t)pr.x;
        const int32_t* fb = d.faces + 3 * (int64_t)pr.y;
        const bool hit = sat_tri_tri(ld3(d.verts, fa[0]), ld3(d.verts, fa[1]), ld3(d.verts, fa[2]),
                                     ld3(d.verts, fb[0]), ld3(d.verts, fb[1]), ld3(d.verts, fb[2]));
        if (hit) {
            const uint32_t slot = wave_append(&d.ctr[C_NP], true);
            if (slot < (uint32_t)d.pair_cap) d.np[slot] = pr;
            else atomicOr(&d.ctr[C_OVF], 8u);
        }
    }
}

void launch_narrow(DevState& d, hipStream_t s) {
    hipMemsetAsync(d.ctr + C_NP, 0, sizeof(uint32_t), s);
    hipLaunchKernelGGL(k_narrow, dim3(persistent_grid() * 2), dim3(kBlock), 0, s, d);
}

// ---------------------------------------------------------------- pair-set digest
__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__global__ void __launch_bounds__(kBlock) k_digest(const int2* __restrict__ pairs, const uint32_t* __restrict__ cnt,
                                                   uint64_t cap, uint64_t* out) {
    const uint32_t n = (uint32_t)min((uint64_t)*cnt, cap);
    uint64_t acc = 0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int2 p = pairs[k];
        acc += mix64(((uint64_t)(uint32_t)p.x << 32) | (uint32_t)p.y);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if (lane_id() == 0 && acc) atomicAdd((unsigned long long*)out, (unsigned long long)acc);
}

void launch_digest(DevState& d, int filtered, hipStream_t s) {
    uint64_t* out = ctr64(d.ctr) + C64_DIGEST;
    hipMemsetAsync(out, 0, sizeof(uint64_t), s);
    hipLaunchKernelGGL(k_digest, dim3(1024), dim3(kBlock), 0, s, filtered ? d.np : d.bp,
                       d.ctr + (filtered ? C_NP : C_BP), (uint64_t)d.pair_cap, out);
}

}  // namespace bvhcd
