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
// Layout and scheduling on gfx950:
//  * Every node visit is one 32-byte record (two 16-byte loads): box + rangex/escape
//    (internal) or box + parent/escape (leaf).
//  * Fronts are SoA (node id, query leaf) so that ordering is a radix sort on the node id
//    with the query as payload: no per-node counters, no atomics, and the bucket starts
//    (the reference's FlOrderLog offsets) come from binary searches on the sorted ids.
//  * Work is distributed grid-stride (consecutive entries per wave64, coalesced SoA loads).
//  * Appends go through a per-wave LDS staging buffer: lanes take slots with a ballot
//    prefix, and one global atomic reserves space for a whole buffer when it is flushed.
//    Capacity is checked before every global write (overflow bit instead of a write).
#include <cstdio>
#include <string>
#include <vector>

#include "bvhcd_internal.hpp"
#include "kernels.hpp"

namespace bvhcd {

constexpr int kWaves = kBlock / 64;
constexpr uint32_t kStage = 256;   // staged entries per wave, single-list kernels
// Per list when a kernel stages three lists. 128 would fit 8 waves/SIMD in LDS, but every flush
// is one atomic on the list's single counter and those serialise (measured: GENERATE 2.9 ms at
// 256, 5.5 ms at 128 on the 2M-triangle cloth), so fewer, larger flushes win.
constexpr uint32_t kStage3 = 256;

__device__ inline uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

// Per-wave LDS words shared between lanes are reached through pointers typed as LDS: their
// accesses are then ds_read / ds_write. Through generic pointers they would be flat accesses,
// which also count against the vector-memory counter, so waiting for one of them would wait for
// every global load in flight as well.
#define BVHCD_LDS __attribute__((address_space(3)))
typedef volatile BVHCD_LDS int2 lds_int2;
typedef volatile BVHCD_LDS uint32_t lds_u32;

// ---------------------------------------------------------------- output sinks
// Front output (SoA node / query), segmented like the pair lists: kPairSegs counters
// kCntStride words apart, segment g at [g * seg_cap, g * seg_cap + count). The producing kernel
// appends segment-wise; front_finish then gathers the segments into the current buffer.
struct FrontSink {
    uint32_t* cnt;
    uint32_t* node;
    int32_t* query;
    int64_t seg_cap;
    uint32_t ovf_bit;
    // wave-ordered production (null: off): chunk records and per-wave totals
    uint4* chunks = nullptr;
    uint32_t* wave_tot = nullptr;
    uint32_t chunk_seg_cap = 0;
};
struct PairSink {  // int2 list with one counter
    uint32_t* count;
    int2* out;
    uint32_t cap;
    uint32_t ovf_bit;
    __device__ void write(uint32_t o, int2 v) const { out[o] = v; }
};
struct SegPairSink {  // segmented pair list: kPairSegs counters kCntStride words apart
    uint32_t* cnt;
    int2* base;
    int64_t seg_cap;
    uint32_t ovf_bit;
};

// The staging word pair of a wave and list: sfill[0] = staged entries, sfill[1] = rotor that
// picks the segment of the next flush of a segmented list.
template <class Sink>
__device__ void stage_flush(const Sink& sk, uint32_t* ovf, lds_int2* sbuf, lds_u32*, uint32_t f,
                            uint64_t active) {
    const int lane = lane_id();
    const int leader = __ffsll((long long)active) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(sk.count, f);
    base = __shfl(base, leader);
    const uint32_t rank = (uint32_t)__popcll(active & lanes_below());
    const uint32_t nact = (uint32_t)__popcll(active);
    bool over = false;
    for (uint32_t i = rank; i < f; i += nact) {
        const uint32_t o = base + i;
        const int2 v = make_int2(sbuf[i].x, sbuf[i].y);
        if (o < sk.cap) sk.write(o, v);
        else over = true;
    }
    if (over) atomicOr(ovf, sk.ovf_bit);
}
// Segmented lists: successive flushes of a wave rotate over the segments, so the appends
// spread over kPairSegs counters and every segment fills evenly whichever workgroups produce
// the entries. The rotor start spreads the waves of one XCD (workgroups b with equal b % 8,
// which a per-XCD split makes neighbours in the data) over all segments too.
template <int W = kWaves>
__device__ inline uint32_t rotor_start() {
    const uint32_t b = blockIdx.x, w = threadIdx.x >> 6;
    return (b / kXcds) * W + w + (b % kXcds) * kXcds;
}
// One flush takes its slots in the rotor's next segment; what does not fit there (the segment
// filled up) is given back to the counter and tries the following segments, so a list overflows
// only once all its segments are full, and every segment counter ends at the entries its segment
// holds. Entries no segment took are counted in word 2 of the list's counter block (the list's
// requested length is the segment counts plus that). Wave-ordered production (chunk records):
// one 64-bit add on the segment counter takes both the slots (low word) and the chunk's seq in
// the segment (high word), and every seq gets its record, which keeps only what its segment
// holds; a record that does not fit drops its chunk (overflow bit) so that the gathered front
// stays dense. sfill[2] counts the wave's recorded entries.
struct SegTake {
    uint32_t g, base, fit;
};
__device__ inline SegTake seg_take(uint32_t* cnt, int64_t seg_cap, const FrontSink* chunked, uint32_t* ovf,
                                   lds_u32* sfill, uint32_t want, uint64_t active) {
    const int leader = __ffsll((long long)active) - 1;
    uint32_t g = 0, base = 0, fit = 0;
    if (lane_id() == leader) {
        g = sfill[1] % kPairSegs;
        sfill[1] = sfill[1] + 1;
        if (chunked) {
            const unsigned long long old =
                atomicAdd(reinterpret_cast<unsigned long long*>(cnt + g * kCntStride), (1ull << 32) | want);
            base = (uint32_t)old;
            const uint32_t seq = (uint32_t)(old >> 32);
            fit = base < (uint64_t)seg_cap ? (uint32_t)min<int64_t>(want, seg_cap - base) : 0u;
            if (seq < chunked->chunk_seg_cap) {
                const uint32_t wave = blockIdx.x * kWaves + (threadIdx.x >> 6);
                chunked->chunks[(int64_t)g * chunked->chunk_seg_cap + seq] = make_uint4(base, wave, sfill[2], fit);
            } else {
                fit = 0u;
                atomicOr(ovf, chunked->ovf_bit);
            }
            sfill[2] = sfill[2] + fit;
        } else {
            base = atomicAdd(cnt + g * kCntStride, want);
            fit = base < (uint64_t)seg_cap ? (uint32_t)min<int64_t>(want, seg_cap - base) : 0u;
        }
        if (fit < want) atomicSub(cnt + g * kCntStride, want - fit);  // slots past the segment's end
    }
    return SegTake{(uint32_t)__shfl((int)g, leader), (uint32_t)__shfl((int)base, leader),
                   (uint32_t)__shfl((int)fit, leader)};
}
__device__ inline void stage_flush(const SegPairSink& sk, uint32_t* ovf, lds_int2* sbuf, lds_u32* sfill,
                                   uint32_t f, uint64_t active) {
    const uint32_t rank = (uint32_t)__popcll(active & lanes_below());
    const uint32_t nact = (uint32_t)__popcll(active);
    uint32_t done = 0;
    for (int a = 0; a < kPairSegs && done < f; ++a) {
        const SegTake t = seg_take(sk.cnt, sk.seg_cap, nullptr, ovf, sfill, f - done, active);
        int2* out = sk.base + (int64_t)t.g * sk.seg_cap + t.base;
        for (uint32_t i = rank; i < t.fit; i += nact) out[i] = make_int2(sbuf[done + i].x, sbuf[done + i].y);
        done += t.fit;
    }
    if (done < f && lane_id() == __ffsll((long long)active) - 1) {
        atomicOr(ovf, sk.ovf_bit);
        atomicAdd(sk.cnt + kDroppedWord, f - done);
    }
}
__device__ inline void stage_flush(const FrontSink& sk, uint32_t* ovf, lds_int2* sbuf, lds_u32* sfill,
                                   uint32_t f, uint64_t active) {
    const uint32_t rank = (uint32_t)__popcll(active & lanes_below());
    const uint32_t nact = (uint32_t)__popcll(active);
    uint32_t done = 0;
    for (int a = 0; a < kPairSegs && done < f; ++a) {
        const SegTake t = seg_take(sk.cnt, sk.seg_cap, sk.chunks ? &sk : nullptr, ovf, sfill, f - done, active);
        const int64_t o = (int64_t)t.g * sk.seg_cap + t.base;
        for (uint32_t i = rank; i < t.fit; i += nact) {
            sk.node[o + i] = (uint32_t)sbuf[done + i].y;
            sk.query[o + i] = sbuf[done + i].x;
        }
        done += t.fit;
    }
    if (done < f && lane_id() == __ffsll((long long)active) - 1) {
        atomicOr(ovf, sk.ovf_bit);
        atomicAdd(sk.cnt + kDroppedWord, f - done);
    }
}

// Called by the currently active lanes of a wave; `pred` selects the lanes that emit `v`.
template <class Sink>
__device__ inline void stage_push(const Sink& sk, uint32_t* ovf, lds_int2* sbuf, lds_u32* sfill,
                                  uint32_t cap, bool pred, int2 v) {
    const uint64_t active = __ballot(1);
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    uint32_t f = *sfill;
    const uint32_t cnt = (uint32_t)__popcll(m);
    if (f + cnt > cap) {
        stage_flush(sk, ovf, sbuf, sfill, f, active);
        f = 0;
    }
    if (pred) {
        const uint32_t slot = f + (uint32_t)__popcll(m & lanes_below());
        sbuf[slot].x = v.x;
        sbuf[slot].y = v.y;
    }
    if (lane_id() == __ffsll((long long)active) - 1) *sfill = f + cnt;
}

template <class Sink>
__device__ inline void stage_final(const Sink& sk, uint32_t* ovf, lds_int2* sbuf, lds_u32* sfill) {
    const uint32_t f = *sfill;
    if (f) stage_flush(sk, ovf, sbuf, sfill, f, __ballot(1));
    if (lane_id() == 0) *sfill = 0;
}

// Per-block staging storage for up to three lists; fill[list][wave] = {staged, rotor, recorded
// entries of a wave-ordered production}.
struct Stages {
    int2 buf[3][kWaves][kStage3];
    uint32_t fill[3][kWaves][3];
};
__device__ inline void stages_init(Stages& st) {
    const int w = threadIdx.x >> 6;
    if (lane_id() < 3) {
        st.fill[lane_id()][w][0] = 0;
        st.fill[lane_id()][w][1] = rotor_start();
        st.fill[lane_id()][w][2] = 0;
    }
}
// End of a wave's production: its recorded total, every wave of the launch writing its own.
__device__ inline void wave_total(const FrontSink& sk, lds_u32* sfill) {
    if (sk.chunks && lane_id() == __ffsll((long long)__ballot(1)) - 1)
        sk.wave_tot[blockIdx.x * kWaves + (threadIdx.x >> 6)] = sfill[2];
}

// Output context of a traversal thread.
struct Out {
    FrontSink fi, fe;
    SegPairSink bp;
    uint32_t* ovf;
    lds_int2* sb[3];
    lds_u32* sf[3];
    uint32_t scap;  // staging entries per wave and list
    const int32_t* lfaces;
    int skip_adj;   // drop pairs of triangles sharing a vertex
    // Broad-phase pairs are kept as leaf (sorted-slot) ids, query first (qi < j): the narrow
    // phase then reads leaf-ordered faces with locality, and maps to triangle ids itself.
    __device__ void pair_push(bool pred, int2 v) const { stage_push(bp, ovf, sb[2], sf[2], scap, pred, v); }
    __device__ void pair_final() const { stage_final(bp, ovf, sb[2], sf[2]); }
    __device__ void pair(int32_t qi, int32_t j) const {
        if (skip_adj && faces_share_vertex(lfaces, qi, j)) return;
        pair_push(true, make_int2(qi, j));
    }
    __device__ void int_entry(int32_t qi, int32_t node) const {
        stage_push(fi, ovf, sb[0], sf[0], scap, true, make_int2(qi, node));
    }
    __device__ void ext_entry(int32_t qi, int32_t leaf) const {
        stage_push(fe, ovf, sb[1], sf[1], scap, true, make_int2(qi, leaf));
    }
    __device__ void final_flush() const {
        stage_final(fi, ovf, sb[0], sf[0]);
        stage_final(fe, ovf, sb[1], sf[1]);
        pair_final();
        wave_total(fi, sf[0]);
        wave_total(fe, sf[1]);
    }
};

// Segmented pair list of kind 0 (broad phase) or 1 (narrow phase).
__device__ inline SegPairSink pair_segments(const DevState& d, int kind, uint32_t ovf_bit) {
    return SegPairSink{d.pcnt + kind * kPairSegs * kCntStride, kind ? d.np : d.bp, d.pair_seg_cap, ovf_bit};
}

// Pair-only kernels (KEEP, batched traversal) stage one list: a third of the LDS, so 8 waves
// per SIMD instead of 6 hide more of the dependent node loads. W = waves per workgroup.
template <int W = kWaves>
struct StagesBPT {
    int2 buf[W][kStage];
    uint32_t fill[W][2];
};
using StagesBP = StagesBPT<>;
template <int W>
__device__ inline void stages_init(StagesBPT<W>& st) {
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        st.fill[w][0] = 0;
        st.fill[w][1] = rotor_start<W>();
    }
}
template <int W>
__device__ inline Out make_out(StagesBPT<W>& st, const DevState& d) {
    Out o;
    o.fi = FrontSink{nullptr, nullptr, nullptr, 0u, 0u};
    o.fe = o.fi;
    o.bp = pair_segments(d, 0, 4u);
    o.ovf = &d.ctr[C_OVF];
    const int w = threadIdx.x >> 6;
    o.sb[0] = o.sb[1] = nullptr;
    o.sf[0] = o.sf[1] = nullptr;
    o.sb[2] = (lds_int2*)st.buf[w];
    o.sf[2] = (lds_u32*)st.fill[w];
    o.scap = kStage;
    o.lfaces = d.lfaces;
    o.skip_adj = d.skip_adj;
    return o;
}

__device__ inline Out make_out(Stages& st, const FrontSink& fi, const FrontSink& fe, const DevState& d) {
    Out o;
    o.fi = fi;
    o.fe = fe;
    o.bp = pair_segments(d, 0, 4u);
    o.ovf = &d.ctr[C_OVF];
    const int w = threadIdx.x >> 6;
    for (int k = 0; k < 3; ++k) {
        o.sb[k] = (lds_int2*)st.buf[k][w];
        o.sf[k] = (lds_u32*)st.fill[k][w];
    }
    o.scap = kStage3;
    o.lfaces = d.lfaces;
    o.skip_adj = d.skip_adj;
    return o;
}

__device__ inline bool node_overlap(const DevState& d, int32_t nd, const Box& q) {
    const float4 lo = d.inode[nd].lo, hi = d.inode[nd].hi;
    return box_overlap(q, lo, hi);
}

// Stackless DFS from encoded node `st` until `stop` (lca-escape encoding). With FRONT, the
// non-overlapping internal nodes and all reached leaves are appended to the next front.
template <bool FRONT>
__device__ uint32_t stackless(const DevState& d, const Out& out, int32_t qi, const Box& q, int32_t st, int32_t stop) {
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
            j = f2i(lo.w);                    // leftmost leaf of the whole left spine
            const int32_t t = d.leaf_par[j];  // bottom of the spine
            bool reached = false;
            for (;;) {
                ++visits;
                if (!box_overlap(q, lo, hi)) break;
                if (nd == t) {
                    reached = true;
                    break;
                }
                ++nd;
                lo = d.inode[nd].lo;
                hi = d.inode[nd].hi;
            }
            if (!reached) {
                if (FRONT) out.int_entry(qi, nd);
                st = f2i(hi.w);
                continue;
            }
        } else {
            j = st >> 1;
        }
        const float4 llo = d.leaf[j].lo, lhi = d.leaf[j].hi;
        ++visits;
        if (box_overlap(q, llo, lhi)) out.pair(qi, j);
        if (FRONT) out.ext_entry(qi, j);
        st = f2i(lhi.w);
    }
    return visits;
}

// Block-reduced visit count, one 64-bit atomic per block (blocks of up to 16 waves).
__device__ inline void add_visits(const DevState& d, uint32_t v) {
    __shared__ uint64_t red[16];
    uint64_t x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off);
    if (lane_id() == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[i];
        if (t) atomicAdd((unsigned long long*)&ctr64(d.ctr)[C64_VISITS], (unsigned long long)t);
    }
}

// ---------------------------------------------------------------- host-side front views
struct FrontIn {
    const uint32_t* node[2];
    const int32_t* query[2];
    const uint32_t* size[2];
    uint32_t cap[2];
};

// Segmented output into the spare buffer of `kind` (its counters cleared by front_begin).
static FrontSink next_sink(DevState& d, int kind) {
    const int nb = d.fcur[kind] ^ 1;
    return FrontSink{d.fcnt + kind * kPairSegs * kCntStride, d.fnode[kind][nb], d.fquery[kind][nb],
                     d.front_cap[kind] / kPairSegs, kind ? 2u : 1u};
}
static void front_begin(DevState& d, hipStream_t s) { fill_u32(d.fcnt, 2 * kPairSegs * kCntStride, 0u, s); }

// Gathers the segments of the spare buffer into the current buffer of `kind`, after its first
// `base` entries (base read from *base_word before any block writes, 0 when null); the current
// buffer's size becomes base + total, clamped to the capacity (overflow bit otherwise).
__global__ void __launch_bounds__(kBlock) k_front_compact(DevState d, int kind, const uint32_t* __restrict__ src_node,
                                                          const int32_t* __restrict__ src_query,
                                                          uint32_t* __restrict__ dst_node,
                                                          int32_t* __restrict__ dst_query, const uint32_t* base_word,
                                                          uint32_t* dst_size) {
    __shared__ uint32_t pre[kPairSegs + 1];
    const int64_t seg_cap = d.front_cap[kind] / kPairSegs;
    const uint32_t total = seg_prefix(d.fcnt + kind * kPairSegs * kCntStride, (uint32_t)seg_cap, pre);
    const uint32_t base = base_word ? *base_word : 0u;
    const uint32_t cap = (uint32_t)min(d.front_cap[kind], (int64_t)0xFFFFFFFF);
    const uint32_t room = base < cap ? cap - base : 0u;
    const uint32_t m = min(total, room);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) {
        const int64_t src = seg_slot(pre, k, seg_cap);
        dst_node[base + k] = src_node[src];
        dst_query[base + k] = src_query[src];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *dst_size = base + m;
        if (m < total) atomicOr(&d.ctr[C_OVF], kind ? 2u : 1u);
        // capacity that would have held every requested entry (an over-estimate after an
        // overflow: slots a full segment refused were asked for again elsewhere)
        const uint32_t* cnt = d.fcnt + kind * kPairSegs * kCntStride;
        uint32_t sum = cnt[kDroppedWord];
        for (int g = 0; g < kPairSegs; ++g) sum += cnt[g * kCntStride];
        d.ctr[kind ? C_FREQ_E : C_FREQ_I] = base + sum;
    }
}
// append: keep the current buffer's entries and add the new ones after them.
static void front_finish(DevState& d, int kind, bool append, hipStream_t s) {
    const int c = d.fcur[kind], nb = c ^ 1;
    uint32_t* size = d.fsize + 2 * kind + c;
    uint32_t* base_word = nullptr;
    if (append) {  // the base is snapshotted: block 0 rewrites *size while others still read
        base_word = d.ctr + (kind ? C_KEPT_E : C_KEPT_I);
        hipMemcpyAsync(base_word, size, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    }
    hipLaunchKernelGGL(k_front_compact, dim3(2048), dim3(kBlock), 0, s, d, kind, d.fnode[kind][nb],
                       d.fquery[kind][nb], d.fnode[kind][c], d.fquery[kind][c], base_word, size);
}
// Wave-ordered production: the sinks record chunks, `waves` producing waves (the launch's).
static bool wave_ordered_ok(const DevState& d, int64_t waves) {
    return d.fchunk[0] && d.fchunk[1] && waves <= d.wave_cap;
}
static void record_chunks(DevState& d, FrontSink& sk, int kind) {
    sk.chunks = d.fchunk[kind];
    sk.wave_tot = d.wave_tot[kind];
    sk.chunk_seg_cap = (uint32_t)d.chunk_seg_cap;
}

// Orders a wave-ordered production of `kind` into its current buffer. The chunk records are
// first copied into lists by producing wave (k_chunk_count, a scan of the counts, k_chunk_list,
// each with its offset in the segments, so the next kernel finds the entries in one load); then one
// workgroup per producing wave w loads its chunks straight from the segments (chunk c holds the
// wave's entries c.z .. c.z + count - 1) and orders the run by query with a counting sort in LDS
// on q - 64 w (the wave's queries are 64 w .. 64 w + 63), writing it once at the scan of the
// wave totals. The order within a query is not kept (fronts are multisets). A run longer than
// kWaveRunLds entries, or one past the capacity, is written in production order (clamped to the
// capacity; the overflow bit is set and the front stays a dense prefix). The size becomes the
// wave totals' sum, clamped to the capacity.
__global__ void __launch_bounds__(kBlock) k_chunk_count(DevState d, int kind) {
    __shared__ uint32_t pre[kPairSegs + 1];
    const uint32_t nchunks =
        seg_prefix(d.fcnt + kind * kPairSegs * kCntStride + 1, (uint32_t)d.chunk_seg_cap, pre);  // chunk seqs
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x)
        atomicAdd(&d.wave_nch[d.fchunk[kind][seg_slot(pre, c, d.chunk_seg_cap)].y], 1u);
}
__global__ void __launch_bounds__(kBlock) k_chunk_list(DevState d, int kind) {
    __shared__ uint32_t pre[kPairSegs + 1];
    const uint32_t nchunks =
        seg_prefix(d.fcnt + kind * kPairSegs * kCntStride + 1, (uint32_t)d.chunk_seg_cap, pre);
    const int64_t seg_cap = d.front_cap[kind] / kPairSegs;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nchunks; c += gridDim.x * blockDim.x) {
        const int64_t slot = seg_slot(pre, c, d.chunk_seg_cap);
        const uint4 ch = d.fchunk[kind][slot];
        const uint64_t src = (uint64_t)((slot / d.chunk_seg_cap) * seg_cap + ch.x);  // in the segments
        d.clist[d.wave_cbase[ch.y] + atomicSub(&d.wave_nch[ch.y], 1u) - 1u] =
            make_uint4((uint32_t)src, (uint32_t)(src >> 32), ch.z, ch.w);
    }
}
constexpr uint32_t kWaveRunLds = 4096;
constexpr int kOrderBlock = 512;  // LDS allows 4 workgroups per CU: 8 waves each fill it
__global__ void __launch_bounds__(kOrderBlock) k_wave_order(DevState d, int kind, uint32_t waves,
                                                       const uint32_t* __restrict__ src_node,
                                                       const int32_t* __restrict__ src_query,
                                                       uint32_t* __restrict__ dst_node,
                                                       int32_t* __restrict__ dst_query, uint32_t* dst_size) {
    __shared__ uint32_t lnode[kWaveRunLds];
    __shared__ uint32_t lkey[kWaveRunLds];  // (q - 64 w) << 16 | rank among the run's entries of q
    __shared__ uint32_t cnt[64];
    const uint32_t w = blockIdx.x;
    const uint32_t cap = (uint32_t)min(d.front_cap[kind], (int64_t)0xFFFFFFFF);
    const uint32_t b0 = d.wave_base[w], m = d.wave_tot[kind][w];
    const uint32_t c0 = d.wave_cbase[w], c1 = w + 1 < waves ? d.wave_cbase[w + 1] : d.ctr[C_WAVE_CHUNKS];
    const int lane = lane_id(), wid = threadIdx.x >> 6;
    if (w == 0 && threadIdx.x == 0) {
        const uint32_t total = d.ctr[C_WAVE_TOTAL];
        *dst_size = min(total, cap);
        if (total > cap) atomicOr(&d.ctr[C_OVF], kind ? 2u : 1u);
        const uint32_t* fc = d.fcnt + kind * kPairSegs * kCntStride;
        uint32_t sum = fc[kDroppedWord];
        for (int g = 0; g < kPairSegs; ++g) sum += fc[g * kCntStride];
        d.ctr[kind ? C_FREQ_E : C_FREQ_I] = sum;
    }
    if (m > kWaveRunLds || (uint64_t)b0 + m > cap) {
        for (uint32_t c = c0 + wid; c < c1; c += kOrderBlock / 64) {
            const uint4 ch = d.clist[c];
            const int64_t src = (int64_t)((uint64_t)ch.x | (uint64_t)ch.y << 32);
            for (uint32_t i = lane; i < ch.w; i += 64) {
                const uint32_t o = b0 + ch.z + i;
                if (o < cap) {
                    dst_node[o] = src_node[src + i];
                    dst_query[o] = src_query[src + i];
                }
            }
        }
        return;
    }
    if (threadIdx.x < 64) cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int32_t q0 = (int32_t)(w * 64u);
    for (uint32_t c = c0 + wid; c < c1; c += kOrderBlock / 64) {
        const uint4 ch = d.clist[c];
        const int64_t src = (int64_t)((uint64_t)ch.x | (uint64_t)ch.y << 32);
        constexpr int kPerLane = kStage3 / 64;  // a chunk is one staging flush
        uint32_t nd[kPerLane];
        int32_t q[kPerLane];
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            const uint32_t i = lane + 64u * k;
            if (i < ch.w) {
                nd[k] = src_node[src + i];
                q[k] = src_query[src + i];
            }
        }
#pragma unroll
        for (int k = 0; k < kPerLane; ++k) {
            const uint32_t i = lane + 64u * k, j = ch.z + i;
            if (i < ch.w && j < m) {
                const uint32_t qo = min((uint32_t)(q[k] - q0), 63u);
                lnode[j] = nd[k];
                lkey[j] = qo << 16 | atomicAdd(&cnt[qo], 1u);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the 64 query counts
        const uint32_t v = cnt[threadIdx.x];
        uint32_t x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if ((int)threadIdx.x >= off) x += y;
        }
        cnt[threadIdx.x] = x - v;
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < m; j += kOrderBlock) {
        const uint32_t key = lkey[j], qo = key >> 16;
        const uint32_t o = b0 + cnt[qo] + (key & 0xFFFFu);
        dst_node[o] = lnode[j];
        dst_query[o] = q0 + (int32_t)qo;
    }
}

static void front_finish_waves(DevState& d, int kind, int64_t waves, hipStream_t s) {
    const int c = d.fcur[kind], nb = c ^ 1;  // the production is in the spare buffer's segments
    launch_exscan(d, d.wave_tot[kind], d.wave_base, waves, d.ctr + C_WAVE_TOTAL, s);
    fill_u32(d.wave_nch, waves, 0u, s);
    hipLaunchKernelGGL(k_chunk_count, dim3(512), dim3(kBlock), 0, s, d, kind);
    launch_exscan(d, d.wave_nch, d.wave_cbase, waves, d.ctr + C_WAVE_CHUNKS, s);
    hipLaunchKernelGGL(k_chunk_list, dim3(512), dim3(kBlock), 0, s, d, kind);
    hipLaunchKernelGGL(k_wave_order, dim3((unsigned)waves), dim3(kOrderBlock), 0, s, d, kind, (uint32_t)waves,
                       d.fnode[kind][nb], d.fquery[kind][nb], d.fnode[kind][c], d.fquery[kind][c],
                       d.fsize + 2 * kind + c);
}

static FrontIn cur_front(DevState& d) {
    FrontIn f;
    for (int k = 0; k < 2; ++k) {
        const int b = d.fcur[k];
        f.node[k] = d.fnode[k][b];
        f.query[k] = d.fquery[k][b];
        f.size[k] = d.fsize + 2 * k + b;
        f.cap[k] = (uint32_t)std::min<int64_t>(d.front_cap[k], 0xFFFFFFFFll);
    }
    return f;
}
__host__ __device__ inline FrontSink null_sink() { return FrontSink{nullptr, nullptr, nullptr, 0u, 0u}; }

// One 64-lane launch (a kernel, not memsets: the batched frame graph records it): the per-step
// pair counters, the visit counter and the segment counters of both pair lists.
__global__ void k_reset_step(DevState d) { zero_step_counters(d, threadIdx.x); }
static void reset_step_counters(DevState& d, hipStream_t s) {
    static_assert(kPairSegs == 64, "one lane per segment counter");
    hipLaunchKernelGGL(k_reset_step, dim3(1), dim3(kPairSegs), 0, s, d);
}

static unsigned stream_grid() { return 2048; }

// ---------------------------------------------------------------- launch-shape tuning
// The reference picks one block size per kernel from the occupancy API
// (system/CudaDevice/CudaDevice.cu:60-94). Here the grid of the grid-stride and persistent
// kernels, KEEP's entries per thread and the workgroup size of KEEP and the narrow phase (2, 4
// or 8 wave64s, their LDS sized to match) are tuned per kernel and size class ceil(log2 n), by
// timing candidates on live data (bvhcd_autotune / bvhcd_batch_autotune). The table is
// process-wide and not locked.
namespace {
constexpr int kClasses = 40;
const unsigned kDefaultGrid[TK_COUNT] = {2048u, 4096u, 2048u};
unsigned g_grid[TK_COUNT][kClasses] = {};
struct TuneTrial {
    int kernel, cls;
    unsigned grid;
    float ms;
};
std::vector<TuneTrial> g_trials;
int size_class(int64_t n) {
    int c = 0;
    while (c < kClasses - 1 && ((int64_t)1 << c) < n) ++c;
    return c;
}
}  // namespace

unsigned tuned_shape(int kernel, int64_t n) {
    const unsigned g = g_grid[kernel][size_class(n)];
    return g ? g : kDefaultGrid[kernel];
}
void set_tuned_shape(int kernel, int64_t n, unsigned sh) { g_grid[kernel][size_class(n)] = sh; }
void record_tuning(int kernel, int64_t n, unsigned sh, float ms) {
    g_trials.push_back({kernel, size_class(n), sh, ms});
}
std::string tuning_report() {
    static const char* names[TK_COUNT] = {"k_keep", "k_narrow", "k_batch_cd"};
    std::string r = "kernel size_class grid unroll block ms_per_launch chosen\n";
    char line[160];
    for (const TuneTrial& t : g_trials) {
        snprintf(line, sizeof(line), "%s 2^%d %u %u %u %.4f %s\n", names[t.kernel], t.cls, shape_grid(t.grid),
                 shape_unroll(t.grid), 64u * shape_waves(t.grid), t.ms, g_grid[t.kernel][t.cls] == t.grid ? "*" : "");
        r += line;
    }
    return r;
}

// ---------------------------------------------------------------- pure traversal / generate
// Lane of the wave holding the k-th (0-based) set bit of m (k < popcount(m)).
__device__ inline int kth_set_lane(uint64_t m, uint32_t k) {
    int lo = 0, hi = 63;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((uint32_t)__popcll(m & ((2ull << mid) - 1ull)) > k) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// The stackless walk of stackless<FRONT>, one node test per lane and step, with the wave's
// lanes sharing the work: a query's walk from `cur` to `stop` is, at an overlapping internal
// node nd, nd's subtree followed by the walk from nd's escape to `stop`. So a lane at such a node
// can hand the second part to an idle lane (its own walk then stops at nd's escape) — the walks
// of 64 queries end at very different times (the longest is ~1.8x the mean on the folded 2M
// cloth), and without this the wave waits for its longest one. The nodes tested, the entries
// and the pairs are those of the 64 queries' own walks; only their order changes.
template <bool FRONT>
__global__ void __launch_bounds__(kBlock) k_traverse(int32_t n, DevState d, FrontSink fi, FrontSink fe) {
    __shared__ Stages st;
    stages_init(st);
    const Out out = make_out(st, fi, fe, d);
    const int lane = lane_id();
    const int32_t nq = d.inter ? d.nq : n - 1;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // self-CD: query leaf i walks the leaves after it (from lca[i + 1], its record's escape);
    // inter-object: every query walks the whole tree from its root lca[0]
    bool act = i < nq;
    int32_t qi = i, cur = -1, stop = -1, spine = -1;  // spine: bottom of the spine being descended
    Box qb{};
    if (act) {
        const float4 lo = d.qleaf[i].lo, hi = d.qleaf[i].hi;
        qb = rec_box(lo, hi);
        cur = d.inter ? d.leaf_lca[0] : f2i(hi.w);
        act = cur != stop;
    }
    uint32_t v = 0;
    // every node is tested at most once per query walk; the bound only guards against a
    // corrupted tree (flag 16, reported as a device error, instead of a hang)
    const int64_t limit = 64 * (2 * d.maxT + 16);
    for (int64_t round = 0; __ballot(act) != 0ull; ++round) {
        if (round > limit) {
            if (act) atomicOr(&d.ctr[C_OVF], 16u);
            break;
        }
        bool emit_int = false, emit_ext = false, emit_pair = false, can_give = false;
        int32_t node = 0, give = -1;
        if (act) {
            ++v;
            if (!(cur & 1)) {
                const int32_t nd = cur >> 1;
                const float4 lo = d.inode[nd].lo, hi = d.inode[nd].hi;
                if (spine < 0) spine = d.leaf_par[f2i(lo.w)];  // a new subtree: its left spine's bottom
                node = nd;
                if (!box_overlap(qb, lo, hi)) {
                    emit_int = FRONT;
                    cur = f2i(hi.w);
                    spine = -1;
                } else {
                    give = f2i(hi.w);  // what follows nd's subtree
                    can_give = give != stop;
                    cur = nd == spine ? ((f2i(lo.w) << 1) | 1) : cur + 2;  // leftmost leaf, or nd + 1
                    if (nd == spine) spine = -1;
                }
            } else {
                const int32_t j = cur >> 1;
                const float4 llo = d.leaf[j].lo, lhi = d.leaf[j].hi;
                node = j;
                emit_pair = box_overlap(qb, llo, lhi);
                emit_ext = FRONT;
                cur = f2i(lhi.w);
                spine = -1;
            }
            act = cur != stop;
        }
        if (FRONT) {
            stage_push(out.fi, out.ovf, out.sb[0], out.sf[0], out.scap, emit_int, make_int2(qi, node));
            stage_push(out.fe, out.ovf, out.sb[1], out.sf[1], out.scap, emit_ext, make_int2(qi, node));
        }
        bool pair = emit_pair;
        if (out.skip_adj && pair) pair = !faces_share_vertex(out.lfaces, qi, node);
        out.pair_push(pair, make_int2(qi, node));
        // idle lanes take the rest of other lanes' walks
        const uint64_t idle = __ballot(!act), donors = __ballot(can_give);
        const uint32_t pairs = (uint32_t)min(__popcll(idle), __popcll(donors));
        if (pairs >= 8u || (pairs > 0u && __popcll(idle) >= 32)) {
            const uint32_t ri = (uint32_t)__popcll(idle & lanes_below());
            const uint32_t rd = (uint32_t)__popcll(donors & lanes_below());
            const int src = (!act && ri < pairs) ? kth_set_lane(donors, ri) : lane;
            const int32_t g_q = __shfl(qi, src), g_cur = __shfl(give, src), g_stop = __shfl(stop, src);
            const float g0 = __shfl(qb.mnx, src), g1 = __shfl(qb.mny, src), g2 = __shfl(qb.mnz, src);
            const float g3 = __shfl(qb.mxx, src), g4 = __shfl(qb.mxy, src), g5 = __shfl(qb.mxz, src);
            if (can_give && rd < pairs) stop = give;  // the donor ends at nd's escape
            if (!act && ri < pairs) {
                act = true;
                qi = g_q;
                cur = g_cur;
                stop = g_stop;
                spine = -1;
                qb = Box{g0, g1, g2, g3, g4, g5};
            }
        }
    }
    if (FRONT) {
        out.final_flush();
    } else {
        out.pair_final();
    }
    add_visits(d, v);
}

// Queries of a traversal: the leaves before the last one (self-CD) or the other object's leaves.
static int64_t query_count(const DevState& d, int32_t n) { return d.inter ? d.nq : n - 1; }

void launch_pure_cd(DevState& d, int32_t n, hipStream_t s) {
    reset_step_counters(d, s);
    const int64_t nq = query_count(d, n);
    if (nq > 0)
        hipLaunchKernelGGL(k_traverse<false>, dim3(grid_for(nq)), dim3(kBlock), 0, s, n, d, null_sink(), null_sink());
}

// ---------------------------------------------------------------- batched scenes
// Self-CD over many scenes in one tree: the sort keys carry the scene id above the 30 Morton
// bits, so every scene is one subtree and its leaves are contiguous. Query x of the scene whose
// last leaf is b walks the leaves (x, b] only: from lca[x + 1] to that subtree's escape
// lca[b + 1]. Persistent: each wave takes 64 queries at a time from a global counter until
// all are taken (uneven scenes and uneven per-query work balance themselves).
__global__ void __launch_bounds__(kBlock) k_batch_cd(int32_t n, DevState d, const uint64_t* __restrict__ keys,
                                                     const int32_t* __restrict__ scene_last, uint32_t* queue) {
    __shared__ StagesBP st;
    stages_init(st);
    const Out out = make_out(st, d);
    const int lane = lane_id();
    uint32_t v = 0;
    for (;;) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(queue, 64u);
        base = __shfl(base, 0);
        if (base >= (uint32_t)n) break;  // wave-uniform exit: the queue is drained
        const int32_t x = (int32_t)(base + lane);
        if (x < n) {
            const int32_t last = scene_last[(int32_t)(keys[x] >> 30)];
            if (x < last) {
                const float4 lo = d.leaf[x].lo, hi = d.leaf[x].hi;
                v += stackless<false>(d, out, x, rec_box(lo, hi), f2i(hi.w), d.leaf_lca[last + 1]);
            }
        }
    }
    out.pair_final();
    add_visits(d, v);
}

// Counters are cleared with kernels (no memset nodes) so that the sequence can be captured in
// a graph; `grid` workgroups stay resident and drain the query queue.
void launch_batch_cd(DevState& d, int32_t n, const uint64_t* keys, const int32_t* scene_last, uint32_t* queue,
                     unsigned grid, hipStream_t s) {
    reset_step_counters(d, s);
    fill_u32(queue, 1, 0u, s);
    if (n >= 2) hipLaunchKernelGGL(k_batch_cd, dim3(grid), dim3(kBlock), 0, s, n, d, keys, scene_last, queue);
}

// ---------------------------------------------------------------- ordering
// beg[b] = first position of bucket b in the sorted keys (b in [0, nb]); key count from *dn.
__global__ void __launch_bounds__(kBlock) k_bucket_begin(const uint32_t* __restrict__ keys, const uint32_t* dn,
                                                         uint32_t cap, int32_t nb, uint32_t* __restrict__ beg) {
    const uint32_t m = min(*dn, cap);
    for (int32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= nb; b += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = m;  // first index with keys[idx] >= b
        while (lo < hi) {
            const uint32_t mid = lo + ((hi - lo) >> 1);
            if (keys[mid] < (uint32_t)b) lo = mid + 1;
            else hi = mid;
        }
        beg[b] = lo;
    }
}

static int bits_for(int64_t v) {
    int b = 1;
    while (((int64_t)1 << b) <= v) ++b;
    return b;
}

// Fronts of the learned schedule are not node-ordered; they are sorted by query leaf instead
// (stable, so each query keeps its emission order): neighbouring queries read neighbouring
// nodes, which KEEP's per-XCD split then keeps inside one L2.
static void sort_front_by_query(DevState& d, int kind, int32_t n, hipStream_t s) {
    const int c = d.fcur[kind], o = c ^ 1;
    const bool in_tmp = radix_sort_u32_dev(d, reinterpret_cast<uint32_t*>(d.fquery[kind][c]),
                                           reinterpret_cast<int32_t*>(d.fnode[kind][c]),
                                           reinterpret_cast<uint32_t*>(d.fquery[kind][o]),
                                           reinterpret_cast<int32_t*>(d.fnode[kind][o]), d.front_cap[kind],
                                           d.fsize + 2 * kind + c, bits_for(n - 1), s);
    if (in_tmp) {
        hipMemcpyAsync(d.fsize + 2 * kind + o, d.fsize + 2 * kind + c, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        d.fcur[kind] = o;
    }
}
void launch_query_order(DevState& d, int32_t n, hipStream_t s) {
    if (n < 2) return;
    sort_front_by_query(d, 0, n, s);
    sort_front_by_query(d, 1, n, s);
}

// Counting sort by node id (the reference's FlOrderLog::prepare + pureReorderLooseFrontsWithLog,
// FlOrderLog.cu:7-26, BvttFrontReorderKernels.cu:12-24): per-node counts, an exclusive scan
// into the bucket starts intbeg[0..n-1] (internal ids) / extbeg[0..n] (leaves), one scatter
// into the spare buffer. Lanes holding runs of one node id (fronts produced from node-ordered
// fronts keep them together) take their slots with one atomic per run. The order inside a
// bucket is that of the atomics, as in the reference.
__device__ inline uint32_t run_atomic(uint32_t* ctr, bool valid, uint32_t key) {
    const int lane = lane_id();
    const uint64_t vmask = __ballot(valid);
    const uint32_t prev = __shfl_up(key, 1);
    const bool prev_valid = lane > 0 ? ((vmask >> (lane - 1)) & 1ull) != 0 : false;
    const bool start = valid && !(prev_valid && prev == key);
    const uint64_t starts = __ballot(start);
    const uint64_t le = (2ull << lane) - 1ull;  // lanes at or below this one
    const uint64_t upto = starts & le;
    const int s0 = upto ? 63 - __clzll((long long)upto) : lane;  // this lane's run start
    uint32_t base = 0;
    if (start) {  // the run ends at the next start or the next invalid lane
        const uint64_t stop = (starts | ~vmask) & ~le;
        const int e = stop ? __ffsll((long long)stop) - 1 : 64;
        base = atomicAdd(ctr + key, (uint32_t)(e - lane));
    }
    base = __shfl(base, s0);
    return base + (uint32_t)(lane - s0);
}

// The same with every set of equal keys in the wave taken together, one round per distinct key:
// internal fronts put nodes with large subtrees into the fronts of many queries, and their
// counters would otherwise take one contended atomic per entry.
// The rounds only group the lanes: each group's lowest lane issues its atomic and keeps the old
// value, and one shuffle after the last round hands the bases out, so the atomics of all rounds
// are in flight together instead of one round trip per distinct key.
__device__ inline uint32_t match_atomic(uint32_t* ctr, bool valid, uint32_t key) {
    const int lane = lane_id();
    uint64_t todo = __ballot(valid);
    uint32_t base = 0, rank = 0;
    int lead = lane;
    while (todo) {  // wave-uniform
        const int first = __ffsll((long long)todo) - 1;
        const uint32_t k = (uint32_t)__shfl((int)key, first);
        const uint64_t grp = __ballot(key == k) & todo;
        if (lane == first) base = atomicAdd(ctr + k, (uint32_t)__popcll(grp));
        if ((grp >> lane) & 1ull) {
            lead = first;
            rank = (uint32_t)__popcll(grp & lanes_below());
        }
        todo &= ~grp;
    }
    return (uint32_t)__shfl((int)base, lead) + rank;
}
template <bool MATCH>
__device__ inline uint32_t slot_atomic(uint32_t* ctr, bool valid, uint32_t key) {
    return MATCH ? match_atomic(ctr, valid, key) : run_atomic(ctr, valid, key);
}

template <bool MATCH>
__global__ void __launch_bounds__(kBlock) k_order_count(const uint32_t* __restrict__ node, const uint32_t* size,
                                                        uint32_t cap, int32_t nb, uint32_t* __restrict__ cnt) {
    const uint32_t m = min(*size, cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k0 = blockIdx.x * blockDim.x; k0 < m; k0 += stride) {  // wave-uniform trip count
        const uint32_t k = k0 + threadIdx.x;
        const bool valid = k < m;
        const uint32_t key = valid ? min(node[k], (uint32_t)(nb - 1)) : 0u;
        slot_atomic<MATCH>(cnt, valid, key);
    }
}

// BALANCE (REORDER, internal front): inside each bucket the entries whose node does not overlap
// its query go first (slots from the bucket start up), the overlapping ones last (from the bucket
// end down), as the reference's reorderIntLooseIntraFrontsWithLog does with its back counts
// (BvttFrontReorderKernels.cu:27-63).
template <bool BALANCE, bool MATCH>
__global__ void __launch_bounds__(kBlock) k_order_scatter(DevState d, const uint32_t* __restrict__ node,
                                                          const int32_t* __restrict__ query, const uint32_t* size,
                                                          uint32_t cap, int32_t nb, const uint32_t* __restrict__ beg,
                                                          uint32_t* __restrict__ cur, uint32_t* __restrict__ back,
                                                          uint32_t* __restrict__ dst_node,
                                                          int32_t* __restrict__ dst_query, uint32_t* dst_size) {
    const uint32_t m = min(*size, cap);
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst_size = m;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k0 = blockIdx.x * blockDim.x; k0 < m; k0 += stride) {  // wave-uniform trip count
        const uint32_t k = k0 + threadIdx.x;
        const bool valid = k < m;
        uint32_t nd = 0;
        int32_t q = 0;
        if (valid) {
            nd = node[k];
            q = query[k];
        }
        const uint32_t key = min(nd, (uint32_t)(nb - 1));
        bool ov = false;
        if (BALANCE && valid) {
            const float4 qlo = d.qleaf[q].lo, qhi = d.qleaf[q].hi;
            ov = node_overlap(d, (int32_t)key, rec_box(qlo, qhi));
        }
        const uint32_t r = slot_atomic<MATCH>(cur, valid && !ov, key);
        uint32_t o = beg[key] + r;
        if (BALANCE) {
            const uint32_t rb = slot_atomic<MATCH>(back, valid && ov, key);
            if (ov) o = beg[key + 1] - 1u - rb;
        }
        if (valid) {
            dst_node[o] = nd;
            dst_query[o] = q;
        }
    }
}

// One counting sort of a front kind by node id into its spare buffer, which becomes current;
// beg[0..nb] = bucket starts (beg[nb] = size).
static void order_front(DevState& d, int kind, int32_t nb, uint32_t* cnt, uint32_t* beg, bool balance,
                        hipStream_t s) {
    const int c = d.fcur[kind], o = c ^ 1;
    const uint32_t cap = (uint32_t)std::min<int64_t>(d.front_cap[kind], 0xFFFFFFFFll);
    const unsigned g = std::min<unsigned>(grid_for(d.front_cap[kind]), 8192u);
    fill_u32(cnt, nb, 0u, s);
    // internal fronts take equal keys per wave together, external ones (leaves near each query,
    // few repeats) per run of equal neighbours
    if (kind == 0)
        hipLaunchKernelGGL(k_order_count<true>, dim3(g), dim3(kBlock), 0, s, d.fnode[kind][c], d.fsize + 2 * kind + c,
                           cap, nb, cnt);
    else
        hipLaunchKernelGGL(k_order_count<false>, dim3(g), dim3(kBlock), 0, s, d.fnode[kind][c],
                           d.fsize + 2 * kind + c, cap, nb, cnt);
    launch_exscan(d, cnt, beg, nb, beg + nb, s);
    fill_u32(cnt, nb, 0u, s);
    if (balance) {
        fill_u32(d.intback, nb, 0u, s);
        hipLaunchKernelGGL((k_order_scatter<true, true>), dim3(g), dim3(kBlock), 0, s, d, d.fnode[kind][c],
                           d.fquery[kind][c], d.fsize + 2 * kind + c, cap, nb, beg, cnt, d.intback, d.fnode[kind][o],
                           d.fquery[kind][o], d.fsize + 2 * kind + o);
    } else if (kind == 0) {
        hipLaunchKernelGGL((k_order_scatter<false, true>), dim3(g), dim3(kBlock), 0, s, d, d.fnode[kind][c],
                           d.fquery[kind][c], d.fsize + 2 * kind + c, cap, nb, beg, cnt, nullptr, d.fnode[kind][o],
                           d.fquery[kind][o], d.fsize + 2 * kind + o);
    } else {
        hipLaunchKernelGGL((k_order_scatter<false, false>), dim3(g), dim3(kBlock), 0, s, d, d.fnode[kind][c],
                           d.fquery[kind][c], d.fsize + 2 * kind + c, cap, nb, beg, cnt, nullptr, d.fnode[kind][o],
                           d.fquery[kind][o], d.fsize + 2 * kind + o);
    }
    d.fcur[kind] = o;
}

// A freshly generated internal front is radix-sorted instead (3 passes of 8 bits, then the bucket
// starts by binary search): there every query's front holds the few nodes with the largest
// subtrees, and the counting sort's atomics on those nodes' counters contend even when each wave
// takes its equal keys together (measured on the folded 2M-triangle cloth, 45 M entries: counting
// 4.7 ms, radix 1.9 ms; the other fronts: counting 1.0-1.4 ms, radix 1.2-1.9 ms).
static void radix_order_front(DevState& d, int kind, int32_t nb, uint32_t* beg, hipStream_t s) {
    const int c = d.fcur[kind], o = c ^ 1;
    const bool in_tmp = radix_sort_u32_dev(d, d.fnode[kind][c], d.fquery[kind][c], d.fnode[kind][o], d.fquery[kind][o],
                                           d.front_cap[kind], d.fsize + 2 * kind + c, bits_for(nb - 1), s);
    if (in_tmp) {
        hipMemcpyAsync(d.fsize + 2 * kind + o, d.fsize + 2 * kind + c, sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
        d.fcur[kind] = o;
    }
    const int cc = d.fcur[kind];
    hipLaunchKernelGGL(k_bucket_begin, dim3(grid_for(nb + 1)), dim3(kBlock), 0, s, d.fnode[kind][cc],
                       d.fsize + 2 * kind + cc, (uint32_t)std::min<int64_t>(d.front_cap[kind], 0xFFFFFFFFll), nb, beg);
}

void launch_order_fronts(DevState& d, int32_t n, hipStream_t s, bool generated) {
    if (generated) radix_order_front(d, 0, n - 1, d.intbeg, s);
    else order_front(d, 0, n - 1, d.intcnt, d.intbeg, false, s);
    order_front(d, 1, n, d.extcnt, d.extbeg, false, s);
}

// Without node ordering, the front comes out ordered by query without a global sort: the
// production is gathered in wave order (each wave's 64 consecutive queries keep their entries
// together, the waves following in query order), then each wave's run is sorted by query.
void launch_generate(DevState& d, int32_t n, hipStream_t s, bool order) {
    reset_step_counters(d, s);
    front_begin(d, s);
    FrontSink fi = next_sink(d, 0), fe = next_sink(d, 1);
    const int64_t nq = query_count(d, n);
    const unsigned grid = nq > 0 ? grid_for(nq) : 0u;
    const bool waves = !order && nq > 0 && wave_ordered_ok(d, (int64_t)grid * kWaves);
    if (waves) {
        record_chunks(d, fi, 0);
        record_chunks(d, fe, 1);
    }
    if (nq > 0) hipLaunchKernelGGL(k_traverse<true>, dim3(grid), dim3(kBlock), 0, s, n, d, fi, fe);
    if (waves) {
        front_finish_waves(d, 0, (int64_t)grid * kWaves, s);
        front_finish_waves(d, 1, (int64_t)grid * kWaves, s);
    } else {
        front_finish(d, 0, false, s);
        front_finish(d, 1, false, s);
        if (order && n >= 2) launch_order_fronts(d, n, s, true);
        if (!order) launch_query_order(d, n, s);
    }
    d.front_by_query = order ? 0 : 1;
}

// ---------------------------------------------------------------- UPDATE (prune / sprout)
// One entry per lane and loop iteration (grid-stride, so the wave's entries stay neighbours):
// pruned and external entries are settled at once, an overlapping internal entry starts a sprout
// walk (the stackless walk of its subtree, one node test per step, as in k_traverse), and the
// wave steps its walks until all have ended, lanes without a walk taking over the tails of
// others' (from an overlapping node's escape on): the wave no longer waits for its longest
// sprout with the other lanes idle.
__global__ void __launch_bounds__(kBlock) k_update(DevState d, FrontIn in, FrontSink fi, FrontSink fe) {
    __shared__ Stages st;
    stages_init(st);
    const Out out = make_out(st, fi, fe, d);
    const int lane = lane_id();
    const uint32_t ni = min(*in.size[0], in.cap[0]), ne = min(*in.size[1], in.cap[1]), total = ni + ne;
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t visits = 0;
    const int64_t limit = 64 * (2 * d.maxT + 16);
    // wave-uniform trip count: every lane of a wave takes part in the walk rounds
    for (uint32_t k0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); k0 < total; k0 += stride) {
        const uint32_t k = k0 + lane;
        bool walk = false;
        int32_t qi = 0, cur = -1, stop = -1, spine = -1;
        Box qb{};
        if (k < ni) {
            const int32_t q = in.query[0][k];
            int32_t node = (int32_t)in.node[0][k];
            const float4 qlo = d.qleaf[q].lo, qhi = d.qleaf[q].hi;
            const Box b = rec_box(qlo, qhi);
            const float4 lo = d.inode[node].lo, hi = d.inode[node].hi;
            ++visits;
            if (!box_overlap(b, lo, hi)) {  // prune: climb the left spine while not overlapping
                const int32_t t = max(d.leaf_lca[f2i(lo.w)] >> 1, 0);
                for (node--; node >= t; node--) {
                    ++visits;
                    if (node_overlap(d, node, b)) break;
                }
                bool keep = true;
                if (node < t && node + 1 > 0) {
                    ++visits;
                    keep = node_overlap(d, d.int_par[node + 1], b);  // else a left sibling's entry covers it
                }
                if (keep) out.int_entry(q, node + 1);
            } else {  // sprout: walk the subtree from its left child to its escape
                qi = q;
                qb = b;
                cur = (d.int_lc[node] << 1) | (int32_t)(d.int_mark[node] & 1u);
                stop = f2i(hi.w);
                walk = cur != stop;
            }
        } else if (k < total) {
            const int32_t q = in.query[1][k - ni];
            const int32_t j = (int32_t)in.node[1][k - ni];
            const float4 qlo = d.qleaf[q].lo, qhi = d.qleaf[q].hi;
            const Box b = rec_box(qlo, qhi);
            const float4 llo = d.leaf[j].lo, lhi = d.leaf[j].hi;
            ++visits;
            if (box_overlap(b, llo, lhi)) {
                out.ext_entry(q, j);
                out.pair(q, j);
            } else {
                const int32_t gfa = f2i(llo.w);
                ++visits;
                if (node_overlap(d, gfa, b)) {
                    out.ext_entry(q, j);
                } else if (!(d.leaf_mark[j] & 4u)) {  // right child: the left sibling's entry climbs
                    const int32_t top = max(d.leaf_lca[j] >> 1, 0);
                    int32_t node;
                    for (node = gfa - 1; node >= top; node--) {
                        ++visits;
                        if (node_overlap(d, node, b)) break;
                    }
                    bool keep = true;
                    if (node < top && node + 1 > 0) {
                        ++visits;
                        keep = node_overlap(d, d.int_par[node + 1], b);
                    }
                    if (keep) out.int_entry(q, node + 1);
                }
            }
        }
        // the sprout walks of this iteration's entries, shared by the wave
        for (int64_t round = 0; __ballot(walk) != 0ull; ++round) {
            if (round > limit) {  // only a corrupted tree gets here (flag 16, a device error)
                if (walk) atomicOr(&d.ctr[C_OVF], 16u);
                break;
            }
            bool can_give = false;
            int32_t give = -1;
            if (walk) {
                ++visits;
                if (!(cur & 1)) {
                    const int32_t nd = cur >> 1;
                    const float4 lo = d.inode[nd].lo, hi = d.inode[nd].hi;
                    if (spine < 0) spine = d.leaf_par[f2i(lo.w)];
                    if (!box_overlap(qb, lo, hi)) {
                        out.int_entry(qi, nd);
                        cur = f2i(hi.w);
                        spine = -1;
                    } else {
                        give = f2i(hi.w);
                        can_give = give != stop;
                        cur = nd == spine ? ((f2i(lo.w) << 1) | 1) : cur + 2;
                        if (nd == spine) spine = -1;
                    }
                } else {
                    const int32_t j = cur >> 1;
                    const float4 llo = d.leaf[j].lo, lhi = d.leaf[j].hi;
                    if (box_overlap(qb, llo, lhi)) out.pair(qi, j);
                    out.ext_entry(qi, j);
                    cur = f2i(lhi.w);
                    spine = -1;
                }
                walk = cur != stop;
            }
            const uint64_t idle_m = __ballot(!walk), donors = __ballot(can_give);
            const uint32_t pairs = (uint32_t)min(__popcll(idle_m), __popcll(donors));
            if (pairs >= 8u || (pairs > 0u && __popcll(idle_m) >= 32)) {
                const uint32_t ri = (uint32_t)__popcll(idle_m & lanes_below());
                const uint32_t rd = (uint32_t)__popcll(donors & lanes_below());
                const int src = (!walk && ri < pairs) ? kth_set_lane(donors, ri) : lane;
                const int32_t g_q = __shfl(qi, src), g_cur = __shfl(give, src), g_stop = __shfl(stop, src);
                const float g0 = __shfl(qb.mnx, src), g1 = __shfl(qb.mny, src), g2 = __shfl(qb.mnz, src);
                const float g3 = __shfl(qb.mxx, src), g4 = __shfl(qb.mxy, src), g5 = __shfl(qb.mxz, src);
                if (can_give && rd < pairs) stop = give;
                if (!walk && ri < pairs) {
                    walk = true;
                    qi = g_q;
                    cur = g_cur;
                    stop = g_stop;
                    spine = -1;
                    qb = Box{g0, g1, g2, g3, g4, g5};
                }
            }
        }
    }
    out.final_flush();
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
    const float snap1 = d.snapshot[i] + 1.0f;  // a float sum, as in the reference expression
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

void launch_update(DevState& d, int32_t n, int quality, hipStream_t s, bool order) {
    if (!order) quality = 0;  // the quality check reads the ordering's bucket offsets
    reset_step_counters(d, s);
    front_begin(d, s);
    const FrontIn in = cur_front(d);
    FrontSink fi = next_sink(d, 0), fe = next_sink(d, 1);
    hipLaunchKernelGGL(k_update, dim3(stream_grid()), dim3(kBlock), 0, s, d, in, fi, fe);
    front_finish(d, 0, false, s);  // into the input buffers, which k_update is done with
    front_finish(d, 1, false, s);
    if (quality && n >= 2) {
        hipMemsetAsync(d.intrange, 0, sizeof(int32_t) * n, s);
        fill_i32(d.rtroot, n, INT32_MAX, s);
        hipLaunchKernelGGL(k_quality, dim3(grid_for(n - 1)), dim3(kBlock), 0, s, n, d);
    }
    if (order && n >= 2) launch_order_fronts(d, n, s, false);
    if (!order) launch_query_order(d, n, s);
    d.front_by_query = order ? 0 : 1;
}

// ---------------------------------------------------------------- KEEP
// Descents below internal front entries that now overlap their query are shared by the wave:
// instead of one lane walking a subtree while the others wait, the (query, node) pairs still to
// be tested go on a per-wave LDS stack and every lane takes one per round (a round is one
// dependent load: the node's record, and for an overlapping internal node its children). The
// nodes tested and the pairs found are those of the lane's own walk; a stack that is full
// leaves the lane to walk that subtree itself.
constexpr uint32_t kDescStack = 512;
template <int W>
struct DescStack {
    int2 item[W][kDescStack];  // (query, node << 1 | leaf)
    uint32_t top[W];
};

// Pushes (q, node enc) for the lanes with pred set; returns false for a lane whose item did not
// fit (it must then test that node itself). Called by all active lanes of the wave.
__device__ inline bool desc_push(lds_int2* stk, lds_u32* top, bool pred, int32_t q, int32_t enc) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return true;
    const uint32_t t = *top;
    const uint32_t rank = (uint32_t)__popcll(m & lanes_below());
    const bool fits = t + rank < kDescStack;
    if (pred && fits) {
        stk[t + rank].x = q;
        stk[t + rank].y = enc;
    }
    const uint32_t pushed = (uint32_t)__popcll(__ballot(pred && fits));
    if (lane_id() == __ffsll((long long)__ballot(1)) - 1) *top = t + pushed;
    return fits || !pred;
}

// Tests node `enc` against q's box: an overlapping leaf is a pair, an overlapping internal node
// pushes its two children (or, without room, is walked by this lane). Returns the node tests.
__device__ inline uint32_t desc_test(const DevState& d, const Out& out, lds_int2* stk, lds_u32* top,
                                     bool active, int32_t q, int32_t enc) {
    uint32_t visits = 0;
    bool leaf_hit = false, int_hit = false;
    int32_t nd = enc >> 1;
    Box qb{};
    float4 hi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) {
        const float4 qlo = d.qleaf[q].lo, qhi = d.qleaf[q].hi;
        qb = rec_box(qlo, qhi);
        const NodeRec* r = (enc & 1) ? &d.leaf[nd] : &d.inode[nd];
        const float4 lo = r->lo;
        hi = r->hi;
        ++visits;
        const bool hit = box_overlap(qb, lo, hi);
        leaf_hit = hit && (enc & 1);
        int_hit = hit && !(enc & 1);
    }
    bool emit = leaf_hit;
    if (out.skip_adj && emit) emit = !faces_share_vertex(out.lfaces, q, nd);
    out.pair_push(emit, make_int2(q, nd));
    int4 tp = make_int4(0, 0, 0, 0);
    if (int_hit) tp = d.itopo[nd];
    const int32_t lc = (tp.x << 1) | (tp.w & 1), rc = (tp.y << 1) | ((tp.w >> 1) & 1);
    const bool ok_l = desc_push(stk, top, int_hit, q, lc);
    const bool ok_r = desc_push(stk, top, int_hit, q, rc);
    if (int_hit && !(ok_l && ok_r)) {  // no room: walk what did not fit here (one call site)
        int32_t start = lc, end = f2i(hi.w);  // both children: nd's subtree below nd
        if (ok_l) start = rc;                 // the right child alone
        else if (ok_r) end = (lc & 1) ? d.leaf_lca[(lc >> 1) + 1] : f2i(d.inode[lc >> 1].hi.w);  // the left alone
        visits += stackless<false>(d, out, q, qb, start, end);
    }
    return visits;
}

// Works the wave's stack off, one item per lane and round, while it holds at least `least` items
// (64 between loop iterations: only full rounds, the rest waits for more; 1 at the end). Called
// by all active lanes.
__device__ inline uint32_t desc_drain(const DevState& d, const Out& out, lds_int2* stk, lds_u32* top,
                                      uint32_t least) {
    uint32_t visits = 0;
    const int lane = lane_id();
    const uint32_t nact = (uint32_t)__popcll(__ballot(1));
    const uint32_t rank = (uint32_t)__popcll(__ballot(1) & lanes_below());
    // every node is tested at most once per query walk; the bound only guards against a
    // corrupted tree (flag 16, reported as a device error, instead of a hang)
    const int64_t limit = 2 * d.maxT + 16;
    for (int64_t rounds = 0;; ++rounds) {
        const uint32_t t = *top;
        if (t < least || t == 0) break;
        if (rounds > limit) {
            if (lane == __ffsll((long long)__ballot(1)) - 1) {
                atomicOr(&d.ctr[C_OVF], 16u);
                *top = 0;
            }
            break;
        }
        const uint32_t take = min(t, nact);
        const bool active = rank < take;
        int32_t q = 0, enc = 0;
        if (active) {
            q = stk[t - 1 - rank].x;
            enc = stk[t - 1 - rank].y;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == __ffsll((long long)__ballot(1)) - 1) *top = t - take;
        __builtin_amdgcn_wave_barrier();
        visits += desc_test(d, out, stk, top, active, q, enc);
    }
    return visits;
}

// KEEP over the current front, U entries per thread per loop iteration with all their loads
// issued together: the kernel is bound by the latency of dependent loads (entry, then the two
// node records), so more loads in flight per wave is what speeds it up.
// implicit: the current front's shared-vertex head (C_ADJ entries after the last partition) is
// not tested when the last refit found every coordinate finite: its pairs are broad-phase hits
// whatever the positions (both boxes hold the shared vertex) and are counted, not listed.
// W = waves per workgroup (the tuner picks it with the grid and U).
template <int U, int W>
__global__ void __launch_bounds__(64 * W) k_keep(DevState d, FrontIn in, int implicit) {
    __shared__ StagesBPT<W> st;
    __shared__ DescStack<W> ds;
    stages_init(st);
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) ds.top[w] = 0;
    lds_int2* stk = (lds_int2*)ds.item[w];
    lds_u32* stop = (lds_u32*)&ds.top[w];
    const Out out = make_out(st, d);
    uint32_t ni = min(*in.size[0], in.cap[0]), ne = min(*in.size[1], in.cap[1]);
    // with skip_adj the head's pairs are filtered out anyway: skipped and not counted
    const uint32_t head = implicit ? min(d.ctr[C_ADJ], ne) : 0u;
    const uint32_t skip = d.skip_adj || d.ctr[C_NO_IMPLICIT] == 0u ? head : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) d.ctr[C_IMPLICIT] = d.skip_adj ? 0u : skip;
    ne -= skip;
    const uint32_t total = ni + ne;
    const uint32_t* ext_node = in.node[1] + skip;
    const int32_t* ext_query = in.query[1] + skip;
    const XcdLoop it = xcd_loop(total);
    const uint32_t stride = it.stride;
    uint32_t visits = 0;
    // entries of iteration k0 into (q, nd); -1 past the end
    auto load_entries = [&](uint32_t k0, int32_t* q, int32_t* nd) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = k0 + u * stride;
            q[u] = -1;
            nd[u] = 0;
            if (k >= it.end) continue;
            if (k < ni) {
                q[u] = in.query[0][k];
                nd[u] = (int32_t)in.node[0][k];
            } else {
                q[u] = ext_query[k - ni];
                nd[u] = (int32_t)ext_node[k - ni];
            }
        }
    };
    // software pipeline: the next iteration's entries are loaded while this one's node records
    // are in flight, so each iteration waits for one round trip instead of two
    int32_t qn[U], ndn[U];
    load_entries(it.begin, qn, ndn);
    // one loop iteration: test the entries, push the children of now-overlapping nodes, stage pairs
    auto keep_step = [&](uint32_t k0) {
        int32_t q[U], nd[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            q[u] = qn[u];
            nd[u] = ndn[u];
        }
        float4 qlo[U], qhi[U], lo[U], hi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            qlo[u] = qhi[u] = lo[u] = hi[u] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (q[u] >= 0) {
                const NodeRec* r = k0 + u * stride < ni ? &d.inode[nd[u]] : &d.leaf[nd[u]];
                qlo[u] = d.qleaf[q[u]].lo;
                qhi[u] = d.qleaf[q[u]].hi;
                lo[u] = r->lo;
                hi[u] = r->hi;
            }
        }
        load_entries(k0 + U * stride, qn, ndn);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool internal = k0 + u * stride < ni;
            const Box qb = rec_box(qlo[u], qhi[u]);
            const bool hit = q[u] >= 0 && box_overlap(qb, lo[u], hi[u]);
            visits += q[u] >= 0 ? 1u : 0u;
            // the node now overlaps: its children go on the wave's stack
            const bool desc = internal && hit;
            int4 tp = make_int4(0, 0, 0, 0);
            if (desc) tp = d.itopo[nd[u]];
            const int32_t lc = (tp.x << 1) | (tp.w & 1), rc = (tp.y << 1) | ((tp.w >> 1) & 1);
            const bool ok_l = desc_push(stk, stop, desc, q[u], lc);
            const bool ok_r = desc_push(stk, stop, desc, q[u], rc);
            if (desc && !(ok_l && ok_r)) {  // no room on the stack: this lane walks the subtree
                const int32_t start = ok_l ? rc : lc;
                visits += stackless<false>(d, out, q[u], qb, start, f2i(hi[u].w));
            }
            bool emit = hit && !internal;
            if (out.skip_adj && emit) emit = !faces_share_vertex(out.lfaces, q[u], nd[u]);
            out.pair_push(emit, make_int2(q[u], nd[u]));
        }
    };
    // The wave stays in the loop until all its lanes are past the end (their entries read as
    // -1), so every drain of the stack runs with the whole wave, from one call site.
    for (uint32_t k0 = it.begin;; k0 += U * stride) {
        const bool last = __ballot(k0 < it.end) == 0;
        if (!last) keep_step(k0);
        visits += desc_drain(d, out, stk, stop, last ? 1u : 64u);
        if (last) break;
    }
    out.pair_final();
    add_visits(d, visits);
}

// ---------------------------------------------------------------- shared-vertex partition
// Reorders the current external front, stably, into [entries whose two triangles share a vertex |
// the rest] and records the head length in C_ADJ. Two passes over tiles of kAdjTile entries
// (count, exclusive scan, scatter into the spare buffer), run by KEEP once per front version.

__global__ void __launch_bounds__(kBlock) k_adj_count(DevState d, FrontIn in, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t wsum[kWaves];
    const uint32_t ne = min(*in.size[1], in.cap[1]);
    const uint32_t b0 = blockIdx.x * kAdjTile;
    uint32_t c = 0;
    for (uint32_t i = b0 + threadIdx.x; i < min(b0 + kAdjTile, ne); i += kBlock)
        c += leaves_share_vertex(d, in.query[1][i], (int32_t)in.node[1][i]) ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if (lane_id() == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kWaves; ++w) t += wsum[w];
        cnt[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kBlock) k_adj_scatter(DevState d, FrontIn in, const uint32_t* __restrict__ off,
                                                        uint32_t* __restrict__ dst_node, int32_t* __restrict__ dst_query,
                                                        uint32_t* dst_size) {
    __shared__ uint32_t wa[kWaves], wv[kWaves];
    const uint32_t ne = min(*in.size[1], in.cap[1]);
    const uint32_t total_adj = d.ctr[C_ADJ];
    const uint32_t b0 = blockIdx.x * kAdjTile;
    const int w = threadIdx.x >> 6;
    uint32_t run_a = off[blockIdx.x], run_n = total_adj + (b0 - off[blockIdx.x]);
    for (uint32_t t0 = b0; t0 < min(b0 + kAdjTile, ne); t0 += kBlock) {
        const uint32_t i = t0 + threadIdx.x;
        const bool valid = i < ne;
        uint32_t node = 0;
        int32_t query = 0;
        if (valid) {
            node = in.node[1][i];
            query = in.query[1][i];
        }
        const bool adj = valid && leaves_share_vertex(d, query, (int32_t)node);
        const uint64_t ma = __ballot(adj), mv = __ballot(valid);
        if (lane_id() == 0) {
            wa[w] = (uint32_t)__popcll(ma);
            wv[w] = (uint32_t)__popcll(mv);
        }
        __syncthreads();
        uint32_t before_a = 0, before_v = 0, tile_a = 0, tile_v = 0;
        for (int k = 0; k < kWaves; ++k) {
            if (k < w) {
                before_a += wa[k];
                before_v += wv[k];
            }
            tile_a += wa[k];
            tile_v += wv[k];
        }
        if (valid) {
            const uint64_t below = lanes_below();
            const uint32_t o = adj ? run_a + before_a + (uint32_t)__popcll(ma & below)
                                   : run_n + (before_v - before_a) + (uint32_t)__popcll(mv & ~ma & below);
            dst_node[o] = node;
            dst_query[o] = query;
        }
        run_a += tile_a;
        run_n += tile_v - tile_a;
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst_size = ne;
}

void launch_adj_partition(DevState& d, int32_t n, hipStream_t s) {
    const FrontIn in = cur_front(d);
    const unsigned tiles = (unsigned)std::max<int64_t>(1, (d.front_cap[1] + kAdjTile - 1) / kAdjTile);
    hipLaunchKernelGGL(k_adj_count, dim3(tiles), dim3(kBlock), 0, s, d, in, d.adj_cnt);
    launch_exscan(d, d.adj_cnt, d.adj_off, tiles, d.ctr + C_ADJ, s);
    const int nb = d.fcur[1] ^ 1;
    hipLaunchKernelGGL(k_adj_scatter, dim3(tiles), dim3(kBlock), 0, s, d, in, d.adj_off, d.fnode[1][nb],
                       d.fquery[1][nb], d.fsize + 2 + nb);
    d.fcur[1] = nb;
    d.adj_version = d.front_version;
}

void launch_keep(DevState& d, int32_t n, hipStream_t s, bool counters_zeroed) {
    // inter-object fronts pair triangles of two objects: no shared vertices, no partition
    const int implicit = d.inter ? 0 : 1;
    if (implicit && d.adj_version != d.front_version) launch_adj_partition(d, n, s);
    if (!counters_zeroed) reset_step_counters(d, s);
    const unsigned sh = tuned_shape(TK_KEEP, n), g = shape_grid(sh), w = shape_waves(sh);
    const FrontIn in = cur_front(d);
#define BVHCD_KEEP(U, W) hipLaunchKernelGGL((k_keep<U, W>), dim3(g), dim3(64 * W), 0, s, d, in, implicit)
    switch (shape_unroll(sh) * 16 + w) {
        case 4 * 16 + 8: BVHCD_KEEP(4, 8); break;
        case 4 * 16 + 2: BVHCD_KEEP(4, 2); break;
        case 4 * 16 + 4: BVHCD_KEEP(4, 4); break;
        case 2 * 16 + 8: BVHCD_KEEP(2, 8); break;
        case 2 * 16 + 2: BVHCD_KEEP(2, 2); break;
        case 2 * 16 + 4: BVHCD_KEEP(2, 4); break;
        case 1 * 16 + 8: BVHCD_KEEP(1, 8); break;
        case 1 * 16 + 2: BVHCD_KEEP(1, 2); break;
        default: BVHCD_KEEP(1, 4); break;
    }
#undef BVHCD_KEEP
}

// ---------------------------------------------------------------- incremental front restructure
// The intent of BvttFrontLooseIntra::restructure (collision/bvtt_front/BvttFrontLooseIntra.cu:330-400,
// separateFronts :172-211, BvttFrontRestructureKernels.cu:12-90; unreachable in the shipped
// intra path): after a subtree rebuild keep the front entries it does not touch and re-walk
// only what it invalidates, instead of regenerating the whole front.
// Each query's front is a cut of the leaves to its right. For a rebuilt subtree R, leaves [a, b]:
//  * an entry strictly inside R of a query outside [a, b] is one piece of that query's cut of R;
//    the piece whose leftmost leaf is a becomes the walk (query, R), the others are dropped;
//  * an entry strictly inside R of a query inside [a, b] is dropped; that query's part inside
//    R is re-walked from lca[q + 1] to R's escape lca[b + 1];
//  * all other entries (outside every rebuilt subtree, at an ancestor, or at R) stay.
// Classification runs on the old topology (rt_of_leaf and the old rangex), before emission.
__global__ void __launch_bounds__(kBlock) k_restr_classify(DevState d, FrontIn in, FrontSink fi, FrontSink fe,
                                                           PairSink tasks) {
    __shared__ Stages st;
    stages_init(st);
    const Out out = make_out(st, fi, fe, d);
    const uint32_t ni = min(*in.size[0], in.cap[0]), ne = min(*in.size[1], in.cap[1]), total = ni + ne;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gridDim.x * blockDim.x) {
        bool task = false;
        int32_t q, r;
        if (k < ni) {
            q = in.query[0][k];
            const int32_t x = (int32_t)in.node[0][k];
            const int32_t lx = d.int_rx[x];
            r = d.rt_of_leaf[lx];
            const bool inside = r >= 0 && x > r;  // x < r: an ancestor of r
            stage_push(fi, out.ovf, out.sb[0], out.sf[0], out.scap, !inside, make_int2(q, x));
            task = inside && d.rt_of_leaf[q] != r && lx == d.int_rx[r];
        } else {
            q = in.query[1][k - ni];
            const int32_t j = (int32_t)in.node[1][k - ni];
            r = d.rt_of_leaf[j];
            const bool inside = r >= 0;
            stage_push(fe, out.ovf, out.sb[1], out.sf[1], out.scap, !inside, make_int2(q, j));
            task = inside && d.rt_of_leaf[q] != r && j == d.int_rx[r];
        }
        stage_push(tasks, out.ovf, out.sb[2], out.sf[2], out.scap, task, make_int2(q, r));
    }
    stage_final(fi, out.ovf, out.sb[0], out.sf[0]);
    stage_final(fe, out.ovf, out.sb[1], out.sf[1]);
    stage_final(tasks, out.ovf, out.sb[2], out.sf[2]);
}

void launch_front_restr_prepare(DevState& d, int32_t n, hipStream_t s) {
    front_begin(d, s);
    const FrontIn in = cur_front(d);
    FrontSink fi = next_sink(d, 0), fe = next_sink(d, 1);
    hipMemsetAsync(d.ctr + C_TASKS, 0, sizeof(uint32_t), s);
    // walks beyond the capacity are counted, not written (ovf bit 0): the caller compares
    // the count with the capacity and regenerates the front instead
    const PairSink tasks{d.ctr + C_TASKS, d.rtasks, (uint32_t)std::min<int64_t>(d.rtask_cap, 0xFFFFFFFFll), 0u};
    if (n >= 2) hipLaunchKernelGGL(k_restr_classify, dim3(stream_grid()), dim3(kBlock), 0, s, d, in, fi, fe, tasks);
    front_finish(d, 0, false, s);
    front_finish(d, 1, false, s);
}

// Walks of the repaired front: (query, rebuilt root) walks first, then every query inside a
// rebuilt run walks its part of that run. Entries append after the kept ones.
__global__ void __launch_bounds__(kBlock) k_restr_walk(DevState d, uint32_t ntasks, int32_t n, FrontSink fi,
                                                       FrontSink fe) {
    __shared__ Stages st;
    stages_init(st);
    const Out out = make_out(st, fi, fe, d);
    const uint32_t total = ntasks + (uint32_t)n;
    uint32_t visits = 0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gridDim.x * blockDim.x) {
        int32_t q, start, stop;
        if (k < ntasks) {
            const int2 t = d.rtasks[k];
            q = t.x;
            start = t.y << 1;
            stop = d.leaf_lca[d.int_ry[t.y] + 1];
        } else {
            q = (int32_t)(k - ntasks);
            const int32_t r = d.rt_of_leaf[q];
            if (r < 0) continue;
            const int32_t b = d.int_ry[r];
            if (q >= b) continue;
            start = d.leaf_lca[q + 1];
            stop = d.leaf_lca[b + 1];
        }
        const float4 lo = d.leaf[q].lo, hi = d.leaf[q].hi;
        visits += stackless<true>(d, out, q, rec_box(lo, hi), start, stop);
    }
    out.final_flush();
    add_visits(d, visits);
}

void launch_front_restr_apply(DevState& d, int32_t n, uint32_t ntasks, hipStream_t s) {
    reset_step_counters(d, s);
    if (n < 2) return;
    // pairs of the kept entries (KEEP over the kept part only)
    hipMemcpyAsync(d.ctr + C_KEPT_I, d.fsize + d.fcur[0], sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    hipMemcpyAsync(d.ctr + C_KEPT_E, d.fsize + 2 + d.fcur[1], sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    FrontIn kept = cur_front(d);
    kept.size[0] = d.ctr + C_KEPT_I;
    kept.size[1] = d.ctr + C_KEPT_E;
    hipLaunchKernelGGL((k_keep<1, kWaves>), dim3(stream_grid()), dim3(kBlock), 0, s, d, kept, 0);
    // the walks write segments into the spare buffers, then append after the kept entries
    front_begin(d, s);
    FrontSink fi = next_sink(d, 0), fe = next_sink(d, 1);
    hipLaunchKernelGGL(k_restr_walk, dim3(stream_grid()), dim3(kBlock), 0, s, d, ntasks, n, fi, fe);
    front_finish(d, 0, true, s);
    front_finish(d, 1, true, s);
}

// ---------------------------------------------------------------- REORDER (balance)
// Pairs as in KEEP from the unchanged cut; then the internal front is counting-sorted by node
// with each bucket's non-overlapping entries first. Its bucket starts go to scratch (`cursor`):
// intbeg keeps the last ordering's offsets, which the quality check reads (the front need not
// have been node-ordered since: an incremental restructure appends entries).
void launch_reorder(DevState& d, int32_t n, hipStream_t s) {
    launch_keep(d, n, s);
    if (n < 2) return;
    order_front(d, 0, n - 1, d.intcnt, d.cursor, true, s);
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

// Narrow-phase test of a self-CD broad-phase pair of leaves `lv`; `pr` receives the pair in
// canonical order (the triangle with the lower id first, and as P of the test).
__device__ inline bool narrow_self(const int32_t* __restrict__ prim, const int32_t* __restrict__ lfaces,
                                   const uint8_t* __restrict__ finite, const float* __restrict__ verts, int2 lv,
                                   int2& pr) {
    int32_t ta = prim[lv.x], tb = prim[lv.y], la = lv.x, lb = lv.y;
    if (ta > tb) {
        const int32_t t = ta; ta = tb; tb = t;
        const int32_t l = la; la = lb; lb = l;
    }
    pr = make_int2(ta, tb);
    const int32_t* fa = lfaces + 3 * (int64_t)la;
    const int32_t* fb = lfaces + 3 * (int64_t)lb;
    const int32_t a0 = fa[0], a1 = fa[1], a2 = fa[2], b0 = fb[0], b1 = fb[1], b2 = fb[2];
    // A shared vertex has bitwise-equal translated coordinates on both sides, so while every
    // product stays finite (all |coordinates| <= 1e8) its projection lies in both intervals on
    // every axis: the SAT always reports a hit, and the test is skipped.
    const bool share = a0 == b0 || a0 == b1 || a0 == b2 || a1 == b0 || a1 == b1 || a1 == b2 || a2 == b0 ||
                       a2 == b1 || a2 == b2;
    if (share && finite[la] && finite[lb]) return true;
    return sat_tri_tri(ld3(verts, a0), ld3(verts, a1), ld3(verts, a2), ld3(verts, b0), ld3(verts, b1),
                       ld3(verts, b2));
}

template <int W>
__global__ void __launch_bounds__(64 * W) k_narrow(DevState d) {
    __shared__ int2 sbuf[W][kStage];
    __shared__ uint32_t sfill[W][2];
    __shared__ uint32_t pre[kPairSegs + 1];
    const int w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        sfill[w][0] = 0;
        sfill[w][1] = rotor_start<W>();
    }
    const SegPairSink sink = pair_segments(d, 1, 8u);
    lds_int2* const sb = (lds_int2*)sbuf[w];
    lds_u32* const sf = (lds_u32*)sfill[w];
    const uint32_t npairs = seg_prefix(d.pcnt, (uint32_t)d.pair_seg_cap, pre);
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < npairs; k += gridDim.x * blockDim.x) {
        const int2 lv = d.bp[seg_slot(pre, k, d.pair_seg_cap)];  // leaves
        if (d.inter) {  // (query leaf of the other object, leaf of this tree): this tree's triangle is P
            const int32_t* fa = d.lfaces + 3 * (int64_t)lv.y;
            const int32_t* fb = d.qlfaces + 3 * (int64_t)lv.x;
            const bool hit = sat_tri_tri(ld3(d.verts, fa[0]), ld3(d.verts, fa[1]), ld3(d.verts, fa[2]),
                                         ld3(d.qverts, fb[0]), ld3(d.qverts, fb[1]), ld3(d.qverts, fb[2]));
            stage_push(sink, &d.ctr[C_OVF], sb, sf, kStage, hit,
                       make_int2(d.prim_idx[lv.y], d.qprim[lv.x]));
            continue;
        }
        int2 pr;
        const bool hit = narrow_self(d.prim_idx, d.lfaces, d.tri_finite, d.verts, lv, pr);
        stage_push(sink, &d.ctr[C_OVF], sb, sf, kStage, hit, pr);
    }
    stage_final(sink, &d.ctr[C_OVF], sb, sf);
}

void launch_narrow(DevState& d, hipStream_t s, bool counters_zeroed) {
    // kernels, not memsets: the batched frame graph records this
    if (!counters_zeroed) fill_u32(d.pcnt + kPairSegs * kCntStride, kPairSegs * kCntStride, 0u, s);
    const unsigned sh = tuned_shape(TK_NARROW, d.maxT), g = shape_grid(sh);
    switch (shape_waves(sh)) {
        case 8: hipLaunchKernelGGL(k_narrow<8>, dim3(g), dim3(512), 0, s, d); break;
        case 2: hipLaunchKernelGGL(k_narrow<2>, dim3(g), dim3(128), 0, s, d); break;
        default: hipLaunchKernelGGL(k_narrow<4>, dim3(g), dim3(256), 0, s, d); break;
    }
}

// Totals of both segmented pair lists (requested counts: what the segments hold plus what none
// took, so an overflowing list reports what it needed) and their largest segment: ctr[C_BP],
// ctr[C_NP], ctr[C_BP_SEGMAX], ctr[C_NP_SEGMAX]. Implicit (shared-vertex) pairs are added to
// both totals; with a narrow phase they are hits.
__global__ void k_pair_totals(DevState d, int narrow) {
    const int kind = threadIdx.x >> 6, g = lane_id();
    const uint32_t* cnt = d.pcnt + kind * kPairSegs * kCntStride;
    uint32_t v = cnt[g * kCntStride] + (g == 0 ? cnt[kDroppedWord] : 0u), m = cnt[g * kCntStride];
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_down(v, off);
        m = max(m, __shfl_down(m, off));
    }
    if (g == 0) {
        const uint32_t impl = d.ctr[C_IMPLICIT];
        if (kind == 1) d.ctr[C_IMPL_NP] = narrow ? impl : 0u;
        d.ctr[kind ? C_NP : C_BP] = v + (kind == 0 || narrow ? impl : 0u);
        d.ctr[kind ? C_NP_SEGMAX : C_BP_SEGMAX] = m;
    }
}
void launch_pair_totals(DevState& d, hipStream_t s, bool narrow) {
    hipLaunchKernelGGL(k_pair_totals, dim3(1), dim3(128), 0, s, d, narrow ? 1 : 0);
}
void clear_np(DevState& d, hipStream_t s) {
    fill_u32(d.pcnt + kPairSegs * kCntStride, kPairSegs * kCntStride, 0u, s);
}

// Pair k of a list, in the virtual order [listed segments | implicit head of the external front];
// leaf-id pairs (the broad-phase list, the implicit ones) mapped to triangle ids, lower first.
__device__ inline int2 list_pair(const DevState& d, int kind, const uint32_t* pre, uint32_t listed, uint32_t k) {
    int2 p;
    if (k < listed) {
        p = (kind ? d.np : d.bp)[seg_slot(pre, k, d.pair_seg_cap)];
        if (kind == 1) return p;
    } else {
        const int c = d.fcur[1];
        p = make_int2(d.fquery[1][c][k - listed], (int32_t)d.fnode[1][c][k - listed]);
    }
    if (d.inter) return make_int2(d.prim_idx[p.y], d.qprim[p.x]);  // (this tree's triangle, query's)
    const int32_t a = d.prim_idx[p.x], b = d.prim_idx[p.y];
    return a < b ? make_int2(a, b) : make_int2(b, a);
}

// Contiguous copy of pairs [first, first + count) of a list into dst.
__global__ void __launch_bounds__(kBlock) k_pair_gather(DevState d, int kind, int2* __restrict__ dst, uint64_t first,
                                                        uint64_t count) {
    __shared__ uint32_t pre[kPairSegs + 1];
    const uint32_t listed = seg_prefix(d.pcnt + kind * kPairSegs * kCntStride, (uint32_t)d.pair_seg_cap, pre);
    const uint64_t total = (uint64_t)listed + d.ctr[kind ? C_IMPL_NP : C_IMPLICIT];
    const uint64_t end = min(first + count, total);
    for (uint64_t k = first + blockIdx.x * blockDim.x + threadIdx.x; k < end; k += gridDim.x * blockDim.x)
        dst[k - first] = list_pair(d, kind, pre, listed, (uint32_t)k);
}
void launch_pair_gather(DevState& d, int kind, int2* dst, int64_t first, int64_t count, hipStream_t s) {
    hipLaunchKernelGGL(k_pair_gather, dim3(2048), dim3(kBlock), 0, s, d, kind, dst, (uint64_t)first,
                       (uint64_t)count);
}

// ---------------------------------------------------------------- pair-set digest
__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__global__ void __launch_bounds__(kBlock) k_digest(DevState d, int kind, uint64_t* out) {
    __shared__ uint32_t pre[kPairSegs + 1];
    const uint32_t listed = seg_prefix(d.pcnt + kind * kPairSegs * kCntStride, (uint32_t)d.pair_seg_cap, pre);
    const uint32_t n = listed + d.ctr[kind ? C_IMPL_NP : C_IMPLICIT];
    uint64_t acc = 0;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int2 p = list_pair(d, kind, pre, listed, k);
        acc += mix64(((uint64_t)(uint32_t)p.x << 32) | (uint32_t)p.y);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off);
    if (lane_id() == 0 && acc) atomicAdd((unsigned long long*)out, (unsigned long long)acc);
}

void launch_digest(DevState& d, int filtered, hipStream_t s) {
    uint64_t* out = ctr64(d.ctr) + C64_DIGEST;
    fill_u32(reinterpret_cast<uint32_t*>(out), 2, 0u, s);
    hipLaunchKernelGGL(k_digest, dim3(1024), dim3(kBlock), 0, s, d, filtered ? 1 : 0, out);
}

}  // namespace bvhcd
