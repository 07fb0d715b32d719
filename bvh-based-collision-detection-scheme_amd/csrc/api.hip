// C ABI of the gfx950 collision-detection pipeline: context, per-frame maintenance, scheduler.
//
// Mirrors LBvhFixedDeformable::maintain (collision/lbvh/LBvhFixedDeformable.cu:68-87) and
// BvttFrontLooseIntra::maintain (collision/bvtt_front/BvttFrontLooseIntra.cu:69-126) of
// littlemine/BVH-based-Collision-Detection-Scheme (Source/Project/Mine/), with the plain
// schedule of BenchmarkLogic::maintainScheme (system/Logic/BenchmarkLogic.hpp:102-117).
// All work is stream-ordered on the context stream; only get_pairs, pair_digest, get_stats,
// the export calls, timing read-back and the restructure decision wait for the device.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bvhcd_internal.hpp"
#include "kernels.hpp"
#include "learned.hpp"

using namespace bvhcd;

namespace {
// stage ids of bvhcd_stats.stage_ms / stage_bytes
enum Stage { S_BUILD = 0, S_REFIT = 1, S_RESTR = 2, S_FRONT = 3, S_ORDER = 4, S_NARROW = 5, S_SCHED = 6, S_TOTAL = 7 };
constexpr int kEv = 16;   // begin/end event per stage
constexpr int kLog = 8;   // device log words per frame
// Learned schedule: A front this close to its capacity is replaced by a fresh one at the next
// check.
constexpr double kLearnedCapacityShare = 0.75;
constexpr int kLearnedMaxBackoff = 3;
}  // namespace

struct bvhcd_ctx {
    bvhcd_config cfg{};
    DevState d;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::string err;
    int64_t T = 0, V = 0;
    bool built = false, front_valid = false, restr_pending = false;
    bool learned_frame = false;  // set by the learned scheduler for its decision frames
    int learned_backoff = 0;     // decision interval = bvh_cycle << backoff (rejected selections)
    bool repair_pending = false; // front classified for the incremental front restructure
    uint32_t repair_tasks = 0;   // its (query, subtree) walks
    int64_t frame = 0;
    int last_bvh = -1, last_front = -1;
    int64_t restr_sub = 0, restr_leaf = 0;
    // timing: kEv events per recorded frame, device counter log, executed ops
    int64_t tcap = 0, tn = 0;
    std::vector<hipEvent_t> tev;
    std::vector<uint8_t> tused;
    uint64_t* dlog = nullptr;
    std::vector<int32_t> tops;
    std::vector<int64_t> trestr;
    std::vector<void*> allocs;
    std::vector<void*> list_allocs;  // fronts, pair lists and what their capacities size
    LearnedState* learned = nullptr;
    float* boxes = nullptr;      // box mode input (6 floats per triangle), allocated on first use
    float* own_verts = nullptr;  // the context's vertex buffers (d.verts may borrow the caller's);
    float* own_verts2 = nullptr; // host uploads alternate between the two (vbuf: the current one)
    int vbuf = 0;
    hipStream_t copy_stream = nullptr;  // host uploads, overlapping the steps still running
    hipEvent_t copied[2] = {nullptr, nullptr};
    hipEvent_t consumed[2] = {nullptr, nullptr};  // the last step that read each buffer is done
    // learned schedule: the time of every step is taken with two events; at front-cycle points
    // the last KEEP frame's time is compared with the first one after the last rebuild
    hipEvent_t lev[2] = {nullptr, nullptr};
    double keep_base = -1.0;     // ms of the first KEEP frame after the last rebuild
    double keep_excess = 0.0;    // ms above keep_base accumulated since that rebuild
    double rebuild_ms = -1.0;    // ms of the last rebuild frame (BUILD + GENERATE)
    bool rebuild_timed = false;  // the last step was a rebuild whose time is still to be read
    bool totals_pending = false; // pair totals of the last step not summed yet (done on demand)
    bool step_zeroed = false;    // this step's refit zeroed the step counters for its KEEP
};

namespace {

bool is_learned(int sched) { return sched == BVHCD_SCHED_LEARNED || sched == BVHCD_SCHED_LEARNED_NO_NET; }

int fail(bvhcd_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail((c), BVHCD_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
bool dalloc(bvhcd_ctx* c, T*& p, size_t n) {
    void* q = nullptr;
    if (n == 0) n = 1;
    if (hipMalloc(&q, n * sizeof(T)) != hipSuccess) return false;
    c->allocs.push_back(q);
    p = static_cast<T*>(q);
    return true;
}

void free_all(bvhcd_ctx* c) {
    for (void* p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    for (void* p : c->list_allocs) (void)hipFree(p);
    c->list_allocs.clear();
}

// Fronts, pair lists and what their capacities size (the sort's digit table, the shared-vertex
// partition tiles, the wave-ordered production's chunk records), in their own allocation list so
// that a context can grow them (cfg.auto_grow). Their contents are not kept.
bool alloc_lists(bvhcd_ctx* c, int64_t ci, int64_t ce, int64_t cp) {
    DevState& d = c->d;
    const int64_t T = c->d.maxT;
    for (void* p : c->list_allocs) (void)hipFree(p);
    c->list_allocs.clear();
    auto lalloc = [c](auto*& p, size_t n) {
        void* q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(*p)) != hipSuccess) return false;
        c->list_allocs.push_back(q);
        p = static_cast<std::remove_reference_t<decltype(p)>>(q);
        return true;
    };
    // fronts are produced in kPairSegs equal segments, then gathered: capacities are multiples
    d.front_cap[0] = kPairSegs * ((std::max<int64_t>(ci, 1) + kPairSegs - 1) / kPairSegs);
    d.front_cap[1] = kPairSegs * ((std::max<int64_t>(ce, 1) + kPairSegs - 1) / kPairSegs);
    // pair lists: kPairSegs equal segments (an explicit capacity is split evenly)
    d.pair_seg_cap = std::max<int64_t>((std::max<int64_t>(cp, 1) + kPairSegs - 1) / kPairSegs, 1);
    d.pair_cap = d.pair_seg_cap * kPairSegs;
    bool ok = true;
    for (int k = 0; k < 2; ++k)
        for (int b = 0; b < 2; ++b) {
            ok = ok && lalloc(d.fnode[k][b], (size_t)d.front_cap[k]);
            ok = ok && lalloc(d.fquery[k][b], (size_t)d.front_cap[k]);
        }
    // digit table of the largest sort (front ordering) and the scan block sums
    const int64_t max_sort = std::max<int64_t>(T, std::max(d.front_cap[0], d.front_cap[1]));
    d.sort_hist_n = (size_t)(256 * ((max_sort + 4095) / 4096));
    d.scan_tmp_n = (size_t)((std::max<int64_t>((int64_t)d.sort_hist_n, T + 2) + 4095) / 4096 + 1);
    ok = ok && lalloc(d.sort_hist, d.sort_hist_n) && lalloc(d.scan_tmp, d.scan_tmp_n);
    ok = ok && lalloc(d.bp, (size_t)d.pair_cap) && lalloc(d.np, (size_t)d.pair_cap) &&
         lalloc(d.pair_tmp, (size_t)d.pair_cap);
    ok = ok && lalloc(d.adj_cnt, (size_t)(d.front_cap[1] / kAdjTile + 2)) &&
         lalloc(d.adj_off, (size_t)(d.front_cap[1] / kAdjTile + 2));
    // wave-ordered GENERATE: one chunk record per staging flush (every flush but a wave's last
    // holds more than 128 entries; the last ones may all land in one segment)
    d.chunk_seg_cap = std::max(d.front_cap[0], d.front_cap[1]) / kPairSegs / 128 + d.wave_cap + 16;
    for (int k = 0; k < 2; ++k) ok = ok && lalloc(d.fchunk[k], (size_t)(kPairSegs * d.chunk_seg_cap));
    ok = ok && lalloc(d.clist, (size_t)(kPairSegs * d.chunk_seg_cap));
    // the lists start empty
    d.fcur[0] = d.fcur[1] = 0;
    d.adj_version = -1;
    d.front_version++;
    return ok && hipMemsetAsync(d.fsize, 0, 4 * sizeof(uint32_t), c->stream) == hipSuccess;
}

int allocate(bvhcd_ctx* c) {
    DevState& d = c->d;
    const int64_t T = std::max<int64_t>(c->cfg.max_tris, 1), V = std::max<int64_t>(c->cfg.max_verts, 1);
    d.maxT = T;
    d.maxV = V;
    bool ok = true;
    ok &= dalloc(c, d.faces, 3 * T);
    ok &= dalloc(c, d.verts, 3 * V);
    c->own_verts = d.verts;
    ok &= dalloc(c, c->own_verts2, 3 * V);
    ok &= dalloc(c, d.codes, T);
    ok &= dalloc(c, d.keys_tmp, T);
    ok &= dalloc(c, d.keys_tmp2, T);
    ok &= dalloc(c, d.prim_idx, T);
    ok &= dalloc(c, d.vals_tmp, T);
    ok &= dalloc(c, d.prim_map, T);
    ok &= dalloc(c, d.inode, T);
    ok &= dalloc(c, d.leaf, T);
    ok &= dalloc(c, d.int_lc, T);
    ok &= dalloc(c, d.int_rc, T);
    ok &= dalloc(c, d.int_par, T);
    ok &= dalloc(c, d.int_mark, T);
    ok &= dalloc(c, d.int_rx, T);
    ok &= dalloc(c, d.int_ry, T);
    ok &= dalloc(c, d.itopo, T);
    ok &= dalloc(c, d.tri_finite, T);
    ok &= dalloc(c, d.lfaces, 3 * T);
    ok &= dalloc(c, d.sa_ref, T);
    ok &= dalloc(c, d.sa_ref_leaf, T);
    ok &= dalloc(c, d.built_frame, T);
    ok &= dalloc(c, d.rp_cstart, (size_t)refit_chunks((int32_t)T) + 1);
    ok &= dalloc(c, d.rp_cinfo, (size_t)refit_chunks((int32_t)T));
    ok &= dalloc(c, d.rp_ent, T);
    ok &= dalloc(c, d.rp_tent, T);
    ok &= dalloc(c, d.rp_tlist, (size_t)T + kRefitTopLds);
    ok &= dalloc(c, d.rp_tpos, T);
    ok &= dalloc(c, d.rp_tlev, 1024);
    ok &= dalloc(c, d.rp_tinfo, 6);
    ok &= dalloc(c, d.rp_seg, kRefitTopGrid);
    ok &= dalloc(c, d.rp_tsz, T);
    ok &= dalloc(c, d.rp_tseg, T);
    ok &= dalloc(c, d.rp_tdone, 1);
    ok &= dalloc(c, d.rp_tdeep, 1);
    ok &= dalloc(c, d.rp_tcount, 1);
    ok &= dalloc(c, d.rp_tcb, 6 * (size_t)refit_tcb_boxes((int64_t)T));
    ok &= dalloc(c, d.rp_cbad, (size_t)refit_chunks((int32_t)T));
    ok &= dalloc(c, d.leaf_par, T);
    ok &= dalloc(c, d.leaf_mark, T);
    ok &= dalloc(c, d.leaf_lca, T + 1);
    ok &= dalloc(c, d.s_a, T);
    ok &= dalloc(c, d.s_b, T);
    ok &= dalloc(c, d.s_par, T);
    ok &= dalloc(c, d.s_lc, T);
    ok &= dalloc(c, d.s_rc, T);
    ok &= dalloc(c, d.s_lpar, T);
    ok &= dalloc(c, d.s_top, T);
    ok &= dalloc(c, d.depths, T + 1);
    ok &= dalloc(c, d.offsets, T + 1);
    ok &= dalloc(c, d.tkmap, T);
    ok &= dalloc(c, d.scene_box, 8);
    ok &= dalloc(c, d.scene_box_u, 8);
    ok &= dalloc(c, d.fsize, 4);
    ok &= dalloc(c, d.intcnt, T + 1);
    ok &= dalloc(c, d.extcnt, T + 2);
    ok &= dalloc(c, d.intbeg, T + 2);
    ok &= dalloc(c, d.extbeg, T + 2);
    ok &= dalloc(c, d.intback, T + 1);
    ok &= dalloc(c, d.cursor, T + 1);
    ok &= dalloc(c, d.snapshot, T);
    ok &= dalloc(c, d.intrange, T + 1);
    ok &= dalloc(c, d.rtroot, T + 1);
    ok &= dalloc(c, d.intmark, T + 1);
    ok &= dalloc(c, d.rt_of_leaf, T + 1);
    ok &= dalloc(c, d.r_leaf_lo, T + 1);
    ok &= dalloc(c, d.r_leaf_hi, T + 1);
    ok &= dalloc(c, d.r_split_lo, T);
    ok &= dalloc(c, d.r_split_hi, T);
    ok &= dalloc(c, d.r_base, T);
    ok &= dalloc(c, d.r_keep_par, T);
    ok &= dalloc(c, d.r_keep_mark, T);
    ok &= dalloc(c, d.r_keys, T);
    ok &= dalloc(c, d.r_keys2, T);
    ok &= dalloc(c, d.ctr, 32);
    if (!ok) return fail(c, BVHCD_E_OOM, "device allocation failed (fixed arrays)");
    // capacities: explicit, or a multiple of T bounded by free HBM and the 32-bit counters
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const int64_t lim = 0x7FFFFFFFll;
    int64_t ci = c->cfg.int_front_cap, ce = c->cfg.ext_front_cap, cp = c->cfg.pair_cap;
    if (ci <= 0) ci = 48 * T + (1 << 20);
    if (ce <= 0) ce = 24 * T + (1 << 20);
    if (cp <= 0) cp = 24 * T + (1 << 20);
    ci = std::min(ci, lim); ce = std::min(ce, lim); cp = std::min(cp, lim);
    // 2 buffers x 8 B for each front, 3 x 8 B for pairs (broad, narrow, read-out copy)
    const double need = 16.0 * (double)(ci + ce) + 24.0 * (double)cp;
    const double budget = 0.7 * (double)free_b;
    if (need > budget && (c->cfg.int_front_cap <= 0 || c->cfg.ext_front_cap <= 0 || c->cfg.pair_cap <= 0)) {
        const double f = budget / need;
        if (c->cfg.int_front_cap <= 0) ci = (int64_t)(ci * f);
        if (c->cfg.ext_front_cap <= 0) ce = (int64_t)(ce * f);
        if (c->cfg.pair_cap <= 0) cp = (int64_t)(cp * f);
    }
    d.rtask_cap = std::min<int64_t>(4 * T + (1 << 20), lim);
    ok &= dalloc(c, d.rtasks, d.rtask_cap);
    ok &= dalloc(c, d.pcnt, 2 * kPairSegs * kCntStride);
    ok &= dalloc(c, d.fcnt, 2 * kPairSegs * kCntStride);
    d.wave_cap = (int64_t)grid_for(std::max<int64_t>(T - 1, 1)) * (kBlock / 64);
    for (int k = 0; k < 2; ++k) ok &= dalloc(c, d.wave_tot[k], (size_t)d.wave_cap);
    ok &= dalloc(c, d.wave_base, (size_t)d.wave_cap);
    ok &= dalloc(c, d.wave_nch, (size_t)d.wave_cap);
    ok &= dalloc(c, d.wave_cbase, (size_t)d.wave_cap);
    if (!ok) return fail(c, BVHCD_E_OOM, "device allocation failed (fixed arrays)");
    ok = alloc_lists(c, ci, ce, cp);
    if (!ok) return fail(c, BVHCD_E_OOM, "device allocation failed (fronts / pairs)");
    HIPCHK(c, hipMemsetAsync(d.ctr, 0, 32 * sizeof(uint32_t), c->stream));
    HIPCHK(c, hipMemsetAsync(d.pcnt, 0, 2 * kPairSegs * kCntStride * sizeof(uint32_t), c->stream));
    HIPCHK(c, hipMemsetAsync(d.fsize, 0, 4 * sizeof(uint32_t), c->stream));
    d.qleaf = d.leaf;  // self-CD: the tree's own leaves are the queries
    d.qprim = d.prim_idx;
    return BVHCD_OK;
}

// ---------------------------------------------------------------- timing
bool timing_on(const bvhcd_ctx* c) { return c->tcap > 0 && c->tn < c->tcap; }
void mark(bvhcd_ctx* c, int stage, bool end) {
    if (!timing_on(c)) return;
    const int64_t k = c->tn * kEv + 2 * stage + (end ? 1 : 0);
    (void)hipEventRecord(c->tev[k], c->stream);
    c->tused[k] = 1;
}
void begin(bvhcd_ctx* c, int stage) { mark(c, stage, false); }
void end(bvhcd_ctx* c, int stage) { mark(c, stage, true); }

// Sets *bad when a face vertex id lies outside [0, V).
__global__ void k_check_faces(int64_t n, const int32_t* __restrict__ f, int32_t V, uint32_t* bad) {
    bool any = false;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        any |= f[i] < 0 || f[i] >= V;
    if (any) atomicOr(bad, 1u);
}

__global__ void k_log_frame(DevState d, int cur_int, int cur_ext, uint64_t* out) {
    if (threadIdx.x != 0) return;
    out[0] = ctr64(d.ctr)[C64_VISITS];
    out[1] = d.ctr[C_BP];
    out[2] = d.ctr[C_NP];
    out[3] = d.fsize[cur_int];
    out[4] = d.fsize[2 + cur_ext];
    out[5] = d.ctr[C_OVF];
    out[6] = d.ctr[C_IMPLICIT];
    out[7] = d.ctr[C_IMPL_NP];
}

// Pair totals (C_BP, C_NP and the implicit narrow-phase count) are summed from the segment
// counters only when something reads them, not on every step.
void ensure_totals(bvhcd_ctx* c) {
    if (!c->totals_pending) return;
    launch_pair_totals(c->d, c->stream, c->cfg.narrow_phase != 0);
    c->totals_pending = false;
}

void log_frame(bvhcd_ctx* c, int bvh_op, int front_op) {
    if (!timing_on(c)) return;
    ensure_totals(c);
    hipLaunchKernelGGL(k_log_frame, dim3(1), dim3(64), 0, c->stream, c->d, c->d.fcur[0], c->d.fcur[1],
                       c->dlog + kLog * c->tn);
    c->tops.push_back(bvh_op);
    c->tops.push_back(front_op);
    c->trestr.push_back(c->restr_leaf);
    c->tn++;
}

// Algorithmic HBM bytes per stage of one frame:
//   B_build  = 24V + 16T (box + codes) + 4T + 16kT (k = 4 sort passes) + 4T + 24I + 12T + 24(T+I)
//   B_refit  = 12V + 40T + 88I
//   B_restr  = refit + 16 (k+3) L_r over the restructured leaves L_r (keys, sort, emission)
//   B_front  = 32 N_in + 36 N_visit + 12 N_out + 8 P_bp
//   B_order  = 16 N_out + 16T
//   B_narrow = 104 P_bp + 8 P_np
// Algorithmic bytes per stage (read-once / write-once compulsory traffic). The front stage's
// event brackets include the node-id sort of GENERATE / UPDATE / REORDER (16 bytes per entry
// and 16 per node over the passes' reads and writes), so its bytes count there; the
// order+quality stage only times the quality snapshot (16 bytes per node).
// Implicit shared-vertex pairs (impl of bp, impl_np of np) are neither read as front entries nor
// written or re-read as pairs, so they carry no bytes.
void frame_bytes(double T, double V, int bvh_op, int front_op, double visits, double fi0, double fe0, double fi1,
                 double fe1, double bp, double np, double restr_leaves, bool ordered, double impl, double impl_np,
                 double* out) {
    const double I = T > 1 ? T - 1 : 0;
    bp -= impl;
    np -= impl_np;
    for (int i = 0; i < 8; ++i) out[i] = 0;
    if (bvh_op == BVHCD_BVH_BUILD)
        out[S_BUILD] = 24 * V + 16 * T + (4 * T + 16.0 * 4 * T) + (4 * T + 24 * I + 12 * T + 24 * (T + I));
    if (bvh_op == BVHCD_BVH_REFIT || bvh_op == BVHCD_BVH_BUILD) out[S_REFIT] = 12 * V + 40 * T + 88 * I;
    if (bvh_op == BVHCD_BVH_RESTRUCTURE) out[S_RESTR] = 12 * V + 40 * T + 88 * I + 16.0 * 7 * restr_leaves;
    if (front_op == BVHCD_FRONT_RESTRUCTURE) out[S_RESTR] += 16 * (fi0 + fe0);  // front classification
    const bool whole = front_op == BVHCD_FRONT_PURE_BVH_CD || front_op == BVHCD_FRONT_GENERATE;
    const double nin = whole ? T : fi0 + fe0 - impl;
    const double nout = (front_op == BVHCD_FRONT_KEEP || front_op == BVHCD_FRONT_PURE_BVH_CD) ? 0 : fi1 + fe1;
    out[S_FRONT] = 32 * nin + 36 * visits + 12 * nout + 8 * bp;
    if (ordered && (front_op == BVHCD_FRONT_GENERATE || front_op == BVHCD_FRONT_UPDATE))
        out[S_FRONT] += 16 * nout + 16 * T;
    if (front_op == BVHCD_FRONT_REORDER) out[S_FRONT] += 16 * fi1 + 16 * T;  // the internal front only
    if (ordered && front_op == BVHCD_FRONT_GENERATE) out[S_ORDER] = 16 * T;
    out[S_NARROW] = 104 * bp + 8 * np;
    double tot = 0;
    for (int i = 0; i < 7; ++i) tot += out[i];
    out[S_TOTAL] = tot;
}

// ---------------------------------------------------------------- BVH maintenance
int do_build(bvhcd_ctx* c) {
    DevState& d = c->d;
    const int32_t n = (int32_t)c->T;
    hipStream_t s = c->stream;
    launch_scene_box(d, n, s);
    launch_codes(d, n, d.codes, d.prim_idx, s);
    radix_sort_u32(d, d.codes, d.prim_idx, d.keys_tmp, d.vals_tmp, n, 30, s);
    launch_prim_map(d, n, s);
    if (n >= 2) {
        launch_emit_full(d, n, s);
    } else {
        // single leaf: no internal nodes; lca = {leaf 0, end}
        const int32_t lca[2] = {1, -1};
        HIPCHK(c, hipMemcpyAsync(d.leaf_lca, lca, sizeof(lca), hipMemcpyHostToDevice, s));
        HIPCHK(c, hipStreamSynchronize(s));
        HIPCHK(c, hipMemsetAsync(d.leaf_par, 0xff, sizeof(int32_t), s));
        fill_u32(d.leaf_mark, 1, 3u, s);
    }
    launch_refit_prep(d, n, s);
    return BVHCD_OK;
}

int build_and_refit(bvhcd_ctx* c) {
    begin(c, S_BUILD);
    int st = do_build(c);
    if (st) return st;
    end(c, S_BUILD);
    begin(c, S_REFIT);
    launch_refit(c->d, (int32_t)c->T, c->stream);
    end(c, S_REFIT);
    if (c->T >= 2) update_sa_ref(c->d, (int32_t)c->T, 1, (int32_t)c->frame, c->stream);
    return BVHCD_OK;
}

// Rebuilds the marked subtrees. repair: keep the front through an incremental front
// restructure (applied by the next front op) instead of regenerating it; afterwards
// c->front_valid says whether the front survived.
int apply_restructure(bvhcd_ctx* c, int64_t nsub, int64_t nleaf, bool repair) {
    if (c->repair_pending) c->repair_pending = c->front_valid = false;  // one repair at a time
    const bool prep = repair && c->front_valid;
    begin(c, S_RESTR);
    int st = restructure_subtrees(c->d, (int32_t)c->T, c->stream, prep);
    if (st) return fail(c, st, "restructure failed");
    update_sa_ref(c->d, (int32_t)c->T, 0, (int32_t)c->frame, c->stream);
    end(c, S_RESTR);
    c->restr_sub = nsub;
    c->restr_leaf = nleaf;
    c->front_valid = false;
    if (prep) {
        uint32_t cnt = 0;
        HIPCHK(c, hipMemcpyAsync(&cnt, c->d.ctr + C_TASKS, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if ((int64_t)cnt <= c->d.rtask_cap) {  // else the walks did not fit: regenerate
            c->repair_pending = true;
            c->repair_tasks = cnt;
            c->front_valid = true;
        }
    }
    return BVHCD_OK;
}

// Restructure frame: select subtrees (reference heuristic, or the learned scheduler), then
// rebuild (> I/2 flagged internal nodes), refit (< I/8) or restructure the selection.
int run_restructure(bvhcd_ctx* c, int& bvh_op, bool& topo_changed, bool use_learned) {
    DevState& d = c->d;
    const int32_t n = (int32_t)c->T;
    const int64_t I = n - 1;
    begin(c, S_SCHED);
    int64_t nsub = 0, nint = 0, nleaf = 0;
    const bool learned = use_learned && learned_ready(c->learned);
    if (learned) {
        const int bc = c->cfg.bvh_cycle > 0 ? c->cfg.bvh_cycle : 12;
        bool rejected = false;
        int st = learned_select(c->learned, d, n, c->frame, bc << c->learned_backoff, c->stream, &nsub, &nint, &nleaf,
                                &rejected, &c->err);
        if (st < 0) return st;
        // a decision that changes nothing (no subtree worth it, or a selection that cannot pay
        // for its front repair) makes the next one come later (up to 8 cycles); an applied
        // selection resets the interval
        if (rejected || nsub == 0) c->learned_backoff = std::min(c->learned_backoff + 1, kLearnedMaxBackoff);
        else c->learned_backoff = 0;
    } else {
        launch_restr_decision(d, n, c->stream);
        uint32_t cnt[3];
        HIPCHK(c, hipMemcpyAsync(cnt, d.ctr + C_RT_SUB, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        nsub = cnt[0]; nint = cnt[1]; nleaf = cnt[2];
    }
    end(c, S_SCHED);
    if (!learned && nint > (I >> 1)) {
        int st = build_and_refit(c);
        if (st) return st;
        bvh_op = BVHCD_BVH_BUILD;
        topo_changed = true;
    } else if ((!learned && nint < (I >> 3)) || nsub == 0) {
        begin(c, S_REFIT);
        launch_refit(d, n, c->stream, true);
        c->step_zeroed = true;
        end(c, S_REFIT);
        bvh_op = BVHCD_BVH_REFIT;
    } else {
        // the learned schedule repairs the front; the plain one regenerates it, as the
        // reference's intra front does after a restructure
        int st = apply_restructure(c, nsub, nleaf, learned);
        if (st) return st;
        topo_changed = !c->front_valid;
    }
    return BVHCD_OK;
}

int run_bvh(bvhcd_ctx* c, int& bvh_op, bool& topo_changed) {
    topo_changed = false;
    const bool learned_frame = c->learned_frame;
    c->learned_frame = false;
    if (learned_frame && c->built && c->T >= 2) {
        bvh_op = BVHCD_BVH_RESTRUCTURE;
        c->restr_sub = c->restr_leaf = 0;
        return run_restructure(c, bvh_op, topo_changed, true);
    }
    if (bvh_op == BVHCD_BVH_RESTRUCTURE && (!c->built || !c->restr_pending)) bvh_op = BVHCD_BVH_REFIT;
    if (bvh_op != BVHCD_BVH_BUILD && bvh_op != BVHCD_BVH_NONE && !c->built)
        return fail(c, BVHCD_E_NOT_BUILT, "no BVH yet: first step must BUILD");
    if (bvh_op == BVHCD_BVH_NONE && !c->built) return fail(c, BVHCD_E_NOT_BUILT, "no BVH yet");
    c->restr_sub = c->restr_leaf = 0;
    switch (bvh_op) {
        case BVHCD_BVH_BUILD: {
            int st = build_and_refit(c);
            if (st) return st;
            c->built = true;
            topo_changed = true;
            break;
        }
        case BVHCD_BVH_REFIT:
            begin(c, S_REFIT);
            launch_refit(c->d, (int32_t)c->T, c->stream, true);
            c->step_zeroed = true;
            end(c, S_REFIT);
            break;
        case BVHCD_BVH_RESTRUCTURE: {
            int st = run_restructure(c, bvh_op, topo_changed, is_learned(c->cfg.scheduler));
            if (st) return st;
            break;
        }
        case BVHCD_BVH_NONE:
            break;
        default:
            return fail(c, BVHCD_E_INVALID, "bad bvh_op");
    }
    return BVHCD_OK;
}

int run_front(bvhcd_ctx* c, int& front_op) {
    DevState& d = c->d;
    const int32_t n = (int32_t)c->T;
    hipStream_t s = c->stream;
    if (c->repair_pending) {
        if (front_op == BVHCD_FRONT_PURE_BVH_CD || !c->front_valid) c->repair_pending = c->front_valid = false;
        else front_op = BVHCD_FRONT_RESTRUCTURE;
    }
    if (front_op != BVHCD_FRONT_PURE_BVH_CD && !c->front_valid) front_op = BVHCD_FRONT_GENERATE;
    c->restr_pending = false;  // the quality flag lives for one frame
    // The learned schedule reads neither the node-id ordering's bucket offsets nor the
    // front-quality heuristic: its fronts are ordered by query leaf (neighbouring queries
    // together, the best locality for KEEP and for the narrow phase over KEEP's pairs), which
    // its GENERATE produces without a global sort. (Node-ordered fronts made KEEP's descents
    // about 4 % faster on the folded 2M-triangle cloth but cost a 3-pass radix sort, 2.8 ms
    // more per rebuild.)
    const bool order = !is_learned(c->cfg.scheduler);  // node order, else query order
    const bool quality = c->cfg.enable_restructure != 0 && order;
    if (front_op != BVHCD_FRONT_KEEP && front_op != BVHCD_FRONT_PURE_BVH_CD) d.front_version++;
    switch (front_op) {
        case BVHCD_FRONT_PURE_BVH_CD:
            begin(c, S_FRONT);
            launch_pure_cd(d, n, s);
            end(c, S_FRONT);
            break;
        case BVHCD_FRONT_GENERATE:
            begin(c, S_FRONT);
            launch_generate(d, n, s, order);  // traversal + ordering
            end(c, S_FRONT);
            begin(c, S_ORDER);
            if (quality) launch_snapshot(d, n, s);
            end(c, S_ORDER);
            c->front_valid = true;
            break;
        case BVHCD_FRONT_UPDATE:
            begin(c, S_FRONT);
            launch_update(d, n, quality, s, order);  // prune/sprout + quality + ordering
            end(c, S_FRONT);
            if (quality) c->restr_pending = true;
            break;
        case BVHCD_FRONT_REORDER:
            begin(c, S_FRONT);
            launch_reorder(d, n, s);
            end(c, S_FRONT);
            break;
        case BVHCD_FRONT_KEEP:
            begin(c, S_FRONT);
            launch_keep(d, n, s, c->step_zeroed);
            end(c, S_FRONT);
            break;
        case BVHCD_FRONT_RESTRUCTURE:
            begin(c, S_FRONT);
            launch_front_restr_apply(d, n, c->repair_tasks, s);
            end(c, S_FRONT);
            c->repair_pending = false;
            break;
        default:
            return fail(c, BVHCD_E_INVALID, "bad front_op");
    }
    begin(c, S_NARROW);
    // every front op starts with the step-counter reset, which also zeroes the narrow-phase
    // segment counters (without a narrow phase they stay zero)
    if (c->cfg.narrow_phase) launch_narrow(d, s, true);
    c->totals_pending = true;
    end(c, S_NARROW);
    return BVHCD_OK;
}

// cfg.auto_grow: after the step's front maintenance, a list that overflowed is grown to 1.25x
// what the frame asked for (C_FREQ_I / C_FREQ_E for the fronts, the requested broad-phase count
// for the pair lists) and the front is generated anew on the same tree, whose pairs are then the
// frame's; up to four times, while device memory lasts. `front_op` = PURE_BVH_CD is redone as is.
int grow_and_retry(bvhcd_ctx* c, int& front_op) {
    DevState& d = c->d;
    for (int attempt = 0; attempt < 4; ++attempt) {
        uint32_t ctr[C_NUM];
        launch_pair_totals(d, c->stream, c->cfg.narrow_phase != 0);  // requested pair counts
        HIPCHK(c, hipMemcpyAsync(ctr, d.ctr, sizeof(ctr), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const uint32_t ovf = ctr[C_OVF];
        if ((ovf & 15u) == 0u || (ovf & 16u)) return BVHCD_OK;  // fits (or a device error, reported later)
        auto grown = [](int64_t cap, int64_t need) {
            return std::min<int64_t>(std::max<int64_t>(cap, need + need / 4 + kPairSegs), 0x7FFFFFFFll);
        };
        const int64_t ci = (ovf & 1u) ? grown(d.front_cap[0], ctr[C_FREQ_I]) : d.front_cap[0];
        const int64_t ce = (ovf & 2u) ? grown(d.front_cap[1], ctr[C_FREQ_E]) : d.front_cap[1];
        // the pairs of a regenerated front are all listed (no implicit head)
        const int64_t cp = (ovf & 12u) ? grown(d.pair_cap, std::max<int64_t>(ctr[C_BP], ctr[C_NP])) : d.pair_cap;
        if (!alloc_lists(c, ci, ce, cp)) {
            alloc_lists(c, d.front_cap[0], d.front_cap[1], d.pair_cap);  // the sizes that fitted before
            return fail(c, BVHCD_E_OOM, "auto_grow: device memory exhausted");
        }
        c->front_valid = false;
        c->repair_pending = false;
        if (front_op != BVHCD_FRONT_PURE_BVH_CD) front_op = BVHCD_FRONT_GENERATE;
        const int st = run_front(c, front_op);
        c->step_zeroed = false;
        if (st) return st;
    }
    return BVHCD_OK;  // still overflowing: reported by the read-back calls
}

int do_step(bvhcd_ctx* c, int bvh_op, int front_op) {
    if (c->T <= 0) return fail(c, BVHCD_E_INVALID, "set_topology first");
    if (c->d.pbox && c->cfg.narrow_phase)
        return fail(c, BVHCD_E_INVALID, "box mode: the narrow phase needs vertices (disable it or update_vertices)");
    if (bvh_op < 0 || bvh_op > BVHCD_BVH_NONE || front_op < 0 || front_op > BVHCD_FRONT_KEEP)
        return fail(c, BVHCD_E_INVALID, "bad op");
    begin(c, S_TOTAL);
    bool topo = false;
    c->step_zeroed = false;
    // Once kernels that read an own vertex buffer are enqueued, the next host upload into that
    // buffer must wait for them, also when a later part of the step fails.
    auto mark_consumed = [c]() {
        if (c->d.verts == c->own_verts || c->d.verts == c->own_verts2)
            (void)hipEventRecord(c->consumed[c->d.verts == c->own_verts2 ? 1 : 0], c->stream);
    };
    int st = run_bvh(c, bvh_op, topo);
    if (st) { mark_consumed(); return st; }
    if (topo) c->front_valid = false;
    st = run_front(c, front_op);
    c->step_zeroed = false;
    if (!st && c->cfg.auto_grow) st = grow_and_retry(c, front_op);
    mark_consumed();
    if (st) return st;
    end(c, S_TOTAL);
    c->last_bvh = bvh_op;
    c->last_front = front_op;
    log_frame(c, bvh_op, front_op);
    c->frame++;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, BVHCD_E_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return BVHCD_OK;
}

// The reference's plain schemes (BenchmarkLogic::maintainScheme, BenchmarkLogic.hpp:90-156, and
// getSchemeOpt, :56-87) at frame f counted from the first frame: bc = BVH rebuild cycle,
// fc = front cycle. The learned schemes take the PLAIN table as their starting point.
void plain_schedule(int sched, int64_t f, int bc, int fc, int& bvh_op, int& front_op) {
    const int64_t fr = f % fc;
    // front cycle: GENERATE on the first frame, UPDATE at later cycle starts, then KEEP; with
    // the divergent mark the last frame of each cycle REORDERs
    auto cycle = [&](bool divergent) {
        if (fr == 0) front_op = f == 0 ? BVHCD_FRONT_GENERATE : BVHCD_FRONT_UPDATE;
        else if (divergent && fr == fc - 1) front_op = BVHCD_FRONT_REORDER;
        else front_op = BVHCD_FRONT_KEEP;
    };
    switch (sched) {
        case BVHCD_SCHED_GENERATE:
            bvh_op = BVHCD_BVH_BUILD;
            front_op = BVHCD_FRONT_PURE_BVH_CD;
            break;
        case BVHCD_SCHED_FRONT_GENERATE:
            bvh_op = BVHCD_BVH_BUILD;
            front_op = BVHCD_FRONT_GENERATE;
            break;
        case BVHCD_SCHED_STATIC_NODIV:
            bvh_op = (f % bc) ? BVHCD_BVH_REFIT : BVHCD_BVH_BUILD;
            cycle(false);
            if (bvh_op == BVHCD_BVH_BUILD) front_op = BVHCD_FRONT_GENERATE;
            break;
        case BVHCD_SCHED_REFIT_ONLY_FRONT:
        case BVHCD_SCHED_REFIT_ONLY_FRONT_NODIV:
        case BVHCD_SCHED_ARCSIM_STATIC:
            bvh_op = f ? BVHCD_BVH_REFIT : BVHCD_BVH_BUILD;
            cycle(sched != BVHCD_SCHED_REFIT_ONLY_FRONT_NODIV);
            break;
        case BVHCD_SCHED_ARCSIM_PERIODIC:
            bvh_op = (f % bc) ? BVHCD_BVH_REFIT : BVHCD_BVH_BUILD;
            front_op = fr == 0 ? BVHCD_FRONT_GENERATE : BVHCD_FRONT_KEEP;
            break;
        default:  // STATIC_MANDATORY with divergent mark
            bvh_op = (f % bc) ? BVHCD_BVH_REFIT : BVHCD_BVH_BUILD;
            cycle(true);
            if (bvh_op == BVHCD_BVH_BUILD) front_op = BVHCD_FRONT_GENERATE;
            if ((f + 1) % bc == 0) front_op = BVHCD_FRONT_KEEP;
            break;
    }
}

int check_overflow(bvhcd_ctx* c) {
    uint32_t ovf = 0;
    HIPCHK(c, hipMemcpyAsync(&ovf, c->d.ctr + C_OVF, sizeof(ovf), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (ovf) {
        char buf[160];
        if (ovf & 16u) return fail(c, BVHCD_E_DEVICE, "traversal step bound exceeded (corrupted tree)");
        snprintf(buf, sizeof(buf), "capacity exceeded (flags 0x%x: 1 int front, 2 ext front, 4 bp pairs, 8 np pairs)",
                 ovf);
        return fail(c, BVHCD_E_CAPACITY, buf);
    }
    return BVHCD_OK;
}

void timing_free(bvhcd_ctx* c) {
    for (auto& e : c->tev) (void)hipEventDestroy(e);
    c->tev.clear();
    c->tused.clear();
    if (c->dlog) {
        (void)hipFree(c->dlog);
        c->dlog = nullptr;
    }
    c->tcap = c->tn = 0;
    c->tops.clear();
    c->trestr.clear();
}

}  // namespace

extern "C" {

namespace {
void release_upload(bvhcd_ctx* c) {
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (int b = 0; b < 2; ++b) {
        if (c->copied[b]) (void)hipEventDestroy(c->copied[b]);
        if (c->consumed[b]) (void)hipEventDestroy(c->consumed[b]);
        c->copied[b] = c->consumed[b] = nullptr;
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    c->copy_stream = nullptr;
}
}  // namespace

int bvhcd_create(const bvhcd_config* cfg, bvhcd_ctx** out) {
    if (!cfg || !out || cfg->max_tris < 1 || cfg->max_verts < 1 || cfg->max_tris > (1ll << 30) ||
        cfg->max_verts > (1ll << 30) || cfg->scheduler < BVHCD_SCHED_PLAIN ||
        cfg->scheduler > BVHCD_SCHED_LEARNED_NO_NET)
        return BVHCD_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0 || cfg->device < 0 || cfg->device >= ndev)
        return BVHCD_E_DEVICE;
    bvhcd_ctx* c = new (std::nothrow) bvhcd_ctx();
    if (!c) return BVHCD_E_OOM;
    c->cfg = *cfg;
    c->d.skip_adj = cfg->skip_adjacent ? 1 : 0;
    if (c->cfg.bvh_cycle <= 0) c->cfg.bvh_cycle = 12;
    if (c->cfg.front_cycle <= 0) c->cfg.front_cycle = 6;
    if (hipSetDevice(cfg->device) != hipSuccess || noalloc_scratch_init(cfg->device) != BVHCD_OK) {
        delete c;
        return BVHCD_E_DEVICE;
    }
    if (cfg->stream) {
        c->stream = static_cast<hipStream_t>(cfg->stream);
    } else {
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            delete c;
            return BVHCD_E_DEVICE;
        }
        c->own_stream = true;
    }
    int st = allocate(c);
    bool ok = st == 0 && hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) == hipSuccess;
    for (int b = 0; b < 2 && ok; ++b) {
        ok &= hipEventCreateWithFlags(&c->copied[b], hipEventDisableTiming) == hipSuccess;
        ok &= hipEventCreateWithFlags(&c->consumed[b], hipEventDisableTiming) == hipSuccess;
        ok &= ok && hipEventRecord(c->consumed[b], c->stream) == hipSuccess;
    }
    if (!ok) {
        release_upload(c);
        free_all(c);
        if (c->own_stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return st ? st : BVHCD_E_DEVICE;
    }
    *out = c;
    return BVHCD_OK;
}

int bvhcd_destroy(bvhcd_ctx* c) {
    if (!c) return BVHCD_E_INVALID;
    (void)hipStreamSynchronize(c->stream);
    release_upload(c);
    for (hipEvent_t& e : c->lev)
        if (e) (void)hipEventDestroy(e);
    learned_destroy(c->learned);
    timing_free(c);
    free_all(c);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BVHCD_OK;
}

const char* bvhcd_last_error(const bvhcd_ctx* c) { return c ? c->err.c_str() : "null context"; }

int bvhcd_set_topology(bvhcd_ctx* c, const int32_t* faces, int64_t T, int64_t V, int where) {
    if (!c) return BVHCD_E_INVALID;
    if (!faces || T < 1 || V < 1 || T > c->d.maxT || V > c->d.maxV) return fail(c, BVHCD_E_INVALID, "bad topology size");
    if (where != BVHCD_MEM_HOST && where != BVHCD_MEM_DEVICE) return fail(c, BVHCD_E_INVALID, "bad memory kind");
    c->T = 0;  // no steps until a valid topology is in place
    // Every vertex id must lie in [0, V): the kernels gather vertices through them unchecked.
    if (where == BVHCD_MEM_HOST) {
        for (int64_t k = 0; k < 3 * T; ++k)
            if (faces[k] < 0 || faces[k] >= V) return fail(c, BVHCD_E_INVALID, "face vertex id out of range");
    }
    HIPCHK(c, hipMemcpyAsync(c->d.faces, faces, sizeof(int32_t) * 3 * T,
                             where == BVHCD_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    if (where == BVHCD_MEM_DEVICE) {
        uint32_t bad = 0;
        HIPCHK(c, hipMemsetAsync(c->d.ctr + C_BADFACE, 0, sizeof(uint32_t), c->stream));
        hipLaunchKernelGGL(k_check_faces, dim3(std::min<unsigned>(grid_for(3 * T), 4096)), dim3(kBlock), 0, c->stream,
                           3 * T, (const int32_t*)c->d.faces, (int32_t)V, c->d.ctr + C_BADFACE);
        HIPCHK(c, hipMemcpyAsync(&bad, c->d.ctr + C_BADFACE, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (bad) return fail(c, BVHCD_E_INVALID, "face vertex id out of range");
    } else {
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->T = T;
    c->V = V;
    c->built = c->front_valid = c->restr_pending = c->learned_frame = c->repair_pending = false;
    c->keep_base = -1.0;
    c->keep_excess = 0.0;
    c->rebuild_ms = -1.0;
    c->rebuild_timed = false;
    c->d.front_version++;
    c->frame = 0;
    c->d.fcur[0] = c->d.fcur[1] = 0;
    HIPCHK(c, hipMemsetAsync(c->d.fsize, 0, 4 * sizeof(uint32_t), c->stream));
    HIPCHK(c, hipMemsetAsync(c->d.ctr, 0, 32 * sizeof(uint32_t), c->stream));
    // the learned schedule's decision buffers are sized here, not on its first decision frame
    if (c->learned && c->cfg.scheduler == BVHCD_SCHED_LEARNED && learned_reserve(c->learned, (int32_t)T))
        return fail(c, BVHCD_E_OOM, "learned decision buffers");
    return BVHCD_OK;
}

int bvhcd_update_vertices(bvhcd_ctx* c, const float* xyz, int64_t V, int where) {
    if (!c) return BVHCD_E_INVALID;
    if (!xyz || V != c->V) return fail(c, BVHCD_E_INVALID, "vertex count differs from set_topology");
    if (where != BVHCD_MEM_HOST && where != BVHCD_MEM_DEVICE && where != BVHCD_MEM_DEVICE_BORROW)
        return fail(c, BVHCD_E_INVALID, "bad memory kind");
    c->d.pbox = nullptr;  // vertex mode
    if (where == BVHCD_MEM_DEVICE_BORROW) {  // read in place by the following steps
        c->d.verts = const_cast<float*>(xyz);
        return BVHCD_OK;
    }
    if (where == BVHCD_MEM_DEVICE) {  // stream-ordered after the steps that read the buffer
        c->d.verts = c->vbuf ? c->own_verts2 : c->own_verts;
        HIPCHK(c, hipMemcpyAsync(c->d.verts, xyz, sizeof(float) * 3 * V, hipMemcpyDeviceToDevice, c->stream));
        return BVHCD_OK;
    }
    // Host memory: the upload goes on the copy stream into the buffer the latest step did not
    // read (once the step before it, which did, is done), so it overlaps the steps still
    // running; the call returns when the copy is complete (the caller may reuse xyz).
    const int b = c->vbuf ^ 1;
    float* dst = b ? c->own_verts2 : c->own_verts;
    HIPCHK(c, hipStreamWaitEvent(c->copy_stream, c->consumed[b], 0));
    HIPCHK(c, hipMemcpyAsync(dst, xyz, sizeof(float) * 3 * V, hipMemcpyHostToDevice, c->copy_stream));
    HIPCHK(c, hipEventRecord(c->copied[b], c->copy_stream));
    HIPCHK(c, hipEventSynchronize(c->copied[b]));
    c->vbuf = b;
    c->d.verts = dst;
    return BVHCD_OK;
}

int bvhcd_update_boxes(bvhcd_ctx* c, const float* boxes, int64_t T, int where) {
    if (!c) return BVHCD_E_INVALID;
    if (!boxes || T != c->T || T < 1) return fail(c, BVHCD_E_INVALID, "box count differs from set_topology");
    if (!c->boxes) {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(float) * 6 * (size_t)c->d.maxT) != hipSuccess)
            return fail(c, BVHCD_E_OOM, "box buffer");
        c->allocs.push_back(p);
        c->boxes = static_cast<float*>(p);
    }
    HIPCHK(c, hipMemcpyAsync(c->boxes, boxes, sizeof(float) * 6 * T,
                             where == BVHCD_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    if (where != BVHCD_MEM_DEVICE) HIPCHK(c, hipStreamSynchronize(c->stream));
    c->d.pbox = c->boxes;
    return BVHCD_OK;
}

int bvhcd_schedule_ops(int sched, int64_t frame, int bvh_cycle, int front_cycle, int* bvh_op, int* front_op) {
    if (!bvh_op || !front_op || frame < 0 || sched < BVHCD_SCHED_PLAIN || sched > BVHCD_SCHED_LEARNED_NO_NET)
        return BVHCD_E_INVALID;
    plain_schedule(sched, frame, bvh_cycle > 0 ? bvh_cycle : 12, front_cycle > 0 ? front_cycle : 6, *bvh_op,
                   *front_op);
    return BVHCD_OK;
}

int bvhcd_step(bvhcd_ctx* c, int bvh_op, int front_op) {
    if (!c) return BVHCD_E_INVALID;
    return do_step(c, bvh_op, front_op);
}

int bvhcd_step_scheduled(bvhcd_ctx* c) {
    if (!c) return BVHCD_E_INVALID;
    int bvh_op, front_op;
    const int sched = c->cfg.scheduler;
    plain_schedule(sched, c->frame, c->cfg.bvh_cycle, c->cfg.front_cycle, bvh_op, front_op);
    const bool net = sched == BVHCD_SCHED_LEARNED && learned_ready(c->learned);
    if (net || sched == BVHCD_SCHED_LEARNED_NO_NET) {
        // Learned schedule: the periodic global maintenance of the plain schedule (rebuild +
        // regenerate every bvh_cycle frames, UPDATE / REORDER every front_cycle) is replaced by
        // decisions. Every frame refits and keeps the front; every bvh_cycle frames (and on the
        // heuristic's restructure flag) the MLP scores the candidate subtrees, and the chosen
        // ones are rebuilt with an incremental front restructure. At front-cycle points the
        // whole tree is rebuilt and its front generated anew when keeping them has cost as much
        // as a rebuild (below), or a front nears its capacity. The pair set is the same for
        // every schedule.
        bvh_op = c->built ? BVHCD_BVH_REFIT : BVHCD_BVH_BUILD;
        front_op = c->built ? BVHCD_FRONT_KEEP : BVHCD_FRONT_GENERATE;
        const int bc = c->cfg.bvh_cycle > 0 ? c->cfg.bvh_cycle : 12;
        const int fc = c->cfg.front_cycle > 0 ? c->cfg.front_cycle : 6;
        if (!c->lev[0]) {
            HIPCHK(c, hipEventCreate(&c->lev[0]));
            HIPCHK(c, hipEventCreate(&c->lev[1]));
        }
        if (c->rebuild_timed) {  // the rebuild frame just stepped: its time is the rebuild cost
            float ms = 0.f;
            HIPCHK(c, hipEventSynchronize(c->lev[1]));
            HIPCHK(c, hipEventElapsedTime(&ms, c->lev[0], c->lev[1]));
            c->rebuild_ms = ms;
            c->rebuild_timed = false;
        }
        if (net && c->built &&
            (c->frame % ((int64_t)bc << c->learned_backoff) == 0 || (c->cfg.enable_restructure && c->restr_pending)))
            c->learned_frame = true;
        if (c->built && c->frame % fc == 0 && c->last_front == BVHCD_FRONT_KEEP) {
            uint32_t sz[4] = {0, 0, 0, 0};
            HIPCHK(c, hipMemcpyAsync(sz, c->d.fsize, sizeof(sz), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            const double fi = (double)sz[c->d.fcur[0]], fe = (double)sz[2 + c->d.fcur[1]];
            // The previous frame was a KEEP frame: its time, against the first such frame after
            // the last rebuild, is what keeping the aging tree and front costs per frame. Once
            // that excess, summed over the frames since the rebuild, reaches what a rebuild frame
            // costs beyond a KEEP frame, keeping has cost as much as rebuilding would have: the
            // tree is rebuilt and its front generated anew (self-calibrating, times measured on
            // the device with the context's events).
            float ms = 0.f;
            HIPCHK(c, hipEventElapsedTime(&ms, c->lev[0], c->lev[1]));
            if (c->keep_base < 0) c->keep_base = ms;
            c->keep_excess += std::max(0.0, (double)ms - c->keep_base) * fc;
            const bool full = fi > kLearnedCapacityShare * (double)c->d.front_cap[0] ||
                              fe > kLearnedCapacityShare * (double)c->d.front_cap[1];
            if (full || (c->rebuild_ms > 0 && c->keep_excess >= c->rebuild_ms - c->keep_base)) {
                bvh_op = BVHCD_BVH_BUILD;
                front_op = BVHCD_FRONT_GENERATE;
                c->learned_frame = false;  // the rebuild replaces this frame's subtree decision
            }
        }
        if (bvh_op == BVHCD_BVH_BUILD) {
            c->keep_base = -1.0;
            c->keep_excess = 0.0;
            c->rebuild_timed = true;
        }
        HIPCHK(c, hipEventRecord(c->lev[0], c->stream));
        const int st = do_step(c, bvh_op, front_op);
        HIPCHK(c, hipEventRecord(c->lev[1], c->stream));
        return st;
    }
    if (bvh_op == BVHCD_BVH_REFIT && c->cfg.enable_restructure && c->restr_pending) bvh_op = BVHCD_BVH_RESTRUCTURE;
    return do_step(c, bvh_op, front_op);
}

int bvhcd_get_pairs(bvhcd_ctx* c, int32_t* out, int64_t cap, int64_t* num, int where, int filtered) {
    if (!c || !num) return BVHCD_E_INVALID;
    ensure_totals(c);
    uint32_t cnt = 0;
    HIPCHK(c, hipMemcpyAsync(&cnt, c->d.ctr + (filtered ? C_NP : C_BP), sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    int st = check_overflow(c);
    *num = cnt;
    if (st) return st;
    if (out && cap > 0) {
        const int64_t k = std::min<int64_t>(cap, cnt);
        if (k > 0) {
            // segments, then the implicit shared-vertex pairs, gathered (leaf ids mapped to
            // triangle ids) straight into a device buffer, or through the read-out buffer to the
            // host in pieces of its size (implicit pairs are not stored, so the total can exceed it)
            if (where == BVHCD_MEM_DEVICE) {
                launch_pair_gather(c->d, filtered ? 1 : 0, reinterpret_cast<int2*>(out), 0, k, c->stream);
            } else {
                const int64_t piece = std::max<int64_t>(1, c->d.pair_cap);
                for (int64_t first = 0; first < k; first += piece) {
                    const int64_t m = std::min<int64_t>(piece, k - first);
                    launch_pair_gather(c->d, filtered ? 1 : 0, c->d.pair_tmp, first, m, c->stream);
                    HIPCHK(c, hipMemcpyAsync(out + 2 * first, c->d.pair_tmp, sizeof(int2) * m, hipMemcpyDeviceToHost,
                                             c->stream));
                    HIPCHK(c, hipStreamSynchronize(c->stream));
                }
            }
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return BVHCD_OK;
}

int bvhcd_pair_digest(bvhcd_ctx* c, int filtered, uint64_t* digest, int64_t* num) {
    if (!c || !digest) return BVHCD_E_INVALID;
    ensure_totals(c);
    launch_digest(c->d, filtered, c->stream);
    uint64_t dg = 0;
    uint32_t cnt = 0;
    HIPCHK(c, hipMemcpyAsync(&dg, ctr64(c->d.ctr) + C64_DIGEST, sizeof(dg), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&cnt, c->d.ctr + (filtered ? C_NP : C_BP), sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
    int st = check_overflow(c);
    *digest = dg;
    if (num) *num = cnt;
    return st;
}

int bvhcd_set_timing(bvhcd_ctx* c, int frames) {
    if (!c || frames < 0) return BVHCD_E_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    timing_free(c);
    if (frames == 0) return BVHCD_OK;
    c->tev.resize((size_t)frames * kEv);
    for (auto& e : c->tev)
        if (hipEventCreate(&e) != hipSuccess) return fail(c, BVHCD_E_DEVICE, "hipEventCreate");
    c->tused.assign((size_t)frames * kEv, 0);
    if (hipMalloc(&c->dlog, sizeof(uint64_t) * kLog * (size_t)frames) != hipSuccess)
        return fail(c, BVHCD_E_OOM, "timing log");
    HIPCHK(c, hipMemsetAsync(c->dlog, 0, sizeof(uint64_t) * kLog * (size_t)frames, c->stream));
    c->tcap = frames;
    c->tn = 0;
    return BVHCD_OK;
}

int bvhcd_timing_counts(bvhcd_ctx* c, int64_t cap, int64_t* counts, int64_t* n) {
    if (!c || !n) return BVHCD_E_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *n = c->tn;
    const int64_t k = std::min<int64_t>(cap, c->tn);
    if (k <= 0 || !counts) return BVHCD_OK;
    std::vector<uint64_t> log((size_t)kLog * k);
    HIPCHK(c, hipMemcpy(log.data(), c->dlog, sizeof(uint64_t) * log.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < log.size(); ++i) counts[i] = (int64_t)log[i];
    return BVHCD_OK;
}

int bvhcd_log_text(bvhcd_ctx* c, char* buf, int64_t cap, int64_t* len) {
    if (!c || !len) return BVHCD_E_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<uint64_t> log((size_t)kLog * c->tn);
    if (c->tn > 0)
        HIPCHK(c, hipMemcpy(log.data(), c->dlog, sizeof(uint64_t) * log.size(), hipMemcpyDeviceToHost));
    static const char* kFrontMsg[] = {"pure bvhcd. ", "generate front. ", "update front. ", "balance front. ",
                                      "preserve front. ", "restructure front(in restr) "};
    static const char* kFrontSection[] = {"broad_phase_cd_bvh", "broad_phase_cd_gen", "broad_phase_cd_update",
                                          "broad_phase_cd_balance", "broad_phase_cd_preserve",
                                          "broad_phase_restr_front"};
    std::string msgs, gpu;
    char line[256];
    for (int64_t f = 0; f < c->tn; ++f) {
        const uint64_t* L = &log[kLog * f];
        const int bo = c->tops[2 * f], fo = c->tops[2 * f + 1];
        if (bo == BVHCD_BVH_RESTRUCTURE && c->trestr[f] > 0) {
            snprintf(line, sizeof(line), "%lld leaves in restructured subtrees. restructure\n", (long long)c->trestr[f]);
            msgs += line;
        }
        if (fo >= 0 && fo <= BVHCD_FRONT_RESTRUCTURE) msgs += std::string(kFrontMsg[fo]) + "\n";
        snprintf(line, sizeof(line), "front (%llu, %llu) cpNum %llu actualCpNum %llu\n", (unsigned long long)L[3],
                 (unsigned long long)L[4], (unsigned long long)L[1], (unsigned long long)L[2]);
        msgs += line;
        float total = 0;
        for (int st = 0; st < 7; ++st) {
            const int64_t k = f * kEv + 2 * st;
            float t = 0;
            if (!(c->tused[k] && c->tused[k + 1])) continue;
            (void)hipEventElapsedTime(&t, c->tev[k], c->tev[k + 1]);
            const char* name = st == S_BUILD ? "construct_bvh" : st == S_REFIT ? "refit_bvh" : st == S_RESTR ? "restr_bvh"
                             : st == S_FRONT ? (fo >= 0 && fo <= 5 ? kFrontSection[fo] : "broad_phase")
                             : st == S_ORDER ? "broad_phase_front_ordering" : st == S_NARROW ? "narrow_phase"
                             : "restructure_schedule";
            snprintf(line, sizeof(line), "@%s: %g\n\n", name, (double)t);
            gpu += line;
            total += t;
        }
        snprintf(line, sizeof(line), "#cd_total_time: %g\n###: 0\n", (double)total);
        gpu += line;
    }
    const std::string text = "\n" + msgs + "\n" + gpu + "\n";
    *len = (int64_t)text.size();
    if (buf && cap > 0) {
        const int64_t k = std::min<int64_t>(cap - 1, (int64_t)text.size());
        std::memcpy(buf, text.data(), (size_t)k);
        buf[k] = 0;
    }
    return BVHCD_OK;
}

int bvhcd_timing_frames(bvhcd_ctx* c, int64_t cap, double* ms, double* bytes, int32_t* ops, int64_t* n) {
    if (!c || !n) return BVHCD_E_INVALID;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *n = c->tn;
    if (c->tn == 0) return BVHCD_OK;
    std::vector<uint64_t> log((size_t)kLog * c->tn);
    HIPCHK(c, hipMemcpy(log.data(), c->dlog, sizeof(uint64_t) * log.size(), hipMemcpyDeviceToHost));
    double fi0 = 0, fe0 = 0;
    for (int64_t f = 0; f < c->tn && f < cap; ++f) {
        const uint64_t* L = &log[kLog * f];
        const int bo = c->tops[2 * f], fo = c->tops[2 * f + 1];
        if (ms) {
            for (int s = 0; s < 8; ++s) {
                const int64_t k = f * kEv + 2 * s;
                float t = 0;
                if (c->tused[k] && c->tused[k + 1]) (void)hipEventElapsedTime(&t, c->tev[k], c->tev[k + 1]);
                ms[8 * f + s] = t;
            }
        }
        if (bytes)
            frame_bytes((double)c->T, (double)c->V, bo, fo, (double)L[0], fi0, fe0, (double)L[3], (double)L[4],
                        (double)L[1], (double)L[2], (double)c->trestr[f], !is_learned(c->cfg.scheduler),
                        (double)L[6], (double)L[7], bytes + 8 * f);
        if (ops) { ops[2 * f] = bo; ops[2 * f + 1] = fo; }
        fi0 = (double)L[3];
        fe0 = (double)L[4];
    }
    return BVHCD_OK;
}

int bvhcd_get_stats(bvhcd_ctx* c, bvhcd_stats* out) {
    if (!c || !out) return BVHCD_E_INVALID;
    ensure_totals(c);
    std::memset(out, 0, sizeof(*out));
    uint32_t sz[4] = {0, 0, 0, 0}, ctr[16];
    uint64_t c64[2];
    HIPCHK(c, hipMemcpyAsync(sz, c->d.fsize, sizeof(sz), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ctr, c->d.ctr, sizeof(ctr), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(c64, ctr64(c->d.ctr), sizeof(c64), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    out->frame = c->frame;
    out->last_bvh_op = c->last_bvh;
    out->last_front_op = c->last_front;
    out->int_front = sz[c->d.fcur[0]];
    out->ext_front = sz[2 + c->d.fcur[1]];
    out->bp_pairs = ctr[C_BP];
    out->np_pairs = ctr[C_NP];
    out->required_int_front = std::max<int64_t>(out->int_front, ctr[C_FREQ_I]);
    out->required_ext_front = std::max<int64_t>(out->ext_front, ctr[C_FREQ_E]);
    // the requested broad-phase count (flushes move on to segments with room, so the pair
    // capacity is a plain total)
    out->required_pairs = (int64_t)ctr[C_BP];
    out->node_visits = (int64_t)c64[C64_VISITS];
    out->implicit_pairs = ctr[C_IMPLICIT];
    out->restructured_subtrees = c->restr_sub;
    out->restructured_leaves = c->restr_leaf;
    if (c->tcap > 0 && c->tn > 0) {
        const int64_t f = c->tn - 1;
        for (int s = 0; s < 8; ++s) {
            const int64_t k = f * kEv + 2 * s;
            float t = 0;
            if (c->tused[k] && c->tused[k + 1]) (void)hipEventElapsedTime(&t, c->tev[k], c->tev[k + 1]);
            out->stage_ms[s] = t;
        }
    }
    frame_bytes((double)c->T, (double)c->V, c->last_bvh, c->last_front, (double)c64[C64_VISITS], 0, 0,
                (double)out->int_front, (double)out->ext_front, (double)out->bp_pairs, (double)out->np_pairs,
                (double)c->restr_leaf, !is_learned(c->cfg.scheduler), (double)ctr[C_IMPLICIT],
                (double)ctr[C_IMPL_NP], out->stage_bytes);
    if (ctr[C_OVF]) {
        if (ctr[C_OVF] & 16u) return fail(c, BVHCD_E_DEVICE, "traversal step bound exceeded (corrupted tree)");
        out->required_int_front = std::max<int64_t>(out->int_front, c->d.front_cap[0] + 1);
        char buf[160];
        snprintf(buf, sizeof(buf), "capacity exceeded (flags 0x%x: 1 int front, 2 ext front, 4 bp pairs, 8 np pairs)",
                 ctr[C_OVF]);
        return fail(c, BVHCD_E_CAPACITY, buf);
    }
    return BVHCD_OK;
}

int bvhcd_export_scene_box(bvhcd_ctx* c, float* box6) {
    if (!c || !box6) return BVHCD_E_INVALID;
    HIPCHK(c, hipMemcpyAsync(box6, c->d.scene_box, 6 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return BVHCD_OK;
}

int bvhcd_export_bvh(bvhcd_ctx* c, uint32_t* codes, int32_t* prim_idx, int32_t* prim_map, int32_t* int_lc,
                     int32_t* int_rc, int32_t* int_par, uint32_t* int_mark, int32_t* int_rx, int32_t* int_ry,
                     int32_t* leaf_par, uint32_t* leaf_mark, int32_t* leaf_lca, float* int_box, float* leaf_box) {
    if (!c) return BVHCD_E_INVALID;
    if (!c->built) return fail(c, BVHCD_E_NOT_BUILT, "no BVH");
    const int64_t n = c->T, I = n - 1;
    DevState& d = c->d;
    hipStream_t s = c->stream;
    auto cp = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        if (!dst || bytes == 0) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    HIPCHK(c, cp(codes, d.codes, 4 * n));
    HIPCHK(c, cp(prim_idx, d.prim_idx, 4 * n));
    HIPCHK(c, cp(prim_map, d.prim_map, 4 * n));
    HIPCHK(c, cp(int_lc, d.int_lc, 4 * I));
    HIPCHK(c, cp(int_rc, d.int_rc, 4 * I));
    HIPCHK(c, cp(int_par, d.int_par, 4 * I));
    HIPCHK(c, cp(int_mark, d.int_mark, 4 * I));
    HIPCHK(c, cp(int_rx, d.int_rx, 4 * I));
    HIPCHK(c, cp(int_ry, d.int_ry, 4 * I));
    HIPCHK(c, cp(leaf_par, d.leaf_par, 4 * n));
    HIPCHK(c, cp(leaf_mark, d.leaf_mark, 4 * n));
    HIPCHK(c, cp(leaf_lca, d.leaf_lca, 4 * (n + 1)));
    std::vector<NodeRec> tmp;
    auto boxes = [&](float* dst, const NodeRec* src, int64_t m) -> int {
        if (!dst || m <= 0) return BVHCD_OK;
        tmp.resize(m);
        HIPCHK(c, hipMemcpyAsync(tmp.data(), src, sizeof(NodeRec) * m, hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        for (int64_t i = 0; i < m; ++i) {
            dst[6 * i + 0] = tmp[i].lo.x; dst[6 * i + 1] = tmp[i].lo.y; dst[6 * i + 2] = tmp[i].lo.z;
            dst[6 * i + 3] = tmp[i].hi.x; dst[6 * i + 4] = tmp[i].hi.y; dst[6 * i + 5] = tmp[i].hi.z;
        }
        return BVHCD_OK;
    };
    int st = boxes(int_box, d.inode, I);
    if (st) return st;
    st = boxes(leaf_box, d.leaf, n);
    if (st) return st;
    HIPCHK(c, hipStreamSynchronize(s));
    return BVHCD_OK;
}

int bvhcd_export_front(bvhcd_ctx* c, int which, int32_t* out, int64_t cap, int64_t* num) {
    if (!c || !num || which < 0 || which > 1) return BVHCD_E_INVALID;
    uint32_t sz[4];
    HIPCHK(c, hipMemcpyAsync(sz, c->d.fsize, sizeof(sz), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int b = c->d.fcur[which];
    const int64_t n = sz[2 * which + b];
    *num = n;
    if (out && cap > 0) {
        const int64_t k = std::min<int64_t>(std::min<int64_t>(cap, n), c->d.front_cap[which]);
        if (k > 0) {
            std::vector<uint32_t> node((size_t)k);
            std::vector<int32_t> query((size_t)k);
            HIPCHK(c, hipMemcpyAsync(node.data(), c->d.fnode[which][b], sizeof(uint32_t) * k, hipMemcpyDeviceToHost,
                                     c->stream));
            HIPCHK(c, hipMemcpyAsync(query.data(), c->d.fquery[which][b], sizeof(int32_t) * k,
                                     hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (int64_t i = 0; i < k; ++i) {
                out[2 * i] = query[i];
                out[2 * i + 1] = (int32_t)node[i];
            }
        }
    }
    return check_overflow(c);
}

// ---- learned scheduler entry points (learned.hip holds the kernels)
int bvhcd_mlp_set_weights(bvhcd_ctx* c, const float* p, int64_t n) {
    if (!c || !p || n != BVHCD_MLP_NPARAM) return BVHCD_E_INVALID;
    if (!c->learned) c->learned = learned_create(c->stream);
    if (!c->learned) return fail(c, BVHCD_E_OOM, "learned state");
    const int st = learned_set_weights(c->learned, p, c->stream, &c->err);
    if (st) return st;
    if (c->T > 0 && c->cfg.scheduler == BVHCD_SCHED_LEARNED && learned_reserve(c->learned, (int32_t)c->T))
        return fail(c, BVHCD_E_OOM, "learned decision buffers");
    return BVHCD_OK;
}
int bvhcd_mlp_get_weights(bvhcd_ctx* c, float* p, int64_t n) {
    if (!c || !p || n != BVHCD_MLP_NPARAM) return BVHCD_E_INVALID;
    if (!c->learned) return fail(c, BVHCD_E_INVALID, "no weights");
    return learned_get_weights(c->learned, p, c->stream, &c->err);
}
int bvhcd_mlp_features(bvhcd_ctx* c, float* feats, int32_t* roots, int64_t cap, int64_t* rows) {
    if (!c || !rows) return BVHCD_E_INVALID;
    if (!c->built || !c->front_valid) return fail(c, BVHCD_E_NOT_BUILT, "features need a BVH and a front");
    if (!c->learned) c->learned = learned_create(c->stream);
    if (!c->learned) return fail(c, BVHCD_E_OOM, "learned state");
    return learned_features(c->learned, c->d, (int32_t)c->T, c->frame, feats, roots, cap, rows, c->stream, &c->err);
}
int bvhcd_mlp_forward(bvhcd_ctx* c, const float* feats, int64_t rows, float* out, int where) {
    if (!c || !feats || !out || rows < 0) return BVHCD_E_INVALID;
    if (!c->learned) return fail(c, BVHCD_E_INVALID, "no weights");
    return learned_forward(c->learned, feats, rows, out, where, c->stream, &c->err);
}
int bvhcd_mlp_forward_noalloc(bvhcd_ctx* c, const float* feats, int64_t rows, float* out) {
    if (!c || !feats || !out || rows < 0) return BVHCD_E_INVALID;
    if (!c->learned) return fail(c, BVHCD_E_INVALID, "no weights");
    return learned_forward_noalloc(c->learned, c->cfg.device, feats, rows, out, c->stream, &c->err);
}
int bvhcd_mlp_train_step(bvhcd_ctx* c, const float* feats, const float* targets, int64_t rows, float lr, float* loss,
                         int where) {
    if (!c || !feats || !targets || rows <= 0) return BVHCD_E_INVALID;
    if (!c->learned) return fail(c, BVHCD_E_INVALID, "no weights");
    return learned_train_step(c->learned, feats, targets, rows, lr, loss, where, c->stream, &c->err);
}
int bvhcd_mlp_grad(bvhcd_ctx* c, const float* feats, const float* targets, int64_t rows, float* grad, float* loss,
                   int where) {
    if (!c || !feats || !targets || !grad || rows <= 0) return BVHCD_E_INVALID;
    if (!c->learned) return fail(c, BVHCD_E_INVALID, "no weights");
    return learned_grad(c->learned, feats, targets, rows, grad, loss, where, c->stream, &c->err);
}
// ---- autotuner: grid sizes of KEEP and the narrow phase on the context's current data
int bvhcd_autotune(bvhcd_ctx* c) {
    if (!c) return BVHCD_E_INVALID;
    if (!c->built || !c->front_valid || c->repair_pending)
        return fail(c, BVHCD_E_NOT_BUILT, "autotune needs a built BVH and a front (run a step first)");
    DevState& d = c->d;
    const int32_t n = (int32_t)c->T;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    std::vector<unsigned> cand[2];
    for (unsigned g : {1024u, 2048u, 4096u, 8192u, 16384u, 32768u})
        for (unsigned u : {1u, 2u, 4u})
            for (unsigned w : {2u, 4u, 8u}) cand[TK_KEEP].push_back(make_shape(g, u, w));
    for (unsigned g : {1024u, 2048u, 4096u, 8192u, 16384u})
        for (unsigned w : {2u, 4u, 8u}) cand[TK_NARROW].push_back(make_shape(g, 1u, w));
    for (int k : {(int)TK_KEEP, (int)TK_NARROW}) {
        float best_ms = 1e30f;
        unsigned best = tuned_shape(k, k == TK_KEEP ? c->T : d.maxT);
        for (unsigned g : cand[k]) {
            set_tuned_shape(k, k == TK_KEEP ? c->T : d.maxT, g);
            auto run = [&]() {
                if (k == TK_KEEP) launch_keep(d, n, c->stream);  // rewrites the same pair list
                else launch_narrow(d, c->stream);
            };
            run();  // warm-up
            (void)hipEventRecord(e0, c->stream);
            for (int r = 0; r < 3; ++r) run();
            (void)hipEventRecord(e1, c->stream);
            if (hipEventSynchronize(e1) != hipSuccess) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                return fail(c, BVHCD_E_DEVICE, "autotune run failed");
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            record_tuning(k, k == TK_KEEP ? c->T : d.maxT, g, ms / 3.0f);
            if (ms < best_ms) {
                best_ms = ms;
                best = g;
            }
        }
        set_tuned_shape(k, k == TK_KEEP ? c->T : d.maxT, best);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // leave a consistent result: the current front's pairs and their narrow phase
    launch_keep(d, n, c->stream);
    if (c->cfg.narrow_phase) launch_narrow(d, c->stream);
    else clear_np(d, c->stream);
    launch_pair_totals(d, c->stream, c->cfg.narrow_phase != 0);
    c->totals_pending = false;
    return check_overflow(c);
}

int bvhcd_autotune_set_block(int kernel, int64_t n, uint32_t grid, uint32_t unroll, uint32_t block) {
    if (kernel < 0 || kernel >= TK_COUNT || n < 1 || grid < 1 || grid >= (1u << 24)) return BVHCD_E_INVALID;
    const bool ok = unroll == 1 || ((unroll == 2 || unroll == 4) && kernel == TK_KEEP);
    const bool bok = block == 256 || ((block == 128 || block == 512) && kernel != TK_BATCH_CD);
    if (!ok || !bok) return BVHCD_E_INVALID;
    set_tuned_shape(kernel, n, make_shape(grid, unroll, block / 64));
    return BVHCD_OK;
}
int bvhcd_autotune_set(int kernel, int64_t n, uint32_t grid, uint32_t unroll) {
    return bvhcd_autotune_set_block(kernel, n, grid, unroll, 256);
}

int bvhcd_autotune_report(char* buf, int64_t cap) {
    const std::string r = tuning_report();
    if (buf && cap > 0) {
        const size_t k = std::min<size_t>((size_t)cap - 1, r.size());
        std::memcpy(buf, r.data(), k);
        buf[k] = 0;
    }
    return (int)r.size();
}

int bvhcd_mlp_set_repair_cost(bvhcd_ctx* c, float cost) {
    if (!c || !(cost >= 0.f)) return BVHCD_E_INVALID;
    if (!c->learned) c->learned = learned_create(c->stream);
    if (!c->learned) return fail(c, BVHCD_E_OOM, "learned state");
    learned_set_repair_cost(c->learned, cost);
    c->learned_backoff = 0;
    return BVHCD_OK;
}

int bvhcd_mlp_set_threshold(bvhcd_ctx* c, float t) {
    if (!c || !(t == t)) return BVHCD_E_INVALID;
    if (!c->learned) c->learned = learned_create(c->stream);
    if (!c->learned) return fail(c, BVHCD_E_OOM, "learned state");
    learned_set_threshold(c->learned, t);
    return BVHCD_OK;
}
int bvhcd_restructure(bvhcd_ctx* c, const int32_t* roots, int64_t k) {
    if (!c || (k > 0 && !roots) || k < 0) return BVHCD_E_INVALID;
    if (!c->built) return fail(c, BVHCD_E_NOT_BUILT, "restructure needs a BVH");
    if (c->T < 2 || k == 0) return BVHCD_OK;
    int64_t nsub = 0, nint = 0, nleaf = 0;
    int st = mark_given_roots(c->d, (int32_t)c->T, roots, k, c->stream, &nsub, &nint, &nleaf, &c->err);
    if (st) return st;
    const int64_t tcap = c->tcap;  // outside a frame: keep it out of the stage timing
    c->tcap = 0;
    st = apply_restructure(c, nsub, nleaf, true);  // the next step repairs the front
    c->tcap = tcap;
    if (st) return st;
    c->restr_pending = false;
    return check_overflow(c);
}


}  // extern "C"

// ---------------------------------------------------------------- inter-object front
// BvttFrontLooseInter (collision/bvtt_front/BvttFrontLooseInter.cu:16-260) with LBvhRigid
// (collision/lbvh/LBvhRigid.cu): a second object's triangles queried against a context's tree.
// The second object ("rigid": its topology is fixed, BUILD orders its triangles, REFIT updates
// their boxes) is held in a context of its own that shares the first one's stream; its sorted
// leaves are the queries, each walking the whole tree of the first context (the reference's
// generate kernel leaves the last query out, BvttFrontGenerateKernels.cu:95; here every
// triangle is a query). The front, pair lists and counters belong to this object; the front
// kernels run on a view of the first context's tree that points at them.
struct bvhcd_inter {
    bvhcd_ctx* def = nullptr;    // the tree queried (not owned)
    bvhcd_ctx* rigid = nullptr;  // the querying object (owned)
    std::vector<void*> allocs;
    std::string err;
    DevState own;                // front, pairs, counters and ordering offsets of this object
    bool front_valid = false;
    int64_t def_version = -1;    // topology version of def the front was made on
    int last_front = -1;
    bool narrow = true;
};

namespace {

int ifail(bvhcd_inter* x, int code, const std::string& msg) {
    if (x) x->err = msg;
    return code;
}
#define IHIPCHK(x, expr)                                                                       \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return ifail((x), BVHCD_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <class T>
bool ialloc(bvhcd_inter* x, T*& p, size_t n) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return false;
    x->allocs.push_back(q);
    p = static_cast<T*>(q);
    return true;
}

// The first context's tree with this object's front, pairs and counters, and the second
// object's sorted leaves as queries.
DevState inter_view(bvhcd_inter* x) {
    DevState v = x->def->d;
    const DevState& o = x->own;
    const DevState& r = x->rigid->d;
    for (int k = 0; k < 2; ++k) {
        v.front_cap[k] = o.front_cap[k];
        v.fcur[k] = o.fcur[k];
        for (int b = 0; b < 2; ++b) {
            v.fnode[k][b] = o.fnode[k][b];
            v.fquery[k][b] = o.fquery[k][b];
        }
    }
    v.fsize = o.fsize;
    v.fcnt = o.fcnt;
    v.front_version = o.front_version;
    v.adj_version = o.front_version;
    v.intbeg = o.intbeg;
    v.extbeg = o.extbeg;
    v.bp = o.bp;
    v.np = o.np;
    v.pair_tmp = o.pair_tmp;
    v.pair_cap = o.pair_cap;
    v.pair_seg_cap = o.pair_seg_cap;
    v.pcnt = o.pcnt;
    v.ctr = o.ctr;
    v.skip_adj = 0;
    v.inter = 1;
    v.fchunk[0] = v.fchunk[1] = nullptr;  // inter fronts are sorted by query after production
    v.wave_tot[0] = v.wave_tot[1] = nullptr;
    v.wave_base = nullptr;
    v.wave_nch = v.wave_cbase = nullptr;
    v.clist = nullptr;
    v.wave_cap = 0;
    v.qleaf = r.leaf;
    v.qprim = r.prim_idx;
    v.qlfaces = r.lfaces;
    v.qverts = r.verts;
    v.nq = (int32_t)x->rigid->T;
    return v;
}
// Host-side front state the launchers changed (buffer flips, swapped arrays).
void inter_keep_view(bvhcd_inter* x, const DevState& v) {
    DevState& o = x->own;
    for (int k = 0; k < 2; ++k) {
        o.fcur[k] = v.fcur[k];
        for (int b = 0; b < 2; ++b) {
            o.fnode[k][b] = v.fnode[k][b];
            o.fquery[k][b] = v.fquery[k][b];
        }
    }
    o.front_version = v.front_version;
}

int inter_check_overflow(bvhcd_inter* x) {
    uint32_t ovf = 0;
    hipStream_t s = x->def->stream;
    IHIPCHK(x, hipMemcpyAsync(&ovf, x->own.ctr + C_OVF, sizeof(ovf), hipMemcpyDeviceToHost, s));
    IHIPCHK(x, hipStreamSynchronize(s));
    if (ovf & 16u) return ifail(x, BVHCD_E_DEVICE, "traversal step bound exceeded (corrupted tree)");
    if (ovf) return ifail(x, BVHCD_E_CAPACITY, "inter front or pair capacity exceeded");
    return BVHCD_OK;
}

}  // namespace

extern "C" {

int bvhcd_inter_create(bvhcd_ctx* def, const bvhcd_config* cfg, bvhcd_inter** out) {
    if (!def || !cfg || !out || cfg->max_tris < 1 || cfg->max_verts < 1) return BVHCD_E_INVALID;
    *out = nullptr;
    bvhcd_inter* x = new (std::nothrow) bvhcd_inter();
    if (!x) return BVHCD_E_OOM;
    x->def = def;
    x->narrow = cfg->narrow_phase != 0;
    bvhcd_config rc = *cfg;  // the second object's tree: no front or pair lists of its own to speak of
    rc.int_front_cap = rc.ext_front_cap = rc.pair_cap = kPairSegs;
    rc.scheduler = BVHCD_SCHED_PLAIN;
    rc.enable_restructure = 0;
    rc.device = def->cfg.device;
    rc.stream = def->stream;
    int st = bvhcd_create(&rc, &x->rigid);
    if (st != BVHCD_OK) {
        delete x;
        return st;
    }
    DevState& o = x->own;
    const int64_t Tq = cfg->max_tris, Td = def->d.maxT, lim = 0x7FFFFFFFll;
    int64_t ci = cfg->int_front_cap > 0 ? cfg->int_front_cap : 48 * Tq + (1 << 20);
    int64_t ce = cfg->ext_front_cap > 0 ? cfg->ext_front_cap : 24 * Tq + (1 << 20);
    int64_t cp = cfg->pair_cap > 0 ? cfg->pair_cap : 24 * Tq + (1 << 20);
    ci = std::min(ci, lim); ce = std::min(ce, lim); cp = std::min(cp, lim);
    o.front_cap[0] = kPairSegs * ((ci + kPairSegs - 1) / kPairSegs);
    o.front_cap[1] = kPairSegs * ((ce + kPairSegs - 1) / kPairSegs);
    o.pair_seg_cap = std::max<int64_t>((cp + kPairSegs - 1) / kPairSegs, 1);
    o.pair_cap = o.pair_seg_cap * kPairSegs;
    bool ok = true;
    for (int k = 0; k < 2; ++k)
        for (int b = 0; b < 2; ++b) {
            ok &= ialloc(x, o.fnode[k][b], o.front_cap[k]);
            ok &= ialloc(x, o.fquery[k][b], o.front_cap[k]);
        }
    ok &= ialloc(x, o.fsize, 4);
    ok &= ialloc(x, o.fcnt, 2 * kPairSegs * kCntStride);
    ok &= ialloc(x, o.intbeg, Td + 2);
    ok &= ialloc(x, o.extbeg, Td + 2);
    ok &= ialloc(x, o.bp, o.pair_cap);
    ok &= ialloc(x, o.np, o.pair_cap);
    ok &= ialloc(x, o.pair_tmp, o.pair_cap);
    ok &= ialloc(x, o.pcnt, 2 * kPairSegs * kCntStride);
    ok &= ialloc(x, o.ctr, 32);
    // the front ordering sorts this object's fronts with the first context's sort scratch: it
    // must hold the larger of the two
    ok &= std::max(o.front_cap[0], o.front_cap[1]) <= std::max<int64_t>(Td, std::max(def->d.front_cap[0],
                                                                                       def->d.front_cap[1]));
    hipStream_t s = def->stream;
    ok = ok && hipMemsetAsync(o.ctr, 0, 32 * sizeof(uint32_t), s) == hipSuccess &&
         hipMemsetAsync(o.pcnt, 0, 2 * kPairSegs * kCntStride * sizeof(uint32_t), s) == hipSuccess &&
         hipMemsetAsync(o.fsize, 0, 4 * sizeof(uint32_t), s) == hipSuccess && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) {
        bvhcd_inter_destroy(x);
        return BVHCD_E_OOM;
    }
    *out = x;
    return BVHCD_OK;
}

int bvhcd_inter_destroy(bvhcd_inter* x) {
    if (!x) return BVHCD_E_INVALID;
    if (x->def) (void)hipStreamSynchronize(x->def->stream);
    for (void* p : x->allocs) (void)hipFree(p);
    if (x->rigid) bvhcd_destroy(x->rigid);
    delete x;
    return BVHCD_OK;
}

const char* bvhcd_inter_last_error(const bvhcd_inter* x) {
    if (!x) return "null inter";
    return !x->err.empty() ? x->err.c_str() : x->rigid ? x->rigid->err.c_str() : "";
}

int bvhcd_inter_set_topology(bvhcd_inter* x, const int32_t* faces, int64_t T, int64_t V, int where) {
    if (!x) return BVHCD_E_INVALID;
    x->front_valid = false;
    x->err.clear();
    return bvhcd_set_topology(x->rigid, faces, T, V, where);
}

int bvhcd_inter_update_vertices(bvhcd_inter* x, const float* xyz, int64_t V, int where) {
    if (!x) return BVHCD_E_INVALID;
    x->err.clear();
    return bvhcd_update_vertices(x->rigid, xyz, V, where);
}

int bvhcd_inter_step(bvhcd_inter* x, int rigid_op, int front_op) {
    if (!x) return BVHCD_E_INVALID;
    x->err.clear();
    bvhcd_ctx* def = x->def;
    bvhcd_ctx* rg = x->rigid;
    if (!def->built || def->T < 1) return ifail(x, BVHCD_E_NOT_BUILT, "step the queried context first");
    if (rg->T < 1) return ifail(x, BVHCD_E_INVALID, "set_topology first");
    if ((rigid_op != BVHCD_BVH_BUILD && rigid_op != BVHCD_BVH_REFIT) || front_op < BVHCD_FRONT_PURE_BVH_CD ||
        front_op > BVHCD_FRONT_KEEP)
        return ifail(x, BVHCD_E_INVALID, "bad op");
    if (rigid_op == BVHCD_BVH_REFIT && !rg->built) return ifail(x, BVHCD_E_NOT_BUILT, "first inter step must BUILD");
    hipStream_t s = def->stream;
    // the querying object's tree: build (codes, sort, emission) or refit its boxes
    if (rigid_op == BVHCD_BVH_BUILD) {
        int st = build_and_refit(rg);
        if (st) return ifail(x, st, "inter build: " + rg->err);
        rg->built = true;
        x->front_valid = false;
    } else {
        launch_refit(rg->d, (int32_t)rg->T, s, false);
    }
    if (x->def_version != def->d.topo_version) x->front_valid = false;
    if (front_op != BVHCD_FRONT_PURE_BVH_CD && !x->front_valid) front_op = BVHCD_FRONT_GENERATE;
    DevState v = inter_view(x);
    const int32_t n = (int32_t)def->T;
    if (front_op != BVHCD_FRONT_KEEP && front_op != BVHCD_FRONT_PURE_BVH_CD) v.front_version++;
    switch (front_op) {
        case BVHCD_FRONT_PURE_BVH_CD: launch_pure_cd(v, n, s); break;
        case BVHCD_FRONT_GENERATE:
            launch_generate(v, n, s, true);
            x->front_valid = true;
            x->def_version = def->d.topo_version;
            break;
        case BVHCD_FRONT_UPDATE: launch_update(v, n, 0, s, true); break;
        case BVHCD_FRONT_REORDER: launch_reorder(v, n, s); break;
        default: launch_keep(v, n, s, false); break;
    }
    if (x->narrow) launch_narrow(v, s, false);
    else clear_np(v, s);
    launch_pair_totals(v, s, x->narrow);
    inter_keep_view(x, v);
    x->last_front = front_op;
    // the next host upload into the querying object's vertex buffer waits for these kernels
    if (rg->d.verts == rg->own_verts || rg->d.verts == rg->own_verts2)
        IHIPCHK(x, hipEventRecord(rg->consumed[rg->d.verts == rg->own_verts2 ? 1 : 0], s));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return ifail(x, BVHCD_E_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    return BVHCD_OK;
}

int bvhcd_inter_get_pairs(bvhcd_inter* x, int32_t* out, int64_t cap, int64_t* num, int where, int filtered) {
    if (!x || !num) return BVHCD_E_INVALID;
    hipStream_t s = x->def->stream;
    uint32_t cnt = 0;
    IHIPCHK(x, hipMemcpyAsync(&cnt, x->own.ctr + (filtered ? C_NP : C_BP), sizeof(cnt), hipMemcpyDeviceToHost, s));
    const int st = inter_check_overflow(x);
    *num = cnt;
    if (st || !out || cap <= 0 || cnt == 0) return st;
    DevState v = inter_view(x);
    const int64_t k = std::min<int64_t>(cap, cnt);
    if (where == BVHCD_MEM_DEVICE) {
        launch_pair_gather(v, filtered ? 1 : 0, reinterpret_cast<int2*>(out), 0, k, s);
    } else {
        launch_pair_gather(v, filtered ? 1 : 0, v.pair_tmp, 0, k, s);
        IHIPCHK(x, hipMemcpyAsync(out, v.pair_tmp, sizeof(int2) * k, hipMemcpyDeviceToHost, s));
    }
    IHIPCHK(x, hipStreamSynchronize(s));
    return BVHCD_OK;
}

int bvhcd_inter_export_front(bvhcd_inter* x, int which, int32_t* out, int64_t cap, int64_t* n) {
    if (!x || !n || which < 0 || which > 1) return BVHCD_E_INVALID;
    hipStream_t s = x->def->stream;
    uint32_t sz = 0;
    IHIPCHK(x, hipMemcpyAsync(&sz, x->own.fsize + 2 * which + x->own.fcur[which], sizeof(sz), hipMemcpyDeviceToHost,
                              s));
    IHIPCHK(x, hipStreamSynchronize(s));
    *n = sz;
    const int64_t k = std::min<int64_t>(cap, sz);
    if (!out || k <= 0) return BVHCD_OK;
    std::vector<uint32_t> node((size_t)k);
    std::vector<int32_t> query((size_t)k);
    IHIPCHK(x, hipMemcpy(node.data(), x->own.fnode[which][x->own.fcur[which]], 4 * k, hipMemcpyDeviceToHost));
    IHIPCHK(x, hipMemcpy(query.data(), x->own.fquery[which][x->own.fcur[which]], 4 * k, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < k; ++i) {
        out[2 * i] = query[i];
        out[2 * i + 1] = (int32_t)node[i];
    }
    return BVHCD_OK;
}

int bvhcd_inter_stats(bvhcd_inter* x, int64_t* out) {
    if (!x || !out) return BVHCD_E_INVALID;
    hipStream_t s = x->def->stream;
    uint32_t sz[4], ctr[4];
    uint64_t visits = 0;
    IHIPCHK(x, hipMemcpyAsync(sz, x->own.fsize, sizeof(sz), hipMemcpyDeviceToHost, s));
    IHIPCHK(x, hipMemcpyAsync(ctr, x->own.ctr, sizeof(ctr), hipMemcpyDeviceToHost, s));
    IHIPCHK(x, hipMemcpyAsync(&visits, ctr64(x->own.ctr) + C64_VISITS, sizeof(visits), hipMemcpyDeviceToHost, s));
    IHIPCHK(x, hipStreamSynchronize(s));
    out[0] = sz[x->own.fcur[0]];
    out[1] = sz[2 + x->own.fcur[1]];
    out[2] = ctr[C_BP];
    out[3] = ctr[C_NP];
    out[4] = (int64_t)visits;
    out[5] = x->last_front;
    return inter_check_overflow(x);
}

int bvhcd_inter_export_rigid(bvhcd_inter* x, uint32_t* codes, int32_t* prim_idx) {
    if (!x) return BVHCD_E_INVALID;
    if (!x->rigid->built) return ifail(x, BVHCD_E_NOT_BUILT, "no tree yet");
    const int64_t T = x->rigid->T;
    if (codes) IHIPCHK(x, hipMemcpy(codes, x->rigid->d.codes, 4 * T, hipMemcpyDeviceToHost));
    if (prim_idx) IHIPCHK(x, hipMemcpy(prim_idx, x->rigid->d.prim_idx, 4 * T, hipMemcpyDeviceToHost));
    return BVHCD_OK;
}

}  // extern "C"
