// bvhcd_oracle.cpp — CPU oracle for the BVH / BVTT-front collision-detection path.
//
// TEST INFRASTRUCTURE ONLY. This file restates the reference algorithm on the host so that
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg can check and time the
// HIP pipeline. Nothing in the product library links, loads or calls it.
//
// Reference files followed (paths relative to Source/Project/Mine/ of
// littlemine/BVH-based-Collision-Detection-Scheme):
//   AABB ops                base/Bv.h:10-89, collision/lbvh/BvhBV.h:54-65
//   Morton code             utility/CudaDeviceUtils.cu:82-130, collision/lbvh/LBvhBuildKernels.cu:55-91
//   sort + prim map         collision/lbvh/LBvhFixedDeformable.cu:355-363, utility/CudaKernelUtils.cu:28-32
//   split metric            collision/lbvh/BvhExtNode.cu:128-134
//   emission (Apetrei)      collision/lbvh/LBvhBuildKernels.cu:135-172
//   DFS reorder             collision/lbvh/LBvhFixedDeformable.cu:365-380, LBvhBuildKernels.cu:174-217
//   refit                   collision/lbvh/LBvhUpdateKernels.cu:35-74, BvhIntNode.h:59-74
//   stackless traversal     collision/bvtt_front/BvhCDKernels.cu:17-49, BvttFrontGenerateKernels.cu:51-90
//   front UPDATE            collision/bvtt_front/BvttFrontPruneSproutKernels.cu:12-93
//   front KEEP              collision/bvtt_front/BvttFrontKeepKernels.cu:11-58
//   front REORDER, ordering collision/bvtt_front/BvttFrontReorderKernels.cu:12-63, BvttFrontLooseIntra.cu:140-170,285-312
//   order log               collision/auxiliary/FlOrderLog.cu:7-26, FlOrderLog.cpp:113-120
//   quality + snapshot      collision/bvtt_front/BvttFrontGenerateKernels.cu:26-49, BvttFrontLooseIntra.cu:213-228
//   restructure decision    collision/lbvh/LBvhFixedDeformable.cu:210-267, LBvhRestructureKernels.cu:9-35
//   restructure (intent)    collision/lbvh/LBvhFixedDeformable.cu:269-352, LBvhRestructureKernels.cu:37-215
//   plain schedule          system/Logic/BenchmarkLogic.hpp:90-156
//   front op overrides      collision/bvtt_front/BvttFrontLooseIntra.cu:69-126
//   narrow phase            collision/narrow_phase/narrow_phase.cu:8-24,
//                           Externals/inc/gProximity/cuda_intersect_tritri.h:247-266,349-433 (operation order)
//
// Build: g++ -O2 -ffp-contract=off -fPIC -shared -pthread (see oracle/Makefile).
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <numeric>
#include <thread>
#include <vector>


namespace {

struct Box {
    float mn[3], mx[3];
};

inline Box empty_box() { return {{FLT_MAX, FLT_MAX, FLT_MAX}, {-FLT_MAX, -FLT_MAX, -FLT_MAX}}; }

inline void combine(Box& b, const float* v) {
    for (int k = 0; k < 3; ++k) {
        b.mn[k] = std::fmin(b.mn[k], v[k]);
        b.mx[k] = std::fmax(b.mx[k], v[k]);
    }
}
inline Box merge(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; ++k) {
        r.mn[k] = std::fmin(a.mn[k], b.mn[k]);
        r.mx[k] = std::fmax(a.mx[k], b.mx[k]);
    }
    return r;
}
// base/Bv.h:40-45: closed intervals, false as soon as one axis separates.
inline bool overlaps(const Box& a, const Box& b) {
    for (int k = 0; k < 3; ++k)
        if (b.mn[k] > a.mx[k] || b.mx[k] < a.mn[k]) return false;
    return true;
}

inline Box tri_box(const int32_t* faces, const float* verts, int64_t t) {
    Box b = empty_box();
    for (int k = 0; k < 3; ++k) combine(b, verts + 3 * (int64_t)faces[3 * t + k]);
    return b;
}

inline uint32_t expand_bits(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
inline uint32_t morton3d(float x, float y, float z) {
    x = std::fmin(std::fmax(x * 1024.0f, 0.0f), 1023.0f);
    y = std::fmin(std::fmax(y * 1024.0f, 0.0f), 1023.0f);
    z = std::fmin(std::fmax(z * 1024.0f, 0.0f), 1023.0f);
    return expand_bits((uint32_t)x) * 4u + expand_bits((uint32_t)y) * 2u + expand_bits((uint32_t)z);
}
// AABB::center (base/Bv.h:65-67): float sum times the double 0.5, rounded back to float.
inline uint32_t tri_code(const Box& b, const Box& s) {
    float c[3], q[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = (float)((double)(b.mn[k] + b.mx[k]) * 0.5);
        float off = c[k] - s.mn[k];
        float ext = s.mx[k] - s.mn[k];
        q[k] = off / ext;
    }
    return morton3d(q[0], q[1], q[2]);
}

inline int delta_of(const uint32_t* c, int64_t i, int64_t n) {
    if (i >= n - 1) return 33;
    uint32_t x = c[i] ^ c[i + 1];
    return x ? 32 - __builtin_clz(x) : 0;
}

template <class F>
void parallel_for(int64_t n, int threads, F&& fn) {
    if (threads <= 0) threads = (int)std::thread::hardware_concurrency();
    if (threads < 1) threads = 1;
    if (n < 4096 || threads == 1) {
        fn(0, n, 0);
        return;
    }
    std::vector<std::thread> ts;
    // interleaved chunks balance uneven per-item work
    std::atomic<int64_t> next{0};
    const int64_t chunk = 1024;
    for (int t = 0; t < threads; ++t)
        ts.emplace_back([&, t] {
            for (;;) {
                int64_t a = next.fetch_add(chunk);
                if (a >= n) break;
                fn(a, std::min(n, a + chunk), t);
            }
        });
    for (auto& th : ts) th.join();
}

// ---------------------------------------------------------------- narrow phase (SAT)
struct V3 {
    float x, y, z;
};
inline V3 sub(const V3& a, const V3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 cross(const V3& a, const V3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross_fma(const V3& a, const V3& b) {
    return {std::fma(a.y, b.z, -(a.z * b.y)), std::fma(a.z, b.x, -(a.x * b.z)),
            std::fma(a.x, b.y, -(a.y * b.x))};
}
inline float dot_fma(const V3& a, const V3& b) { return std::fma(a.z, b.z, std::fma(a.y, b.y, a.x * b.x)); }

template <bool FMA>
inline bool axis_overlap(const V3& ax, const V3* p, const V3* q) {
    float P[3], Q[3];
    for (int k = 0; k < 3; ++k) {
        P[k] = FMA ? dot_fma(ax, p[k]) : dot(ax, p[k]);
        Q[k] = FMA ? dot_fma(ax, q[k]) : dot(ax, q[k]);
    }
    float mx1 = std::fmax(std::fmax(P[0], P[1]), P[2]);
    float mn1 = std::fmin(std::fmin(P[0], P[1]), P[2]);
    float mx2 = std::fmax(std::fmax(Q[0], Q[1]), Q[2]);
    float mn2 = std::fmin(std::fmin(Q[0], Q[1]), Q[2]);
    if (mn1 > mx2) return false;
    if (mn2 > mx1) return false;
    return true;
}

// 17-axis separating-axis test. P, Q: 3 vertices each (9 floats).
template <bool FMA>
bool sat_tt(const float* Pv, const float* Qv) {
    V3 P1{Pv[0], Pv[1], Pv[2]}, P2{Pv[3], Pv[4], Pv[5]}, P3{Pv[6], Pv[7], Pv[8]};
    V3 Q1{Qv[0], Qv[1], Qv[2]}, Q2{Qv[3], Qv[4], Qv[5]}, Q3{Qv[6], Qv[7], Qv[8]};
    V3 p[3] = {{0.f, 0.f, 0.f}, sub(P2, P1), sub(P3, P1)};
    V3 q[3] = {sub(Q1, P1), sub(Q2, P1), sub(Q3, P1)};
    V3 e[3] = {sub(p[1], p[0]), sub(p[2], p[1]), sub(p[0], p[2])};
    V3 f[3] = {sub(q[1], q[0]), sub(q[2], q[1]), sub(q[0], q[2])};
    auto X = [](const V3& a, const V3& b) { return FMA ? cross_fma(a, b) : cross(a, b); };
    V3 n = X(e[0], e[1]), m = X(f[0], f[1]);
    if (!axis_overlap<FMA>(n, p, q)) return false;
    if (!axis_overlap<FMA>(m, p, q)) return false;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (!axis_overlap<FMA>(X(e[i], f[j]), p, q)) return false;
    for (int k = 0; k < 3; ++k)
        if (!axis_overlap<FMA>(X(e[k], n), p, q)) return false;
    for (int k = 0; k < 3; ++k)
        if (!axis_overlap<FMA>(X(f[k], m), p, q)) return false;
    return true;
}

inline void gather_tri(const int32_t* faces, const float* verts, int64_t t, float* out9) {
    for (int k = 0; k < 3; ++k) {
        const float* v = verts + 3 * (int64_t)faces[3 * t + k];
        out9[3 * k] = v[0]; out9[3 * k + 1] = v[1]; out9[3 * k + 2] = v[2];
    }
}
// Canonical order: the lower original triangle id is P.
inline bool narrow_canonical(const int32_t* faces, const float* verts, int32_t a, int32_t b) {
    float P[9], Q[9];
    if (a > b) std::swap(a, b);
    gather_tri(faces, verts, a, P);
    gather_tri(faces, verts, b, Q);
    return sat_tt<false>(P, Q);
}

// ---------------------------------------------------------------- BVH
struct Bvh {
    int64_t n = 0;                        // leaves
    std::vector<uint32_t> codes;          // sorted codes (slot order)
    std::vector<int32_t> prim_idx;        // slot -> triangle
    std::vector<int32_t> prim_map;        // triangle -> slot
    std::vector<int32_t> lc, rc, par, rx, ry;  // internal, preorder
    std::vector<uint32_t> mark;
    std::vector<int32_t> lpar, llca;      // leaves (llca has n+1)
    std::vector<uint32_t> lmark;
    std::vector<Box> ibox, lbox;
    Box scene;
};

// Apetrei bottom-up emission over sorted codes c[0..n-1] (LBvhBuildKernels.cu:135-172),
// simulated sequentially; then DFS reorder (LBvhFixedDeformable.cu:365-380,
// LBvhBuildKernels.cu:174-217). Output in preorder into bvh topology arrays.
void build_topology(Bvh& b) {
    const int64_t n = b.n;
    const uint32_t* c = b.codes.data();
    const int64_t I = n > 0 ? n - 1 : 0;
    b.lc.assign(I, -1); b.rc.assign(I, -1); b.par.assign(I, -1); b.rx.assign(I, -1); b.ry.assign(I, -1);
    b.mark.assign(I, 0);
    b.lpar.assign(n, -1); b.lmark.assign(n, 3); b.llca.assign(n + 1, -1);
    if (n <= 1) {
        if (n == 1) b.llca[0] = 1;
        return;
    }
    std::vector<int> metric(n);
    for (int64_t i = 0; i < n; ++i) metric[i] = delta_of(c, i, n);
    std::vector<int32_t> ulc(I, -1), urc(I, -1), upar(I, -1), urx(I, -1), ury(I, -1);
    std::vector<uint32_t> umark(I, 0), flag(I, 0);
    std::vector<int32_t> lca(n + 1, -1);
    std::vector<uint32_t> depths(n, 0);
    for (int64_t idx = 0; idx < n; ++idx) {
        int64_t l = idx - 1, r = idx;
        bool mk = l >= 0 ? metric[l] < metric[r] : false;
        int64_t cur = mk ? l : r;
        b.lpar[idx] = (int32_t)cur;
        if (mk) { urc[cur] = (int32_t)idx; ury[cur] = (int32_t)idx; umark[cur] |= 2u; b.lmark[idx] = 7; }
        else    { ulc[cur] = (int32_t)idx; urx[cur] = (int32_t)idx; umark[cur] |= 1u; b.lmark[idx] = 3; }
        while (flag[cur]++ == 1) {
            umark[cur] &= 7u;
            l = urx[cur] - 1; r = ury[cur];
            lca[l + 1] = (int32_t)cur;
            depths[l + 1]++;
            mk = l >= 0 ? metric[l] < metric[r] : false;
            if (l + 1 == 0 && r == n - 1) { upar[cur] = -1; umark[cur] &= ~4u; break; }
            int64_t p = mk ? l : r;
            upar[cur] = (int32_t)p;
            if (mk) { urc[p] = (int32_t)cur; ury[p] = (int32_t)r; umark[p] &= ~2u; umark[cur] |= 4u; }
            else    { ulc[p] = (int32_t)cur; urx[p] = (int32_t)(l + 1); umark[p] &= ~1u; umark[cur] &= ~4u; }
            cur = p;
        }
    }
    // DFS (preorder) ids: exclusive scan of left-chain lengths, then walk down lc.
    std::vector<uint32_t> offsets(n, 0);
    for (int64_t i = 1; i < n; ++i) offsets[i] = offsets[i - 1] + depths[i - 1];
    std::vector<int32_t> tk(I, -1);
    for (int64_t idx = 0; idx < n; ++idx) {
        int32_t node = lca[idx];
        uint32_t depth = depths[idx], id = offsets[idx];
        if (node != -1)
            for (; depth--; node = ulc[node]) tk[node] = (int32_t)id++;
    }
    for (int64_t idx = 0; idx < n; ++idx) {
        b.lpar[idx] = tk[b.lpar[idx]];
        b.llca[idx] = lca[idx] != -1 ? tk[lca[idx]] << 1 : (int32_t)(idx << 1 | 1);
    }
    b.llca[n] = -1;
    for (int64_t k = 0; k < I; ++k) {
        int32_t nid = tk[k];
        uint32_t m = umark[k];
        b.lc[nid] = (m & 1u) ? ulc[k] : tk[ulc[k]];
        b.rc[nid] = (m & 2u) ? urc[k] : tk[urc[k]];
        b.mark[nid] = m;
        b.par[nid] = upar[k] != -1 ? tk[upar[k]] : -1;
        b.rx[nid] = urx[k];
        b.ry[nid] = ury[k];
    }
}

void refit_bvh(Bvh& b, const int32_t* faces, const float* verts, int threads) {
    const int64_t n = b.n, I = n > 0 ? n - 1 : 0;
    b.lbox.resize(n);
    b.ibox.resize(I);
    parallel_for(n, threads, [&](int64_t a, int64_t e, int) {
        for (int64_t s = a; s < e; ++s) b.lbox[s] = tri_box(faces, verts, b.prim_idx[s]);
    });
    // children have larger preorder ids than their parent
    for (int64_t i = I - 1; i >= 0; --i) {
        uint32_t m = b.mark[i];
        const Box& l = (m & 1u) ? b.lbox[b.lc[i]] : b.ibox[b.lc[i]];
        const Box& r = (m & 2u) ? b.lbox[b.rc[i]] : b.ibox[b.rc[i]];
        b.ibox[i] = merge(l, r);
    }
}

Box scene_box(int64_t T, const int32_t* faces, const float* verts) {
    Box s = empty_box();
    for (int64_t t = 0; t < T; ++t) {
        Box bb = tri_box(faces, verts, t);
        s = merge(s, bb);
    }
    return s;
}

void full_build(Bvh& b, int64_t T, const int32_t* faces, const float* verts, int threads) {
    b.n = T;
    b.scene = scene_box(T, faces, verts);
    std::vector<uint32_t> codes(T);
    parallel_for(T, threads, [&](int64_t a, int64_t e, int) {
        for (int64_t t = a; t < e; ++t) codes[t] = tri_code(tri_box(faces, verts, t), b.scene);
    });
    b.prim_idx.resize(T);
    std::iota(b.prim_idx.begin(), b.prim_idx.end(), 0);
    std::stable_sort(b.prim_idx.begin(), b.prim_idx.end(),
                     [&](int32_t x, int32_t y) { return codes[x] < codes[y]; });
    b.codes.resize(T);
    b.prim_map.resize(T);
    for (int64_t s = 0; s < T; ++s) {
        b.codes[s] = codes[b.prim_idx[s]];
        b.prim_map[b.prim_idx[s]] = (int32_t)s;
    }
    build_topology(b);
    refit_bvh(b, faces, verts, threads);
}

// ---------------------------------------------------------------- front
struct Entry {
    int32_t q, node;
};
struct Pair {
    int32_t a, b;
};

struct OrderLog {  // FlOrderLog (auxiliary/FlOrderLog.h:150-175)
    std::vector<uint32_t> intcnt, extcnt, intbeg, extbeg, intback;
    void setup(int64_t n) {
        intcnt.assign(n, 0); extcnt.assign(n + 1, 0); intbeg.assign(n, 0); extbeg.assign(n + 1, 0);
        intback.assign(n, 0);
    }
    void clear() {
        std::fill(intcnt.begin(), intcnt.end(), 0u);
        std::fill(extcnt.begin(), extcnt.end(), 0u);
        std::fill(intback.begin(), intback.end(), 0u);
    }
    void prepare() {
        uint32_t acc = 0;
        for (size_t i = 0; i < intcnt.size(); ++i) { intbeg[i] = acc; acc += intcnt[i]; }
        acc = 0;
        for (size_t i = 0; i < extcnt.size(); ++i) { extbeg[i] = acc; acc += extcnt[i]; }
    }
};

struct Sink {  // per-thread outputs
    std::vector<Entry> fi, fe;
    std::vector<Pair> pairs;
    int64_t visits = 0;
};

struct Ctx {
    int64_t T = 0, V = 0;
    std::vector<int32_t> faces;
    std::vector<float> verts;
    Bvh bvh;
    bool built = false, front_valid = false;
    std::vector<Entry> fi, fe;   // current fronts
    OrderLog log;
    std::vector<float> snapshot;
    // restructure log (auxiliary/BvhRestrLog.h:42-80)
    std::vector<int32_t> intrange, rtroot;
    bool restr_pending = false;
    std::vector<Pair> bp, np;    // last step pairs (canonical a < b)
    int64_t frame = 0;
    int bvh_cycle = 12, front_cycle = 6;
    bool enable_restr = true;
    int scheme = 0;              // scheme_ops id used by orc_step_scheduled
    int64_t topo_version = 0;    // bumped whenever the BVH topology changes (inter fronts rely on it)
    int last_bvh_op = -1, last_front_op = -1;
    int64_t visits = 0, restr_subtrees = 0, restr_leaves = 0;
    int narrow = 1;
    // incremental front restructure: subtree root of each leaf in a rebuilt run (-1 none),
    // and the (query, subtree root) walks to redo once the rebuilt subtrees are in place
    bool repair_pending = false;
    std::vector<int32_t> run_root;
    std::vector<Entry> tasks;
};

inline void emit_pair(Sink& s, const Bvh& b, int64_t i, int64_t j) {
    int32_t x = b.prim_idx[i], y = b.prim_idx[j];
    s.pairs.push_back(x < y ? Pair{x, y} : Pair{y, x});
}

// Stackless DFS from encoded node `st` until `stop` (A.6). FRONT: record int/ext entries.
template <bool FRONT>
void stackless(const Bvh& b, int64_t qi, int32_t st, int32_t stop, Sink& s, OrderLog* log,
               std::vector<std::atomic<uint32_t>>* icnt, std::vector<std::atomic<uint32_t>>* ecnt) {
    const Box& bv = b.lbox[qi];
    (void)log;
    while (st != stop) {
        int32_t t = st & 1;
        st >>= 1;
        int32_t lbd;
        if (!t) {
            lbd = b.rx[st];
            t = b.lpar[lbd];
            for (; st <= t; ++st) {
                ++s.visits;
                if (!overlaps(b.ibox[st], bv)) break;
            }
        } else {
            t = st - 1;
            lbd = st;
        }
        if (st == t + 1) {
            ++s.visits;
            if (overlaps(b.lbox[lbd], bv)) emit_pair(s, b, qi, lbd);
            if (FRONT) {
                s.fe.push_back({(int32_t)qi, lbd});
                (*ecnt)[lbd].fetch_add(1, std::memory_order_relaxed);
            }
            st = b.llca[lbd + 1];
        } else {
            if (FRONT) {
                s.fi.push_back({(int32_t)qi, st});
                (*icnt)[st].fetch_add(1, std::memory_order_relaxed);
            }
            st = b.llca[b.ry[st] + 1];
        }
    }
}

template <class F>
void run_sinks(int64_t n, int threads, std::vector<Sink>& sinks, F&& fn) {
    int nt = threads <= 0 ? (int)std::thread::hardware_concurrency() : threads;
    if (nt < 1) nt = 1;
    sinks.assign(nt, Sink());
    parallel_for(n, nt, [&](int64_t a, int64_t e, int t) {
        for (int64_t i = a; i < e; ++i) fn(i, sinks[t]);
    });
}

void gather_sinks(std::vector<Sink>& sinks, std::vector<Entry>* fi, std::vector<Entry>* fe,
                  std::vector<Pair>& pairs, int64_t& visits) {
    pairs.clear();
    if (fi) fi->clear();
    if (fe) fe->clear();
    visits = 0;
    for (auto& s : sinks) {
        pairs.insert(pairs.end(), s.pairs.begin(), s.pairs.end());
        if (fi) fi->insert(fi->end(), s.fi.begin(), s.fi.end());
        if (fe) fe->insert(fe->end(), s.fe.begin(), s.fe.end());
        visits += s.visits;
    }
}

using Counters = std::vector<std::atomic<uint32_t>>;

void counters_to_log(const Counters& ic, const Counters& ec, OrderLog& log) {
    for (size_t i = 0; i < log.intcnt.size(); ++i) log.intcnt[i] = ic[i].load();
    for (size_t i = 0; i < log.extcnt.size(); ++i) log.extcnt[i] = ec[i].load();
}

// FlOrderLog::prepare + clear + pureReorderLooseFrontsWithLog (BvttFrontLooseIntra.cu:140-170):
// a counting sort of each front by node id.
void reorder_fronts(Ctx& c) {
    c.log.prepare();
    c.log.clear();
    std::vector<Entry> ni(c.fi.size()), ne(c.fe.size());
    for (const auto& e : c.fi) ni[c.log.intbeg[e.node] + c.log.intcnt[e.node]++] = e;
    for (const auto& e : c.fe) ne[c.log.extbeg[e.node] + c.log.extcnt[e.node]++] = e;
    c.fi.swap(ni);
    c.fe.swap(ne);
}

void calc_snapshot(Ctx& c) {  // frontSnapshot (BvttFrontGenerateKernels.cu:26-32)
    const Bvh& b = c.bvh;
    const int64_t I = b.n - 1;
    c.snapshot.assign(I > 0 ? I : 0, 0.f);
    for (int64_t i = 0; i < I; ++i) {
        int32_t x = b.rx[i], y = b.ry[i];
        uint32_t ic = c.log.intbeg[i + y - x] - c.log.intbeg[i];
        uint32_t ec = c.log.extbeg[y + 1] - c.log.extbeg[x];
        c.snapshot[i] = (float)(1.0 * ic / (ec + 1));
    }
}

void check_quality(Ctx& c) {  // checkFrontQuality (BvttFrontGenerateKernels.cu:34-49)
    const Bvh& b = c.bvh;
    const int64_t I = b.n - 1;
    c.intrange.assign(I + 1, 0);
    c.rtroot.assign(b.n, INT32_MAX);
    for (int64_t i = 0; i < I; ++i) {
        int32_t x = b.rx[i], y = b.ry[i];
        int intCnt = (int)(c.log.intbeg[i + y - x] - c.log.intbeg[i]);
        int extCnt = (int)(c.log.extbeg[y + 1] - c.log.extbeg[x]);
        // `_snapshot[idx] + 1` is a float sum in the reference; the rest is double.
        const float snap1 = c.snapshot[i] + 1.0f;
        if (1.0 * intCnt / extCnt > (double)snap1 - 0.7 * intCnt / c.log.intbeg[I]) {
            c.intrange[i] += 1;
            c.intrange[i + y - x] -= 1;
            c.rtroot[x] = std::min<int32_t>(c.rtroot[x], (int32_t)i);
        }
    }
    c.restr_pending = true;
}

void generate(Ctx& c, int threads) {
    const Bvh& b = c.bvh;
    const int64_t n = b.n;
    Counters ic(n), ec(n + 1);
    for (auto& x : ic) x = 0;
    for (auto& x : ec) x = 0;
    std::vector<Sink> sinks;
    if (n >= 2)
        run_sinks(n - 1, threads, sinks, [&](int64_t i, Sink& s) {
            stackless<true>(b, i, b.llca[i + 1], -1, s, &c.log, &ic, &ec);
        });
    gather_sinks(sinks, &c.fi, &c.fe, c.bp, c.visits);
    counters_to_log(ic, ec, c.log);
    reorder_fronts(c);
    if (c.enable_restr) calc_snapshot(c);
    c.front_valid = true;
}

void pure_cd(Ctx& c, int threads) {
    const Bvh& b = c.bvh;
    std::vector<Sink> sinks;
    if (b.n >= 2)
        run_sinks(b.n - 1, threads, sinks, [&](int64_t i, Sink& s) {
            stackless<false>(b, i, b.llca[i + 1], -1, s, nullptr, nullptr, nullptr);
        });
    gather_sinks(sinks, nullptr, nullptr, c.bp, c.visits);
}

// maintainInt/ExtLooseIntraFrontsWithLog (BvttFrontPruneSproutKernels.cu:12-93)
void update(Ctx& c, int threads) {
    const Bvh& b = c.bvh;
    const int64_t n = b.n;
    Counters ic(n), ec(n + 1);
    for (auto& x : ic) x = 0;
    for (auto& x : ec) x = 0;
    std::vector<Sink> sinks;
    const int64_t ni = (int64_t)c.fi.size(), ne = (int64_t)c.fe.size();
    run_sinks(ni + ne, threads, sinks, [&](int64_t k, Sink& s) {
        if (k < ni) {
            Entry e = c.fi[k];
            int32_t st = e.node;
            const Box& bv = b.lbox[e.q];
            ++s.visits;
            if (!overlaps(b.ibox[st], bv)) {  // prune: climb the left spine
                int32_t t = b.llca[b.rx[st]] >> 1;
                for (st--; st >= t; st--) {
                    ++s.visits;
                    if (overlaps(b.ibox[st], bv)) break;
                }
                if (st < t && st + 1 > 0) {
                    ++s.visits;
                    if (!overlaps(b.ibox[b.par[st + 1]], bv)) return;
                }
                s.fi.push_back({e.q, st + 1});
                ic[st + 1].fetch_add(1, std::memory_order_relaxed);
                return;
            }
            // sprout: stackless DFS inside the subtree
            int32_t stop = b.llca[b.ry[st] + 1];
            int32_t start = (b.lc[st] << 1) | (int32_t)(b.mark[st] & 1u);
            stackless<true>(b, e.q, start, stop, s, nullptr, &ic, &ec);
        } else {
            Entry e = c.fe[k - ni];
            int32_t st = e.node;
            const Box& bv = b.lbox[e.q];
            ++s.visits;
            if (!overlaps(b.lbox[st], bv)) {
                int32_t gfa = b.lpar[st];
                ++s.visits;
                if (overlaps(b.ibox[gfa], bv)) {
                    s.fe.push_back(e);
                    ec[st].fetch_add(1, std::memory_order_relaxed);
                    return;
                }
                if ((b.lmark[st] & 4u) == 4u) return;
                int32_t leaf = st;
                int32_t top = b.llca[leaf] >> 1;
                for (st = gfa - 1; st >= top; st--) {
                    ++s.visits;
                    if (overlaps(b.ibox[st], bv)) break;
                }
                if (st < top && st + 1 > 0) {
                    ++s.visits;
                    if (!overlaps(b.ibox[b.par[st + 1]], bv)) return;
                }
                s.fi.push_back({e.q, st + 1});
                ic[st + 1].fetch_add(1, std::memory_order_relaxed);
            } else {
                s.fe.push_back(e);
                ec[st].fetch_add(1, std::memory_order_relaxed);
                emit_pair(s, b, e.q, st);
            }
        }
    });
    gather_sinks(sinks, &c.fi, &c.fe, c.bp, c.visits);
    counters_to_log(ic, ec, c.log);  // counts of the new front; beg still holds the old offsets
    if (c.enable_restr) check_quality(c);
    reorder_fronts(c);
}

// keepInt/ExtLooseIntraFronts (BvttFrontKeepKernels.cu:11-58)
void keep(Ctx& c, int threads) {
    const Bvh& b = c.bvh;
    std::vector<Sink> sinks;
    const int64_t ni = (int64_t)c.fi.size(), ne = (int64_t)c.fe.size();
    run_sinks(ni + ne, threads, sinks, [&](int64_t k, Sink& s) {
        if (k < ni) {
            Entry e = c.fi[k];
            ++s.visits;
            if (!overlaps(b.ibox[e.node], b.lbox[e.q])) return;
            int32_t start = (b.lc[e.node] << 1) | (int32_t)(b.mark[e.node] & 1u);
            int32_t stop = b.llca[b.ry[e.node] + 1];
            stackless<false>(b, e.q, start, stop, s, nullptr, nullptr, nullptr);
        } else {
            Entry e = c.fe[k - ni];
            ++s.visits;
            if (overlaps(b.lbox[e.q], b.lbox[e.node])) emit_pair(s, b, e.q, e.node);
        }
    });
    int64_t v;
    gather_sinks(sinks, nullptr, nullptr, c.bp, v);
    c.visits = v;
}

// balance: reorderIntLooseIntraFrontsWithLog + copied ext front (BvttFrontLooseIntra.cu:285-312).
// Internal entries are re-bucketed by node: non-overlapping at the bucket front, overlapping at
// the back. Counts keep the full per-node totals afterwards.
void reorder_balance(Ctx& c, int threads) {
    const Bvh& b = c.bvh;
    c.log.prepare();
    const int64_t ni = (int64_t)c.fi.size();
    std::vector<uint8_t> ov(ni, 0);
    parallel_for(ni, threads, [&](int64_t a, int64_t e, int) {
        for (int64_t k = a; k < e; ++k) ov[k] = overlaps(b.ibox[c.fi[k].node], b.lbox[c.fi[k].q]) ? 1 : 0;
    });
    std::vector<uint32_t> front(c.log.intcnt.size(), 0), back(c.log.intcnt.size(), 0);
    std::vector<Entry> nf(ni);
    for (int64_t k = 0; k < ni; ++k) {
        int32_t node = c.fi[k].node;
        if (!ov[k]) nf[c.log.intbeg[node] + front[node]++] = c.fi[k];
        else nf[c.log.intbeg[node + 1] - 1 - back[node]++] = c.fi[k];
    }
    keep(c, threads);  // pairs of this frame from the (unchanged) front cut
    c.fi.swap(nf);
}

// ---------------------------------------------------------------- restructure
// Decision (LBvhFixedDeformable.cu:210-267, LBvhRestructureKernels.cu:9-35) -> returns tag
// 0 refit, 1 restructure, 2 build; fills the outermost flagged subtree roots.
int restructure_decision(Ctx& c, std::vector<int32_t>& roots, int64_t& num_int) {
    const Bvh& b = c.bvh;
    const int64_t I = b.n - 1;
    std::vector<int32_t> intmark(I);
    int32_t acc = 0;
    for (int64_t i = 0; i < I; ++i) { acc += c.intrange[i]; intmark[i] = acc; }
    roots.clear();
    num_int = 0;
    for (int64_t x = 0; x < b.n; ++x) {
        int32_t r = c.rtroot[x];
        if (r != INT32_MAX && intmark[r] == 1) {
            roots.push_back(r);
            num_int += b.ry[r] - b.rx[r];
        }
    }
    if (num_int > (I >> 1)) return 2;
    if (num_int < (I >> 3)) return 0;
    return 1;
}

// Rebuild one subtree (preorder root R, leaves [a, b]) from fresh Morton codes of its leaves,
// with the range ends as hard boundaries (restrIntNodes bdmark, LBvhRestructureKernels.cu:122-166),
// keeping R's id, parent link and right-child bit; internal ids R..R+(b-a)-1 in preorder.
void restructure_subtree(Bvh& bv, int32_t R, const Box& scene, const int32_t* faces, const float* verts) {
    const int32_t a = bv.rx[R], e = bv.ry[R];
    const int32_t m = e - a + 1;
    std::vector<std::pair<uint32_t, int32_t>> items(m);
    for (int32_t k = 0; k < m; ++k) {
        int32_t tri = bv.prim_idx[a + k];
        items[k] = {tri_code(tri_box(faces, verts, tri), scene), tri};
    }
    std::stable_sort(items.begin(), items.end(),
                     [](const auto& x, const auto& y) { return x.first < y.first; });
    for (int32_t k = 0; k < m; ++k) {
        bv.codes[a + k] = items[k].first;
        bv.prim_idx[a + k] = items[k].second;
        bv.prim_map[items[k].second] = a + k;
    }
    // local Apetrei climb with boundary sentinels
    const int32_t Il = m - 1;
    std::vector<int> metric(m);
    for (int32_t k = 0; k < m; ++k) {
        if (k == m - 1) metric[k] = 33;
        else {
            uint32_t x = bv.codes[a + k] ^ bv.codes[a + k + 1];
            metric[k] = x ? 32 - __builtin_clz(x) : 0;
        }
    }
    std::vector<int32_t> ulc(Il), urc(Il), upar(Il), urx(Il), ury(Il), lca(m, -1), lpar(m);
    std::vector<uint32_t> umark(Il, 0), flag(Il, 0), depths(m, 0), lmark(m);
    for (int32_t idx = 0; idx < m; ++idx) {
        int32_t l = idx - 1, r = idx;
        bool mk = l >= 0 ? metric[l] < metric[r] : false;
        int32_t cur = mk ? l : r;
        lpar[idx] = cur;
        if (mk) { urc[cur] = idx; ury[cur] = idx; umark[cur] |= 2u; lmark[idx] = 7; }
        else    { ulc[cur] = idx; urx[cur] = idx; umark[cur] |= 1u; lmark[idx] = 3; }
        while (flag[cur]++ == 1) {
            umark[cur] &= 7u;
            l = urx[cur] - 1; r = ury[cur];
            lca[l + 1] = cur;
            depths[l + 1]++;
            mk = l >= 0 ? metric[l] < metric[r] : false;
            if (l + 1 == 0 && r == m - 1) { upar[cur] = -1; break; }
            int32_t p = mk ? l : r;
            upar[cur] = p;
            if (mk) { urc[p] = cur; ury[p] = r; umark[p] &= ~2u; umark[cur] |= 4u; }
            else    { ulc[p] = cur; urx[p] = l + 1; umark[p] &= ~1u; umark[cur] &= ~4u; }
            cur = p;
        }
    }
    std::vector<uint32_t> off(m, 0);
    for (int32_t k = 1; k < m; ++k) off[k] = off[k - 1] + depths[k - 1];
    std::vector<int32_t> tk(Il);
    for (int32_t idx = 0; idx < m; ++idx) {
        int32_t node = lca[idx];
        uint32_t d = depths[idx], id = off[idx];
        if (node != -1)
            for (; d--; node = ulc[node]) tk[node] = R + (int32_t)id++;
    }
    const int32_t keep_par = bv.par[R];
    const uint32_t keep_right = bv.mark[R] & 4u;
    for (int32_t k = 0; k < Il; ++k) {
        int32_t nid = tk[k];
        uint32_t mm = umark[k];
        bv.lc[nid] = (mm & 1u) ? a + ulc[k] : tk[ulc[k]];
        bv.rc[nid] = (mm & 2u) ? a + urc[k] : tk[urc[k]];
        bv.mark[nid] = mm & 3u;
        bv.mark[nid] |= (upar[k] == -1) ? keep_right : (mm & 4u);
        bv.par[nid] = upar[k] != -1 ? tk[upar[k]] : keep_par;
        bv.rx[nid] = a + urx[k];
        bv.ry[nid] = a + ury[k];
    }
    for (int32_t idx = 0; idx < m; ++idx) {
        bv.lpar[a + idx] = tk[lpar[idx]];
        bv.lmark[a + idx] = lmark[idx];
        if (idx > 0) bv.llca[a + idx] = lca[idx] != -1 ? tk[lca[idx]] << 1 : ((a + idx) << 1 | 1);
    }
}

// ---------------------------------------------------------------- incremental front restructure
// The intent of the reference's front restructure (BvttFrontLooseIntra::restructure,
// collision/bvtt_front/BvttFrontLooseIntra.cu:330-400, with separateFronts :172-211 and
// BvttFrontRestructureKernels.cu:12-90; unreachable in the shipped intra path): keep the front
// entries the rebuilt subtrees do not touch and re-traverse only what they invalidate.
// Every query's front is a cut of the leaves to its right, so for a rebuilt subtree R with
// leaves [a, b]:
//  * an entry strictly inside R of a query q outside [a, b] is one piece of q's cut of R; the
//    piece covering leaf a stands for the whole cut and becomes the walk (q, R);
//  * an entry strictly inside R of a query q inside [a, b] is dropped; q's part inside R is
//    re-walked from lca[q + 1] up to R's escape lca[b + 1];
//  * every other entry (outside all rebuilt subtrees, or R itself) stays valid.
// Prepare runs on the old topology, apply after the rebuild and refit.
void front_restr_prepare(Ctx& c, const std::vector<int32_t>& roots) {
    const Bvh& b = c.bvh;
    c.run_root.assign(b.n, -1);
    for (int32_t R : roots)
        for (int32_t j = b.rx[R]; j <= b.ry[R]; ++j) c.run_root[j] = R;
    std::vector<Entry> fi, fe;
    c.tasks.clear();
    for (const Entry& e : c.fi) {
        const int32_t r = c.run_root[b.rx[e.node]];
        if (r < 0 || e.node == r) { fi.push_back(e); continue; }  // outside, ancestor or the root itself
        if (e.node < r) { fi.push_back(e); continue; }            // ancestor of r (its range holds r's)
        if (c.run_root[e.q] != r && b.rx[e.node] == b.rx[r]) c.tasks.push_back({e.q, r});
    }
    for (const Entry& e : c.fe) {
        const int32_t r = c.run_root[e.node];
        if (r < 0) { fe.push_back(e); continue; }
        if (c.run_root[e.q] != r && e.node == b.rx[r]) c.tasks.push_back({e.q, r});
    }
    c.fi.swap(fi);
    c.fe.swap(fe);
    c.repair_pending = true;
}

void front_restr_apply(Ctx& c, int threads) {
    const Bvh& b = c.bvh;
    const int64_t n = b.n;
    // pairs of the entries that stayed (as KEEP), then the walks append entries and pairs
    keep(c, threads);
    std::vector<Pair> kept_pairs;
    kept_pairs.swap(c.bp);
    const int64_t kept_visits = c.visits;
    Counters ic(n), ec(n + 1);
    for (auto& x : ic) x = 0;
    for (auto& x : ec) x = 0;
    const int64_t nt = (int64_t)c.tasks.size();
    std::vector<Sink> sinks;
    run_sinks(nt + n, threads, sinks, [&](int64_t k, Sink& s) {
        if (k < nt) {
            const Entry t = c.tasks[k];
            stackless<true>(b, t.q, t.node << 1, b.llca[b.ry[t.node] + 1], s, nullptr, &ic, &ec);
        } else {
            const int64_t q = k - nt;
            const int32_t r = c.run_root[q];
            if (r >= 0 && q < b.ry[r]) stackless<true>(b, q, b.llca[q + 1], b.llca[b.ry[r] + 1], s, nullptr, &ic, &ec);
        }
    });
    std::vector<Entry> nfi, nfe;
    int64_t v = 0;
    gather_sinks(sinks, &nfi, &nfe, c.bp, v);
    c.bp.insert(c.bp.end(), kept_pairs.begin(), kept_pairs.end());
    c.visits = kept_visits + v;
    c.fi.insert(c.fi.end(), nfi.begin(), nfi.end());
    c.fe.insert(c.fe.end(), nfe.begin(), nfe.end());
    // per-node counts of the new front (the offsets keep the last ordering's values)
    std::fill(c.log.intcnt.begin(), c.log.intcnt.end(), 0u);
    std::fill(c.log.extcnt.begin(), c.log.extcnt.end(), 0u);
    for (const Entry& e : c.fi) c.log.intcnt[e.node]++;
    for (const Entry& e : c.fe) c.log.extcnt[e.node]++;
    c.repair_pending = false;
    c.tasks.clear();
}

int do_restructure(Ctx& c, int threads, int& tag) {
    std::vector<int32_t> roots;
    int64_t num_int = 0;
    tag = restructure_decision(c, roots, num_int);
    c.restr_pending = false;
    c.restr_subtrees = 0;
    c.restr_leaves = 0;
    if (tag == 2) {
        full_build(c.bvh, c.T, c.faces.data(), c.verts.data(), threads);
        return 0;
    }
    if (tag == 0) {
        refit_bvh(c.bvh, c.faces.data(), c.verts.data(), threads);
        return 0;
    }
    Box scene = scene_box(c.T, c.faces.data(), c.verts.data());
    for (int32_t R : roots) {
        restructure_subtree(c.bvh, R, scene, c.faces.data(), c.verts.data());
        c.restr_subtrees++;
        c.restr_leaves += c.bvh.ry[R] - c.bvh.rx[R] + 1;
    }
    refit_bvh(c.bvh, c.faces.data(), c.verts.data(), threads);
    return 0;
}

void narrow(Ctx& c, int threads) {
    c.np.clear();
    if (!c.narrow) return;
    const int64_t P = (int64_t)c.bp.size();
    std::vector<uint8_t> hit(P, 0);
    parallel_for(P, threads, [&](int64_t a, int64_t e, int) {
        for (int64_t k = a; k < e; ++k)
            hit[k] = narrow_canonical(c.faces.data(), c.verts.data(), c.bp[k].a, c.bp[k].b) ? 1 : 0;
    });
    for (int64_t k = 0; k < P; ++k)
        if (hit[k]) c.np.push_back(c.bp[k]);
}

// BVH op then front op (LBvhFixedDeformable::maintain, BvttFrontLooseIntra::maintain).
int step(Ctx& c, int bvh_op, int front_op, int threads) {
    if (c.T <= 0) return -1;
    c.restr_subtrees = c.restr_leaves = 0;  // counts of this step's restructure only
    int tag = 0;
    bool topo_changed = false;
    if (bvh_op == 2) {  // restructure
        if (!c.built || !c.restr_pending) bvh_op = 1;
    }
    if (bvh_op != 0 && !c.built) return -5;
    switch (bvh_op) {
        case 0:
            full_build(c.bvh, c.T, c.faces.data(), c.verts.data(), threads);
            c.built = true;
            topo_changed = true;
            c.restr_pending = false;
            break;
        case 1:
            refit_bvh(c.bvh, c.faces.data(), c.verts.data(), threads);
            break;
        case 2:
            do_restructure(c, threads, tag);
            topo_changed = tag != 0;
            break;
        case 3:
            break;
        default:
            return -1;
    }
    c.last_bvh_op = bvh_op == 2 ? (tag == 2 ? 0 : tag == 0 ? 1 : 2) : bvh_op;
    if (topo_changed) {
        c.front_valid = false;
        c.topo_version++;
    }
    if (c.repair_pending && (!c.front_valid || front_op == 0)) {
        c.repair_pending = false;  // a pure traversal frame or a new topology drops the repair
        c.front_valid = false;
    }
    if (c.repair_pending) front_op = 5;
    // The restructure flag lives for one frame (BvttFrontLooseIntra.cu:70 clears it at every
    // front maintain; checkQuality sets it again).
    c.restr_pending = false;
    if (front_op != 0 && !c.front_valid) front_op = 1;  // a new topology needs a new front
    switch (front_op) {
        case 0: pure_cd(c, threads); break;
        case 1: generate(c, threads); break;
        case 2: update(c, threads); break;
        case 3: reorder_balance(c, threads); break;
        case 4: keep(c, threads); break;
        case 5: front_restr_apply(c, threads); break;
        default: return -1;
    }
    c.last_front_op = front_op;
    narrow(c, threads);
    c.frame++;
    return 0;
}

// BenchmarkLogic::maintainScheme, STATIC_MANDATORY with divergent mark (BenchmarkLogic.hpp:102-117).
void plain_schedule(int64_t f, int bvh_cycle, int front_cycle, int& bvh_op, int& front_op) {
    bvh_op = (f % bvh_cycle) ? 1 : 0;
    int64_t fr = f % front_cycle;
    if (fr == 0) front_op = f == 0 ? 1 : 2;
    else if (fr == front_cycle - 1) front_op = 3;
    else front_op = 4;
    if (bvh_op == 0) front_op = 1;
    if ((f + 1) % bvh_cycle == 0) front_op = 4;
}

// The other schemes, case by case. Scheme ids follow the library's BVHCD_SCHED_* values:
// 0 STATIC_MANDATORY divergent, 2 GENERATE, 3 FRONT_GENERATE, 4 STATIC_MANDATORY without the
// divergent mark, 5 / 6 REFIT_ONLY_FRONT with / without it, 7 / 8 getSchemeOpt 0 / 1.
// Ops: bvh 0 BUILD 1 REFIT; front 0 PURE_BVH_CD 1 GENERATE 2 UPDATE 3 REORDER 4 KEEP.
void scheme_ops(int scheme, int64_t f, int bc, int fc, int& bvh_op, int& front_op) {
    switch (scheme) {
        case 2:  // GENERATE (BenchmarkLogic.hpp:94-97)
            bvh_op = 0; front_op = 0; return;
        case 3:  // FRONT_GENERATE (:98-101)
            bvh_op = 0; front_op = 1; return;
        case 4:  // STATIC_MANDATORY, no divergent mark (:118-126)
            bvh_op = (f % bc) ? 1 : 0;
            front_op = (f % fc == 0) ? (f == 0 ? 1 : 2) : 4;
            if (bvh_op == 0) front_op = 1;
            return;
        case 5:  // REFIT_ONLY_FRONT, divergent mark (:130-140; its missing break reaches
                 // unreachable code, the ops it computed are taken)
        case 7:  // getSchemeOpt scheme 0 (:62-70)
            bvh_op = f ? 1 : 0;
            if (f % fc == 0) front_op = f == 0 ? 1 : 2;
            else if (f % fc == fc - 1) front_op = 3;
            else front_op = 4;
            return;
        case 6:  // REFIT_ONLY_FRONT, no divergent mark (:142-149)
            bvh_op = f ? 1 : 0;
            front_op = (f % fc == 0) ? (f == 0 ? 1 : 2) : 4;
            return;
        case 8:  // getSchemeOpt scheme 1 (:72-77)
            bvh_op = (f % bc) ? 1 : 0;
            front_op = (f % fc == 0) ? 1 : 4;
            return;
        default:
            plain_schedule(f, bc, fc, bvh_op, front_op);
    }
}

}  // namespace

// ====================================================================== C API
extern "C" {

uint32_t orc_expand_bits(uint32_t v) { return expand_bits(v); }
uint32_t orc_morton3d(float x, float y, float z) { return morton3d(x, y, z); }

void orc_tri_boxes(int64_t T, const int32_t* faces, const float* verts, float* boxes) {
    for (int64_t t = 0; t < T; ++t) {
        Box b = tri_box(faces, verts, t);
        std::memcpy(boxes + 6 * t, b.mn, 12);
        std::memcpy(boxes + 6 * t + 3, b.mx, 12);
    }
}
void orc_scene_box(int64_t T, const int32_t* faces, const float* verts, float* box6) {
    Box s = scene_box(T, faces, verts);
    std::memcpy(box6, s.mn, 12);
    std::memcpy(box6 + 3, s.mx, 12);
}
void orc_morton_codes(int64_t T, const int32_t* faces, const float* verts, const float* box6, uint32_t* codes) {
    Box s;
    std::memcpy(s.mn, box6, 12);
    std::memcpy(s.mx, box6 + 3, 12);
    for (int64_t t = 0; t < T; ++t) codes[t] = tri_code(tri_box(faces, verts, t), s);
}
void orc_sort(int64_t T, const uint32_t* codes, uint32_t* sorted, int32_t* prim_idx, int32_t* prim_map) {
    std::vector<int32_t> idx(T);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return codes[x] < codes[y]; });
    for (int64_t s = 0; s < T; ++s) {
        sorted[s] = codes[idx[s]];
        prim_idx[s] = idx[s];
        prim_map[idx[s]] = (int32_t)s;
    }
}

static void export_topology(const Bvh& b, int32_t* int_lc, int32_t* int_rc, int32_t* int_par, uint32_t* int_mark,
                            int32_t* int_rx, int32_t* int_ry, int32_t* leaf_par, uint32_t* leaf_mark,
                            int32_t* leaf_lca) {
    const size_t I = b.lc.size();
    if (int_lc) std::memcpy(int_lc, b.lc.data(), 4 * I);
    if (int_rc) std::memcpy(int_rc, b.rc.data(), 4 * I);
    if (int_par) std::memcpy(int_par, b.par.data(), 4 * I);
    if (int_mark) std::memcpy(int_mark, b.mark.data(), 4 * I);
    if (int_rx) std::memcpy(int_rx, b.rx.data(), 4 * I);
    if (int_ry) std::memcpy(int_ry, b.ry.data(), 4 * I);
    if (leaf_par) std::memcpy(leaf_par, b.lpar.data(), 4 * b.n);
    if (leaf_mark) std::memcpy(leaf_mark, b.lmark.data(), 4 * b.n);
    if (leaf_lca) std::memcpy(leaf_lca, b.llca.data(), 4 * (b.n + 1));
}

// Topology of sorted codes by the Apetrei climb + DFS reorder.
void orc_build_topology(int64_t n, const uint32_t* sorted_codes, int32_t* int_lc, int32_t* int_rc, int32_t* int_par,
                        uint32_t* int_mark, int32_t* int_rx, int32_t* int_ry, int32_t* leaf_par,
                        uint32_t* leaf_mark, int32_t* leaf_lca) {
    Bvh b;
    b.n = n;
    b.codes.assign(sorted_codes, sorted_codes + n);
    build_topology(b);
    export_topology(b, int_lc, int_rc, int_par, int_mark, int_rx, int_ry, leaf_par, leaf_mark, leaf_lca);
}

// Refit: leaf boxes given in slot order (6 floats each); internal boxes out.
void orc_refit_boxes(int64_t n, const int32_t* int_lc, const int32_t* int_rc, const uint32_t* int_mark,
                     const float* leaf_box, float* int_box) {
    for (int64_t i = n - 2; i >= 0; --i) {
        const float* l = (int_mark[i] & 1u) ? leaf_box + 6 * (int64_t)int_lc[i] : int_box + 6 * (int64_t)int_lc[i];
        const float* r = (int_mark[i] & 2u) ? leaf_box + 6 * (int64_t)int_rc[i] : int_box + 6 * (int64_t)int_rc[i];
        for (int k = 0; k < 3; ++k) {
            int_box[6 * i + k] = std::fmin(l[k], r[k]);
            int_box[6 * i + 3 + k] = std::fmax(l[3 + k], r[3 + k]);
        }
    }
}

int orc_sat(const float* P9, const float* Q9) { return sat_tt<false>(P9, Q9) ? 1 : 0; }
int orc_sat_fma(const float* P9, const float* Q9) { return sat_tt<true>(P9, Q9) ? 1 : 0; }

// Narrow-phase decision of each pair (a, b) under the four evaluation variants that the
// reference's build could have used: bit 0 = lower triangle id as P (this build's canonical
// order), bit 1 = higher id as P (the reference passes the lower sorted leaf's triangle first,
// narrow_phase.cu:11-19, which is either one), bits 2 / 3 = the same two orders with the
// products fused into FMAs (nvcc's default contraction of cuda_intersect_tritri.h:247-266,
// 349-433). A mask of 0 or 15 means the four agree.
void orc_sat_variants(int64_t n, const int32_t* pairs, const int32_t* faces, const float* verts, int threads,
                      uint8_t* out) {
    parallel_for(n, threads, [&](int64_t a, int64_t e, int) {
        for (int64_t k = a; k < e; ++k) {
            int32_t x = pairs[2 * k], y = pairs[2 * k + 1];
            if (x > y) std::swap(x, y);
            float P[9], Q[9];
            gather_tri(faces, verts, x, P);
            gather_tri(faces, verts, y, Q);
            uint8_t m = 0;
            m |= sat_tt<false>(P, Q) ? 1 : 0;
            m |= sat_tt<false>(Q, P) ? 2 : 0;
            m |= sat_tt<true>(P, Q) ? 4 : 0;
            m |= sat_tt<true>(Q, P) ? 8 : 0;
            out[k] = m;
        }
    });
}

// Brute-force all-pairs over unordered {a, b}, a != b: AABB overlap and (narrow=1) SAT hit.
int64_t orc_bruteforce_pairs(int64_t T, const int32_t* faces, const float* verts, int narrow_on, int32_t* out,
                             int64_t cap, int threads) {
    std::vector<Box> boxes(T);
    for (int64_t t = 0; t < T; ++t) boxes[t] = tri_box(faces, verts, t);
    // sweep along x to prune: sort by min x
    std::vector<int32_t> ord(T);
    std::iota(ord.begin(), ord.end(), 0);
    std::sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) {
        return boxes[a].mn[0] < boxes[b].mn[0] || (boxes[a].mn[0] == boxes[b].mn[0] && a < b);
    });
    int nt = threads <= 0 ? (int)std::thread::hardware_concurrency() : threads;
    std::vector<std::vector<Pair>> res(nt);
    // boxes whose min x is NaN never prune; fall back to full scan for them
    parallel_for(T, nt, [&](int64_t a0, int64_t e0, int tid) {
        for (int64_t ii = a0; ii < e0; ++ii) {
            int32_t i = ord[ii];
            const Box& bi = boxes[i];
            for (int64_t jj = ii + 1; jj < T; ++jj) {
                int32_t j = ord[jj];
                const Box& bj = boxes[j];
                if (bj.mn[0] > bi.mx[0]) break;
                if (!overlaps(bi, bj)) continue;
                if (narrow_on && !narrow_canonical(faces, verts, i, j)) continue;
                res[tid].push_back(i < j ? Pair{i, j} : Pair{j, i});
            }
        }
    });
    int64_t n = 0;
    for (auto& r : res)
        for (auto& p : r) {
            if (n < cap && out) { out[2 * n] = p.a; out[2 * n + 1] = p.b; }
            ++n;
        }
    return n;
}

// ---- pipeline context
void* orc_create(int64_t T, int64_t V, const int32_t* faces, int bvh_cycle, int front_cycle, int enable_restr,
                 int narrow_on) {
    Ctx* c = new Ctx();
    c->T = T;
    c->V = V;
    c->faces.assign(faces, faces + 3 * T);
    c->verts.assign(3 * V, 0.f);
    c->bvh_cycle = bvh_cycle > 0 ? bvh_cycle : 12;
    c->front_cycle = front_cycle > 0 ? front_cycle : 6;
    c->enable_restr = enable_restr != 0;
    c->narrow = narrow_on;
    c->log.setup(T);
    return c;
}
void orc_destroy(void* p) { delete static_cast<Ctx*>(p); }
void orc_update_vertices(void* p, const float* xyz) {
    Ctx* c = static_cast<Ctx*>(p);
    std::memcpy(c->verts.data(), xyz, 12 * c->V);
}
int orc_step(void* p, int bvh_op, int front_op, int threads) { return step(*static_cast<Ctx*>(p), bvh_op, front_op, threads); }

// Plain scheduler: restructure flag from the last quality check overrides a scheduled REFIT.
int orc_step_scheduled(void* p, int threads) {
    Ctx* c = static_cast<Ctx*>(p);
    int bvh_op, front_op;
    scheme_ops(c->scheme, c->frame, c->bvh_cycle, c->front_cycle, bvh_op, front_op);
    // a pending restructure flag turns a REFIT into a RESTRUCTURE (LBvhFixedDeformable.cu:72-78)
    if (bvh_op == 1 && c->enable_restr && c->restr_pending) bvh_op = 2;
    return step(*c, bvh_op, front_op, threads);
}
void orc_plain_schedule(int64_t f, int bvh_cycle, int front_cycle, int* bvh_op, int* front_op) {
    plain_schedule(f, bvh_cycle, front_cycle, *bvh_op, *front_op);
}
void orc_scheme_ops(int scheme, int64_t f, int bvh_cycle, int front_cycle, int* bvh_op, int* front_op) {
    scheme_ops(scheme, f, bvh_cycle, front_cycle, *bvh_op, *front_op);
}
void orc_set_scheme(void* p, int scheme) { static_cast<Ctx*>(p)->scheme = scheme; }

int64_t orc_get_pairs(void* p, int filtered, int32_t* out, int64_t cap) {
    Ctx* c = static_cast<Ctx*>(p);
    const auto& v = filtered ? c->np : c->bp;
    int64_t n = (int64_t)v.size();
    for (int64_t k = 0; k < n && k < cap && out; ++k) { out[2 * k] = v[k].a; out[2 * k + 1] = v[k].b; }
    return n;
}
int64_t orc_export_front(void* p, int which, int32_t* out, int64_t cap) {
    Ctx* c = static_cast<Ctx*>(p);
    const auto& v = which ? c->fe : c->fi;
    int64_t n = (int64_t)v.size();
    for (int64_t k = 0; k < n && k < cap && out; ++k) { out[2 * k] = v[k].q; out[2 * k + 1] = v[k].node; }
    return n;
}
void orc_export_bvh(void* p, uint32_t* codes, int32_t* prim_idx, int32_t* prim_map, int32_t* int_lc,
                    int32_t* int_rc, int32_t* int_par, uint32_t* int_mark, int32_t* int_rx, int32_t* int_ry,
                    int32_t* leaf_par, uint32_t* leaf_mark, int32_t* leaf_lca, float* int_box, float* leaf_box) {
    const Bvh& b = static_cast<Ctx*>(p)->bvh;
    if (codes) std::memcpy(codes, b.codes.data(), 4 * b.n);
    if (prim_idx) std::memcpy(prim_idx, b.prim_idx.data(), 4 * b.n);
    if (prim_map) std::memcpy(prim_map, b.prim_map.data(), 4 * b.n);
    export_topology(b, int_lc, int_rc, int_par, int_mark, int_rx, int_ry, leaf_par, leaf_mark, leaf_lca);
    if (int_box)
        for (size_t i = 0; i < b.ibox.size(); ++i) {
            std::memcpy(int_box + 6 * i, b.ibox[i].mn, 12);
            std::memcpy(int_box + 6 * i + 3, b.ibox[i].mx, 12);
        }
    if (leaf_box)
        for (size_t i = 0; i < b.lbox.size(); ++i) {
            std::memcpy(leaf_box + 6 * i, b.lbox[i].mn, 12);
            std::memcpy(leaf_box + 6 * i + 3, b.lbox[i].mx, 12);
        }
}
// stats: [frame, last_bvh_op, last_front_op, int_front, ext_front, bp, np, visits,
//         restr_subtrees, restr_leaves, restr_pending]
void orc_stats(void* p, int64_t* out) {
    Ctx* c = static_cast<Ctx*>(p);
    out[0] = c->frame; out[1] = c->last_bvh_op; out[2] = c->last_front_op;
    out[3] = (int64_t)c->fi.size(); out[4] = (int64_t)c->fe.size();
    out[5] = (int64_t)c->bp.size(); out[6] = (int64_t)c->np.size(); out[7] = c->visits;
    out[8] = c->restr_subtrees; out[9] = c->restr_leaves; out[10] = c->restr_pending ? 1 : 0;
}
// Restructure log of the last quality check: intrange (n entries, I+1 used) and rtroot (n).
// The quality heuristic on caller-given preorder ranges and ordering offsets (known-answer
// tests): do_snapshot = 1 writes frontSnapshot into snapshot[n-1]; then checkFrontQuality
// against snapshot[] fills intrange[n] and rtroot[n].
void orc_quality_eval(int64_t n, const int32_t* rx, const int32_t* ry, const uint32_t* intbeg,
                      const uint32_t* extbeg, float* snapshot, int do_snapshot, int32_t* intrange,
                      int32_t* rtroot) {
    Ctx c;
    c.T = n;
    c.bvh.n = n;
    c.bvh.rx.assign(rx, rx + (n - 1));
    c.bvh.ry.assign(ry, ry + (n - 1));
    c.log.setup(n);
    std::memcpy(c.log.intbeg.data(), intbeg, 4 * n);
    std::memcpy(c.log.extbeg.data(), extbeg, 4 * (n + 1));
    if (do_snapshot) {
        calc_snapshot(c);
        std::memcpy(snapshot, c.snapshot.data(), 4 * (n - 1));
    } else {
        c.snapshot.assign(snapshot, snapshot + (n - 1));
    }
    check_quality(c);
    std::memcpy(intrange, c.intrange.data(), 4 * n);
    std::memcpy(rtroot, c.rtroot.data(), 4 * n);
}

// The restructure decision (LBvhFixedDeformable.cu:210-267) on caller-given marks: returns the
// tag (0 refit, 1 restructure, 2 build), the outermost flagged roots and their internal nodes.
int orc_restr_decision(int64_t n, const int32_t* rx, const int32_t* ry, const int32_t* intrange,
                       const int32_t* rtroot, int32_t* roots, int64_t* nroots, int64_t* num_int) {
    Ctx c;
    c.T = n;
    c.bvh.n = n;
    c.bvh.rx.assign(rx, rx + (n - 1));
    c.bvh.ry.assign(ry, ry + (n - 1));
    c.intrange.assign(intrange, intrange + n);
    c.rtroot.assign(rtroot, rtroot + n);
    std::vector<int32_t> r;
    const int tag = restructure_decision(c, r, *num_int);
    *nroots = (int64_t)r.size();
    if (roots) std::memcpy(roots, r.data(), 4 * r.size());
    return tag;
}

void orc_export_restr_log(void* p, int32_t* intrange, int32_t* rtroot) {
    Ctx* c = static_cast<Ctx*>(p);
    if (intrange) std::memcpy(intrange, c->intrange.data(), 4 * c->intrange.size());
    if (rtroot) std::memcpy(rtroot, c->rtroot.data(), 4 * c->rtroot.size());
}

// Explicit-root restructure (same rebuild as the heuristic path, roots given by the caller):
// validated disjoint preorder roots, one shared scene box, then a full refit. A valid front is
// prepared for the incremental front restructure, which the next step applies (front op 5).
// Returns 0, or -1 for a bad root / overlapping ranges.
int orc_restructure_roots(void* p, const int32_t* roots, int64_t k, int threads) {
    Ctx* c = static_cast<Ctx*>(p);
    if (!c->built || c->T < 2) return -5;
    std::vector<std::pair<int32_t, int32_t>> rg;
    for (int64_t i = 0; i < k; ++i) {
        if (roots[i] < 0 || roots[i] >= c->T - 1) return -1;
        rg.push_back({c->bvh.rx[roots[i]], c->bvh.ry[roots[i]]});
    }
    std::sort(rg.begin(), rg.end());
    for (size_t i = 1; i < rg.size(); ++i)
        if (rg[i].first <= rg[i - 1].second) return -1;
    if (k == 0) return 0;
    c->topo_version++;
    if (c->repair_pending) c->front_valid = c->repair_pending = false;  // one repair at a time
    if (c->front_valid) front_restr_prepare(*c, std::vector<int32_t>(roots, roots + k));
    Box scene = scene_box(c->T, c->faces.data(), c->verts.data());
    for (int64_t i = 0; i < k; ++i) restructure_subtree(c->bvh, roots[i], scene, c->faces.data(), c->verts.data());
    refit_bvh(c->bvh, c->faces.data(), c->verts.data(), threads);
    c->restr_pending = false;
    return 0;
}


}  // extern "C"

// ====================================================================== inter-object front
// A rigid object's triangles queried against the deformable BVH (BvttFrontLooseInter,
// collision/bvtt_front/BvttFrontLooseInter.cu:16-260). The rigid object gets its own LBVH
// (LBvhRigid, collision/lbvh/LBvhRigid.cu: build = codes, sort, emission; refit = boxes), whose
// sorted primitives are the queries: query q is the rigid leaf slot q, with that leaf's box.
// Each query walks the whole deformable tree from its root (genLooseInterFrontsWithLog,
// BvttFrontGenerateKernels.cu:92-125, which stops one query short: `idx >= travPrimSize - 1`;
// here every rigid primitive is a query). Pairs are (deformable triangle, rigid triangle).
namespace {

struct InterCtx {
    Ctx* def = nullptr;
    int64_t T = 0, V = 0;
    std::vector<int32_t> faces;
    std::vector<float> verts;
    Bvh rb;
    bool built = false, front_valid = false;
    int64_t def_version = -1;
    std::vector<Entry> fi, fe;  // (rigid slot, deformable node)
    std::vector<Pair> bp, np;   // (deformable triangle, rigid triangle)
    int64_t visits = 0;
    int last_front_op = -1;
};

struct InterSink {
    std::vector<Entry> fi, fe;
    std::vector<std::pair<int32_t, int32_t>> pairs;  // (deformable leaf, rigid slot)
    int64_t visits = 0;
};

// Stackless walk of the deformable tree for one query box (A.6), from encoded `st` to `stop`.
template <bool FRONT>
void inter_walk(const Bvh& b, const Box& bv, int32_t q, int32_t st, int32_t stop, InterSink& s) {
    while (st != stop) {
        int32_t t = st & 1;
        st >>= 1;
        int32_t lbd;
        if (!t) {
            lbd = b.rx[st];
            t = b.lpar[lbd];
            for (; st <= t; ++st) {
                ++s.visits;
                if (!overlaps(b.ibox[st], bv)) break;
            }
        } else {
            t = st - 1;
            lbd = st;
        }
        if (st == t + 1) {
            ++s.visits;
            if (overlaps(b.lbox[lbd], bv)) s.pairs.push_back({lbd, q});
            if (FRONT) s.fe.push_back({q, lbd});
            st = b.llca[lbd + 1];
        } else {
            if (FRONT) s.fi.push_back({q, st});
            st = b.llca[b.ry[st] + 1];
        }
    }
}

template <class F>
void inter_run(InterCtx& x, int64_t n, int threads, F&& fn, bool keep_front) {
    const int nt = threads <= 0 ? (int)std::thread::hardware_concurrency() : threads;
    std::vector<InterSink> sinks(nt < 1 ? 1 : nt);
    parallel_for(n, (int)sinks.size(), [&](int64_t a, int64_t e, int t) {
        for (int64_t k = a; k < e; ++k) fn(k, sinks[t]);
    });
    const Bvh& b = x.def->bvh;
    if (keep_front) {
        x.fi.clear();
        x.fe.clear();
    }
    x.bp.clear();
    x.visits = 0;
    for (auto& s : sinks) {
        if (keep_front) {
            x.fi.insert(x.fi.end(), s.fi.begin(), s.fi.end());
            x.fe.insert(x.fe.end(), s.fe.begin(), s.fe.end());
        }
        for (auto& pr : s.pairs) x.bp.push_back({b.prim_idx[pr.first], x.rb.prim_idx[pr.second]});
        x.visits += s.visits;
    }
}

// genLooseInterFrontsWithLog / pureBvhInterCD (BvhCDKernels.cu:51-75).
template <bool FRONT>
void inter_traverse(InterCtx& x, int threads) {
    const Bvh& b = x.def->bvh;
    const int32_t root = b.llca[0];
    inter_run(x, x.rb.n, threads, [&](int64_t q, InterSink& s) {
        inter_walk<FRONT>(b, x.rb.lbox[q], (int32_t)q, root, -1, s);
    }, FRONT);
}

// maintainInt/ExtLooseInterFrontsWithLog (BvttFrontPruneSproutKernels.cu:203-275).
void inter_update(InterCtx& x, int threads) {
    const Bvh& b = x.def->bvh;
    const std::vector<Entry> fi = x.fi, fe = x.fe;
    const int64_t ni = (int64_t)fi.size(), ne = (int64_t)fe.size();
    inter_run(x, ni + ne, threads, [&](int64_t k, InterSink& s) {
        if (k < ni) {
            const Entry e = fi[k];
            int32_t st = e.node;
            const Box& bv = x.rb.lbox[e.q];
            ++s.visits;
            if (!overlaps(b.ibox[st], bv)) {  // prune: climb the left spine
                const int32_t t = b.llca[b.rx[st]] >> 1;
                for (st--; st >= t; st--) {
                    ++s.visits;
                    if (overlaps(b.ibox[st], bv)) break;
                }
                if (st < t && st + 1 > 0) {
                    ++s.visits;
                    if (!overlaps(b.ibox[b.par[st + 1]], bv)) return;
                }
                s.fi.push_back({e.q, st + 1});
                return;
            }
            inter_walk<true>(b, bv, e.q, (b.lc[st] << 1) | (int32_t)(b.mark[st] & 1u), b.llca[b.ry[st] + 1], s);
        } else {
            const Entry e = fe[k - ni];
            int32_t st = e.node;
            const Box& bv = x.rb.lbox[e.q];
            ++s.visits;
            if (overlaps(b.lbox[st], bv)) {
                s.fe.push_back(e);
                s.pairs.push_back({st, e.q});
                return;
            }
            const int32_t gfa = b.lpar[st];
            ++s.visits;
            if (overlaps(b.ibox[gfa], bv)) {
                s.fe.push_back(e);
                return;
            }
            if ((b.lmark[st] & 4u) == 4u) return;
            const int32_t top = b.llca[st] >> 1;
            for (st = gfa - 1; st >= top; st--) {
                ++s.visits;
                if (overlaps(b.ibox[st], bv)) break;
            }
            if (st < top && st + 1 > 0) {
                ++s.visits;
                if (!overlaps(b.ibox[b.par[st + 1]], bv)) return;
            }
            s.fi.push_back({e.q, st + 1});
        }
    }, true);
}

// keepInt/ExtLooseInterFronts (BvttFrontKeepKernels.cu:60-96).
void inter_keep(InterCtx& x, int threads) {
    const Bvh& b = x.def->bvh;
    const int64_t ni = (int64_t)x.fi.size(), ne = (int64_t)x.fe.size();
    inter_run(x, ni + ne, threads, [&](int64_t k, InterSink& s) {
        if (k < ni) {
            const Entry e = x.fi[k];
            const Box& bv = x.rb.lbox[e.q];
            ++s.visits;
            if (!overlaps(b.ibox[e.node], bv)) return;
            inter_walk<false>(b, bv, e.q, (b.lc[e.node] << 1) | (int32_t)(b.mark[e.node] & 1u),
                              b.llca[b.ry[e.node] + 1], s);
        } else {
            const Entry e = x.fe[k - ni];
            ++s.visits;
            if (overlaps(b.lbox[e.node], x.rb.lbox[e.q])) s.pairs.push_back({e.node, e.q});
        }
    }, false);
}

// Narrow phase of inter pairs: the deformable triangle is P, the rigid one Q.
void inter_narrow(InterCtx& x, int threads) {
    x.np.clear();
    if (!x.def->narrow) return;
    std::vector<uint8_t> hit(x.bp.size(), 0);
    parallel_for((int64_t)x.bp.size(), threads, [&](int64_t a, int64_t e, int) {
        for (int64_t k = a; k < e; ++k) {
            float P[9], Q[9];
            gather_tri(x.def->faces.data(), x.def->verts.data(), x.bp[k].a, P);
            gather_tri(x.faces.data(), x.verts.data(), x.bp[k].b, Q);
            hit[k] = sat_tt<false>(P, Q) ? 1 : 0;
        }
    });
    for (size_t k = 0; k < x.bp.size(); ++k)
        if (hit[k]) x.np.push_back(x.bp[k]);
}

}  // namespace

extern "C" {

void* orc_inter_create(void* def, int64_t T, int64_t V, const int32_t* faces) {
    InterCtx* x = new InterCtx();
    x->def = static_cast<Ctx*>(def);
    x->T = T;
    x->V = V;
    x->faces.assign(faces, faces + 3 * T);
    x->verts.assign(3 * V, 0.f);
    return x;
}
void orc_inter_destroy(void* p) { delete static_cast<InterCtx*>(p); }
void orc_inter_update_vertices(void* p, const float* xyz) {
    InterCtx* x = static_cast<InterCtx*>(p);
    std::memcpy(x->verts.data(), xyz, sizeof(float) * 3 * x->V);
}
// rigid_op: 0 BUILD, 1 REFIT; front_op: 0 PURE_BVH_CD, 1 GENERATE, 2 UPDATE, 3 REORDER, 4 KEEP.
// The front is regenerated whenever the deformable tree's topology changed since it was made.
int orc_inter_step(void* p, int rigid_op, int front_op, int threads) {
    InterCtx* x = static_cast<InterCtx*>(p);
    if (!x->def->built || x->T <= 0) return -5;
    if (rigid_op == 0) {
        full_build(x->rb, x->T, x->faces.data(), x->verts.data(), threads);
        x->built = true;
    } else if (rigid_op == 1) {
        if (!x->built) return -5;
        refit_bvh(x->rb, x->faces.data(), x->verts.data(), threads);
    } else {
        return -1;
    }
    if (x->def_version != x->def->topo_version || rigid_op == 0) x->front_valid = false;
    if (front_op != 0 && !x->front_valid) front_op = 1;
    switch (front_op) {
        case 0: inter_traverse<false>(*x, threads); break;
        case 1:
            inter_traverse<true>(*x, threads);
            x->front_valid = true;
            x->def_version = x->def->topo_version;
            break;
        case 2: inter_update(*x, threads); break;
        case 3:  // balance: pairs from the unchanged cut (the order of entries is not observable)
        case 4: inter_keep(*x, threads); break;
        default: return -1;
    }
    x->last_front_op = front_op;
    inter_narrow(*x, threads);
    return 0;
}
int64_t orc_inter_get_pairs(void* p, int filtered, int32_t* out, int64_t cap) {
    InterCtx* x = static_cast<InterCtx*>(p);
    const auto& v = filtered ? x->np : x->bp;
    const int64_t n = (int64_t)v.size();
    for (int64_t k = 0; k < n && k < cap && out; ++k) { out[2 * k] = v[k].a; out[2 * k + 1] = v[k].b; }
    return n;
}
int64_t orc_inter_export_front(void* p, int which, int32_t* out, int64_t cap) {
    InterCtx* x = static_cast<InterCtx*>(p);
    const auto& v = which ? x->fe : x->fi;
    const int64_t n = (int64_t)v.size();
    for (int64_t k = 0; k < n && k < cap && out; ++k) { out[2 * k] = v[k].q; out[2 * k + 1] = v[k].node; }
    return n;
}
// The rigid object's sorted codes and slot -> triangle map.
void orc_inter_export_rigid(void* p, uint32_t* codes, int32_t* prim_idx) {
    InterCtx* x = static_cast<InterCtx*>(p);
    if (codes) std::memcpy(codes, x->rb.codes.data(), 4 * x->rb.codes.size());
    if (prim_idx) std::memcpy(prim_idx, x->rb.prim_idx.data(), 4 * x->rb.prim_idx.size());
}
int64_t orc_inter_last_front_op(void* p) { return static_cast<InterCtx*>(p)->last_front_op; }

}  // extern "C"
