"""CPU tests that pin the oracle (oracle/bvhcd_oracle.cpp) before anything is checked against it.

Each check uses something other than the oracle's own formula: closed-form known answers,
an independent bit-by-bit Morton interleave, numpy's stable argsort, a naive recursive
"leftmost maximal delta" tree builder, brute-force all-pairs overlap, subtree min/max boxes,
and a double-precision separating-axis test.
"""
import json
import os

import numpy as np
import pytest

from conftest import check_cut, load_package, sheet_sphere

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- Morton
def test_expand_bits_known_answers(oracle):
    assert oracle.expand_bits(0x3FF) == 0x09249249
    assert oracle.expand_bits(0) == 0
    assert oracle.expand_bits(1) == 1
    assert oracle.expand_bits(2) == 0b1000
    for v in range(1024):
        ref = 0
        for b in range(10):
            ref |= ((v >> b) & 1) << (3 * b)
        assert oracle.expand_bits(v) == ref


def test_morton3d_known_answers(oracle):
    assert oracle.morton3d(0.0, 0.0, 0.0) == 0
    assert oracle.morton3d(1.0, 1.0, 1.0) == 0x3FFFFFFF      # clamp to 1023
    assert oracle.morton3d(2.0, -1.0, 0.0) == 0x24924924     # x clamps high, y clamps low
    # x owns the highest bit of each triple
    assert oracle.morton3d(0.5, 0.0, 0.0) == 1 << 29
    assert oracle.morton3d(0.0, 0.5, 0.0) == 1 << 28
    assert oracle.morton3d(0.0, 0.0, 0.5) == 1 << 27
    assert oracle.morton3d(float("nan"), 0.0, 0.0) == 0      # 0/0 extent -> 0


def _interleave(ix, iy, iz):
    out = np.zeros_like(ix, dtype=np.uint64)
    for b in range(10):
        out |= ((ix >> b) & 1).astype(np.uint64) << np.uint64(3 * b + 2)
        out |= ((iy >> b) & 1).astype(np.uint64) << np.uint64(3 * b + 1)
        out |= ((iz >> b) & 1).astype(np.uint64) << np.uint64(3 * b)
    return out.astype(np.uint32)


def _codes_numpy(faces, verts, box):
    v = verts[faces]                      # T x 3 x 3
    mn = v.min(axis=1).astype(np.float32)
    mx = v.max(axis=1).astype(np.float32)
    c = ((mn + mx).astype(np.float32) * np.float32(0.5)).astype(np.float32)
    off = (c - box[:3]).astype(np.float32)
    ext = (box[3:] - box[:3]).astype(np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        q = (off / ext).astype(np.float32)
    q = np.fmin(np.fmax(q * np.float32(1024.0), np.float32(0.0)), np.float32(1023.0))
    iq = q.astype(np.uint32)
    return _interleave(iq[:, 0], iq[:, 1], iq[:, 2])


@pytest.mark.parametrize("name,kw,frame", [
    ("spheres", dict(subdiv=3), 17),
    ("cloth256", dict(n=48), 0),        # flat sheet: zero z extent
    ("cloth256", dict(n=48), 299),
])
def test_morton_codes_independent(oracle, name, kw, frame):
    pkg = load_package()
    sc = pkg.Scene.preset(name, **kw)
    f, v = sc.faces(), sc.frame(frame)
    box = oracle.scene_box(f, v)
    np.testing.assert_array_equal(box[:3], v[np.unique(f)].min(axis=0))
    np.testing.assert_array_equal(box[3:], v[np.unique(f)].max(axis=0))
    np.testing.assert_array_equal(oracle.morton_codes(f, v, box), _codes_numpy(f, v, box))


def test_sort_is_stable_argsort(oracle):
    rng = np.random.default_rng(7)
    codes = rng.integers(0, 64, size=5000, dtype=np.uint32)  # many duplicates
    s, idx, pm = oracle.sort_codes(codes)
    ref = np.argsort(codes, kind="stable")
    np.testing.assert_array_equal(idx, ref)
    np.testing.assert_array_equal(s, codes[ref])
    np.testing.assert_array_equal(pm[idx], np.arange(codes.size))


# ---------------------------------------------------------------- topology
def naive_topology(c):
    """Recursive builder: split of [a, b] = leftmost k in [a, b-1] with maximal delta(k);
    internal nodes numbered in preorder; encodings as in the reference (A.5)."""
    n = len(c)
    I = n - 1

    def delta(k):
        x = int(c[k]) ^ int(c[k + 1])
        return x.bit_length()

    lc, rc, par, mark, rx, ry = ([0] * I for _ in range(6))
    lpar, lmark = [0] * n, [0] * n
    lca = [None] * (n + 1)
    nid = [0]
    stack = [(0, n - 1, -1, False)]  # range, parent id, is right child
    while stack:
        a, b, p, right = stack.pop()
        if a == b:
            lpar[a] = p
            lmark[a] = 7 if right else 3
            continue
        me = nid[0]
        nid[0] += 1
        best, k = -1, a
        for j in range(a, b):
            d = delta(j)
            if d > best:
                best, k = d, j
        rx[me], ry[me], par[me] = a, b, p
        mark[me] = 4 if right else 0
        if lca[a] is None:
            lca[a] = me
        # children: left [a, k], right [k+1, b]; left subtree is numbered first
        if k == a:
            lc[me] = a
            mark[me] |= 1
        else:
            lc[me] = me + 1
        if k + 1 == b:
            rc[me] = b
            mark[me] |= 2
        else:
            rc[me] = me + 1 + (k - a)  # internal nodes in the left subtree: k - a
        stack.append((k + 1, b, me, True))
        stack.append((a, k, me, False))
    enc = [(lca[x] << 1) if lca[x] is not None else (x << 1 | 1) for x in range(n)] + [-1]
    return dict(int_lc=lc, int_rc=rc, int_par=par, int_mark=mark, int_rx=rx, int_ry=ry,
                leaf_par=lpar, leaf_mark=lmark, leaf_lca=enc)


@pytest.mark.parametrize("seed,n,hi", [(1, 2, 4), (2, 3, 2), (3, 17, 8), (4, 64, 4), (5, 300, 1 << 30),
                                       (6, 300, 16), (7, 257, 1), (8, 1000, 200)])
def test_topology_matches_naive_builder(oracle, seed, n, hi):
    rng = np.random.default_rng(seed)
    c = np.sort(rng.integers(0, hi, size=n, dtype=np.uint32))
    got = oracle.build_topology(c)
    ref = naive_topology(c)
    for k, v in ref.items():
        np.testing.assert_array_equal(got[k], np.asarray(v), err_msg=k)


def check_invariants(t, n):
    """Intent of checkBVHIntegrity (LBvhBuildKernels.cu:235-293)."""
    I = n - 1
    lc, rc, par, mark, rx, ry = (t[k] for k in ("int_lc", "int_rc", "int_par", "int_mark", "int_rx", "int_ry"))
    assert t["leaf_lca"][n] == -1
    assert par[0] == -1 and rx[0] == 0 and ry[0] == n - 1
    for i in range(I):
        if not (mark[i] & 1):
            assert lc[i] == i + 1                      # left internal child follows in preorder
            assert rx[lc[i]] == rx[i] and par[lc[i]] == i and not (mark[lc[i]] & 4)
        else:
            assert lc[i] == rx[i] and t["leaf_par"][lc[i]] == i and t["leaf_mark"][lc[i]] == 3
        if not (mark[i] & 2):
            assert ry[rc[i]] == ry[i] and par[rc[i]] == i and (mark[rc[i]] & 4)
            assert rc[i] == i + (ry[i] - rx[i]) - (ry[rc[i]] - rx[rc[i]])
        else:
            assert rc[i] == ry[i] and t["leaf_par"][rc[i]] == i and t["leaf_mark"][rc[i]] == 7
    for x in range(n):
        e = t["leaf_lca"][x]
        if e & 1:
            assert e >> 1 == x and t["leaf_mark"][x] == 7
        else:
            top = e >> 1
            assert rx[top] == x and (top == 0 or (mark[top] & 4))


def test_topology_invariants_random(oracle):
    rng = np.random.default_rng(11)
    for n in (2, 5, 100, 2000):
        c = np.sort(rng.integers(0, 1 << 12, size=n, dtype=np.uint32))
        check_invariants(oracle.build_topology(c), n)


def test_refit_is_subtree_min_max(oracle):
    pkg = load_package()
    sc = pkg.Scene.preset("spheres", subdiv=2)
    f, v = sc.faces(), sc.frame(3)
    ctx = oracle.OracleContext(f, sc.num_verts)
    ctx.update_vertices(v)
    ctx.step(0, 0)
    b = ctx.export_bvh()
    lb = b["leaf_box"]
    np.testing.assert_array_equal(lb, oracle.tri_boxes(f, v)[b["prim_idx"]])
    for i in range(len(b["int_rx"])):
        seg = lb[b["int_rx"][i]: b["int_ry"][i] + 1]
        np.testing.assert_array_equal(b["int_box"][i, :3], seg[:, :3].min(axis=0))
        np.testing.assert_array_equal(b["int_box"][i, 3:], seg[:, 3:].max(axis=0))
    check_invariants(b, sc.num_tris)


# ---------------------------------------------------------------- narrow phase
def tri(*p):
    return np.array(p, np.float32).reshape(3, 3)


def test_sat_special_cases(oracle):
    A = tri(0, 0, 0, 1, 0, 0, 0, 1, 0)
    assert oracle.sat(A, A)                                          # identical
    assert oracle.sat(A, tri(0, 0, 0, -1, 0, 0, 0, -1, 0.3))         # shares one vertex (touching)
    assert oracle.sat(A, tri(1, 0, 0, 0, 1, 0, 1, 1, 0))             # shares an edge
    assert not oracle.sat(A, A + np.float32(5.0))                    # far apart
    assert oracle.sat(A, tri(0.2, 0.2, 0, 0.9, 0.1, 0, 0.1, 0.9, 0))  # coplanar overlapping
    assert not oracle.sat(A, tri(2, 2, 0, 3, 2, 0, 2, 3, 0))         # coplanar disjoint
    assert oracle.sat(A, tri(0.25, 0.25, -1, 0.25, 0.25, 1, 0.3, 0.2, 1))  # piercing edge
    assert not oracle.sat(A, tri(0.25, 0.25, 0.1, 0.25, 0.3, 1, 0.3, 0.2, 1))  # above the plane
    # edge-edge crossing of two perpendicular triangles
    assert oracle.sat(tri(-1, 0, 0, 1, 0, 0, 0, 0, 1), tri(0, -1, 0.5, 0, 1, 0.5, 0, 0, -1))


def test_sat_shared_vertex_always_hits(oracle):
    """Property behind the narrow-phase shortcut: with every |coordinate| <= 1e8, two triangles
    that share a vertex are reported as intersecting, whatever their shape or scale, in either
    argument order and with or without fused multiply-adds."""
    rng = np.random.default_rng(5)
    for scale in (1e-30, 1e-6, 1.0, 1e3, 1e8):
        for _ in range(400):
            P = (rng.normal(size=(3, 3)) * scale).astype(np.float32)
            P = np.clip(P, -1e8, 1e8).astype(np.float32)
            Q = (rng.normal(size=(3, 3)) * scale).astype(np.float32)
            Q = np.clip(Q, -1e8, 1e8).astype(np.float32)
            i, j = rng.integers(0, 3, size=2)
            Q[j] = P[i]
            assert oracle.sat(P, Q) and oracle.sat(Q, P), (P, Q)
            assert oracle.sat(P, Q, fma=True) and oracle.sat(Q, P, fma=True), (P, Q)


def _sat_double(P, Q):
    """Separating-axis test in float64 with the same 17 axes; returns (hit, min margin)."""
    P = P.astype(np.float64)
    Q = Q.astype(np.float64)
    e = [P[1] - P[0], P[2] - P[1], P[0] - P[2]]
    f = [Q[1] - Q[0], Q[2] - Q[1], Q[0] - Q[2]]
    n, m = np.cross(e[0], e[1]), np.cross(f[0], f[1])
    axes = [n, m] + [np.cross(a, b) for a in e for b in f] + [np.cross(a, n) for a in e] + [np.cross(b, m) for b in f]
    margin = np.inf
    for ax in axes:
        p, q = P @ ax, Q @ ax
        sep = max(p.min() - q.max(), q.min() - p.max())
        scale = np.abs(ax).sum() * (np.abs(P).max() + np.abs(Q).max() + 1e-30)
        if sep > 0:
            return False, sep / scale
        margin = min(margin, -sep / scale if scale > 0 else np.inf)
    return True, margin


def test_sat_vs_double_precision(oracle):
    """Random triangle pairs: wherever the fp32 test (any of its four evaluation variants: lower
    or higher id as P, with or without FMA) disagrees with the exact rational test, the pair's
    exact relative margin is within 1e-6 of the boundary; every such pair is listed."""
    import narrow_band
    rng = np.random.default_rng(3)
    listed = []
    for k in range(4000):
        P = rng.normal(size=(3, 3)).astype(np.float32)
        Q = (rng.normal(size=(3, 3)) * 0.7 + rng.normal(size=3) * 0.8).astype(np.float32)
        hits = {oracle.sat(P, Q), oracle.sat(Q, P), oracle.sat(P, Q, fma=True), oracle.sat(Q, P, fma=True)}
        ref, _ = _sat_double(P, Q)
        if hits != {ref}:
            mg = narrow_band.exact_margin(P, Q)
            listed.append((k, float(mg)))
            assert abs(mg) <= 1e-6, (k, float(mg))
            assert (mg > 0) != ref or mg == 0  # the double test agrees with the exact one
    assert len(listed) <= 4, listed


def test_exact_margin_known_answers():
    """exact_margin on hand-made pairs: a unit gap along z between parallel triangles, touching
    triangles (margin 0), and a crossing pair (negative margin)."""
    import narrow_band
    from fractions import Fraction
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    Q = P + np.array([0, 0, 1], np.float32)
    # gap 1 along n = (0,0,1) (|n|_1 = 1), R = 1 after translating by P's first vertex
    assert narrow_band.exact_margin(P, Q) == Fraction(1)
    assert narrow_band.exact_margin(P, P + np.array([0, 0, 0.5], np.float32)) == Fraction(1, 2)
    T = np.array([[1, 0, 0], [2, 0, 0], [1, 1, 0]], np.float32)  # touches P at the vertex (1,0,0)
    assert narrow_band.exact_margin(P, T) == 0
    X = np.array([[0.2, 0.2, -1], [0.2, 0.2, 1], [0.3, 0.3, 1]], np.float32)  # pierces P
    assert narrow_band.exact_margin(P, X) < 0


def test_narrow_band_fixture(oracle):
    """The committed list of variant disagreements (tests/golden/narrow_band.json, written by
    make_band.py over every frame of the two-sphere scene and the 256^2 cloth and three folded
    frames each of the 1024^2 cloth and the 4-layer scene) stays within 1e-6 of the boundary,
    and re-scanning a sample of those frames reproduces it."""
    import narrow_band
    pkg = load_package()
    meta = json.load(open(os.path.join(GOLDEN, "narrow_band.json")))
    assert meta["band"] == 1e-6
    for name, cfg in meta["configs"].items():
        assert cfg["broad_phase_pairs"] > 0, name
        for fr, a, b, mask, mg in cfg["disagreements"]:
            assert mask not in (0, 15) and abs(mg) <= 1e-6, (name, fr, a, b)
    for name, frames in (("spheres", (0, 199)), ("cloth256", (150, 214, 299))):
        sc = pkg.Scene.preset(name)
        f = sc.faces()
        ctx = oracle.OracleContext(f, sc.num_verts, narrow=False)
        listed = meta["configs"][name]["disagreements"]
        for fr in frames:
            v = sc.frame(fr)
            ctx.update_vertices(v)
            ctx.step(0, 0)
            got = narrow_band.band_of_frame(oracle, f, v, ctx.pairs(False))
            want = [tuple(r[1:]) for r in listed if r[0] == fr]
            assert [(a, b, m) for a, b, m, _ in got] == [(a, b, m) for a, b, m, _ in want], (name, fr)
            for (a, b, m, mg), w in zip(got, want):
                assert abs(mg - w[3]) <= 1e-12
        ctx.close()


# ---------------------------------------------------------------- brute force
def _bruteforce_numpy(f, v, narrow, oracle):
    b = oracle.tri_boxes(f, v)
    mn, mx = b[:, :3], b[:, 3:]
    ov = np.all((mn[:, None, :] <= mx[None, :, :]) & (mx[:, None, :] >= mn[None, :, :]), axis=2)
    i, j = np.nonzero(np.triu(ov, 1))
    if narrow:
        keep = [oracle.sat(v[f[a]], v[f[c]]) for a, c in zip(i, j)]
        i, j = i[keep], j[keep]
    return np.stack([i, j], 1)


def test_bruteforce_matches_numpy(oracle):
    pkg = load_package()
    sc = pkg.Scene.preset("spheres", subdiv=2, spacing=1.9)
    f, v = sc.faces(), sc.frame(0)
    for narrow in (False, True):
        got = oracle.pair_set(oracle.bruteforce_pairs(f, v, narrow=narrow))
        ref = oracle.pair_set(_bruteforce_numpy(f, v, narrow, oracle))
        np.testing.assert_array_equal(got, ref)


# ---------------------------------------------------------------- pipeline vs brute force
def front_partitions(fi, fe, b, n):
    """Front cut completeness: per query i, the front subtrees cover leaves (i, n-1] once."""
    cover = np.zeros(n + 1, np.int64)
    rx, ry = b["int_rx"], b["int_ry"]
    q = np.concatenate([fi[:, 0], fe[:, 0]])
    lo = np.concatenate([rx[fi[:, 1]], fe[:, 1]])
    hi = np.concatenate([ry[fi[:, 1]], fe[:, 1]])
    assert np.all(lo > q)
    # sum over queries of covered lengths must equal sum of (n-1-i)
    assert int((hi - lo + 1).sum()) == int(((n - 1) - np.arange(n - 1)).sum())
    # and no overlap within a query: sort by (q, lo) and check contiguity
    order = np.lexsort((lo, q))
    q, lo, hi = q[order], lo[order], hi[order]
    starts = np.r_[True, q[1:] != q[:-1]]
    assert np.all(lo[starts] == q[starts] + 1)
    nxt = ~starts[1:]
    assert np.all(lo[1:][nxt] == hi[:-1][nxt] + 1)
    ends = np.r_[q[1:] != q[:-1], True]
    assert np.all(hi[ends] == n - 1)


@pytest.mark.parametrize("name,kw,frames", [
    ("spheres", dict(subdiv=3, spacing=2.05, speed=0.004), 30),
    ("cloth256", dict(n=40, frames=30), 30),
])
def test_oracle_pipeline_equals_bruteforce(oracle, name, kw, frames):
    pkg = load_package()
    sc = pkg.Scene.preset(name, **kw)
    f = sc.faces()
    ctx = oracle.OracleContext(f, sc.num_verts, bvh_cycle=12, front_cycle=6)
    seen_ops = set()
    for fr in range(frames):
        v = sc.frame(fr)
        ctx.update_vertices(v)
        ctx.step_scheduled()
        st = ctx.stats()
        seen_ops.add((st["last_bvh_op"], st["last_front_op"]))
        bf_bp = oracle.pair_set(oracle.bruteforce_pairs(f, v, narrow=False))
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)), bf_bp, err_msg=f"frame {fr} bp")
        bf_np = oracle.pair_set(oracle.bruteforce_pairs(f, v, narrow=True))
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), bf_np, err_msg=f"frame {fr} np")
        if st["last_front_op"] != 0:
            front_partitions(ctx.front(0), ctx.front(1), ctx.export_bvh(), sc.num_tris)
    fronts = {fo for _, fo in seen_ops}
    assert {1, 2, 3, 4} <= fronts, seen_ops


def test_oracle_explicit_ops_and_restructure(oracle):
    """Every front op after every BVH op gives the brute-force pairs; forced restructure
    keeps the tree valid."""
    pkg = load_package()
    sc = pkg.Scene.preset("cloth256", n=32, frames=20)
    f = sc.faces()
    ctx = oracle.OracleContext(f, sc.num_verts)
    ctx.update_vertices(sc.frame(0))
    ctx.step(0, 1)
    for fr in range(1, 20):
        v = sc.frame(fr)
        ctx.update_vertices(v)
        fo = [2, 4, 3, 2, 0][fr % 5]
        bo = 2 if ctx.stats()["restr_pending"] else 1
        ctx.step(bo, fo)
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)),
                                      oracle.pair_set(oracle.bruteforce_pairs(f, v, True)), err_msg=str(fr))
        check_invariants(ctx.export_bvh(), sc.num_tris)


def _disjoint_roots(b, max_leaves, every):
    rx, ry, par = b["int_rx"], b["int_ry"], b["int_par"]
    cnt = ry - rx + 1
    cand = [i for i in range(len(rx)) if cnt[i] <= max_leaves and (par[i] < 0 or cnt[par[i]] > max_leaves)]
    return np.array(cand[::every], np.int32)


@pytest.mark.parametrize("max_leaves,every", [(64, 2), (500, 3), (100000, 1)])
def test_oracle_front_restructure(oracle, max_leaves, every):
    """Incremental front restructure after an explicit subtree rebuild: the repaired front is
    again a complete cut for every query, its pairs are the brute-force pairs, and later front
    ops (UPDATE / REORDER / KEEP) stay exact."""
    pkg = load_package()
    sc = pkg.Scene.preset("cloth256", n=48, frames=40)
    f = sc.faces()
    n = sc.num_tris
    ctx = oracle.OracleContext(f, sc.num_verts)
    ctx.update_vertices(sc.frame(0))
    ctx.step(0, 1)
    for fr in range(1, 8):
        ctx.update_vertices(sc.frame(fr * 4))
        ctx.step(1, 2 if fr % 3 == 0 else 4)
    before = {w: ctx.front(w) for w in (0, 1)}
    b0 = ctx.export_bvh()
    roots = _disjoint_roots(b0, max_leaves, every)
    assert roots.size >= 1
    ctx.restructure(roots)
    b1 = ctx.export_bvh()
    check_invariants(b1, n)
    # entries untouched by the rebuilt subtrees survive unchanged
    inside = np.zeros(n, bool)
    for r in roots:
        inside[b0["int_rx"][r]:b0["int_ry"][r] + 1] = True
    for fr, fo in [(33, 4), (34, 4), (35, 2), (36, 3), (37, 4), (38, 2)]:
        v = sc.frame(fr)
        ctx.update_vertices(v)
        ctx.step(1, fo)
        st = ctx.stats()
        if fr == 33:
            assert st["last_front_op"] == 5
            fi = ctx.front(0)
            keep_i = before[0][~inside[b0["int_rx"][before[0][:, 1]]] & ~inside[before[0][:, 0]]]
            got = {tuple(x) for x in fi.tolist()}
            assert all(tuple(x) in got for x in keep_i.tolist())
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)),
                                      oracle.pair_set(oracle.bruteforce_pairs(f, v, False)), err_msg=str(fr))
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)),
                                      oracle.pair_set(oracle.bruteforce_pairs(f, v, True)), err_msg=str(fr))
        front_partitions(ctx.front(0), ctx.front(1), ctx.export_bvh(), n)


def test_plain_schedule_table(oracle):
    pkg = load_package()
    got = [oracle.plain_schedule(f) for f in range(26)]
    assert got == [pkg.plain_schedule(f) for f in range(26)]
    # frame 0 build+generate; 1-4 refit+keep; 5 refit+reorder; 6 refit+update; 11 keep; 12 build+generate
    assert got[0] == (0, 1) and got[1] == (1, 4) and got[5] == (1, 3) and got[6] == (1, 2)
    assert got[11] == (1, 4) and got[12] == (0, 1) and got[17] == (1, 3) and got[18] == (1, 2)


# Known answers read off BenchmarkLogic.hpp:56-153 for bvh_cycle 12, front_cycle 6:
# (frame, bvh op, front op); bvh 0 BUILD 1 REFIT, front 0 PURE 1 GENERATE 2 UPDATE 3 REORDER 4 KEEP.
SCHEME_KNOWN = {
    2: [(0, 0, 0), (7, 0, 0), (12, 0, 0)],                                   # GENERATE
    3: [(0, 0, 1), (7, 0, 1), (12, 0, 1)],                                   # FRONT_GENERATE
    4: [(0, 0, 1), (1, 1, 4), (5, 1, 4), (6, 1, 2), (11, 1, 4), (12, 0, 1), (18, 1, 2)],
    5: [(0, 0, 1), (1, 1, 4), (5, 1, 3), (6, 1, 2), (11, 1, 3), (12, 1, 2), (24, 1, 2), (29, 1, 3)],
    6: [(0, 0, 1), (5, 1, 4), (6, 1, 2), (11, 1, 4), (12, 1, 2), (13, 1, 4)],
    7: [(0, 0, 1), (1, 1, 4), (5, 1, 3), (6, 1, 2), (12, 1, 2), (17, 1, 3)],
    8: [(0, 0, 1), (1, 1, 4), (6, 1, 1), (7, 1, 4), (12, 0, 1), (18, 1, 1), (24, 0, 1)],
    0: [(0, 0, 1), (5, 1, 3), (6, 1, 2), (11, 1, 4), (12, 0, 1), (17, 1, 3), (23, 1, 4)],
}


@pytest.mark.parametrize("scheme", sorted(SCHEME_KNOWN))
def test_scheme_tables(oracle, scheme):
    """Every plain scheme: the oracle's restatement matches the known answers, and the
    library's bvhcd_schedule_ops matches the oracle for 60 frames at two cycle settings."""
    pkg = load_package()
    for fr, b, f in SCHEME_KNOWN[scheme]:
        assert oracle.scheme_ops(scheme, fr) == (b, f), (scheme, fr)
    for bc, fc in ((12, 6), (8, 4)):
        want = [oracle.scheme_ops(scheme, fr, bc, fc) for fr in range(60)]
        got = [pkg.plain_schedule(fr, bc, fc, scheduler=scheme) for fr in range(60)]
        assert got == want, (scheme, bc, fc)
    # the learned schedulers start from the plain table
    for sched in (pkg.SCHED_LEARNED, pkg.SCHED_LEARNED_NO_NET):
        assert [pkg.plain_schedule(fr, scheduler=sched) for fr in range(30)] == \
            [oracle.plain_schedule(fr) for fr in range(30)]


# ---------------------------------------------------------------- restructure heuristic
# frontSnapshot / checkFrontQuality (BvttFrontGenerateKernels.cu:26-49) and the decision tiers
# (LBvhFixedDeformable.cu:210-267), pinned on small hand-built trees and fronts. Front counts
# per node come from numpy bincount over front entries; the expected values below are computed
# from preorder subtree ranges with numpy, independently of the oracle's loops.

def _small_tree(oracle, n, seed):
    rng = np.random.default_rng(seed)
    codes = np.sort(rng.choice(1 << 20, size=n, replace=False)).astype(np.uint32)
    return oracle.build_topology(codes)


def _offsets(int_nodes, ext_leaves, n):
    ci = np.bincount(int_nodes, minlength=n - 1).astype(np.int64)
    ce = np.bincount(ext_leaves, minlength=n).astype(np.int64)
    intbeg = np.concatenate([[0], np.cumsum(ci)])            # n entries: intbeg[n-1] = total
    extbeg = np.concatenate([[0], np.cumsum(ce)])            # n + 1 entries
    return ci, ce, intbeg.astype(np.uint32), extbeg.astype(np.uint32)


def _subtree_counts(topo, ci, ce):
    rx, ry = topo["int_rx"].astype(np.int64), topo["int_ry"].astype(np.int64)
    I = rx.size
    ic = np.array([ci[i:i + ry[i] - rx[i]].sum() for i in range(I)], np.int64)  # preorder subtree
    ec = np.array([ce[rx[i]:ry[i] + 1].sum() for i in range(I)], np.int64)
    return ic, ec


def _expected_flags(ic, ec, snap, total, ext_plus_one=False, range_share=None):
    with np.errstate(divide="ignore", invalid="ignore"):
        ratio = ic / (ec + (1 if ext_plus_one else 0))
    share = 0.7 * (range_share if range_share is not None else ic / total)
    return ratio > (snap + np.float32(1)).astype(np.float64) - share


def test_quality_snapshot_and_check_known_answers(oracle):
    """Snapshot = intCnt / (extCnt + 1) as float; a node is flagged when, on the later front,
    intCnt / extCnt > (snapshot + 1) - 0.7 intCnt / total internal entries. The marks are a
    +1 / -1 difference pair over the node's preorder range and rtroot[rangex] = the smallest
    flagged node id. Among the seeded cases, some are such that each plausible slip changes the
    flags: extCnt + 1 in the check, extCnt without + 1 in the snapshot, the 0.7 term on the
    range length (the reference's commented-out variant) instead of the front share."""
    n = 40
    topo = _small_tree(oracle, n, 3)
    rx, ry = topo["int_rx"], topo["int_ry"]
    I = n - 1
    span = (ry - rx).astype(np.float64) / I
    caught = {"ext_plus_one": False, "range_share": False, "snapshot_plus_one": False}
    for seed in range(48):
        rng = np.random.default_rng(seed)
        # generate-time front, then a later one whose internal entries pile up on the left
        ci0, ce0, ib0, eb0 = _offsets(rng.integers(0, I, 30), rng.integers(0, n, 60), n)
        snap, _, _ = oracle.quality_eval(rx, ry, ib0, eb0)
        ic0, ec0 = _subtree_counts(topo, ci0, ce0)
        np.testing.assert_array_equal(snap, (ic0 / (ec0 + 1)).astype(np.float32))
        later_int = np.concatenate([rng.integers(0, I // 3, int(rng.integers(20, 200))), rng.integers(0, I, 20)])
        ci1, ce1, ib1, eb1 = _offsets(later_int, rng.integers(0, n, int(rng.integers(10, 80))), n)
        _, ir, rt = oracle.quality_eval(rx, ry, ib1, eb1, snapshot=snap)
        ic1, ec1 = _subtree_counts(topo, ci1, ce1)
        total = int(ib1[-1])
        want = _expected_flags(ic1, ec1, snap, total)
        want_ir = np.zeros(n, np.int64)
        want_rt = np.full(n, 2**31 - 1, np.int64)
        for i in np.nonzero(want)[0]:
            want_ir[i] += 1
            want_ir[i + ry[i] - rx[i]] -= 1
            want_rt[rx[i]] = min(want_rt[rx[i]], i)
        np.testing.assert_array_equal(ir, want_ir, err_msg=str(seed))
        np.testing.assert_array_equal(rt, want_rt, err_msg=str(seed))
        caught["ext_plus_one"] |= not np.array_equal(want, _expected_flags(ic1, ec1, snap, total, ext_plus_one=True))
        caught["range_share"] |= not np.array_equal(want, _expected_flags(ic1, ec1, snap, total, range_share=span))
        with np.errstate(divide="ignore", invalid="ignore"):
            snap_slip = (ic0 / ec0).astype(np.float32)
        caught["snapshot_plus_one"] |= not np.array_equal(want, _expected_flags(ic1, ec1, snap_slip, total))
    assert all(caught.values()), caught


def test_quality_previous_offsets_matter(oracle):
    """checkFrontQuality reads the offsets of the previous ordering (it runs before the UPDATE's
    re-ordering): with the counts of the new front instead the flags differ on this case."""
    n = 32
    topo = _small_tree(oracle, n, 5)
    rx, ry = topo["int_rx"], topo["int_ry"]
    rng = np.random.default_rng(2)
    _, _, ib0, eb0 = _offsets(rng.integers(0, n - 1, 20), rng.integers(0, n, 40), n)
    snap, _, _ = oracle.quality_eval(rx, ry, ib0, eb0)
    _, _, ib1, eb1 = _offsets(rng.integers(0, 8, 90), rng.integers(0, n, 30), n)
    _, _, ib2, eb2 = _offsets(rng.integers(0, n - 1, 25), rng.integers(0, n, 45), n)
    _, ir1, _ = oracle.quality_eval(rx, ry, ib1, eb1, snapshot=snap)
    _, ir2, _ = oracle.quality_eval(rx, ry, ib2, eb2, snapshot=snap)
    assert not np.array_equal(ir1, ir2)


def _outermost(topo, flagged):
    par = topo["int_par"]
    out = []
    for r in sorted(flagged):
        p, inner = par[r], False
        while p >= 0:
            if p in flagged:
                inner = True
                break
            p = par[p]
        if not inner:
            out.append(r)
    return out


def _marks(topo, flagged, n):
    rx, ry = topo["int_rx"], topo["int_ry"]
    ir = np.zeros(n, np.int32)
    rt = np.full(n, 2**31 - 1, np.int32)
    for i in flagged:
        ir[i] += 1
        ir[i + ry[i] - rx[i]] -= 1
        rt[rx[i]] = min(rt[rx[i]], i)
    return ir, rt


def test_restructure_decision_tiers(oracle):
    """Outermost flagged subtrees (no flagged ancestor) and the three tiers on their internal
    node count m: m > I/2 rebuilds (2), m < I/8 refits (0), else restructures (1); both
    boundaries are strict (m = I >> 1 restructures, m = I >> 3 restructures)."""
    n = 257
    topo = _small_tree(oracle, n, 9)
    rx, ry, par = topo["int_rx"], topo["int_ry"], topo["int_par"]
    I = n - 1
    size = (ry - rx).astype(np.int64)
    rng = np.random.default_rng(4)
    seen = set()
    for trial in range(400):
        k = int(rng.integers(1, 12))
        flagged = set(int(x) for x in rng.choice(I, size=k, replace=False))
        ir, rt = _marks(topo, flagged, n)
        tag, roots, m = oracle.restr_decision(rx, ry, ir, rt)
        outer = _outermost(topo, flagged)
        assert sorted(roots.tolist()) == outer
        assert m == int(size[outer].sum())
        want = 2 if m > (I >> 1) else (0 if m < (I >> 3) else 1)
        assert tag == want, (trial, m)
        seen.add(tag)
    assert seen == {0, 1, 2}
    # exact boundaries: find nodes whose subtree holds exactly I >> 1 or I >> 3 internal nodes
    # when combined; a single node of each size decides the boundary tier
    for target, lo_tag in ((I >> 1, 1), (I >> 3, 1)):
        # greedy disjoint set of subtrees summing to target
        chosen, tot = [], 0
        for r in np.argsort(-size):
            r = int(r)
            if tot + size[r] <= target and all(not (rx[r] <= rx[c] <= ry[r] or rx[c] <= rx[r] <= ry[c]) for c in chosen):
                chosen.append(r)
                tot += int(size[r])
            if tot == target:
                break
        if tot != target:
            continue
        ir, rt = _marks(topo, set(chosen), n)
        tag, _, m = oracle.restr_decision(rx, ry, ir, rt)
        assert m == target and tag == lo_tag


# ---------------------------------------------------------------- inter-object front
def test_oracle_inter_front(oracle):
    """Inter-object front (rigid sphere against the folding sheet's BVH) under GENERATE, KEEP,
    UPDATE and REORDER, with the sheet's tree rebuilt on its own cycle: every frame's pairs equal
    brute force over all (sheet, sphere) triangle pairs, both before and after the narrow phase,
    and each query's front is a complete cut of the sheet's tree."""
    pkg = load_package()
    sc, fa, fb, nv = sheet_sphere(pkg)
    dctx = oracle.OracleContext(fa, nv)
    inter = oracle.OracleInter(dctx, fb, sc.num_verts - nv)
    ops = [1, 4, 4, 2, 4, 3]
    hits = 0
    for fr in range(sc.frames):
        v = sc.frame(fr)
        va, vb = v[:nv], v[nv:]
        dctx.update_vertices(va)
        dctx.step_scheduled()
        inter.update_vertices(vb)
        inter.step(0 if fr == 0 else 1, ops[fr % len(ops)])
        for narrow in (False, True):
            got = oracle.ordered_pair_set(inter.pairs(narrow))
            want = oracle.ordered_pair_set(oracle.inter_bruteforce(fa, va, fb, vb, narrow))
            np.testing.assert_array_equal(got, want, err_msg=f"frame {fr} narrow {narrow}")
        check_cut(inter.front(0), inter.front(1), dctx.export_bvh(), len(fb))
        hits += len(inter.pairs(True))
    assert hits > 0
    codes, idx = inter.rigid_order()
    # the rigid tree is refit, not rebuilt: its order is that of the first frame
    v0 = sc.frame(0)[nv:]
    s0, idx0, _ = oracle.sort_codes(oracle.morton_codes(fb, v0, oracle.scene_box(fb, v0)))
    np.testing.assert_array_equal(idx, idx0)
    np.testing.assert_array_equal(codes, s0)
