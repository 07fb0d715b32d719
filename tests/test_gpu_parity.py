"""GPU parity: the HIP pipeline against the CPU oracle through the C ABI.

Integer stages (Morton codes, sort permutation, topology, lca encoding) must be bit-exact;
boxes equal as floats; per-frame pair sets identical (compared as sorted unordered pairs);
front contents identical as multisets of (query leaf, node) entries.
"""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu

TOPO_KEYS = ["codes", "prim_idx", "prim_map", "int_lc", "int_rc", "int_par", "int_mark", "int_rx", "int_ry",
             "leaf_par", "leaf_mark", "leaf_lca"]


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def _ctx(pkg, sc, **kw):
    ctx = pkg.CollisionContext(sc.num_tris, sc.num_verts, **kw)
    ctx.set_topology(sc.faces(), sc.num_verts)
    return ctx


def _rows(a):
    a = np.asarray(a).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


def compare_bvh(got, ref, tag=""):
    for k in TOPO_KEYS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{tag} {k}")
    # float equality (-0 == +0)
    assert np.array_equal(got["leaf_box"], ref["leaf_box"]), tag
    assert np.array_equal(got["int_box"], ref["int_box"]), tag


@pytest.mark.parametrize("name,kw,frame", [
    ("spheres", dict(subdiv=4), 5),
    ("cloth256", dict(n=64), 0),       # flat: zero z extent, many duplicate codes
    ("cloth256", dict(n=64), 299),
    ("cloth256", dict(), 150),         # 130,050 triangles
])
def test_build_bitexact(pkg, oracle, name, kw, frame):
    print(f"[gpu] build {name} {kw} frame {frame}", flush=True)
    sc = pkg.Scene.preset(name, **kw)
    f, v = sc.faces(), sc.frame(frame)
    ctx = _ctx(pkg, sc)
    ctx.update_vertices(v)
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_PURE_BVH_CD)
    box = ctx.scene_box()
    assert np.array_equal(box, oracle.scene_box(f, v))
    orc = oracle.OracleContext(f, sc.num_verts)
    orc.update_vertices(v)
    orc.step(0, 0)
    compare_bvh(ctx.export_bvh(), orc.export_bvh(), f"{name} {frame}")
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)), oracle.pair_set(orc.pairs(False)))
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), oracle.pair_set(orc.pairs(True)))
    ctx.close()


@pytest.mark.parametrize("name,kw,frames", [
    ("spheres", dict(subdiv=4, spacing=2.1, speed=0.004), 40),
    ("cloth256", dict(n=64, frames=40), 40),
])
def test_plain_schedule_parity(pkg, oracle, name, kw, frames):
    print(f"[gpu] schedule {name} {kw}", flush=True)
    sc = pkg.Scene.preset(name, **kw)
    f = sc.faces()
    ctx = _ctx(pkg, sc)
    orc = oracle.OracleContext(f, sc.num_verts)
    for fr in range(frames):
        v = sc.frame(fr)
        ctx.update_vertices(v)
        orc.update_vertices(v)
        ctx.step_scheduled()
        orc.step_scheduled()
        gs, os_ = ctx.stats(), orc.stats()
        assert (gs["last_bvh_op"], gs["last_front_op"]) == (os_["last_bvh_op"], os_["last_front_op"]), fr
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)), oracle.pair_set(orc.pairs(False)),
                                      err_msg=f"bp frame {fr}")
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), oracle.pair_set(orc.pairs(True)),
                                      err_msg=f"np frame {fr}")
        if gs["last_front_op"] != pkg.FRONT_PURE_BVH_CD:
            for w in (0, 1):
                np.testing.assert_array_equal(_rows(ctx.export_front(w)), _rows(orc.front(w)),
                                              err_msg=f"front {w} frame {fr}")
        if gs["last_bvh_op"] in (pkg.BVH_BUILD, pkg.BVH_RESTRUCTURE):
            compare_bvh(ctx.export_bvh(), orc.export_bvh(), f"frame {fr}")
    ctx.close()


@pytest.mark.parametrize("scheme", [3, 4, 5, 6, 7, 8])
def test_scheme_parity(pkg, oracle, scheme):
    """The reference's other plain schemes (FRONT_GENERATE, STATIC_MANDATORY without the divergent
    mark, REFIT_ONLY_FRONT with and without it, the two ARCSim schemes) through
    bvhcd_step_scheduled: ops, pair sets and fronts equal the oracle's every frame, with the
    restructure heuristic on."""
    print(f"[gpu] scheme {scheme}", flush=True)
    sc = pkg.Scene.preset("cloth256", n=56, frames=40)
    f = sc.faces()
    ctx = _ctx(pkg, sc, scheduler=scheme)
    orc = oracle.OracleContext(f, sc.num_verts, scheme=scheme)
    for fr in range(0, 40, 1):
        v = sc.frame(fr)
        ctx.update_vertices(v)
        orc.update_vertices(v)
        ctx.step_scheduled()
        orc.step_scheduled()
        gs, os_ = ctx.stats(), orc.stats()
        assert (gs["last_bvh_op"], gs["last_front_op"]) == (os_["last_bvh_op"], os_["last_front_op"]), fr
        for narrow in (False, True):
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(narrow)), oracle.pair_set(orc.pairs(narrow)),
                                          err_msg=f"frame {fr} narrow {narrow}")
        for w in (0, 1):
            np.testing.assert_array_equal(_rows(ctx.export_front(w)), _rows(orc.front(w)),
                                          err_msg=f"front {w} frame {fr}")
    ctx.close()


def _scrambled_frames(sc, frames, frac=0.5, per_frame=20, seed=0):
    """A cloth frame whose vertices in one corner patch exchange positions, more of them every
    frame: the tree built at the first frame degrades there, which makes the front-quality
    heuristic flag subtrees (plain schedule with a long BVH cycle)."""
    n = sc.spec.n
    base = sc.frame(20)
    V = sc.num_verts
    x, y = np.arange(V) % n, np.arange(V) // n
    m = np.nonzero((x < int(n * frac)) & (y < int(n * frac)))[0]
    rng = np.random.default_rng(seed)
    out = []
    for fr in range(frames):
        v = base.copy()
        k = min(m.size, fr * per_frame)
        p = m[:k]
        v[p] = base[rng.permutation(p)]
        out.append(v)
    return out


def test_quality_heuristic_flags_parity(pkg, oracle):
    """The reference's restructure heuristic on a case where it fires: UPDATE frames flag
    subtrees (checkFrontQuality), the next REFIT becomes a subtree restructure in the middle
    tier (I/8 <= flagged internal nodes <= I/2) and the front is regenerated. Ops, restructure
    counts, topology after each restructure, fronts and pair sets equal the oracle's every
    frame."""
    print("[gpu] quality heuristic flags; launch shapes:\n" + pkg.autotune_report(), flush=True)
    sc = pkg.Scene.preset("cloth256", n=48, frames=40)
    f = sc.faces()
    frames = _scrambled_frames(sc, 36)
    ctx = _ctx(pkg, sc, bvh_cycle=100)
    orc = oracle.OracleContext(f, sc.num_verts, bvh_cycle=100)
    restructured = 0
    for fr, v in enumerate(frames):
        ctx.update_vertices(v)
        orc.update_vertices(v)
        ctx.step_scheduled()
        orc.step_scheduled()
        try:
            gs = ctx.stats()
        except pkg.BvhcdError as e:
            raise AssertionError(f"frame {fr} (oracle ops {orc.stats()['last_bvh_op']}, "
                                 f"{orc.stats()['last_front_op']}): {e}") from e
        os_ = orc.stats()
        assert (gs["last_bvh_op"], gs["last_front_op"]) == (os_["last_bvh_op"], os_["last_front_op"]), fr
        assert (gs["restructured_subtrees"], gs["restructured_leaves"]) == \
            (os_["restructured_subtrees"], os_["restructured_leaves"]), fr
        if gs["last_bvh_op"] == pkg.BVH_RESTRUCTURE and gs["restructured_subtrees"] > 0:
            restructured += 1
            compare_bvh(ctx.export_bvh(), orc.export_bvh(), f"restructure frame {fr}")
        for narrow in (False, True):
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(narrow)), oracle.pair_set(orc.pairs(narrow)),
                                          err_msg=f"frame {fr} narrow {narrow}")
        for w in (0, 1):
            np.testing.assert_array_equal(_rows(ctx.export_front(w)), _rows(orc.front(w)),
                                          err_msg=f"front {w} frame {fr}")
    assert restructured >= 2
    ctx.close()


def test_explicit_ops_parity(pkg, oracle):
    print("[gpu] explicit ops", flush=True)
    sc = pkg.Scene.preset("cloth256", n=48, frames=25)
    f = sc.faces()
    ctx = _ctx(pkg, sc)
    orc = oracle.OracleContext(f, sc.num_verts)
    ops = [(0, 1)] + [(1, [2, 4, 3, 2, 0, 1][k % 6]) for k in range(1, 25)]
    for fr, (bo, fo) in enumerate(ops):
        v = sc.frame(fr)
        ctx.update_vertices(v)
        orc.update_vertices(v)
        ctx.step(bo, fo)
        orc.step(bo, fo)
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), oracle.pair_set(orc.pairs(True)),
                                      err_msg=str(fr))
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)), oracle.pair_set(orc.pairs(False)),
                                      err_msg=str(fr))
        for w in (0, 1):
            np.testing.assert_array_equal(_rows(ctx.export_front(w)), _rows(orc.front(w)), err_msg=f"{fr} {w}")
        _check_front_order(ctx, fo, fr)
    ctx.close()


def _check_front_order(ctx, fo, fr):
    """The plain schedule's fronts are counting-sorted by node after GENERATE / UPDATE; after
    REORDER the internal front is too, each node's non-overlapping entries before its overlapping
    ones (boxes from the exported tree)."""
    if fo not in (1, 2, 3):
        return
    fi = ctx.export_front(0)
    assert np.all(np.diff(fi[:, 1]) >= 0), f"frame {fr}: internal front not ordered by node"
    if fo in (1, 2):
        fe = ctx.export_front(1)
        assert np.all(np.diff(fe[:, 1]) >= 0), f"frame {fr}: external front not ordered by node"
        return
    b = ctx.export_bvh()
    qb = b["leaf_box"].reshape(-1, 6)[fi[:, 0]]
    nb = b["int_box"].reshape(-1, 6)[fi[:, 1]]
    ov = np.all((qb[:, :3] <= nb[:, 3:]) & (nb[:, :3] <= qb[:, 3:]), axis=1)
    same = fi[1:, 1] == fi[:-1, 1]
    assert ov.any() and (~ov).any(), f"frame {fr}: REORDER case without both kinds of entries"
    assert not np.any(same & ov[:-1] & ~ov[1:]), f"frame {fr}: overlapping entry before a non-overlapping one"


@pytest.mark.parametrize("name,frames", [
    ("cloth1024", (0, 37)),     # 2,093,058 triangles: 4095 top-tree nodes, merged in LDS
    ("cloth1024", (100, 137)),  # folded: ~6.8K top-tree nodes, LDS segments + upper nodes
    ("layers4", (0, 9)),        # 8,372,232 triangles: ~16K+ top-tree nodes, segments + upper nodes
])
def test_refit_large_trees(pkg, oracle, name, frames):
    """Treelet-chunk refit and top-tree merge at full size, after a build and after a later
    refit: leaf boxes equal the per-triangle min/max, internal boxes equal the oracle's refit
    over the exported topology."""
    print(f"[gpu] refit {name} frames {frames}", flush=True)
    sc = pkg.Scene.preset(name)
    f = sc.faces()
    ctx = _ctx(pkg, sc, narrow_phase=False)
    for k, fr in enumerate(frames):
        v = sc.frame(fr)
        ctx.update_vertices(v)
        ctx.step(pkg.BVH_BUILD if k == 0 else pkg.BVH_REFIT, pkg.FRONT_PURE_BVH_CD)
        b = ctx.export_bvh()
        tv = v[f[b["prim_idx"]]]
        lb = np.concatenate([tv.min(1), tv.max(1)], 1)
        assert np.array_equal(b["leaf_box"], lb), f"{name} frame {fr} leaf boxes"
        ref = oracle.refit_boxes({key: b[key] for key in ("int_lc", "int_rc", "int_mark")}, lb)
        assert np.array_equal(b["int_box"], ref), f"{name} frame {fr} internal boxes"
        del tv, lb, ref, b
    ctx.close()


def test_host_upload_overlap(pkg, oracle):
    """Host vertex uploads go through two device buffers on a copy stream and overlap the steps
    still running: 24 frames stepped back to back without a synchronisation log the same
    per-frame counts (the frame log's algorithmic bytes, which carry the pair counts and node
    tests) as the same frames with a synchronisation after every step, and the last frame's
    pairs equal brute force."""
    print("[gpu] host upload overlap", flush=True)
    sc = pkg.Scene.preset("cloth256", n=96)
    frames = [sc.frame(100 + k) for k in range(24)]
    logs = []
    for sync in (False, True):
        ctx = _ctx(pkg, sc)
        ctx.set_timing(len(frames))
        for k, v in enumerate(frames):
            ctx.update_vertices(v)
            ctx.step(pkg.BVH_BUILD if k == 0 else pkg.BVH_REFIT, pkg.FRONT_GENERATE if k == 0 else pkg.FRONT_KEEP)
            if sync:
                ctx.stats()
        ms, by, ops = ctx.timing_frames()
        logs.append(by)
        if not sync:
            ref = oracle.pair_set(oracle.bruteforce_pairs(sc.faces(), frames[-1], True))
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), ref)
        ctx.close()
    np.testing.assert_array_equal(logs[0], logs[1])


def test_refit_deep_top_chain(pkg, oracle):
    """3000 identical point triangles among 4096 form a chain of equal Morton codes: a top tree
    ~2500 levels deep, deeper than the refit's level bins, merged by climbing. Boxes equal the
    oracle's refit after a build and after a refit of moved vertices."""
    print("[gpu] refit deep top-tree chain", flush=True)
    rng = np.random.default_rng(3)
    T = 4096
    v = rng.random((3 * T, 3)).astype(np.float32)
    v[: 3 * 3000] = np.float32(0.25)
    f = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    ctx = pkg.CollisionContext(T, 3 * T)
    ctx.set_topology(f, 3 * T)
    for k in range(2):
        if k:
            v = v + rng.normal(scale=1e-3, size=v.shape).astype(np.float32)
            v[: 3 * 3000] = np.float32(0.25)
        ctx.update_vertices(v)
        ctx.step(pkg.BVH_BUILD if k == 0 else pkg.BVH_REFIT, pkg.FRONT_PURE_BVH_CD)
        b = ctx.export_bvh()
        tv = v[f[b["prim_idx"]]]
        lb = np.concatenate([tv.min(1), tv.max(1)], 1)
        assert np.array_equal(b["leaf_box"], lb)
        ref = oracle.refit_boxes({key: b[key] for key in ("int_lc", "int_rc", "int_mark")}, lb)
        assert np.array_equal(b["int_box"], ref), f"step {k} internal boxes"
    ctx.close()


@pytest.mark.parametrize("T", [1, 2, 3, 7])
def test_tiny_scenes(pkg, oracle, T):
    print(f"[gpu] tiny T={T}", flush=True)
    rng = np.random.default_rng(T)
    v = rng.normal(size=(3 * T, 3)).astype(np.float32) * 0.3
    f = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    ctx = pkg.CollisionContext(T, 3 * T)
    ctx.set_topology(f, 3 * T)
    ctx.update_vertices(v)
    orc = oracle.OracleContext(f, 3 * T)
    orc.update_vertices(v)
    for bo, fo in [(0, 1), (1, 2), (1, 4), (1, 3), (0, 0)]:
        ctx.step(bo, fo)
        orc.step(bo, fo)
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), oracle.pair_set(orc.pairs(True)))
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)),
                                      oracle.pair_set(oracle.bruteforce_pairs(f, v, False)))
    ctx.close()


def test_degenerate_inputs(pkg, oracle):
    print("[gpu] degenerate", flush=True)
    # all triangles at one point (every code equal), plus a NaN vertex
    T = 200
    v = np.zeros((3 * T, 3), np.float32)
    v[5, 1] = np.nan
    f = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
    ctx = pkg.CollisionContext(T, 3 * T)
    ctx.set_topology(f, 3 * T)
    ctx.update_vertices(v)
    ctx.step(0, 1)
    orc = oracle.OracleContext(f, 3 * T)
    orc.update_vertices(v)
    orc.step(0, 1)
    compare_bvh(ctx.export_bvh(), orc.export_bvh(), "degenerate")
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), oracle.pair_set(orc.pairs(True)))
    assert ctx.pairs(False).shape[0] == T * (T - 1) // 2
    ctx.close()


def test_huge_and_shared_vertices(pkg, oracle):
    """Vertex-sharing pairs at extreme coordinates: the narrow phase's shared-vertex shortcut
    only applies below |x| = 1e8; above it the full SAT decides, as in the oracle."""
    print("[gpu] huge coordinates", flush=True)
    rng = np.random.default_rng(9)
    T = 300
    v = (rng.normal(size=(T + 2, 3)) * 0.01).astype(np.float32)
    v[::7] *= np.float32(1e30)   # some vertices far beyond 1e8
    v[3::11] = np.inf
    f = np.stack([np.arange(T), np.arange(T) + 1, np.arange(T) + 2], 1).astype(np.int32)  # strip
    ctx = pkg.CollisionContext(T, T + 2)
    ctx.set_topology(f, T + 2)
    ctx.update_vertices(v)
    ctx.step(0, 1)
    orc = oracle.OracleContext(f, T + 2)
    orc.update_vertices(v)
    orc.step(0, 1)
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)), oracle.pair_set(orc.pairs(False)))
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), oracle.pair_set(orc.pairs(True)))
    ctx.close()


def test_keep_shared_vertex_pairs(pkg, oracle):
    """KEEP counts the front's shared-vertex entries instead of testing them while every
    coordinate is finite: pair sets equal brute force on finite frames, on a frame with a vertex
    beyond 1e8 (where the shortcut is off and the full tests decide) and after it; the host
    read-out still returns every pair when listed + implicit pairs exceed the pair buffer."""
    print("[gpu] keep shared-vertex pairs", flush=True)
    sc = pkg.Scene.preset("cloth256", n=96)
    f = sc.faces()
    ctx = _ctx(pkg, sc)
    v = sc.frame(100)
    ctx.update_vertices(v)
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    for fr, huge in ((101, False), (102, True), (103, False)):
        v = sc.frame(fr)
        if huge:
            v[5] = np.float32(2e8)
        ctx.update_vertices(v)
        ctx.step(pkg.BVH_REFIT, pkg.FRONT_KEEP)
        st = ctx.stats()
        if huge:
            assert st["implicit_pairs"] == 0
        else:
            assert 0 < st["implicit_pairs"] <= st["np_pairs"] <= st["bp_pairs"]
        for narrow in (False, True):
            ref = oracle.pair_set(oracle.bruteforce_pairs(f, v, narrow))
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(narrow)), ref, err_msg=f"frame {fr} {narrow}")
    st = ctx.stats()
    listed = st["bp_pairs"] - st["implicit_pairs"]
    ctx.close()
    # a pair buffer that holds the listed pairs (with room for uneven segments) but not
    # listed + implicit: the host read-out then goes in pieces
    cap = (st["bp_pairs"] - 1) // 64 * 64
    assert 2 * listed < cap < st["bp_pairs"], (listed, cap, st["bp_pairs"])
    ctx = _ctx(pkg, sc, pair_cap=cap)
    ctx.update_vertices(sc.frame(100))
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    ctx.update_vertices(sc.frame(103))
    ctx.step(pkg.BVH_REFIT, pkg.FRONT_KEEP)
    for narrow in (False, True):
        ref = oracle.pair_set(oracle.bruteforce_pairs(f, sc.frame(103), narrow))
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(narrow)), ref, err_msg=f"small buffer {narrow}")
    ctx.close()


@pytest.mark.skip(reason="faulted the GPU with a GENERATE that puts the shared-vertex head in place (code "
                         "since taken out; this test still asserts that behaviour); its sequence includes a "
                         "rebuild with one vertex at 2e8, where two Morton codes make an 18 045-deep chain and "
                         "the fronts overflow about 100x; cause not found, not to be run until it is")
@pytest.mark.parametrize("skip", [False, True])
def test_generate_shared_vertex_head(pkg, oracle, skip):
    """The query-ordered GENERATE of the learned schedules puts the external entries whose
    triangles share a vertex at the front's head and, while every coordinate is finite, holds
    their pairs implicitly there: the head is exactly those entries (each part still ordered by
    query), and pair sets equal brute force on GENERATE frames, on the KEEP frames that reuse the
    head, and on a GENERATE frame with a vertex beyond 1e8 (no implicit pairs). With
    skip_adjacent the vertex-sharing pairs are dropped instead."""
    print(f"[gpu] generate shared-vertex head skip={skip}", flush=True)
    sc = pkg.Scene.preset("cloth256", n=96)
    f = sc.faces()
    ctx = _ctx(pkg, sc, scheduler=pkg.SCHED_LEARNED_NO_NET, skip_adjacent=skip)
    for fr, gen, huge in ((100, True, False), (101, False, False), (102, True, True), (103, True, False),
                          (104, False, False)):
        v = sc.frame(fr)
        if huge:
            v[5] = np.float32(2e8)
        ctx.update_vertices(v)
        ctx.step(pkg.BVH_BUILD if gen else pkg.BVH_REFIT, pkg.FRONT_GENERATE if gen else pkg.FRONT_KEEP)
        st = ctx.stats()
        if skip or huge:
            assert st["implicit_pairs"] == 0, fr
        else:
            assert 0 < st["implicit_pairs"] <= st["np_pairs"] <= st["bp_pairs"], (fr, st)
        if gen:
            prim = ctx.export_bvh()["prim_idx"]
            fe = ctx.export_front(1)
            adj = (f[prim[fe[:, 0]]][:, :, None] == f[prim[fe[:, 1]]][:, None, :]).any(axis=(1, 2))
            h = int(adj.sum())
            assert h > 0 and adj[:h].all() and not adj[h:].any(), f"frame {fr}: head is not the shared-vertex part"
            if not (skip or huge):
                assert st["implicit_pairs"] == h, (fr, st["implicit_pairs"], h)
            for part in (fe[:h], fe[h:]):
                assert np.all(np.diff(part[:, 0]) >= 0), f"frame {fr}: a part is not ordered by query"
        for narrow in (False, True):
            ref = oracle.pair_set(oracle.bruteforce_pairs(f, v, narrow))
            if skip:
                ref = ref[~_share_vertex(f, ref)]
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(narrow)), ref, err_msg=f"frame {fr} {narrow}")
    ctx.close()


def _share_vertex(faces, keys):
    """Which pair keys (oracle.pair_set: a << 32 | b) join triangles sharing a vertex (numpy covertex)."""
    if len(keys) == 0:
        return np.zeros(0, bool)
    a = faces[(keys >> np.uint64(32)).astype(np.int64)]
    b = faces[(keys & np.uint64(0xFFFFFFFF)).astype(np.int64)]
    return (a[:, :, None] == b[:, None, :]).any(axis=(1, 2))


@pytest.mark.parametrize("ops", ["plain", "explicit"])
def test_skip_adjacent_filter(pkg, oracle, ops):
    """skip_adjacent drops exactly the pairs of triangles sharing a vertex, at every emission
    point (traversal, UPDATE sprouts, KEEP, REORDER), from both lists: pair sets equal brute force
    minus the vertex-sharing pairs."""
    print(f"[gpu] skip adjacent {ops}", flush=True)
    sc = pkg.Scene.preset("cloth256", n=64, frames=40)
    f = sc.faces()
    ctx = _ctx(pkg, sc, skip_adjacent=True)
    seq = [(pkg.BVH_BUILD, pkg.FRONT_PURE_BVH_CD), (pkg.BVH_BUILD, pkg.FRONT_GENERATE),
           (pkg.BVH_REFIT, pkg.FRONT_KEEP), (pkg.BVH_REFIT, pkg.FRONT_UPDATE), (pkg.BVH_REFIT, pkg.FRONT_KEEP),
           (pkg.BVH_REFIT, pkg.FRONT_REORDER), (pkg.BVH_REFIT, pkg.FRONT_KEEP)]
    for k in range(len(seq) if ops == "explicit" else 30):
        fr = 5 + k
        v = sc.frame(fr)
        ctx.update_vertices(v)
        if ops == "explicit":
            ctx.step(*seq[k])
        else:
            ctx.step_scheduled()
        st = ctx.stats()
        assert st["implicit_pairs"] == 0
        for narrow in (False, True):
            ref = oracle.pair_set(oracle.bruteforce_pairs(f, v, narrow))
            ref = ref[~_share_vertex(f, ref)]
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(narrow)), ref, err_msg=f"frame {fr} {narrow}")
    ctx.close()


def test_box_mode(pkg, oracle):
    """Per-triangle boxes given directly (the simulator-facing input): broad-phase pairs equal a
    brute-force overlap of those boxes after a build, refits, KEEP / UPDATE fronts and a plain
    schedule; the narrow phase is refused in box mode; vertices switch back to vertex mode."""
    print("[gpu] box mode", flush=True)
    sc = pkg.Scene.preset("cloth256", n=32, frames=40)
    f = sc.faces()
    T = sc.num_tris
    rng = np.random.default_rng(21)

    def boxes_of(fr):
        tv = sc.frame(fr)[f]
        lo, hi = tv.min(1), tv.max(1)
        pad = (rng.random((T, 1)) * 2e-3).astype(np.float32)  # swept-box style inflation
        return np.ascontiguousarray(np.concatenate([lo - pad, hi + pad], 1).astype(np.float32))

    def brute(b):
        lo, hi = b[:, :3], b[:, 3:]
        i, j = np.triu_indices(T, 1)
        ok = np.all((lo[i] <= hi[j]) & (lo[j] <= hi[i]), axis=1)
        return oracle.pair_set(np.stack([i[ok], j[ok]], 1).astype(np.int32))

    ctx = _ctx(pkg, sc, narrow_phase=False)
    ops = [(pkg.BVH_BUILD, pkg.FRONT_GENERATE), (pkg.BVH_REFIT, pkg.FRONT_KEEP), (pkg.BVH_REFIT, pkg.FRONT_UPDATE),
           (pkg.BVH_REFIT, pkg.FRONT_KEEP), (pkg.BVH_BUILD, pkg.FRONT_PURE_BVH_CD)]
    for k, op in enumerate(ops):
        b = boxes_of(3 + k)
        ctx.update_boxes(b)
        ctx.step(*op)
        assert ctx.stats()["implicit_pairs"] == 0
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)), brute(b), err_msg=f"op {k}")
    ctx.close()
    ctx = _ctx(pkg, sc)  # narrow phase on: refused in box mode, fine with vertices again
    ctx.update_boxes(boxes_of(0))
    with pytest.raises(pkg.BvhcdError) as e:
        ctx.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    assert e.value.code == pkg.E_INVALID
    v = sc.frame(0)
    ctx.update_vertices(v)
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)),
                                  oracle.pair_set(oracle.bruteforce_pairs(f, v, True)))
    ctx.close()


def test_borrowed_vertices(pkg, oracle):
    """Vertices read in place from a caller's device buffer (MEM_DEVICE_BORROW) give the same
    pair sets, schedule ops and BVH as copied ones; a later host update switches back to the
    context's own buffer."""
    import torch
    print("[gpu] borrowed vertices", flush=True)
    sc = pkg.Scene.preset("cloth256", n=64, frames=20)
    frames = torch.from_numpy(np.stack([sc.frame(f) for f in range(20)])).cuda()
    a, b = _ctx(pkg, sc), _ctx(pkg, sc)
    nbytes = sc.num_verts * 12
    for f in range(20):
        a.update_vertices(frames.data_ptr() + f * nbytes, pkg.MEM_DEVICE)
        b.update_vertices(frames.data_ptr() + f * nbytes, pkg.MEM_DEVICE_BORROW)
        a.step_scheduled()
        b.step_scheduled()
        sa, sb = a.stats(), b.stats()
        assert (sa["last_bvh_op"], sa["last_front_op"]) == (sb["last_bvh_op"], sb["last_front_op"])
        for narrow in (False, True):
            np.testing.assert_array_equal(oracle.pair_set(a.pairs(narrow)), oracle.pair_set(b.pairs(narrow)),
                                          err_msg=f"frame {f} {narrow}")
    compare_bvh(b.export_bvh(), a.export_bvh(), "borrowed")
    v = sc.frame(19)
    b.update_vertices(v)
    del frames
    torch.cuda.synchronize()
    b.step(pkg.BVH_REFIT, pkg.FRONT_KEEP)
    np.testing.assert_array_equal(oracle.pair_set(b.pairs(True)), oracle.pair_set(oracle.bruteforce_pairs(sc.faces(), v, True)))
    a.close()
    b.close()


def test_capacity_overflow_reported(pkg):
    print("[gpu] overflow", flush=True)
    sc = pkg.Scene.preset("spheres", subdiv=3, spacing=2.0)
    ctx = _ctx(pkg, sc, pair_cap=100)
    ctx.update_vertices(sc.frame(0))
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_PURE_BVH_CD)
    with pytest.raises(pkg.BvhcdError) as e:
        ctx.pairs(True)
    assert e.value.code == pkg.E_CAPACITY
    ctx.close()
    ctx = _ctx(pkg, sc, int_front_cap=10, ext_front_cap=10)
    ctx.update_vertices(sc.frame(0))
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    with pytest.raises(pkg.BvhcdError) as e:
        ctx.stats()
    assert e.value.code == pkg.E_CAPACITY
    ctx.close()


def test_capacities_are_totals(pkg, oracle):
    """Fronts and pair lists are written in segments, but a flush that does not fit its segment
    moves on to others: lists sized 1 % above what a frame needs hold it exactly (pairs equal
    brute force, required sizes equal the lengths), and a capacity below the length is reported as an
    overflow."""
    print("[gpu] capacities are totals", flush=True)
    sc = pkg.Scene.preset("cloth256", n=96)
    f, v = sc.faces(), sc.frame(150)
    probe = _ctx(pkg, sc)
    probe.update_vertices(v)
    probe.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    st = probe.stats()
    fi, fe, bp = st["int_front"], st["ext_front"], len(probe.pairs(False))
    probe.close()
    assert min(fi, fe, bp) > 64 * 256  # several flushes per segment
    ref_bp = oracle.pair_set(oracle.bruteforce_pairs(f, v, False))
    ref_np = oracle.pair_set(oracle.bruteforce_pairs(f, v, True))
    ctx = _ctx(pkg, sc, int_front_cap=fi + fi // 100, ext_front_cap=fe + fe // 100, pair_cap=bp + bp // 100)
    ctx.update_vertices(v)
    for op in (pkg.FRONT_GENERATE, pkg.FRONT_KEEP):
        ctx.step(pkg.BVH_BUILD if op == pkg.FRONT_GENERATE else pkg.BVH_REFIT, op)
        st = ctx.stats()
        assert (st["int_front"], st["ext_front"]) == (fi, fe)
        assert (st["required_int_front"], st["required_ext_front"]) == (fi, fe)
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)), ref_bp)
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), ref_np)
    ctx.close()
    ctx = _ctx(pkg, sc, ext_front_cap=fe - 64)  # rounded up to a multiple of 64: still below fe
    ctx.update_vertices(v)
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    with pytest.raises(pkg.BvhcdError) as e:
        ctx.stats()
    assert e.value.code == pkg.E_CAPACITY
    ctx.close()


@pytest.mark.parametrize("scheduler", ["SCHED_PLAIN", "SCHED_LEARNED_NO_NET"])
def test_auto_grow(pkg, oracle, scheduler):
    """auto_grow: a context created with tiny front and pair capacities grows them on the frames
    that overflow and redoes the front, so every frame's pairs still equal brute force and no
    call reports a capacity error; the grown lists hold what the frames need."""
    print(f"[gpu] auto grow {scheduler}", flush=True)
    sc = pkg.Scene.preset("cloth256", n=80, frames=300)
    f = sc.faces()
    ctx = _ctx(pkg, sc, int_front_cap=64, ext_front_cap=64, pair_cap=64, auto_grow=True,
               scheduler=getattr(pkg, scheduler))
    for fr in range(150, 166):
        v = sc.frame(fr)
        ctx.update_vertices(v)
        ctx.step_scheduled()
        st = ctx.stats()
        assert st["required_int_front"] >= st["int_front"] > 64 and st["ext_front"] > 64
        for narrow in (False, True):
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(narrow)),
                                          oracle.pair_set(oracle.bruteforce_pairs(f, v, narrow)),
                                          err_msg=f"frame {fr} narrow {narrow}")
    ctx.close()


def test_full_spheres_sequence_digests(pkg, oracle):
    """Two icospheres (40,960 triangles), all 200 frames, plain schedule: per-frame digests equal."""
    print("[gpu] spheres 200 frames", flush=True)
    sc = pkg.Scene.preset("spheres")
    f = sc.faces()
    ctx = _ctx(pkg, sc)
    orc = oracle.OracleContext(f, sc.num_verts)
    restr = 0
    for fr in range(sc.frames):
        v = sc.frame(fr)
        ctx.update_vertices(v)
        orc.update_vertices(v)
        ctx.step_scheduled()
        orc.step_scheduled(threads=8)
        assert ctx.digest(True) == oracle.pair_digest(orc.pairs(True)), fr
        assert ctx.digest(False) == oracle.pair_digest(orc.pairs(False)), fr
        restr += ctx.stats()["last_bvh_op"] == pkg.BVH_RESTRUCTURE
    print(f"[gpu] spheres restructure frames: {restr}", flush=True)
    ctx.close()
