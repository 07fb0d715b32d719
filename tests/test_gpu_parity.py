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
