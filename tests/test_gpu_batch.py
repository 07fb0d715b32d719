"""GPU tests of the batched-scene pipeline and the autotuner through the C ABI.

Each scene of a batch must get exactly the pairs the oracle reports for that scene alone
(pure traversal + narrow phase), on BUILD and REFIT frames, eagerly and through the captured
frame graphs (every replay checked against the eager run, with eager work and synchronisation
between replays), and after the launch shapes are autotuned.
"""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def _scenes(pkg, num, seed, frames, t_min=2e3, t_max=2e4):
    specs = pkg.batch_scene_specs(num=num, seed=seed, frames=frames, t_min=t_min, t_max=t_max)
    return [pkg.Scene(s) for s in specs]


def _batch_frame(scenes, f):
    return np.concatenate([s.frame(f) for s in scenes])


def _run_batch(pkg, scenes, frames, use_graph, ops=None):
    b = pkg.Batch([s.num_tris for s in scenes], [s.num_verts for s in scenes])
    b.set_topology(np.concatenate([s.faces() for s in scenes]))
    out = []
    for f in range(frames):
        b.update_vertices(_batch_frame(scenes, f))
        op = ops[f] if ops else (pkg.BVH_BUILD if f % 3 == 0 else pkg.BVH_REFIT)
        b.step(op, use_graph=use_graph)
        b.pair_counts()  # eager work and a synchronisation between graph replays
        out.append([b.pairs(k) for k in range(len(scenes))])
    return b, out


def _set(a):
    a = np.asarray(a).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))]


def test_batch_matches_oracle_per_scene(pkg, oracle):
    print("[gpu] batch vs oracle", flush=True)
    scenes = _scenes(pkg, 7, seed=11, frames=12)
    b, got = _run_batch(pkg, scenes, 6, use_graph=False)
    b.close()
    for k, sc in enumerate(scenes):
        orc = oracle.OracleContext(sc.faces(), sc.num_verts)
        for f in range(6):
            orc.update_vertices(sc.frame(f))
            orc.step(0 if f % 3 == 0 else 1, 0)
            np.testing.assert_array_equal(oracle.pair_set(got[f][k]), oracle.pair_set(orc.pairs(True)),
                                          err_msg=f"scene {k} frame {f}")
        orc.close()


def test_batch_full_size_matches_single_scenes(pkg, oracle):
    """The benchmark's 64-scene batch (2.6 M triangles in one tree, whose top tree is merged in
    LDS segments) gives every scene the pairs of that scene alone in its own context, on a BUILD
    and a REFIT frame through the captured graphs (single contexts are checked against the
    oracle by the parity tests)."""
    print("[gpu] full-size batch vs single scenes", flush=True)
    scenes = [pkg.Scene(s) for s in pkg.batch_scene_specs()]
    b, got = _run_batch(pkg, scenes, 2, use_graph=True, ops=[pkg.BVH_BUILD, pkg.BVH_REFIT])
    b.close()
    for k, sc in enumerate(scenes):
        ctx = pkg.CollisionContext(sc.num_tris, sc.num_verts)
        ctx.set_topology(sc.faces(), sc.num_verts)
        for f in range(2):
            ctx.update_vertices(sc.frame(f))
            ctx.step(pkg.BVH_BUILD if f == 0 else pkg.BVH_REFIT, pkg.FRONT_PURE_BVH_CD)
            np.testing.assert_array_equal(oracle.pair_set(got[f][k]), oracle.pair_set(ctx.pairs(True)),
                                          err_msg=f"scene {k} frame {f}")
        ctx.close()


def test_batch_tiny_and_uneven_scenes(pkg, oracle):
    print("[gpu] batch tiny scenes", flush=True)
    rng = np.random.default_rng(5)
    faces, verts, tps, vps = [], [], [], []
    for T in (1, 2, 3, 700, 1, 40):
        v = rng.normal(size=(3 * T, 3)).astype(np.float32) * 0.2
        f = np.arange(3 * T, dtype=np.int32).reshape(T, 3)
        faces.append(f)
        verts.append(v)
        tps.append(T)
        vps.append(3 * T)
    b = pkg.Batch(tps, vps)
    b.set_topology(np.concatenate(faces))
    b.update_vertices(np.concatenate(verts))
    b.step(pkg.BVH_BUILD, use_graph=False)
    counts = b.pair_counts()
    for k in range(len(tps)):
        ref = oracle.pair_set(oracle.bruteforce_pairs(faces[k], verts[k], True))
        np.testing.assert_array_equal(oracle.pair_set(b.pairs(k)), ref, err_msg=str(k))
        assert counts[k] == len(ref)
    b.close()


def test_batch_graph_replays_equal_eager(pkg):
    print("[gpu] batch graphs", flush=True)
    scenes = _scenes(pkg, 5, seed=12, frames=10)
    ops = [pkg.BVH_BUILD, pkg.BVH_REFIT, pkg.BVH_REFIT, pkg.BVH_BUILD, pkg.BVH_REFIT, pkg.BVH_REFIT,
           pkg.BVH_REFIT, pkg.BVH_BUILD]
    b0, eager = _run_batch(pkg, scenes, len(ops), use_graph=False, ops=ops)
    b0.close()
    b1, graph = _run_batch(pkg, scenes, len(ops), use_graph=True, ops=ops)
    b1.close()
    for f in range(len(ops)):
        for k in range(len(scenes)):
            np.testing.assert_array_equal(_set(graph[f][k]), _set(eager[f][k]), err_msg=f"frame {f} scene {k}")


def test_autotune_keeps_results(pkg, oracle):
    print("[gpu] autotune", flush=True)
    scenes = _scenes(pkg, 4, seed=13, frames=6)
    b, before = _run_batch(pkg, scenes, 2, use_graph=True)
    b.autotune()
    after = [b.pairs(k) for k in range(len(scenes))]
    for k in range(len(scenes)):
        np.testing.assert_array_equal(_set(after[k]), _set(before[-1][k]))
    b.update_vertices(_batch_frame(scenes, 2))
    b.step(pkg.BVH_REFIT, use_graph=True)  # re-captured with the tuned grid
    for k, sc in enumerate(scenes):
        ref = oracle.pair_set(oracle.bruteforce_pairs(sc.faces(), sc.frame(2), True))
        np.testing.assert_array_equal(oracle.pair_set(b.pairs(k)), ref, err_msg=str(k))
    b.close()
    # single context: KEEP + narrow grids
    sc = pkg.Scene.preset("cloth256", n=96)
    ctx = pkg.CollisionContext(sc.num_tris, sc.num_verts)
    ctx.set_topology(sc.faces(), sc.num_verts)
    ctx.update_vertices(sc.frame(100))
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    ref = oracle.pair_set(ctx.pairs(True))
    ctx.autotune()
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), ref)
    ctx.update_vertices(sc.frame(101))
    ctx.step(pkg.BVH_REFIT, pkg.FRONT_KEEP)
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)),
                                  oracle.pair_set(oracle.bruteforce_pairs(sc.faces(), sc.frame(101), True)))
    rep = pkg.autotune_report()
    assert "k_keep" in rep and "k_narrow" in rep and "k_batch_cd" in rep and "*" in rep
    ctx.close()


@pytest.mark.parametrize("grid,unroll,block", [(64, 1, 256), (1024, 2, 128), (3000, 4, 512), (1, 4, 128),
                                              (5, 2, 512), (777, 1, 128)])
def test_keep_launch_shapes(pkg, oracle, grid, unroll, block):
    """KEEP with every unroll factor and workgroup size, and the narrow phase with the same grids
    and workgroups (down to one workgroup, and sizes that leave partial iterations) give the
    brute-force pair sets; the default shapes are restored afterwards."""
    print(f"[gpu] keep/narrow shape grid {grid} unroll {unroll} block {block}", flush=True)
    sc = pkg.Scene.preset("cloth256", n=96)
    ctx = pkg.CollisionContext(sc.num_tris, sc.num_verts)
    ctx.set_topology(sc.faces(), sc.num_verts)
    ctx.update_vertices(sc.frame(100))
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_GENERATE)
    try:
        pkg.autotune_set(pkg.TUNE_KEEP, sc.num_tris, grid, unroll, block)
        pkg.autotune_set(pkg.TUNE_NARROW, sc.num_tris, grid, 1, block)
        for fr in (101, 102):
            ctx.update_vertices(sc.frame(fr))
            ctx.step(pkg.BVH_REFIT, pkg.FRONT_KEEP)
            ref = oracle.pair_set(oracle.bruteforce_pairs(sc.faces(), sc.frame(fr), True))
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)), ref, err_msg=f"frame {fr}")
            ref_bp = oracle.pair_set(oracle.bruteforce_pairs(sc.faces(), sc.frame(fr), False))
            np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(False)), ref_bp, err_msg=f"bp frame {fr}")
    finally:
        pkg.autotune_set(pkg.TUNE_KEEP, sc.num_tris, 2048, 1)
        pkg.autotune_set(pkg.TUNE_NARROW, sc.num_tris, 4096, 1)
        ctx.close()
