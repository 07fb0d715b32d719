"""CPU-only tests of the product library: it loads, exports every entry point declared in
include/*.h, generates deterministic scenes, and fails loudly (no fallback) without a GPU.
Also pins the oracle against the committed golden fixtures."""
import glob
import hashlib
import json
import os
import re

import numpy as np
import pytest

from conftest import ROOT, load_package

GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(bvhcd_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    pkg = load_package()
    L = pkg.lib()
    names = declared_symbols()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_scene_sizes():
    pkg = load_package()
    expect = {"spheres": (40960, 20484), "cloth256": (130050, 65536), "cloth1024": (2093058, 1048576),
              "layers4": (8372232, 4194304)}
    for name, (T, V) in expect.items():
        sc = pkg.Scene.preset(name)
        assert (sc.num_tris, sc.num_verts) == (T, V), name


def test_scene_faces_valid_and_deterministic():
    pkg = load_package()
    sc = pkg.Scene.preset("spheres", subdiv=3)
    f = sc.faces()
    assert f.min() == 0 and f.max() == sc.num_verts - 1
    a = sc.frame(7)
    b = pkg.Scene.preset("spheres", subdiv=3).frame(7, threads=1)
    np.testing.assert_array_equal(a, b)
    assert not np.array_equal(sc.frame(7), sc.frame(8))  # fresh jitter per frame
    cl = pkg.Scene.preset("cloth256", n=32)
    f0 = cl.frame(0)
    assert np.all(f0[:, 2] == 0)                            # flat at frame 0
    flat = pkg.Scene.preset("cloth256", n=32, fold=0.0)
    assert np.allclose(flat.frame(299)[:, 0].max(), 0.5, atol=1e-6)  # x compressed by 0.5


def test_cloth_fold_is_a_bend():
    """With the sinusoids off, the fold only bends the compressed flat sheet about a line:
    every mesh edge keeps its length (the bend is an isometry of the plane), the vertices on
    one side of the line do not move, and past 180 degrees the flap has crossed below the
    unbent side's plane."""
    pkg = load_package()
    n, F = 48, 100
    sc = pkg.Scene.preset("cloth256", n=n, frames=F, amplitude=0.0)
    f = sc.faces()
    edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    for fr in (0, 30, 60, 99):
        v = sc.frame(fr).astype(np.float64)
        s = fr / (F - 1)
        x0 = np.tile(np.arange(n) / (n - 1), n) * (1 - 0.5 * s)
        y0 = np.repeat(np.arange(n) / (n - 1), n)
        flat = np.stack([x0, y0, np.zeros_like(x0)], 1)
        le = np.linalg.norm(v[edges[:, 0]] - v[edges[:, 1]], axis=1)
        lf = np.linalg.norm(flat[edges[:, 0]] - flat[edges[:, 1]], axis=1)
        # a bend never stretches an edge; only edges across the curved strip (radius 0.05,
        # about 2 edge lengths wide here) become chords of the arc, shorter by < 1 %
        assert np.all(le <= lf + 2e-6)
        assert np.all(le >= 0.99 * lf)
        exact = np.abs(le - lf) <= 2e-6
        assert exact.mean() > 0.8
        still = np.all(v == flat.astype(np.float32), axis=1)
        if fr > 0:
            assert 0.2 * n * n < still.sum() < 0.9 * n * n  # one side stays put, the flap moves
    theta = 3.7 * (1 - (1 - 99 / (F - 1)) ** 4)
    assert theta > np.pi and v[:, 2].min() < -0.05      # the flap has passed through z = 0


def test_folding_cloth_has_self_contact(oracle):
    """The folding cloths bring distant parts of the sheet into contact: the narrow phase finds
    pairs of triangles that share no vertex on the later frames (frames 100..299 of the 256^2
    cloth, 150..399 of the 1024^2 cloth), which a height field never has."""
    pkg = load_package()
    for name, frames in (("cloth256", range(100, 300, 33)), ("cloth1024", (150, 399))):
        sc = pkg.Scene.preset(name)
        f = sc.faces()
        ctx = oracle.OracleContext(f, sc.num_verts)
        for fr in frames:
            ctx.update_vertices(sc.frame(fr))
            ctx.step(0, 0)  # BUILD + pure traversal
            p = ctx.pairs(True)
            fa, fb = f[p[:, 0]], f[p[:, 1]]
            shared = (fa[:, :, None] == fb[:, None, :]).any(axis=(1, 2))
            assert (~shared).sum() > 1000, (name, fr, int((~shared).sum()))
        ctx.close()


def test_scene_frames_match_golden_hashes():
    pkg = load_package()
    meta = json.load(open(os.path.join(GOLDEN, "golden.json")))
    for key, h in meta["frame_sha256"].items():
        name, fr = key.split("/")
        got = hashlib.sha256(pkg.Scene.preset(name).frame(int(fr)).tobytes()).hexdigest()
        assert got == h, key


def test_oracle_matches_golden_build(oracle):
    g = np.load(os.path.join(GOLDEN, "golden_build1000.npz"), allow_pickle=False)
    box = oracle.scene_box(g["faces"], g["verts"])
    np.testing.assert_array_equal(box, g["box"])
    codes = oracle.morton_codes(g["faces"], g["verts"], box)
    np.testing.assert_array_equal(codes, g["codes"])
    s, idx, pm = oracle.sort_codes(codes)
    np.testing.assert_array_equal(idx, g["prim_idx"])
    topo = oracle.build_topology(s)
    for k, v in topo.items():
        np.testing.assert_array_equal(v, g[k], err_msg=k)


@pytest.mark.parametrize("seq", ["spheres", "cloth64"])
def test_oracle_matches_golden_digests(oracle, seq):
    pkg = load_package()
    meta = json.load(open(os.path.join(GOLDEN, "golden.json")))[seq]
    name, kw = ("spheres", {}) if seq == "spheres" else ("cloth256", {"n": 64})
    sc = pkg.Scene.preset(name, **kw)
    ctx = oracle.OracleContext(sc.faces(), sc.num_verts)
    last = max(int(k) for k in meta)
    for fr in range(last + 1):
        ctx.update_vertices(sc.frame(fr))
        ctx.step_scheduled()
        if str(fr) in meta:
            d, n = oracle.pair_digest(ctx.pairs(True))
            assert (n, str(d)) == (meta[str(fr)]["pairs"], meta[str(fr)]["digest"]), fr


def test_context_without_gpu_fails_loudly():
    torch = pytest.importorskip("torch")
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    pkg = load_package()
    with pytest.raises(pkg.BvhcdError):
        pkg.CollisionContext(100, 100)
    with pytest.raises(pkg.BvhcdError):
        pkg.Batch([10, 20], [30, 60])


def test_batch_workload_specs():
    """The batched workload: log-uniform triangle counts in [2e3, 2e5], sheet + icosphere."""
    pkg = load_package()
    specs = pkg.batch_scene_specs()
    assert len(specs) == 64
    tris = []
    for s in specs[:16]:
        sc = pkg.Scene(s)
        sheet = 2 * (s.n - 1) ** 2
        assert sc.num_tris == sheet + 1280  # icosphere, subdivision 3
        assert 2e3 <= 2 * (s.n - 2) ** 2 <= 2e5 or 2e3 <= sheet <= 2.1e5
        tris.append(sc.num_tris)
    assert len(set(tris)) > 8
    assert pkg.batch_scene_specs() == specs  # seeded


def test_bench_rejects_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in r.stdout


def test_autotune_set_validates_shapes():
    """Launch-shape override (host-side table only): grid in [1, 2^24), unroll 1/2/4 for KEEP
    and 1 for the other kernels, workgroups of 128 / 256 / 512 threads for KEEP and the narrow
    phase and 256 for the batched traversal."""
    pkg = load_package()
    for args in [(pkg.TUNE_KEEP, 1000, 0, 1), (pkg.TUNE_KEEP, 1000, 1 << 24, 1), (pkg.TUNE_KEEP, 1000, 64, 3),
                 (pkg.TUNE_NARROW, 1000, 64, 2), (pkg.TUNE_BATCH_CD, 1000, 64, 2), (7, 1000, 64, 1),
                 (pkg.TUNE_KEEP, 0, 64, 1), (pkg.TUNE_KEEP, 1000, 64, 1, 64), (pkg.TUNE_KEEP, 1000, 64, 1, 1024),
                 (pkg.TUNE_NARROW, 1000, 64, 1, 384), (pkg.TUNE_BATCH_CD, 1000, 64, 1, 128)]:
        with pytest.raises(pkg.BvhcdError):
            pkg.autotune_set(*args)
    pkg.autotune_set(pkg.TUNE_KEEP, 1000, 512, 4)
    pkg.autotune_set(pkg.TUNE_KEEP, 1000, 2048, 1, 512)
    pkg.autotune_set(pkg.TUNE_NARROW, 1000, 2048, 1, 128)
    pkg.autotune_set(pkg.TUNE_NARROW, 1000, 4096, 1)
    pkg.autotune_set(pkg.TUNE_KEEP, 1000, 2048, 1)


def test_bench_reads_profiled_traffic():
    """bench.py takes roofline.traffic from the committed counter summary: the learned KEEP row
    exists and parses to a byte count; unknown kernels or workloads give None."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    t = bench.profiled_traffic("k_keep", "cloth1024", "learned")
    assert isinstance(t, int) and t > 0
    assert bench.profiled_traffic("k_nothing", "cloth1024", "learned") is None
    assert bench.profiled_traffic("k_keep", "layers4", "learned") is None
