"""GPU tests of input validation and host-buffer reuse at the C ABI.

Face vertex ids outside [0, V) must be refused with BVHCD_E_INVALID (host faces checked on the
host, device faces by a check kernel) before any step can gather through them, for the single
context and for batches; a host vertex array overwritten right after a batch upload must not
change that frame's pairs.
"""
import numpy as np
import pytest

from conftest import load_package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def _bad_faces(sc):
    f = sc.faces().copy()
    g = f.copy()
    g[len(g) // 2, 1] = sc.num_verts          # one past the end
    h = f.copy()
    h[-1, 2] = -1
    return f, (g, h)


def test_set_topology_rejects_bad_vertex_ids(pkg, oracle):
    import torch
    print("[gpu] set_topology bad ids", flush=True)
    sc = pkg.Scene.preset("cloth256", n=24, frames=3)
    good, bads = _bad_faces(sc)
    ctx = pkg.CollisionContext(sc.num_tris, sc.num_verts)
    for bad in bads:
        with pytest.raises(pkg.BvhcdError) as e:
            ctx.set_topology(bad, sc.num_verts)
        assert e.value.code == pkg.E_INVALID
        with pytest.raises(pkg.BvhcdError):  # no step without a valid topology
            ctx.step(pkg.BVH_BUILD, pkg.FRONT_PURE_BVH_CD)
        dev = torch.from_numpy(bad).cuda()
        with pytest.raises(pkg.BvhcdError) as e:
            ctx.set_topology((dev.data_ptr(), bad.shape[0]), sc.num_verts, where=pkg.MEM_DEVICE)
        assert e.value.code == pkg.E_INVALID
    dev = torch.from_numpy(good).cuda()
    ctx.set_topology((dev.data_ptr(), good.shape[0]), sc.num_verts, where=pkg.MEM_DEVICE)
    v = sc.frame(2)
    ctx.update_vertices(v)
    ctx.step(pkg.BVH_BUILD, pkg.FRONT_PURE_BVH_CD)
    np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)),
                                  oracle.pair_set(oracle.bruteforce_pairs(good, v, True)))
    ctx.close()


def test_batch_rejects_bad_vertex_ids(pkg):
    import torch
    print("[gpu] batch bad ids", flush=True)
    scenes = [pkg.Scene.preset("cloth256", n=n, frames=2) for n in (12, 20)]
    faces = np.concatenate([s.faces() for s in scenes])
    b = pkg.Batch([s.num_tris for s in scenes], [s.num_verts for s in scenes])
    bad = faces.copy()
    bad[-1, 0] = scenes[1].num_verts        # local id past the second scene's vertices
    with pytest.raises(pkg.BvhcdError) as e:
        b.set_topology(bad)
    assert e.value.code == pkg.E_INVALID
    dev = torch.from_numpy(bad).cuda()
    st = pkg.lib().bvhcd_batch_set_topology(b._h, pkg.C.c_void_p(dev.data_ptr()), pkg.MEM_DEVICE)
    assert st == pkg.E_INVALID
    with pytest.raises(pkg.BvhcdError):
        b.step(pkg.BVH_BUILD, use_graph=False)
    with pytest.raises(ValueError):
        b.set_topology(faces[:-1])
    b.close()


def test_batch_host_array_reused_after_upload(pkg, oracle):
    """bvhcd_batch_update_vertices returns once the host copy is done: overwriting the array
    right after the call must not leak into the frame."""
    print("[gpu] batch host reuse", flush=True)
    scenes = [pkg.Scene.preset("cloth256", n=n, frames=30) for n in (40, 64)]
    faces = [s.faces() for s in scenes]
    b = pkg.Batch([s.num_tris for s in scenes], [s.num_verts for s in scenes])
    b.set_topology(np.concatenate(faces))
    for f in (5, 29):
        frames = [s.frame(f) for s in scenes]
        host = np.concatenate(frames)
        b.update_vertices(host)
        host[:] = 1e3 * np.random.default_rng(f).standard_normal(host.shape).astype(np.float32)
        b.step(pkg.BVH_BUILD, use_graph=False)
        for k, s in enumerate(scenes):
            want = oracle.pair_set(oracle.bruteforce_pairs(faces[k], frames[k], True))
            np.testing.assert_array_equal(oracle.pair_set(b.pairs(k)), want, err_msg=f"frame {f} scene {k}")
    b.close()


def test_mesh_sequence_and_log_text(pkg, oracle, tmp_path):
    """Frames read from PLY files (one file per frame, printf-named) give the same pairs as the
    generator's frames, and the log text carries one inspectResults line per recorded frame
    with the step's front sizes and pair counts, plus its timed sections."""
    import struct
    print("[gpu] mesh sequence + log text", flush=True)
    sc = pkg.Scene.preset("cloth256", n=40, frames=14)
    f = sc.faces()
    for k in range(sc.frames):
        v = sc.frame(k)
        head = ["ply", "format binary_little_endian 1.0", f"element vertex {len(v)}", "property float x",
                "property float y", "property float z", f"element face {len(f)}",
                "property list uchar int vertex_indices", "end_header"]
        with open(tmp_path / ("cloth%d.ply" % (k + 1)), "wb") as fh:
            fh.write(("\n".join(head) + "\n").encode())
            fh.write(v.astype("<f4").tobytes())
            for t in f:
                fh.write(struct.pack("<Biii", 3, *(int(i) for i in t)))
    seq = pkg.MeshSequence(str(tmp_path / "cloth%d.ply"), 1, sc.frames)
    ctx = pkg.CollisionContext(seq.num_tris, seq.num_verts)
    ctx.set_topology(seq.faces(), seq.num_verts)
    ctx.set_timing(sc.frames)
    rows = []
    for k in range(seq.frames):
        v = seq.frame(k)
        ctx.update_vertices(v)
        ctx.step_scheduled()
        st = ctx.stats()
        rows.append((st["int_front"], st["ext_front"], st["bp_pairs"], st["np_pairs"]))
        np.testing.assert_array_equal(oracle.pair_set(ctx.pairs(True)),
                                      oracle.pair_set(oracle.bruteforce_pairs(f, sc.frame(k), True)), err_msg=str(k))
    text = ctx.log_text()
    assert pkg.front_status(text) == rows
    assert text.count("#cd_total_time") == sc.frames and text.count("###: 0") == sc.frames
    assert text.count("@refit_bvh") >= 1 and text.count("@narrow_phase") == sc.frames
    assert "preserve front. " in text and "generate front. " in text
    ctx.close()
