"""CPU tests of mesh-file ingest (OBJ, PLY), frame-file patterns and the log text helpers.

Files are written here from the seeded generator's meshes in every supported encoding and read
back through the library; the results must equal the generator's faces and vertices exactly
(text coordinates are written with 9 significant digits, which round-trip float32).
"""
import os
import struct

import numpy as np
import pytest

from conftest import load_package


@pytest.fixture(scope="module")
def pkg():
    return load_package()


def _mesh(pkg):
    sc = pkg.Scene.preset("cloth256", n=12, frames=10)
    return sc.faces(), sc.frame(9)


def _write_obj(path, f, v, style):
    with open(path, "w") as fh:
        fh.write("# test mesh\no first\n")
        for x in v:
            fh.write("v %.9g %.9g %.9g\n" % tuple(float(c) for c in x))
        fh.write("vt 0 0\nvn 0 0 1\n")
        nv = len(v)
        for k, t in enumerate(f):
            if style == "plain":
                fh.write("f %d %d %d\n" % tuple(int(i) + 1 for i in t))
            elif style == "slashes":
                form = ["%d/1/1", "%d//1", "%d/1"][k % 3]
                fh.write("f " + " ".join(form % (int(i) + 1) for i in t) + "\n")
            else:  # negative (relative to the vertices read so far: all of them here)
                fh.write("f %d %d %d\n" % tuple(int(i) - nv for i in t))


def _write_ply(path, f, v, fmt, coord="float", count="uchar", index="int"):
    head = ["ply", f"format {fmt} 1.0", "comment test", f"element vertex {len(v)}",
            f"property {coord} x", f"property {coord} y", f"property {coord} z", "property uchar red",
            f"element face {len(f)}", f"property list {count} {index} vertex_indices", "end_header"]
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for x in v:
                fh.write(("%.9g %.9g %.9g 7\n" % tuple(float(c) for c in x)).encode())
            for t in f:
                fh.write(("3 %d %d %d\n" % tuple(int(i) for i in t)).encode())
            return
        e = "<" if fmt == "binary_little_endian" else ">"
        cf = {"float": "f", "double": "d"}[coord]
        cc = {"uchar": "B", "int": "i"}[count]
        ci = {"int": "i", "uint": "I"}[index]
        for x in v:
            fh.write(struct.pack(e + cf * 3 + "B", *(float(c) for c in x), 7))
        for t in f:
            fh.write(struct.pack(e + cc + ci * 3, 3, *(int(i) for i in t)))


@pytest.mark.parametrize("style", ["plain", "slashes", "negative"])
def test_obj_roundtrip(pkg, tmp_path, style):
    f, v = _mesh(pkg)
    p = tmp_path / "m.obj"
    _write_obj(p, f, v, style)
    gf, gv = pkg.load_mesh(str(p))
    np.testing.assert_array_equal(gf, f)
    np.testing.assert_array_equal(gv, v)


@pytest.mark.parametrize("fmt,coord,count,index", [
    ("ascii", "float", "uchar", "int"),
    ("binary_little_endian", "float", "uchar", "int"),
    ("binary_little_endian", "double", "int", "uint"),
    ("binary_big_endian", "float", "uchar", "int"),
])
def test_ply_roundtrip(pkg, tmp_path, fmt, coord, count, index):
    f, v = _mesh(pkg)
    p = tmp_path / "m.ply"
    _write_ply(p, f, v, fmt, coord, count, index)
    gf, gv = pkg.load_mesh(str(p))
    np.testing.assert_array_equal(gf, f)
    np.testing.assert_array_equal(gv, v)


def test_polygons_fan_and_groups(pkg, tmp_path):
    """A quad and a pentagon become fans (0, k, k+1); vertices of later groups keep their global
    ids (meshes concatenated with offsets)."""
    p = tmp_path / "poly.obj"
    p.write_text("o a\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n"
                 "o b\nv 0 0 1\nv 1 0 1\nv 2 1 1\nv 1 2 1\nv 0 1 1\nf 5 6 7 8 9\n")
    f, v = pkg.load_mesh(str(p))
    np.testing.assert_array_equal(f, [[0, 1, 2], [0, 2, 3], [4, 5, 6], [4, 6, 7], [4, 7, 8]])
    assert v.shape == (9, 3) and v[6].tolist() == [2, 1, 1]


def test_bad_files_rejected(pkg, tmp_path):
    bad = {"oob.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n", "zero.obj": "v 0 0 0\nf 0 1 2\n",
           "short.obj": "v 0 0 0\nv 1 0 0\nf 1 2\n", "nohead.ply": "element vertex 1\nend_header\n0 0 0\n",
           "oob.ply": "ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                      "property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n"
                      "0 0 0\n1 0 0\n0 1 0\n3 0 1 3\n",
           "mesh.stl": "solid x\n"}
    for name, text in bad.items():
        p = tmp_path / name
        p.write_text(text)
        with pytest.raises(pkg.BvhcdError):
            pkg.load_mesh(str(p))
    with pytest.raises(pkg.BvhcdError):
        pkg.load_mesh(str(tmp_path / "missing.obj"))


def test_frame_paths_and_sequence(pkg, tmp_path):
    assert pkg.frame_path("funnel/%03d.ply", 7) == "funnel/007.ply"
    assert pkg.frame_path("flowing_cloth/%04d_00.obj", 12) == "flowing_cloth/0012_00.obj"
    assert pkg.frame_path("100%%/f%d.obj", 3) == "100%/f3.obj"
    for bad in ("no_number.obj", "%s.obj", "%d_%d.obj", "%f.obj"):
        with pytest.raises(pkg.BvhcdError):
            pkg.frame_path(bad, 1)
    sc = pkg.Scene.preset("cloth256", n=10, frames=5)
    f = sc.faces()
    for k in range(5):
        _write_obj(tmp_path / ("%04d_00.obj" % (k + 3)), f, sc.frame(k), "plain")
    seq = pkg.MeshSequence(str(tmp_path / "%04d_00.obj"), 3, 5)
    np.testing.assert_array_equal(seq.faces(), f)
    for k in range(5):
        np.testing.assert_array_equal(seq.frame(k), sc.frame(k))
    small = pkg.Scene.preset("cloth256", n=9, frames=2)
    _write_obj(tmp_path / "0008_00.obj", small.faces(), small.frame(1), "plain")
    seq2 = pkg.MeshSequence(str(tmp_path / "%04d_00.obj"), 3, 6)
    with pytest.raises(ValueError):
        seq2.frame(5)


def test_front_status_and_log_file(pkg, tmp_path):
    """front_status reads the inspectResults lines the way the reference's FrontAnalyser does;
    write_log never overwrites an existing file."""
    text = ("\ngenerate front. \nfront (120, 340) cpNum 77 actualCpNum 12\npreserve front. \n"
            "build front(in restr). \nfront (5, 6) cpNum 7 actualCpNum 0\n\n@refit_bvh: 0.5\n\n")
    assert pkg.front_status(text) == [(120, 340, 77, 12), (5, 6, 7, 0)]
    p1 = pkg.write_log(str(tmp_path / "out" / "clothball.txt"), "a")
    p2 = pkg.write_log(str(tmp_path / "out" / "clothball.txt"), "b")
    p3 = pkg.write_log(str(tmp_path / "out" / "clothball.txt"), "c")
    assert [os.path.basename(p) for p in (p1, p2, p3)] == ["clothball.txt", "clothball_.txt", "clothball__.txt"]
    assert open(p2).read() == "b"
