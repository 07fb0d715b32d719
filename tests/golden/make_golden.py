#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the host oracle and the scene generator.

The reference ships no tests or vectors, so these fixtures pin this project's own oracle
(itself pinned by closed forms and brute force in tests/test_oracle.py) against regressions,
and give the GPU tests precomputed answers:

  golden_build1000.npz   1000-triangle case (first 1000 triangles of the 256x256 cloth at frame
                         150, plus 20 duplicated triangles): sorted codes, permutation, topology.
  golden.json            sha256 of generated vertex frames; pair counts and order-independent
                         digests of the narrow-phase pair sets for the two-sphere sequence
                         (frames 0/50/100/199) and a 64x64 folding cloth (frames 0/150/299),
                         plain schedule from frame 0.

Run:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from conftest import load_package  # noqa: E402
import oracle  # noqa: E402


def build_case():
    pkg = load_package()
    sc = pkg.Scene.preset("cloth256")
    f = sc.faces()[:1000].copy()
    v = sc.frame(150)
    f = np.concatenate([f, f[:20]])  # duplicated triangles -> equal codes
    used, inv = np.unique(f, return_inverse=True)  # keep only the referenced vertices
    return inv.reshape(f.shape).astype(np.int32), np.ascontiguousarray(v[used])


def sequence_digests(pkg, name, kw, frames):
    sc = pkg.Scene.preset(name, **kw)
    faces = sc.faces()
    ctx = oracle.OracleContext(faces, sc.num_verts)
    out = {}
    for fr in range(max(frames) + 1):
        ctx.update_vertices(sc.frame(fr))
        ctx.step_scheduled()
        if fr in frames:
            d, n = oracle.pair_digest(ctx.pairs(True))
            out[str(fr)] = {"pairs": n, "digest": str(d)}
    return out


def frame_hash(pkg, name, kw, fr):
    sc = pkg.Scene.preset(name, **kw)
    return hashlib.sha256(sc.frame(fr).tobytes()).hexdigest()


def main():
    pkg = load_package()
    f, v = build_case()
    box = oracle.scene_box(f, v)
    codes = oracle.morton_codes(f, v, box)
    s, idx, pm = oracle.sort_codes(codes)
    topo = oracle.build_topology(s)
    np.savez_compressed(os.path.join(HERE, "golden_build1000.npz"), faces=f, verts=v, box=box, codes=codes,
                        sorted_codes=s, prim_idx=idx, prim_map=pm, **topo)
    meta = {
        "frame_sha256": {
            "spheres/0": frame_hash(pkg, "spheres", {}, 0),
            "spheres/199": frame_hash(pkg, "spheres", {}, 199),
            "cloth256/150": frame_hash(pkg, "cloth256", {}, 150),
            "cloth1024/399": frame_hash(pkg, "cloth1024", {}, 399),
        },
        "spheres": sequence_digests(pkg, "spheres", {}, [0, 50, 100, 199]),
        "cloth64": sequence_digests(pkg, "cloth256", {"n": 64}, [0, 150, 299]),
    }
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
