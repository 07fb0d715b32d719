#!/usr/bin/env python3
"""Writes tests/golden/narrow_band.json: every broad-phase pair of the benchmark scenes whose
narrow-phase decision depends on how the triangle test is evaluated, with its exact margin.

The build evaluates the 17-axis triangle test with the lower triangle id as P and no fused
multiply-adds. The reference passes the triangle of the lower Morton-ranked leaf as P
(narrow_phase.cu:11-19), which is either triangle of the pair, and nvcc contracts products into
FMAs by default. For every broad-phase pair (pairs sharing a vertex excepted: they are hits under
every variant) the oracle evaluates all four variants; a pair whose four decisions do not agree
is listed with its exact relative separating margin (oracle/narrow_band.py), which must lie
within 1e-6 of the boundary.

Frames: every frame of the two-sphere scene (200) and of the 256^2 folding cloth (300); three
folded frames each of the 1024^2 folding cloth and of the 4-layer stress scene.

Run:  python tests/golden/make_band.py      (about ten minutes on 8 cores)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from conftest import load_package  # noqa: E402
import narrow_band  # noqa: E402
import oracle  # noqa: E402

BAND = 1e-6
PLAN = [("spheres", None), ("cloth256", None), ("cloth1024", [150, 275, 399]), ("layers4", [100, 150, 199])]


def scan(pkg, name, frames):
    sc = pkg.Scene.preset(name)
    faces = sc.faces()
    frames = list(range(sc.frames)) if frames is None else frames
    ctx = oracle.OracleContext(faces, sc.num_verts, narrow=False)
    rows, evaluated, t0 = [], 0, time.time()
    for fr in frames:
        v = sc.frame(fr)
        ctx.update_vertices(v)
        ctx.step(0, 0)  # BUILD + pure traversal: the frame's broad-phase pairs
        bp = ctx.pairs(False)
        evaluated += bp.shape[0]
        for a, b, mask, mg in narrow_band.band_of_frame(oracle, faces, v, bp):
            rows.append([fr, a, b, mask, mg])
    ctx.close()
    worst = max((abs(r[4]) for r in rows), default=0.0)
    print(f"{name}: {len(frames)} frames, {evaluated} broad-phase pairs, {len(rows)} variant disagreements, "
          f"max |margin| {worst:.3g}, {time.time() - t0:.0f}s", flush=True)
    assert worst <= BAND, (name, worst)
    return {"frames": frames if len(frames) < 10 else f"0..{frames[-1]}", "broad_phase_pairs": evaluated,
            "disagreements": rows, "max_abs_margin": worst}


def main():
    pkg = load_package()
    out = {"band": BAND,
           "columns": ["frame", "tri_a", "tri_b", "mask (bit0 lower id as P, bit1 higher id as P, +2 bits with FMA)",
                       "exact relative margin"],
           "configs": {name: scan(pkg, name, frames) for name, frames in PLAN}}
    with open(os.path.join(HERE, "narrow_band.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
