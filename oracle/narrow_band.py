"""Exact separating margins of triangle pairs (test infrastructure).

TEST INFRASTRUCTURE ONLY: used by tests/ and tests/golden/make_band.py, never by the product.

The narrow phase decides a pair with the 17-axis separating-axis test in fp32
(X/gProximity/cuda_intersect_tritri.h:247-266,349-433, called by
collision/narrow_phase/narrow_phase.cu:11-19). Its decision can depend on rounding only for
pairs that lie next to the decision boundary. `exact_margin` evaluates the same 17 axes over the
rationals, from the vertices' exact float values, so the margin it returns is the real
geometry's, free of any evaluation order or contraction:

    margin = max over the non-zero axes a of  gap(a) / (|a|_1 * R)

with gap(a) = max(min P.a - max Q.a, min Q.a - max P.a) (positive: a separates P and Q) and R
the largest coordinate magnitude of the six vertices translated by P's first vertex. The pair
is separated iff margin > 0; |margin| is how far, relative to the triangles' size, the pair is
from the boundary.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np


def _sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def exact_margin(P, Q) -> Fraction:
    """Relative separating margin (see the module docstring) of triangles P and Q (3 x 3)."""
    P = np.asarray(P, np.float32).reshape(3, 3)
    Q = np.asarray(Q, np.float32).reshape(3, 3)
    c = [Fraction(float(x)) for x in P[0]]
    p = [_sub([Fraction(float(x)) for x in P[k]], c) for k in range(3)]
    q = [_sub([Fraction(float(x)) for x in Q[k]], c) for k in range(3)]
    e = [_sub(p[1], p[0]), _sub(p[2], p[1]), _sub(p[0], p[2])]
    f = [_sub(q[1], q[0]), _sub(q[2], q[1]), _sub(q[0], q[2])]
    n, m = _cross(e[0], e[1]), _cross(f[0], f[1])
    axes = [n, m] + [_cross(a, b) for a in e for b in f] + [_cross(a, n) for a in e] + [_cross(b, m) for b in f]
    R = max(abs(x) for v in p + q for x in v)
    if R == 0:
        return Fraction(0)
    best = None
    for a in axes:
        l1 = abs(a[0]) + abs(a[1]) + abs(a[2])
        if l1 == 0:
            continue  # a zero axis never separates
        pp = [_dot(a, v) for v in p]
        qq = [_dot(a, v) for v in q]
        gap = max(min(pp) - max(qq), min(qq) - max(pp))
        rel = gap / (l1 * R)
        best = rel if best is None else max(best, rel)
    return best if best is not None else Fraction(0)


def shares_vertex(faces, pairs):
    """Boolean per pair: the two triangles have a vertex id in common."""
    fa, fb = faces[pairs[:, 0]], faces[pairs[:, 1]]
    return (fa[:, :, None] == fb[:, None, :]).any(axis=(1, 2))


def band_of_frame(oracle, faces, verts, pairs, threads=0):
    """Pairs of a frame's broad-phase list whose four fp32 decisions (lower / higher id as P,
    without / with FMA contraction) do not all agree: rows (a, b, mask, margin as float)."""
    pairs = np.asarray(pairs, np.int32).reshape(-1, 2)
    lo, hi = np.minimum(pairs[:, 0], pairs[:, 1]), np.maximum(pairs[:, 0], pairs[:, 1])
    pairs = np.stack([lo, hi], 1)
    # pairs sharing a vertex are hits under every variant (the shared vertex projects to the
    # same value on both sides), so only the others are evaluated
    pairs = pairs[~shares_vertex(faces, pairs)]
    masks = oracle.sat_variants(pairs, faces, verts, threads)
    out = []
    for k in np.nonzero((masks != 0) & (masks != 15))[0]:
        a, b = int(pairs[k, 0]), int(pairs[k, 1])
        mg = exact_margin(verts[faces[a]], verts[faces[b]])
        out.append((a, b, int(masks[k]), float(mg)))
    return out
