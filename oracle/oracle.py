"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product library. The oracle restates the reference algorithm
(littlemine/BVH-based-Collision-Detection-Scheme) on the host; see bvhcd_oracle.cpp for the
reference file:line each function follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    P, I64, I32 = C.c_void_p, C.c_int64, C.c_int
    sig = {
        "orc_expand_bits": (C.c_uint32, [C.c_uint32]),
        "orc_morton3d": (C.c_uint32, [C.c_float, C.c_float, C.c_float]),
        "orc_tri_boxes": (None, [I64, P, P, P]),
        "orc_scene_box": (None, [I64, P, P, P]),
        "orc_morton_codes": (None, [I64, P, P, P, P]),
        "orc_sort": (None, [I64, P, P, P, P]),
        "orc_build_topology": (None, [I64, P] + [P] * 9),
        "orc_refit_boxes": (None, [I64, P, P, P, P, P]),
        "orc_sat": (I32, [P, P]),
        "orc_sat_fma": (I32, [P, P]),
        "orc_sat_variants": (None, [I64, P, P, P, I32, P]),
        "orc_bruteforce_pairs": (I64, [I64, P, P, I32, P, I64, I32]),
        "orc_create": (P, [I64, I64, P, I32, I32, I32, I32]),
        "orc_destroy": (None, [P]),
        "orc_update_vertices": (None, [P, P]),
        "orc_step": (I32, [P, I32, I32, I32]),
        "orc_step_scheduled": (I32, [P, I32]),
        "orc_plain_schedule": (None, [I64, I32, I32, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "orc_scheme_ops": (None, [I32, I64, I32, I32, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "orc_set_scheme": (None, [P, I32]),
        "orc_get_pairs": (I64, [P, I32, P, I64]),
        "orc_export_front": (I64, [P, I32, P, I64]),
        "orc_export_bvh": (None, [P] + [P] * 14),
        "orc_stats": (None, [P, P]),
        "orc_export_restr_log": (None, [P, P, P]),
        "orc_quality_eval": (None, [I64, P, P, P, P, P, I32, P, P]),
        "orc_inter_create": (P, [P, I64, I64, P]),
        "orc_inter_destroy": (None, [P]),
        "orc_inter_update_vertices": (None, [P, P]),
        "orc_inter_step": (I32, [P, I32, I32, I32]),
        "orc_inter_get_pairs": (I64, [P, I32, P, I64]),
        "orc_inter_export_front": (I64, [P, I32, P, I64]),
        "orc_inter_export_rigid": (None, [P, P, P]),
        "orc_inter_last_front_op": (I64, [P]),
        "orc_restr_decision": (I32, [I64, P, P, P, P, P, C.POINTER(I64), C.POINTER(I64)]),
        "orc_restructure_roots": (I32, [P, P, I64, I32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def expand_bits(v):
    return int(lib().orc_expand_bits(v))


def morton3d(x, y, z):
    return int(lib().orc_morton3d(x, y, z))


def tri_boxes(faces, verts):
    f = np.ascontiguousarray(faces, np.int32)
    v = np.ascontiguousarray(verts, np.float32)
    out = np.empty((f.shape[0], 6), np.float32)
    lib().orc_tri_boxes(f.shape[0], _p(f), _p(v), _p(out))
    return out


def scene_box(faces, verts):
    f = np.ascontiguousarray(faces, np.int32)
    v = np.ascontiguousarray(verts, np.float32)
    out = np.empty(6, np.float32)
    lib().orc_scene_box(f.shape[0], _p(f), _p(v), _p(out))
    return out


def morton_codes(faces, verts, box):
    f = np.ascontiguousarray(faces, np.int32)
    v = np.ascontiguousarray(verts, np.float32)
    b = np.ascontiguousarray(box, np.float32)
    out = np.empty(f.shape[0], np.uint32)
    lib().orc_morton_codes(f.shape[0], _p(f), _p(v), _p(b), _p(out))
    return out


def sort_codes(codes):
    c = np.ascontiguousarray(codes, np.uint32)
    n = c.size
    s, idx, pm = np.empty(n, np.uint32), np.empty(n, np.int32), np.empty(n, np.int32)
    lib().orc_sort(n, _p(c), _p(s), _p(idx), _p(pm))
    return s, idx, pm


def build_topology(sorted_codes):
    c = np.ascontiguousarray(sorted_codes, np.uint32)
    n = c.size
    I = max(n - 1, 0)
    a = {"int_lc": np.empty(I, np.int32), "int_rc": np.empty(I, np.int32), "int_par": np.empty(I, np.int32),
         "int_mark": np.empty(I, np.uint32), "int_rx": np.empty(I, np.int32), "int_ry": np.empty(I, np.int32),
         "leaf_par": np.empty(n, np.int32), "leaf_mark": np.empty(n, np.uint32), "leaf_lca": np.empty(n + 1, np.int32)}
    order = ["int_lc", "int_rc", "int_par", "int_mark", "int_rx", "int_ry", "leaf_par", "leaf_mark", "leaf_lca"]
    lib().orc_build_topology(n, _p(c), *[_p(a[k]) for k in order])
    return a


def refit_boxes(topo, leaf_box):
    lb = np.ascontiguousarray(leaf_box, np.float32)
    n = lb.shape[0]
    ib = np.empty((max(n - 1, 0), 6), np.float32)
    lib().orc_refit_boxes(n, _p(topo["int_lc"]), _p(topo["int_rc"]), _p(topo["int_mark"]), _p(lb), _p(ib))
    return ib


def sat(P, Q, fma=False):
    p = np.ascontiguousarray(P, np.float32).ravel()
    q = np.ascontiguousarray(Q, np.float32).ravel()
    return bool((lib().orc_sat_fma if fma else lib().orc_sat)(_p(p), _p(q)))


def sat_variants(pairs, faces, verts, threads=0):
    """Per pair, the decision bits of the four narrow-phase variants (lower / higher triangle id
    as P, without / with fused multiply-adds); 0 or 15 = all agree."""
    p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    f = np.ascontiguousarray(faces, np.int32)
    v = np.ascontiguousarray(verts, np.float32)
    out = np.zeros(p.shape[0], np.uint8)
    if p.shape[0]:
        lib().orc_sat_variants(p.shape[0], _p(p), _p(f), _p(v), threads, _p(out))
    return out


def bruteforce_pairs(faces, verts, narrow=True, threads=0):
    f = np.ascontiguousarray(faces, np.int32)
    v = np.ascontiguousarray(verts, np.float32)
    n = lib().orc_bruteforce_pairs(f.shape[0], _p(f), _p(v), 1 if narrow else 0, None, 0, threads)
    out = np.empty((n, 2), np.int32)
    if n:
        lib().orc_bruteforce_pairs(f.shape[0], _p(f), _p(v), 1 if narrow else 0, _p(out), n, threads)
    return out


class OracleContext:
    """Host restatement of the per-frame pipeline (same ops and schedule as the GPU context)."""

    def __init__(self, faces, num_verts, bvh_cycle=12, front_cycle=6, enable_restructure=True, narrow=True,
                 scheme=0):
        f = np.ascontiguousarray(faces, np.int32)
        self.T = f.shape[0]
        self.V = int(num_verts)
        self._h = lib().orc_create(self.T, self.V, _p(f), bvh_cycle, front_cycle, 1 if enable_restructure else 0,
                                   1 if narrow else 0)
        lib().orc_set_scheme(self._h, int(scheme))

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def update_vertices(self, xyz):
        v = np.ascontiguousarray(xyz, np.float32)
        lib().orc_update_vertices(self._h, _p(v))

    def step(self, bvh_op, front_op, threads=0):
        st = lib().orc_step(self._h, bvh_op, front_op, threads)
        if st:
            raise RuntimeError(f"oracle step failed: {st}")

    def step_scheduled(self, threads=0):
        st = lib().orc_step_scheduled(self._h, threads)
        if st:
            raise RuntimeError(f"oracle step failed: {st}")

    def pairs(self, filtered=True):
        n = lib().orc_get_pairs(self._h, 1 if filtered else 0, None, 0)
        out = np.empty((n, 2), np.int32)
        if n:
            lib().orc_get_pairs(self._h, 1 if filtered else 0, _p(out), n)
        return out

    def front(self, which):
        n = lib().orc_export_front(self._h, which, None, 0)
        out = np.empty((n, 2), np.int32)
        if n:
            lib().orc_export_front(self._h, which, _p(out), n)
        return out

    def export_bvh(self):
        n, I = self.T, max(self.T - 1, 0)
        a = {
            "codes": np.empty(n, np.uint32), "prim_idx": np.empty(n, np.int32), "prim_map": np.empty(n, np.int32),
            "int_lc": np.empty(I, np.int32), "int_rc": np.empty(I, np.int32), "int_par": np.empty(I, np.int32),
            "int_mark": np.empty(I, np.uint32), "int_rx": np.empty(I, np.int32), "int_ry": np.empty(I, np.int32),
            "leaf_par": np.empty(n, np.int32), "leaf_mark": np.empty(n, np.uint32),
            "leaf_lca": np.empty(n + 1, np.int32),
            "int_box": np.empty((I, 6), np.float32), "leaf_box": np.empty((n, 6), np.float32),
        }
        order = ["codes", "prim_idx", "prim_map", "int_lc", "int_rc", "int_par", "int_mark", "int_rx", "int_ry",
                 "leaf_par", "leaf_mark", "leaf_lca", "int_box", "leaf_box"]
        lib().orc_export_bvh(self._h, *[_p(a[k]) for k in order])
        return a

    def restructure(self, roots, threads=0):
        r = np.ascontiguousarray(roots, dtype=np.int32).ravel()
        st = lib().orc_restructure_roots(self._h, _p(r), r.size, threads)
        if st:
            raise ValueError(f"orc_restructure_roots: {st}")

    def stats(self):
        s = np.zeros(11, np.int64)
        lib().orc_stats(self._h, _p(s))
        keys = ["frame", "last_bvh_op", "last_front_op", "int_front", "ext_front", "bp_pairs", "np_pairs",
                "node_visits", "restructured_subtrees", "restructured_leaves", "restr_pending"]
        return dict(zip(keys, (int(x) for x in s)))


class OracleInter:
    """Host restatement of the inter-object front: a rigid object's triangles against the
    deformable BVH of an OracleContext (which must be stepped first each frame)."""

    def __init__(self, deformable, faces, num_verts):
        f = np.ascontiguousarray(faces, np.int32)
        self.T, self.V = f.shape[0], int(num_verts)
        self._def = deformable  # keeps the deformable context alive
        self._h = lib().orc_inter_create(deformable._h, self.T, self.V, _p(f))

    def close(self):
        if self._h:
            lib().orc_inter_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def update_vertices(self, xyz):
        v = np.ascontiguousarray(xyz, np.float32)
        lib().orc_inter_update_vertices(self._h, _p(v))

    def step(self, rigid_op, front_op, threads=0):
        st = lib().orc_inter_step(self._h, rigid_op, front_op, threads)
        if st:
            raise RuntimeError(f"oracle inter step failed: {st}")

    @property
    def last_front_op(self):
        return int(lib().orc_inter_last_front_op(self._h))

    def pairs(self, filtered=True):
        n = lib().orc_inter_get_pairs(self._h, 1 if filtered else 0, None, 0)
        out = np.empty((n, 2), np.int32)
        if n:
            lib().orc_inter_get_pairs(self._h, 1 if filtered else 0, _p(out), n)
        return out

    def front(self, which):
        n = lib().orc_inter_export_front(self._h, which, None, 0)
        out = np.empty((n, 2), np.int32)
        if n:
            lib().orc_inter_export_front(self._h, which, _p(out), n)
        return out

    def rigid_order(self):
        codes, idx = np.empty(self.T, np.uint32), np.empty(self.T, np.int32)
        lib().orc_inter_export_rigid(self._h, _p(codes), _p(idx))
        return codes, idx


def inter_bruteforce(faces_a, verts_a, faces_b, verts_b, narrow=True):
    """All (a, b) with overlapping boxes (and, narrow, an SAT hit with a's triangle as P)."""
    ba, bb = tri_boxes(faces_a, verts_a), tri_boxes(faces_b, verts_b)
    ov = np.all((ba[:, None, :3] <= bb[None, :, 3:]) & (ba[:, None, 3:] >= bb[None, :, :3]), axis=2)
    i, j = np.nonzero(ov)
    if narrow:
        keep = [sat(verts_a[faces_a[x]], verts_b[faces_b[y]]) for x, y in zip(i, j)]
        i, j = i[np.array(keep, bool)], j[np.array(keep, bool)]
    return np.stack([i, j], 1).astype(np.int32)


def ordered_pair_set(pairs) -> np.ndarray:
    """Sorted uint64 keys (a << 32 | b) of an ordered pair list (no swapping)."""
    p = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    return np.sort((p[:, 0].astype(np.uint64) << np.uint64(32)) | p[:, 1].astype(np.uint64))


def quality_eval(rx, ry, intbeg, extbeg, snapshot=None):
    """frontSnapshot (when snapshot is None) and checkFrontQuality on given preorder ranges and
    ordering offsets: returns (snapshot, intrange, rtroot)."""
    n = len(rx) + 1
    rx = np.ascontiguousarray(rx, np.int32)
    ry = np.ascontiguousarray(ry, np.int32)
    ib = np.ascontiguousarray(intbeg, np.uint32)
    eb = np.ascontiguousarray(extbeg, np.uint32)
    assert ib.size == n and eb.size == n + 1
    snap = np.zeros(n - 1, np.float32) if snapshot is None else np.array(snapshot, np.float32)
    ir = np.zeros(n, np.int32)
    rt = np.zeros(n, np.int32)
    lib().orc_quality_eval(n, _p(rx), _p(ry), _p(ib), _p(eb), _p(snap), 1 if snapshot is None else 0, _p(ir), _p(rt))
    return snap, ir, rt


def restr_decision(rx, ry, intrange, rtroot):
    """(tag, outermost roots, internal nodes in them) of the restructure decision."""
    n = len(rx) + 1
    rx = np.ascontiguousarray(rx, np.int32)
    ry = np.ascontiguousarray(ry, np.int32)
    ir = np.ascontiguousarray(intrange, np.int32)
    rt = np.ascontiguousarray(rtroot, np.int32)
    roots = np.zeros(n, np.int32)
    k, m = C.c_int64(), C.c_int64()
    tag = lib().orc_restr_decision(n, _p(rx), _p(ry), _p(ir), _p(rt), _p(roots), C.byref(k), C.byref(m))
    return int(tag), roots[: k.value].copy(), int(m.value)


def plain_schedule(frame, bvh_cycle=12, front_cycle=6):
    b, f = C.c_int(), C.c_int()
    lib().orc_plain_schedule(frame, bvh_cycle, front_cycle, C.byref(b), C.byref(f))
    return b.value, f.value


def scheme_ops(scheme, frame, bvh_cycle=12, front_cycle=6):
    b, f = C.c_int(), C.c_int()
    lib().orc_scheme_ops(int(scheme), frame, bvh_cycle, front_cycle, C.byref(b), C.byref(f))
    return b.value, f.value


def pair_set(pairs) -> np.ndarray:
    """Canonical sorted unique uint64 keys (a << 32 | b, a < b) of a pair list."""
    p = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    a = np.minimum(p[:, 0], p[:, 1])
    b = np.maximum(p[:, 0], p[:, 1])
    return np.sort((a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64))


def pair_digest(pairs):
    """Order-independent digest matching bvhcd_pair_digest: sum of splitmix64 finalizers."""
    k = pair_set(pairs)
    z = k.copy()
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        return int(z.sum(dtype=np.uint64)), int(k.size)
