"""Python host binding of the gfx950 BVH / BVTT-front collision-detection library.

Thin ctypes layer over ``include/bvhcd.h`` (``libbvhcd.so`` built in this directory). The
object model follows the reference's two plugins:

* ``CollisionContext.step(bvh_op, front_op)`` = ``LBvhFixedDeformable::maintain`` followed by
  ``BvttFrontLooseIntra::maintain`` (collision/lbvh/LBvhFixedDeformable.cu:68-87,
  collision/bvtt_front/BvttFrontLooseIntra.cu:69-126 in Source/Project/Mine/ of
  littlemine/BVH-based-Collision-Detection-Scheme);
* ``CollisionContext.pairs()`` = ``applyCpResults`` (BvttFrontLooseIntra.cu:63-67), returning the
  narrow-phase filtered pairs by default;
* ``CollisionContext.stats()`` = ``inspectResults`` (BvttFrontLooseIntra.cu:50-61).

The library must be built (``make`` in this directory, or ``__graft_entry__.build()``); there is
no fallback path: importing works without it, but creating a context raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbvhcd.so")

OK, E_INVALID, E_OOM, E_CAPACITY, E_DEVICE, E_NOT_BUILT = 0, -1, -2, -3, -4, -5
MEM_HOST, MEM_DEVICE, MEM_DEVICE_BORROW = 0, 1, 2
BVH_BUILD, BVH_REFIT, BVH_RESTRUCTURE, BVH_NONE = 0, 1, 2, 3
FRONT_PURE_BVH_CD, FRONT_GENERATE, FRONT_UPDATE, FRONT_REORDER, FRONT_KEEP, FRONT_RESTRUCTURE = 0, 1, 2, 3, 4, 5
SCHED_PLAIN, SCHED_LEARNED, SCHED_GENERATE = 0, 1, 2
SCHED_FRONT_GENERATE, SCHED_STATIC_NODIV, SCHED_REFIT_ONLY_FRONT, SCHED_REFIT_ONLY_FRONT_NODIV = 3, 4, 5, 6
SCHED_ARCSIM_STATIC, SCHED_ARCSIM_PERIODIC, SCHED_LEARNED_NO_NET = 7, 8, 9
SCENE_TWO_SPHERES, SCENE_CLOTH, SCENE_LAYERS, SCENE_SHEET_SPHERE = 0, 1, 2, 3
MLP_NPARAM = 16 * 64 + 64 + 64 * 64 + 64 + 64 + 1

BVH_OP_NAMES = {0: "BUILD", 1: "REFIT", 2: "RESTRUCTURE", 3: "NONE"}
FRONT_OP_NAMES = {0: "PURE_BVH_CD", 1: "GENERATE", 2: "UPDATE", 3: "REORDER", 4: "KEEP", 5: "RESTRUCTURE"}


class BvhcdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bvhcd error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("max_tris", C.c_int64), ("max_verts", C.c_int64),
        ("int_front_cap", C.c_int64), ("ext_front_cap", C.c_int64), ("pair_cap", C.c_int64),
        ("narrow_phase", C.c_int32), ("scheduler", C.c_int32), ("enable_restructure", C.c_int32),
        ("bvh_cycle", C.c_int32), ("front_cycle", C.c_int32), ("device", C.c_int32),
        ("stream", C.c_void_p), ("skip_adjacent", C.c_int32), ("auto_grow", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("frame", C.c_int64), ("last_bvh_op", C.c_int32), ("last_front_op", C.c_int32),
        ("int_front", C.c_int64), ("ext_front", C.c_int64), ("bp_pairs", C.c_int64), ("np_pairs", C.c_int64),
        ("required_int_front", C.c_int64), ("required_ext_front", C.c_int64), ("required_pairs", C.c_int64),
        ("node_visits", C.c_int64), ("restructured_subtrees", C.c_int64), ("restructured_leaves", C.c_int64),
        ("stage_ms", C.c_double * 8), ("stage_bytes", C.c_double * 8), ("implicit_pairs", C.c_int64),
    ]


class SceneParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("n", C.c_int32), ("frames", C.c_int32), ("layers", C.c_int32),
        ("subdiv", C.c_int32), ("amplitude", C.c_float), ("compress", C.c_float), ("sigma", C.c_float),
        ("radius", C.c_float), ("spacing", C.c_float), ("speed", C.c_float),
        ("fold", C.c_float), ("fold_radius", C.c_float),
    ]


_lib = None


def library_built() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    """Load libbvhcd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BvhcdError(E_INVALID, f"{LIB_PATH} is missing: build it first (make -C {_HERE})")
    L = C.CDLL(LIB_PATH)
    P, I64, I32, U64 = C.c_void_p, C.c_int64, C.c_int32, C.c_uint64
    sig = {
        "bvhcd_create": (C.c_int, [C.POINTER(Config), C.POINTER(P)]),
        "bvhcd_destroy": (C.c_int, [P]),
        "bvhcd_last_error": (C.c_char_p, [P]),
        "bvhcd_set_topology": (C.c_int, [P, P, I64, I64, C.c_int]),
        "bvhcd_update_vertices": (C.c_int, [P, P, I64, C.c_int]),
        "bvhcd_step": (C.c_int, [P, C.c_int, C.c_int]),
        "bvhcd_step_scheduled": (C.c_int, [P]),
        "bvhcd_get_pairs": (C.c_int, [P, P, I64, C.POINTER(I64), C.c_int, C.c_int]),
        "bvhcd_pair_digest": (C.c_int, [P, C.c_int, C.POINTER(U64), C.POINTER(I64)]),
        "bvhcd_get_stats": (C.c_int, [P, C.POINTER(Stats)]),
        "bvhcd_set_timing": (C.c_int, [P, C.c_int]),
        "bvhcd_timing_frames": (C.c_int, [P, I64, P, P, P, C.POINTER(I64)]),
        "bvhcd_timing_counts": (C.c_int, [P, I64, P, C.POINTER(I64)]),
        "bvhcd_export_bvh": (C.c_int, [P] + [P] * 14),
        "bvhcd_export_front": (C.c_int, [P, C.c_int, P, I64, C.POINTER(I64)]),
        "bvhcd_export_scene_box": (C.c_int, [P, P]),
        "bvhcd_mlp_set_weights": (C.c_int, [P, P, I64]),
        "bvhcd_mlp_get_weights": (C.c_int, [P, P, I64]),
        "bvhcd_mlp_features": (C.c_int, [P, P, P, I64, C.POINTER(I64)]),
        "bvhcd_mlp_forward": (C.c_int, [P, P, I64, P, C.c_int]),
        "bvhcd_mlp_forward_noalloc": (C.c_int, [P, P, I64, P]),
        "bvhcd_mlp_train_step": (C.c_int, [P, P, P, I64, C.c_float, C.POINTER(C.c_float), C.c_int]),
        "bvhcd_mlp_grad": (C.c_int, [P, P, P, I64, P, C.POINTER(C.c_float), C.c_int]),
        "bvhcd_mlp_set_threshold": (C.c_int, [P, C.c_float]),
        "bvhcd_restructure": (C.c_int, [P, P, I64]),
        "bvhcd_batch_create": (C.c_int, [C.POINTER(Config), I32, P, P, C.POINTER(P)]),
        "bvhcd_batch_destroy": (C.c_int, [P]),
        "bvhcd_batch_set_topology": (C.c_int, [P, P, C.c_int]),
        "bvhcd_batch_update_vertices": (C.c_int, [P, P, C.c_int]),
        "bvhcd_batch_step": (C.c_int, [P, C.c_int, C.c_int]),
        "bvhcd_batch_get_pairs": (C.c_int, [P, I32, P, I64, C.POINTER(I64)]),
        "bvhcd_batch_pair_counts": (C.c_int, [P, P]),
        "bvhcd_batch_last_error": (C.c_char_p, [P]),
        "bvhcd_autotune": (C.c_int, [P]),
        "bvhcd_batch_autotune": (C.c_int, [P]),
        "bvhcd_autotune_report": (C.c_int, [P, I64]),
        "bvhcd_autotune_set": (C.c_int, [C.c_int, I64, C.c_uint32, C.c_uint32]),
        "bvhcd_autotune_set_block": (C.c_int, [C.c_int, I64, C.c_uint32, C.c_uint32, C.c_uint32]),
        "bvhcd_update_boxes": (C.c_int, [P, P, I64, C.c_int]),
        "bvhcd_mlp_set_repair_cost": (C.c_int, [P, C.c_float]),
        "bvhcd_schedule_ops": (C.c_int, [C.c_int, I64, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "bvhcd_log_text": (C.c_int, [P, C.c_char_p, I64, C.POINTER(I64)]),
        "bvhcd_inter_create": (C.c_int, [P, C.POINTER(Config), C.POINTER(P)]),
        "bvhcd_inter_destroy": (C.c_int, [P]),
        "bvhcd_inter_last_error": (C.c_char_p, [P]),
        "bvhcd_inter_set_topology": (C.c_int, [P, P, I64, I64, C.c_int]),
        "bvhcd_inter_update_vertices": (C.c_int, [P, P, I64, C.c_int]),
        "bvhcd_inter_step": (C.c_int, [P, C.c_int, C.c_int]),
        "bvhcd_inter_get_pairs": (C.c_int, [P, P, I64, C.POINTER(I64), C.c_int, C.c_int]),
        "bvhcd_inter_export_front": (C.c_int, [P, C.c_int, P, I64, C.POINTER(I64)]),
        "bvhcd_inter_stats": (C.c_int, [P, P]),
        "bvhcd_inter_export_rigid": (C.c_int, [P, P, P]),
        "bvhcd_mesh_load": (C.c_int, [C.c_char_p, C.POINTER(P)]),
        "bvhcd_mesh_destroy": (None, [P]),
        "bvhcd_mesh_num_verts": (I64, [P]),
        "bvhcd_mesh_num_tris": (I64, [P]),
        "bvhcd_mesh_faces": (C.c_int, [P, P]),
        "bvhcd_mesh_verts": (C.c_int, [P, P]),
        "bvhcd_frame_path": (C.c_int, [C.c_char_p, I32, C.c_char_p, I64]),
        "bvhcd_scene_create": (C.c_int, [C.c_int, C.POINTER(SceneParams), C.POINTER(P)]),
        "bvhcd_scene_destroy": (None, [P]),
        "bvhcd_scene_num_tris": (I64, [P]),
        "bvhcd_scene_num_verts": (I64, [P]),
        "bvhcd_scene_faces": (C.c_int, [P, P]),
        "bvhcd_scene_frame": (C.c_int, [P, I32, P, I32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ----------------------------------------------------------------------------- scenes
@dataclass
class SceneSpec:
    kind: int
    seed: int = 1
    n: int = 0
    frames: int = 1
    layers: int = 1
    subdiv: int = 5
    amplitude: float = 0.0
    compress: float = 0.0
    sigma: float = 0.0
    radius: float = 1.0
    spacing: float = 0.0
    speed: float = 0.0
    fold: float = 0.0
    fold_radius: float = 0.0


# The five workloads of the benchmark suite (seeded synthetic deforming meshes).
PRESETS = {
    # two icospheres (20,480 tris each), centres 2.5 apart, closing 0.01/frame, N(0,1e-3) jitter
    "spheres": SceneSpec(SCENE_TWO_SPHERES, seed=1, frames=200, subdiv=5, radius=1.0, sigma=1e-3,
                         spacing=2.5, speed=0.01),
    # 256x256 folding cloth (130,050 tris), 300 frames; the seeded fold reaches 3.7 rad (212
    # degrees) at the last frame, so the flap passes through the rest of the sheet
    "cloth256": SceneSpec(SCENE_CLOTH, seed=2, n=256, frames=300, amplitude=0.15, compress=0.5, fold=3.7,
                          fold_radius=0.05),
    # 1024x1024 folding cloth (2,093,058 tris), 400 frames, same recipe, another seed
    "cloth1024": SceneSpec(SCENE_CLOTH, seed=4, n=1024, frames=400, amplitude=0.15, compress=0.5, fold=3.7,
                           fold_radius=0.05),
    # 4 stacked 1024x1024 cloths 1e-3 apart (8,372,232 tris), 200 frames
    "layers4": SceneSpec(SCENE_LAYERS, seed=5, n=1024, frames=200, layers=4, amplitude=0.05,
                         compress=0.5, spacing=1e-3),
}


def batch_scene_specs(num=64, seed=3, frames=150, t_min=2e3, t_max=2e5):
    """The batched workload: `num` folding sheets (the 256x256 cloth's recipe, fold included) with
    triangle counts log-uniform in [t_min, t_max] (an n x n sheet has 2 (n-1)^2 triangles;
    n = ceil(sqrt(T / 2)) + 1), each
    with a rigid icosphere (subdivision 3, radius 0.08) on a seeded straight path through it."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(num):
        t = float(np.exp(rng.uniform(np.log(t_min), np.log(t_max))))
        n = int(np.ceil(np.sqrt(t / 2.0))) + 1
        specs.append(SceneSpec(SCENE_SHEET_SPHERE, seed=seed * 1000 + k, n=n, frames=frames, subdiv=3,
                               radius=0.08, amplitude=0.15, compress=0.5, fold=3.7, fold_radius=0.05))
    return specs


class Scene:
    """Seeded synthetic deforming mesh (host generator in the library)."""

    def __init__(self, spec: SceneSpec):
        L = lib()
        p = SceneParams(spec.seed, spec.n, spec.frames, spec.layers, spec.subdiv, spec.amplitude,
                        spec.compress, spec.sigma, spec.radius, spec.spacing, spec.speed, spec.fold,
                        spec.fold_radius)
        h = C.c_void_p()
        st = L.bvhcd_scene_create(spec.kind, C.byref(p), C.byref(h))
        if st != OK:
            raise BvhcdError(st, "scene_create")
        self._h = h
        self.spec = spec
        self.num_tris = int(L.bvhcd_scene_num_tris(h))
        self.num_verts = int(L.bvhcd_scene_num_verts(h))
        self.frames = spec.frames

    @classmethod
    def preset(cls, name: str, **over):
        spec = PRESETS[name]
        spec = SceneSpec(**{**spec.__dict__, **over})
        return cls(spec)

    def faces(self) -> np.ndarray:
        f = np.empty((self.num_tris, 3), dtype=np.int32)
        st = lib().bvhcd_scene_faces(self._h, _ptr(f))
        if st != OK:
            raise BvhcdError(st, "scene_faces")
        return f

    def frame(self, f: int, out: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
        if out is None:
            out = np.empty((self.num_verts, 3), dtype=np.float32)
        st = lib().bvhcd_scene_frame(self._h, int(f), _ptr(out), int(threads))
        if st != OK:
            raise BvhcdError(st, f"scene_frame({f})")
        return out

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.bvhcd_scene_destroy(h)
            self._h = None


# ----------------------------------------------------------------------------- context
class CollisionContext:
    """One BVH + BVTT front over a fixed-topology triangle soup on one GPU."""

    def __init__(self, max_tris, max_verts, *, int_front_cap=0, ext_front_cap=0, pair_cap=0,
                 narrow_phase=True, scheduler=SCHED_PLAIN, enable_restructure=True, bvh_cycle=12,
                 front_cycle=6, device=0, stream=None, skip_adjacent=False, auto_grow=False):
        L = lib()
        cfg = Config(int(max_tris), int(max_verts), int(int_front_cap), int(ext_front_cap), int(pair_cap),
                     1 if narrow_phase else 0, int(scheduler), 1 if enable_restructure else 0,
                     int(bvh_cycle), int(front_cycle), int(device), stream, 1 if skip_adjacent else 0,
                     1 if auto_grow else 0)
        h = C.c_void_p()
        st = L.bvhcd_create(C.byref(cfg), C.byref(h))
        if st != OK:
            raise BvhcdError(st, "bvhcd_create failed (no GPU, or out of memory)")
        self._h = h
        self.T = 0
        self.V = 0

    def _chk(self, st, what):
        if st != OK:
            msg = lib().bvhcd_last_error(self._h)
            raise BvhcdError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            lib().bvhcd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_topology(self, faces, num_verts, where=MEM_HOST):
        if where == MEM_HOST:
            f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
            self._chk(lib().bvhcd_set_topology(self._h, _ptr(f), f.shape[0], int(num_verts), MEM_HOST), "set_topology")
            self.T = f.shape[0]
        else:  # device pointer (int), faces = (ptr, T)
            ptr, T = faces
            self._chk(lib().bvhcd_set_topology(self._h, C.c_void_p(ptr), int(T), int(num_verts), MEM_DEVICE),
                      "set_topology")
            self.T = int(T)
        self.V = int(num_verts)

    def update_vertices(self, xyz, where=MEM_HOST):
        """Host array (copied), or a device pointer: MEM_DEVICE copies it, MEM_DEVICE_BORROW has
        the following steps read it in place (keep it alive and unchanged meanwhile)."""
        if where == MEM_HOST:
            v = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
            self._chk(lib().bvhcd_update_vertices(self._h, _ptr(v), v.shape[0], MEM_HOST), "update_vertices")
        else:
            self._chk(lib().bvhcd_update_vertices(self._h, C.c_void_p(int(xyz)), self.V, int(where)),
                      "update_vertices")

    def update_boxes(self, boxes, where=MEM_HOST):
        """Box mode: (T, 6) float32 per-triangle boxes (min xyz, max xyz) replace the vertices
        until the next update_vertices; the narrow phase must be off."""
        if where == MEM_HOST:
            b = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 6)
            self._chk(lib().bvhcd_update_boxes(self._h, _ptr(b), b.shape[0], MEM_HOST), "update_boxes")
        else:
            self._chk(lib().bvhcd_update_boxes(self._h, C.c_void_p(int(boxes)), self.T, MEM_DEVICE), "update_boxes")

    def step(self, bvh_op, front_op):
        self._chk(lib().bvhcd_step(self._h, int(bvh_op), int(front_op)), "step")

    def step_scheduled(self):
        self._chk(lib().bvhcd_step_scheduled(self._h), "step_scheduled")

    def set_timing(self, frames):
        """Record per-stage event timing for up to `frames` frames (0 = off)."""
        self._chk(lib().bvhcd_set_timing(self._h, int(frames)), "set_timing")

    def timing_frames(self):
        """(ms[n, 8], bytes[n, 8], ops[n, 2]) of the frames recorded since set_timing."""
        n = C.c_int64()
        self._chk(lib().bvhcd_timing_frames(self._h, 0, None, None, None, C.byref(n)), "timing_frames")
        ms = np.zeros((n.value, 8), np.float64)
        by = np.zeros((n.value, 8), np.float64)
        ops = np.zeros((n.value, 2), np.int32)
        if n.value:
            self._chk(lib().bvhcd_timing_frames(self._h, n.value, _ptr(ms), _ptr(by), _ptr(ops), C.byref(n)),
                      "timing_frames")
        return ms, by, ops

    def timing_counts(self):
        """counts[n, 8] of the recorded frames: node tests, bp pairs, np pairs, internal and
        external front length, overflow flags, implicit pairs, implicit narrow-phase hits."""
        n = C.c_int64()
        self._chk(lib().bvhcd_timing_counts(self._h, 0, None, C.byref(n)), "timing_counts")
        out = np.zeros((n.value, 8), np.int64)
        if n.value:
            self._chk(lib().bvhcd_timing_counts(self._h, n.value, _ptr(out), C.byref(n)), "timing_counts")
        return out

    def pairs(self, filtered=True) -> np.ndarray:
        n = C.c_int64()
        self._chk(lib().bvhcd_get_pairs(self._h, None, 0, C.byref(n), MEM_HOST, 1 if filtered else 0), "get_pairs")
        out = np.empty((n.value, 2), dtype=np.int32)
        if n.value:
            self._chk(lib().bvhcd_get_pairs(self._h, _ptr(out), n.value, C.byref(n), MEM_HOST,
                                            1 if filtered else 0), "get_pairs")
        return out

    def log_text(self) -> str:
        """The frames recorded since set_timing in the reference's log file layout
        (bvhcd_log_text)."""
        n = C.c_int64()
        self._chk(lib().bvhcd_log_text(self._h, None, 0, C.byref(n)), "log_text")
        buf = C.create_string_buffer(n.value + 1)
        self._chk(lib().bvhcd_log_text(self._h, buf, n.value + 1, C.byref(n)), "log_text")
        return buf.value.decode()

    def pairs_to_device(self, ptr, cap, filtered=True) -> int:
        """Writes the pairs (int32 pairs of triangle ids) into the device buffer at `ptr` (room for
        `cap` pairs), stream-ordered; returns the total pair count (may exceed cap)."""
        n = C.c_int64()
        self._chk(lib().bvhcd_get_pairs(self._h, C.c_void_p(int(ptr)), int(cap), C.byref(n), MEM_DEVICE,
                                        1 if filtered else 0), "get_pairs")
        return int(n.value)

    def digest(self, filtered=True):
        d = C.c_uint64()
        n = C.c_int64()
        self._chk(lib().bvhcd_pair_digest(self._h, 1 if filtered else 0, C.byref(d), C.byref(n)), "pair_digest")
        return int(d.value), int(n.value)

    def stats(self) -> dict:
        s = Stats()
        self._chk(lib().bvhcd_get_stats(self._h, C.byref(s)), "get_stats")
        out = {f: getattr(s, f) for f, _ in Stats._fields_ if f not in ("stage_ms", "stage_bytes")}
        out["stage_ms"] = list(s.stage_ms)
        out["stage_bytes"] = list(s.stage_bytes)
        return out

    def scene_box(self) -> np.ndarray:
        b = np.empty(6, dtype=np.float32)
        self._chk(lib().bvhcd_export_scene_box(self._h, _ptr(b)), "export_scene_box")
        return b

    def export_bvh(self) -> dict:
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
        self._chk(lib().bvhcd_export_bvh(self._h, *[_ptr(a[k]) for k in order]), "export_bvh")
        return a

    def export_front(self, which) -> np.ndarray:
        n = C.c_int64()
        self._chk(lib().bvhcd_export_front(self._h, int(which), None, 0, C.byref(n)), "export_front")
        out = np.empty((n.value, 2), dtype=np.int32)
        if n.value:
            self._chk(lib().bvhcd_export_front(self._h, int(which), _ptr(out), n.value, C.byref(n)), "export_front")
        return out

    # ---- learned scheduler
    def mlp_set_weights(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._chk(lib().bvhcd_mlp_set_weights(self._h, _ptr(p), p.size), "mlp_set_weights")

    def mlp_get_weights(self):
        p = np.empty(MLP_NPARAM, np.float32)
        self._chk(lib().bvhcd_mlp_get_weights(self._h, _ptr(p), p.size), "mlp_get_weights")
        return p

    def mlp_features(self, cap=1 << 20):
        f = np.empty((cap, 16), np.float32)
        r = np.empty(cap, np.int32)
        rows = C.c_int64()
        self._chk(lib().bvhcd_mlp_features(self._h, _ptr(f), _ptr(r), cap, C.byref(rows)), "mlp_features")
        return f[: rows.value].copy(), r[: rows.value].copy()

    def mlp_forward(self, feats):
        x = np.ascontiguousarray(feats, dtype=np.float32)
        out = np.empty(x.shape[0], np.float32)
        self._chk(lib().bvhcd_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(out), MEM_HOST), "mlp_forward")
        return out

    def mlp_train_step(self, feats, targets, lr):
        x = np.ascontiguousarray(feats, dtype=np.float32)
        y = np.ascontiguousarray(targets, dtype=np.float32)
        loss = C.c_float()
        self._chk(lib().bvhcd_mlp_train_step(self._h, _ptr(x), _ptr(y), x.shape[0], float(lr), C.byref(loss),
                                             MEM_HOST), "mlp_train_step")
        return float(loss.value)

    def mlp_grad(self, feats, targets):
        x = np.ascontiguousarray(feats, dtype=np.float32)
        y = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.empty(MLP_NPARAM, np.float32)
        loss = C.c_float()
        self._chk(lib().bvhcd_mlp_grad(self._h, _ptr(x), _ptr(y), x.shape[0], _ptr(g), C.byref(loss), MEM_HOST),
                  "mlp_grad")
        return g, float(loss.value)

    def mlp_set_threshold(self, t):
        self._chk(lib().bvhcd_mlp_set_threshold(self._h, float(t)), "mlp_set_threshold")

    def mlp_set_repair_cost(self, cost):
        """Cost gate of the learned selection (0 = off: every selection above the threshold)."""
        self._chk(lib().bvhcd_mlp_set_repair_cost(self._h, float(cost)), "mlp_set_repair_cost")

    def restructure(self, roots):
        """Restructure the subtrees under the given preorder internal-node ids now; the next
        step repairs the front (incremental front restructure)."""
        r = np.ascontiguousarray(roots, dtype=np.int32).ravel()
        self._chk(lib().bvhcd_restructure(self._h, _ptr(r) if r.size else None, r.size), "restructure")

    def autotune(self):
        """Tune the grids of KEEP and the narrow phase on the current data (needs a step first)."""
        self._chk(lib().bvhcd_autotune(self._h), "autotune")


TUNE_KEEP, TUNE_NARROW, TUNE_BATCH_CD = 0, 1, 2


def autotune_set(kernel: int, n: int, grid: int, unroll: int = 1, block: int = 256) -> None:
    """Fix one kernel's launch shape for the size class of n (bvhcd_autotune_set_block)."""
    st = lib().bvhcd_autotune_set_block(int(kernel), int(n), int(grid), int(unroll), int(block))
    if st != OK:
        raise BvhcdError(st, "autotune_set")


def autotune_report() -> str:
    n = lib().bvhcd_autotune_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().bvhcd_autotune_report(buf, n + 1)
    return buf.value.decode()


# ----------------------------------------------------------------------------- batched scenes
class Batch:
    """Many independent scenes per launch: one tree whose subtrees are the scenes, a persistent
    self-CD traversal, per-scene narrow-phase hits, and (use_graph) one hipGraph per frame."""

    def __init__(self, tris_per_scene, verts_per_scene, *, pair_cap=0, device=0, stream=None):
        t = np.ascontiguousarray(tris_per_scene, dtype=np.int64)
        v = np.ascontiguousarray(verts_per_scene, dtype=np.int64)
        self.S = int(t.size)
        self.toff = np.concatenate([[0], np.cumsum(t)])
        self.voff = np.concatenate([[0], np.cumsum(v)])
        cfg = Config(1, 1, 0, 0, int(pair_cap), 1, SCHED_PLAIN, 0, 12, 6, int(device), stream)
        h = C.c_void_p()
        st = lib().bvhcd_batch_create(C.byref(cfg), self.S, _ptr(t), _ptr(v), C.byref(h))
        if st != OK:
            raise BvhcdError(st, "bvhcd_batch_create failed")
        self._h = h

    def _chk(self, st, what):
        if st != OK:
            msg = lib().bvhcd_batch_last_error(self._h)
            raise BvhcdError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            lib().bvhcd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_topology(self, faces):
        """faces: concatenated per-scene faces with scene-local vertex ids (host)."""
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        if f.shape[0] != self.toff[-1]:
            raise ValueError(f"batch_set_topology: {f.shape[0]} faces, the batch has {int(self.toff[-1])}")
        self._chk(lib().bvhcd_batch_set_topology(self._h, _ptr(f), MEM_HOST), "batch_set_topology")

    def update_vertices(self, xyz, where=MEM_HOST):
        if where == MEM_HOST:
            v = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
            if v.shape[0] != self.voff[-1]:
                raise ValueError(f"batch_update_vertices: {v.shape[0]} vertices, the batch has {int(self.voff[-1])}")
            self._chk(lib().bvhcd_batch_update_vertices(self._h, _ptr(v), MEM_HOST), "batch_update_vertices")
        elif where != MEM_DEVICE:
            raise ValueError("batch_update_vertices: where must be MEM_HOST or MEM_DEVICE")
        else:
            self._chk(lib().bvhcd_batch_update_vertices(self._h, C.c_void_p(int(xyz)), MEM_DEVICE),
                      "batch_update_vertices")

    def step(self, bvh_op, use_graph=True):
        self._chk(lib().bvhcd_batch_step(self._h, int(bvh_op), 1 if use_graph else 0), "batch_step")

    def pair_counts(self) -> np.ndarray:
        c = np.empty(self.S, np.int64)
        self._chk(lib().bvhcd_batch_pair_counts(self._h, _ptr(c)), "batch_pair_counts")
        return c

    def pairs(self, scene) -> np.ndarray:
        n = C.c_int64()
        self._chk(lib().bvhcd_batch_get_pairs(self._h, int(scene), None, 0, C.byref(n)), "batch_get_pairs")
        out = np.empty((n.value, 2), np.int32)
        if n.value:
            self._chk(lib().bvhcd_batch_get_pairs(self._h, int(scene), _ptr(out), n.value, C.byref(n)),
                      "batch_get_pairs")
        return out

    def autotune(self):
        self._chk(lib().bvhcd_batch_autotune(self._h), "batch_autotune")


def plain_schedule(frame, bvh_cycle=12, front_cycle=6, scheduler=None):
    """The ops a plain scheme schedules at `frame` (bvhcd_schedule_ops): by default the
    reference's BenchmarkLogic::maintainScheme, STATIC_MANDATORY with divergent mark
    (system/Logic/BenchmarkLogic.hpp:102-117)."""
    b, f = C.c_int(), C.c_int()
    sched = SCHED_PLAIN if scheduler is None else int(scheduler)
    st = lib().bvhcd_schedule_ops(sched, int(frame), int(bvh_cycle), int(front_cycle), C.byref(b), C.byref(f))
    if st != OK:
        raise BvhcdError(st, "schedule_ops")
    return b.value, f.value


# ----------------------------------------------------------------------------- inter-object front
class InterFront:
    """A second object's triangles against a CollisionContext's tree (bvhcd_inter_*): the
    reference's BvttFrontLooseInter with its rigid LBVH. Step the queried context first each
    frame; pairs are (queried context's triangle, this object's triangle)."""

    def __init__(self, queried: "CollisionContext", max_tris, max_verts, *, int_front_cap=0, ext_front_cap=0,
                 pair_cap=0, narrow_phase=True):
        cfg = Config(int(max_tris), int(max_verts), int(int_front_cap), int(ext_front_cap), int(pair_cap),
                     1 if narrow_phase else 0, SCHED_PLAIN, 0, 12, 6, 0, None, 0)
        h = C.c_void_p()
        st = lib().bvhcd_inter_create(queried._h, C.byref(cfg), C.byref(h))
        if st != OK:
            raise BvhcdError(st, "bvhcd_inter_create failed")
        self._h = h
        self._queried = queried  # the queried context must outlive this object
        self.T = 0

    def _chk(self, st, what):
        if st != OK:
            msg = lib().bvhcd_inter_last_error(self._h)
            raise BvhcdError(st, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            lib().bvhcd_inter_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_topology(self, faces, num_verts):
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        self._chk(lib().bvhcd_inter_set_topology(self._h, _ptr(f), f.shape[0], int(num_verts), MEM_HOST),
                  "inter_set_topology")
        self.T = f.shape[0]

    def update_vertices(self, xyz, where=MEM_HOST):
        if where == MEM_HOST:
            v = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
            self._chk(lib().bvhcd_inter_update_vertices(self._h, _ptr(v), v.shape[0], MEM_HOST), "inter_update_vertices")
        else:
            ptr, V = xyz
            self._chk(lib().bvhcd_inter_update_vertices(self._h, C.c_void_p(int(ptr)), int(V), int(where)),
                      "inter_update_vertices")

    def step(self, rigid_op, front_op):
        self._chk(lib().bvhcd_inter_step(self._h, int(rigid_op), int(front_op)), "inter_step")

    def pairs(self, filtered=True) -> np.ndarray:
        n = C.c_int64()
        self._chk(lib().bvhcd_inter_get_pairs(self._h, None, 0, C.byref(n), MEM_HOST, 1 if filtered else 0),
                  "inter_get_pairs")
        out = np.empty((n.value, 2), dtype=np.int32)
        if n.value:
            self._chk(lib().bvhcd_inter_get_pairs(self._h, _ptr(out), n.value, C.byref(n), MEM_HOST,
                                                  1 if filtered else 0), "inter_get_pairs")
        return out

    def export_front(self, which) -> np.ndarray:
        n = C.c_int64()
        self._chk(lib().bvhcd_inter_export_front(self._h, int(which), None, 0, C.byref(n)), "inter_export_front")
        out = np.empty((n.value, 2), dtype=np.int32)
        if n.value:
            self._chk(lib().bvhcd_inter_export_front(self._h, int(which), _ptr(out), n.value, C.byref(n)),
                      "inter_export_front")
        return out

    def stats(self) -> dict:
        out = np.zeros(6, np.int64)
        self._chk(lib().bvhcd_inter_stats(self._h, _ptr(out)), "inter_stats")
        keys = ["int_front", "ext_front", "bp_pairs", "np_pairs", "node_visits", "last_front_op"]
        return dict(zip(keys, (int(x) for x in out)))

    def export_rigid(self):
        codes, idx = np.empty(self.T, np.uint32), np.empty(self.T, np.int32)
        self._chk(lib().bvhcd_inter_export_rigid(self._h, _ptr(codes), _ptr(idx)), "inter_export_rigid")
        return codes, idx


# ----------------------------------------------------------------------------- mesh files, logs
def load_mesh(path):
    """(faces int32[T, 3], verts float32[V, 3]) of an OBJ or PLY file (bvhcd_mesh_load): every
    mesh of the file concatenated with vertex-id offsets, as the reference's Scene::addMesh
    (world/Scene.cpp:29-42)."""
    L = lib()
    h = C.c_void_p()
    st = L.bvhcd_mesh_load(os.fsencode(path), C.byref(h))
    if st != OK:
        raise BvhcdError(st, f"mesh_load({path})")
    try:
        f = np.empty((L.bvhcd_mesh_num_tris(h), 3), np.int32)
        v = np.empty((L.bvhcd_mesh_num_verts(h), 3), np.float32)
        L.bvhcd_mesh_faces(h, _ptr(f))
        L.bvhcd_mesh_verts(h, _ptr(v))
    finally:
        L.bvhcd_mesh_destroy(h)
    return f, v


def frame_path(pattern, index):
    """Path of one frame of a sequence named by a printf pattern (bvhcd_frame_path)."""
    buf = C.create_string_buffer(4096)
    st = lib().bvhcd_frame_path(os.fsencode(pattern), int(index), buf, 4096)
    if st != OK:
        raise BvhcdError(st, f"frame_path({pattern!r}, {index})")
    return os.fsdecode(buf.value)


class MeshSequence:
    """Frames first .. first+count-1 of a mesh-file sequence (the reference's Benchmark
    {stIdx, len, inputFileFormat}, setting/CDBenchmarkSettings.cpp:17-60). The topology is the
    first frame's; every later frame must have the same vertex count."""

    def __init__(self, pattern, first, count):
        self.pattern, self.first, self.frames = pattern, int(first), int(count)
        self._faces, v0 = load_mesh(frame_path(pattern, self.first))
        self.num_verts, self.num_tris = v0.shape[0], self._faces.shape[0]

    def faces(self):
        return self._faces.copy()

    def frame(self, f):
        _, v = load_mesh(frame_path(self.pattern, self.first + int(f)))
        if v.shape[0] != self.num_verts:
            raise ValueError(f"frame {f}: {v.shape[0]} vertices, the sequence's topology has {self.num_verts}")
        return v


def write_log(path, text):
    """Writes a log text to `path` as the reference's Logger::record does
    (system/Log/Logger.hpp:68-94): the parent directory is created, and an existing file is
    never overwritten (stem + "_" is tried until the name is free). Returns the path used."""
    parent = os.path.dirname(path)
    if parent and not os.path.exists(parent):
        os.makedirs(parent)
    stem, ext = os.path.splitext(os.path.basename(path))
    while os.path.exists(path):
        stem += "_"
        path = os.path.join(parent, stem + ext)
    with open(path, "w") as fh:
        fh.write(text)
    return path


def front_status(text, keyword="front"):
    """Rows (int front, ext front, cpNum, actualCpNum) of every "front (a, b) cpNum c
    actualCpNum d" line of a log text, as the reference's FrontAnalyser data extractor reads them
    (Tools/FrontAnalyser/DataExtractor/Source.cpp:15-32)."""
    rows = []
    toks = text.split()
    k = 0
    while k < len(toks):
        if toks[k] == keyword and k + 6 < len(toks):
            a = toks[k + 1].strip("(,")
            b = toks[k + 2].strip(")")
            rows.append((int(a), int(b), int(toks[k + 4]), int(toks[k + 6])))
            k += 7
        else:
            k += 1
    return rows
