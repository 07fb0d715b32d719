#!/usr/bin/env python3
"""Trains the learned restructure scheduler's MLP on traces of the plain schedule.

Traces: 20 seeded 256x256 folding-cloth sequences (the 256^2 cloth recipe, seeds 101..120),
each run with the plain schedule (the reference's STATIC_MANDATORY cycles plus its restructure
heuristic). Every 6 frames the plain run is forked: two contexts replay the plain schedule up
to frame f - 1 and then take, at frame f,

  features   the 16 per-subtree features of the candidate subtrees (bvhcd_mlp_features),
             computed from the state the learned schedule would see at its decision;
  branch A   refit to frame f and generate the front;
  branch B   rebuild every candidate (bvhcd_restructure, fresh Morton codes of frame f), refit,
             and generate the front.

Above the candidates both trees are identical, so the queries whose cut enters a candidate are
the same in both branches; the label of a candidate is the fraction of the front entries
inside it (internal entries at its nodes, external entries at its leaves) that branch B saves
over branch A. KEEP streams one entry per front entry, so this is the share of the subtree's
per-frame traversal work that restructuring it saves over refitting it.

Training runs on the GPU through the library's own kernels (bf16 MFMA forward, fp32 backward,
Adam on fp32 master weights; bvhcd_mlp_train_step). The decision threshold (restructure when
the predicted saving exceeds it) is chosen on a held-out sequence (seed 131) by measured frame
time of the learned schedule, among 0.05 ... 0.5 and "never" (an infinite threshold: the
network's decisions are then all rejected). Output: an .npz with `params` (5313 fp32) and
`threshold`.
"""
from __future__ import annotations

import argparse
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_pkg():
    if "bvhcd_amd" in sys.modules:
        return sys.modules["bvhcd_amd"]
    spec = importlib.util.spec_from_file_location("bvhcd_amd", os.path.join(HERE, "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bvhcd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[train] {msg}", flush=True)


def entries_inside(ctx, roots):
    """Front entries inside each candidate subtree: internal entries at its internal nodes
    (preorder ids [r, r + leaves - 1)) plus external entries at its leaves."""
    b = ctx.export_bvh()
    rx, ry = b["int_rx"], b["int_ry"]
    n = len(b["leaf_par"])
    fi, fe = ctx.export_front(0), ctx.export_front(1)
    ci = np.bincount(fi[:, 1], minlength=max(n - 1, 1)).astype(np.int64)
    ce = np.bincount(fe[:, 1], minlength=n).astype(np.int64)
    ci_cum = np.concatenate([[0], np.cumsum(ci)])
    ce_cum = np.concatenate([[0], np.cumsum(ce)])
    a, e = rx[roots], ry[roots]
    m = e - a + 1
    return (ci_cum[roots + m - 1] - ci_cum[roots]) + (ce_cum[e + 1] - ce_cum[a])


def init_params(seed):
    """He-normal weights, zero biases, packed W1[64][16], b1[64], W2[64][64], b2[64], W3[64], b3."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / 16), 64 * 16), np.zeros(64), rng.normal(0, np.sqrt(2 / 64), 64 * 64),
             np.zeros(64), rng.normal(0, np.sqrt(1 / 64), 64), np.zeros(1)]
    return np.concatenate(parts).astype(np.float32)


def plain_ops(pkg, frame):
    bo, fo = pkg.plain_schedule(frame)
    if bo == pkg.BVH_REFIT:
        bo = pkg.BVH_RESTRUCTURE  # taken only when the heuristic flagged subtrees last frame
    return bo, fo


def trace_sequence(pkg, seed, frames, fork_every=6):
    sc = pkg.Scene.preset("cloth256", seed=seed, frames=frames)
    faces = sc.faces()
    verts = [sc.frame(f) for f in range(frames)]
    X, Y = [], []
    for fork in range(fork_every, frames, fork_every):
        # a front generated on a tree up to a rebuild cycle old can be several times larger than
        # a fresh one on the folding cloth: room for it
        cap = 128 * sc.num_tris
        ctxs = [pkg.CollisionContext(sc.num_tris, sc.num_verts, int_front_cap=cap, ext_front_cap=cap, pair_cap=cap // 2)
                for _ in range(2)]
        for c in ctxs:
            c.set_topology(faces, sc.num_verts)
        for f in range(fork):  # both contexts on the plain trajectory up to frame fork - 1
            for c in ctxs:
                c.update_vertices(verts[f])
                c.step(*plain_ops(pkg, f))
        feats, roots = ctxs[0].mlp_features()
        for c in ctxs:
            c.update_vertices(verts[fork])
        ctxs[0].step(pkg.BVH_REFIT, pkg.FRONT_GENERATE)
        ctxs[1].restructure(roots)
        ctxs[1].step(pkg.BVH_REFIT, pkg.FRONT_PURE_BVH_CD)  # drops the pending front repair
        ctxs[1].step(pkg.BVH_NONE, pkg.FRONT_GENERATE)
        a = entries_inside(ctxs[0], roots)
        b = entries_inside(ctxs[1], roots)
        for c in ctxs:
            c.close()
        ok = a > 0
        X.append(feats[ok])
        Y.append(((a - b)[ok] / a[ok]).astype(np.float32))
    return np.concatenate(X), np.concatenate(Y)


def measure_schedule(pkg, sc, frames_host, params=None, threshold=0.0, scheduler=None):
    """Sum of per-frame event times (ms) of a scheduled run, and the per-frame pair digests."""
    ctx = pkg.CollisionContext(sc.num_tris, sc.num_verts,
                               scheduler=pkg.SCHED_LEARNED if params is not None else pkg.SCHED_PLAIN)
    ctx.set_topology(sc.faces(), sc.num_verts)
    if params is not None:
        ctx.mlp_set_weights(params)
        ctx.mlp_set_threshold(threshold)
    ctx.set_timing(len(frames_host))
    digests, restr = [], 0
    for v in frames_host:
        ctx.update_vertices(v)
        ctx.step_scheduled()
        digests.append(ctx.digest(True))
        restr += ctx.stats()["restructured_subtrees"]
    ms, _, _ = ctx.timing_frames()
    ctx.close()
    return float(ms[:, 7].sum()), digests, restr


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "mlp_scheduler.npz"))
    ap.add_argument("--seeds", type=int, default=20)
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--init-seed", type=int, default=0)
    args = ap.parse_args()
    pkg = load_pkg()

    t0 = time.time()
    X, Y = [], []
    for k in range(args.seeds):
        x, y = trace_sequence(pkg, 101 + k, args.frames)
        X.append(x)
        Y.append(y)
        log(f"seed {101 + k}: {len(y)} rows, label mean {y.mean() if len(y) else 0:.4f} "
            f"max {y.max() if len(y) else 0:.4f} ({time.time() - t0:.1f}s)")
    X = np.concatenate(X).astype(np.float32)
    Y = np.concatenate(Y).astype(np.float32)
    log(f"{len(Y)} rows; label mean {Y.mean():.4f} std {Y.std():.4f} min {Y.min():.4f} max {Y.max():.4f}")
    rng = np.random.default_rng(args.init_seed)
    perm = rng.permutation(len(Y))
    nval = max(1, len(Y) // 10)
    val, tr = perm[:nval], perm[nval:]

    ctx = pkg.CollisionContext(16, 48)
    ctx.mlp_set_weights(init_params(args.init_seed))
    best, best_p = np.inf, None
    for step in range(args.steps):
        loss = ctx.mlp_train_step(X[tr], Y[tr], args.lr)
        if step % 250 == 0 or step == args.steps - 1:
            pv = ctx.mlp_forward(X[val])
            vl = float(np.mean((pv - Y[val]) ** 2))
            if vl < best:
                best, best_p = vl, ctx.mlp_get_weights()
            log(f"step {step}: train mse {loss:.6f} val mse {vl:.6f} (label var {Y[val].var():.6f})")
    params = best_p
    ctx.close()

    # threshold on a held-out sequence: measured frame time of the learned schedule
    sc = pkg.Scene.preset("cloth256", seed=131, frames=args.frames)
    frames_host = [sc.frame(f) for f in range(sc.frames)]
    t_plain, d_plain, _ = measure_schedule(pkg, sc, frames_host)
    log(f"held-out plain: {t_plain:.1f} ms")
    best_t, best_ms = 0.0, np.inf
    for thr in (0.05, 0.1, 0.2, 0.3, 0.5, float("inf")):  # inf: the network never restructures
        t_l, d_l, restr = measure_schedule(pkg, sc, frames_host, params, thr)
        assert d_l == d_plain, "learned schedule changed a pair set"
        log(f"held-out learned threshold {thr}: {t_l:.1f} ms ({t_plain / t_l:.3f}x plain), {restr} subtrees rebuilt")
        if t_l < best_ms:
            best_t, best_ms = thr, t_l
    np.savez(args.out, params=params.astype(np.float32), threshold=np.float32(best_t),
             val_mse=np.float32(best), rows=np.int64(len(Y)))
    log(f"wrote {args.out}: threshold {best_t}, val mse {best:.6f} ({time.time() - t0:.1f}s)")


if __name__ == "__main__":
    main()
