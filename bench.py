#!/usr/bin/env python3
"""Benchmark: collision-detection frames/sec on the 1024x1024 folding cloth (2,093,058 triangles).

One step = one frame of the hot path: vertices of frame f (already in HBM, read in place as a
simulator hands its device buffer over: BVHCD_MEM_DEVICE_BORROW) -> BVH maintenance
(build / refit / restructure) -> BVTT front maintenance -> narrow phase -> colliding pairs.
The run replays the seeded 400-frame sequence from frame 0 (warm-up frames first, then the
timed frames), with the plain scheduler (reference STATIC_MANDATORY schedule + restructure
heuristic) and, when weights are available, the learned restructure scheduler.

Output: one JSON line (metric, value = frames/s of the reported scheduler, roofline of the
dominant kernel from live HIP-event timing, CPU baseline from the host oracle on a sample).
"""
from __future__ import annotations

import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, GB/s

STAGES = ["build", "refit", "restructure", "front", "order+quality", "narrow", "schedule", "total"]
FRONT_KERNEL = {0: "k_traverse<false>", 1: "k_traverse<true>", 2: "k_update", 3: "k_keep+k_reorder_keys", 4: "k_keep",
                5: "k_keep+k_restr_walk"}


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "bvhcd_amd", os.path.join(ROOT, "bvh-based-collision-detection-scheme_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bvhcd_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def upload_frames(pkg, sc, first, count, torch):
    """Generate frames [first, first+count) on the host and keep them resident in HBM."""
    dev = torch.empty((count, sc.num_verts, 3), dtype=torch.float32, device="cuda")
    host = np.empty((sc.num_verts, 3), np.float32)
    t0 = time.time()
    for k in range(count):
        sc.frame(first + k, host)
        dev[k].copy_(torch.from_numpy(host))
    torch.cuda.synchronize()
    log(f"generated + uploaded {count} frames in {time.time() - t0:.1f}s")
    return dev


def make_ctx(pkg, sc, faces, scheduler, weights):
    ctx = pkg.CollisionContext(sc.num_tris, sc.num_verts, scheduler=scheduler)
    ctx.set_topology(faces, sc.num_verts)
    if weights is not None:
        ctx.mlp_set_weights(weights[0])
        ctx.mlp_set_threshold(weights[1])
    return ctx


def frame_digests(pkg, sc, faces, frames_dev, count, scheduler, weights):
    """Untimed replay: the narrow-phase pair-set digest of every frame."""
    ctx = make_ctx(pkg, sc, faces, scheduler, weights)
    nbytes = sc.num_verts * 12
    out = []
    for f in range(count):
        ctx.update_vertices(frames_dev.data_ptr() + f * nbytes, pkg.MEM_DEVICE_BORROW)
        ctx.step_scheduled()
        out.append(tuple(ctx.digest(True)))
    ctx.close()
    return out


def run_gpu(pkg, sc, faces, frames_dev, warmup, steps, scheduler, torch, weights=None, digests_upto=0,
            autotune=True, instrument=False):
    """Warm-up frames, then `steps` timed frames. instrument=False: wall clock only (the rate that
    is reported); instrument=True: per-stage hipEvents and counters on every frame (each event is
    a stream barrier, so this pass is slower and is used only for the stage breakdown)."""
    ctx = make_ctx(pkg, sc, faces, scheduler, weights)
    nbytes = sc.num_verts * 12
    base = frames_dev.data_ptr()
    digests = []
    for f in range(warmup):
        ctx.update_vertices(base + f * nbytes, pkg.MEM_DEVICE_BORROW)
        ctx.step_scheduled()
        if f < digests_upto:
            digests.append(ctx.digest(True))
    if autotune:  # launch shapes tuned on this workload before the timed frames (process-wide table)
        ctx.autotune()
        chosen = [ln for ln in pkg.autotune_report().splitlines() if ln.endswith("*")]
        log("tuned launch shapes: " + "; ".join(chosen))
    torch.cuda.synchronize()
    if instrument:
        ctx.set_timing(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        f = warmup + k
        ctx.update_vertices(base + f * nbytes, pkg.MEM_DEVICE_BORROW)
        ctx.step_scheduled()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms, by, ops = ctx.timing_frames() if instrument else (None, None, None)
    st = ctx.stats()
    last_digest = ctx.digest(True)
    ctx.close()
    restructured = int((ops[:, 0] == pkg.BVH_RESTRUCTURE).sum()) if instrument else None
    updates = int((ops[:, 1] == pkg.FRONT_UPDATE).sum()) if instrument else None
    return dict(wall_s=t1 - t0, ms=ms, bytes=by, ops=ops, stats=st, digests=digests, last_digest=last_digest,
                restructured=restructured, updates=updates)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_cloth1024.md")


def profiled_traffic(kernel, workload, schedule):
    """HBM bytes per launch of `kernel` on this workload and schedule, from the committed counter
    summary (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled by what the counter reads for
    a 1 GiB copy); None when the summary holds no row for it."""
    if workload != "cloth1024" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as fh:
        for line in fh:
            cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
            if len(cells) == 6 and cells[0] == kernel and cells[1] == schedule:
                try:
                    return int(cells[5].replace(",", ""))
                except ValueError:
                    return None
    return None


def roofline(res):
    """Dominant (stage, op) group over the timed frames, algorithmic bytes / event time."""
    ms, by, ops = res["ms"], res["bytes"], res["ops"]
    groups = {}
    for f in range(ms.shape[0]):
        for s in range(7):
            if ms[f, s] <= 0:
                continue
            key = (s, int(ops[f, 1]) if s in (3, 4) else -1)
            g = groups.setdefault(key, [0.0, 0.0, 0])
            g[0] += ms[f, s]
            g[1] += by[f, s]
            g[2] += 1
    (s, op), (t, b, n) = max(groups.items(), key=lambda kv: kv[1][0])
    name = {0: "build kernels", 1: "k_refit", 2: "restructure kernels", 4: "ordering kernels",
            5: "k_narrow", 6: "restructure selection"}.get(s, FRONT_KERNEL.get(op, "front"))
    achieved = b / (t * 1e-3) / 1e9
    stage_summary = {}
    for (s2, op2), (t2, b2, n2) in sorted(groups.items()):
        label = STAGES[s2] + (f"/{['PURE', 'GENERATE', 'UPDATE', 'REORDER', 'KEEP', 'RESTRUCTURE'][op2]}" if op2 >= 0 else "")
        stage_summary[label] = {"frames": n2, "avg_ms": round(t2 / n2, 4),
                                "GBps": round(b2 / (t2 * 1e-3) / 1e9, 1),
                                "frac_peak": round(b2 / (t2 * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    return {"kernel": name, "stage": STAGES[s], "launches": n, "avg_ms": t / n,
            "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None}, stage_summary


def copy_ceiling(torch):
    """Measured stream-copy ceiling: read + write bytes per second of a 1 GiB device-to-device
    copy (torch's copy kernel, timed with events on torch's stream)."""
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    del src, dst
    return round(2 * (1 << 30) / (ms * 1e-3) / 1e9, 1)


def cpu_baseline(pkg, sc, faces, budget_s, threads, gpu_digests):
    """Host oracle (restatement of the reference algorithm) on frames 0..S-1, same schedule."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # test infrastructure, used only as the CPU baseline and checker

    orc = oracle.OracleContext(faces, sc.num_verts)
    host = np.empty((sc.num_verts, 3), np.float32)
    done, t_total, match = 0, 0.0, True
    while t_total < budget_s and done < sc.frames:
        sc.frame(done, host)
        orc.update_vertices(host)
        t0 = time.perf_counter()
        orc.step_scheduled(threads=threads)
        t_total += time.perf_counter() - t0
        if done < len(gpu_digests):
            match &= oracle.pair_digest(orc.pairs(True)) == tuple(gpu_digests[done])
        done += 1
    orc.close()
    return done, t_total, match, min(done, len(gpu_digests))


def batch_bench(pkg, torch, steps, warmup, no_autotune):
    """Batched scenes: 64 folding sheets (2e3..2e5 triangles, log-uniform) + icospheres, one
    tree and one captured graph per frame op; BUILD every 12 frames, REFIT otherwise."""
    specs = pkg.batch_scene_specs()
    scenes = [pkg.Scene(s) for s in specs]
    frames = min(specs[0].frames, warmup + steps)
    tps = [s.num_tris for s in scenes]
    vps = [s.num_verts for s in scenes]
    T, V = sum(tps), sum(vps)
    log(f"batch: {len(scenes)} scenes, {T} triangles, {V} vertices, {frames} frames")
    dev = torch.empty((frames, V, 3), dtype=torch.float32, device="cuda")
    for f in range(frames):
        dev[f].copy_(torch.from_numpy(np.concatenate([s.frame(f) for s in scenes])))
    torch.cuda.synchronize()
    b = pkg.Batch(tps, vps)
    b.set_topology(np.concatenate([s.faces() for s in scenes]))
    nb = V * 12

    def frame(f):
        b.update_vertices(dev.data_ptr() + f * nb, pkg.MEM_DEVICE)
        b.step(pkg.BVH_BUILD if f % 12 == 0 else pkg.BVH_REFIT, use_graph=True)

    for f in range(warmup):
        frame(f)
    if not no_autotune:
        b.autotune()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(warmup, frames):
        frame(f)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    counts = b.pair_counts()
    b.close()
    n = frames - warmup
    return {"metric": "batched CD triangles/sec (64 scenes, 2e3-2e5 tris each, exact pairs)",
            "value": round(T * n / dt, 1), "unit": "triangles/s", "n_gpus": 1, "steps": n, "warmup": warmup,
            "ms_per_step": round(1e3 * dt / n, 4), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"batch64: {len(scenes)} sheet+icosphere scenes, {T} triangles", "graph": True},
            "pairs_last_frame": int(counts.sum())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=360)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--workload", default="cloth1024",
                    choices=["cloth1024", "cloth256", "layers4", "spheres", "batch64"],
                    help="batch64: the batched-scene workload (its own JSON line, triangles/s)")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of oracle CPU work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--weights", default=os.path.join(ROOT, "bvh-based-collision-detection-scheme_amd",
                                                      "mlp_scheduler.npz"),
                    help="trained scheduler (params + threshold) written by train_scheduler.py")
    ap.add_argument("--no-verify", action="store_true", help="skip the per-frame pair-set comparison pass")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--host-upload", action="store_true",
                    help="also time the reported schedule with every frame's vertices copied from host memory "
                         "(PCIe-inclusive rate, reported beside value, never as value)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="finish with three 1 GiB device-to-device copies (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on exactly one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import torch

    pkg = load_pkg()
    if args.workload == "batch64":
        print(json.dumps(batch_bench(pkg, torch, min(args.steps, 138), min(args.warmup, 12), args.no_autotune)),
              flush=True)
        return
    sc = pkg.Scene.preset(args.workload)
    steps = min(args.steps, sc.frames - args.warmup)
    warmup = args.warmup
    faces = sc.faces()
    log(f"{args.workload}: {sc.num_tris} triangles, {sc.num_verts} vertices, warmup {warmup}, steps {steps}")
    frames_dev = upload_frames(pkg, sc, 0, warmup + steps, torch)

    ndig = min(warmup, 12)
    plain = run_gpu(pkg, sc, faces, frames_dev, warmup, steps, pkg.SCHED_PLAIN, torch, digests_upto=ndig,
                    autotune=not args.no_autotune)
    plain_fps = steps / plain["wall_s"]
    log(f"plain: {plain_fps:.1f} frames/s ({1e3 * plain['wall_s'] / steps:.3f} ms/frame)")

    learned = None
    weights = None
    identical = None
    if os.path.exists(args.weights):
        z = np.load(args.weights, allow_pickle=False)
        weights = (z["params"], float(z["threshold"]))
        learned = run_gpu(pkg, sc, faces, frames_dev, warmup, steps, pkg.SCHED_LEARNED, torch, weights=weights,
                          autotune=not args.no_autotune)
        log(f"learned: {steps / learned['wall_s']:.1f} frames/s")
        if not args.no_verify:  # untimed replay: pair sets of every frame, learned vs plain
            dp = frame_digests(pkg, sc, faces, frames_dev, warmup + steps, pkg.SCHED_PLAIN, None)
            dl = frame_digests(pkg, sc, faces, frames_dev, warmup + steps, pkg.SCHED_LEARNED, weights)
            identical = dp == dl
            log(f"pair sets identical on all {len(dp)} frames: {identical}")

    rep = learned if learned is not None else plain
    value = steps / rep["wall_s"]
    # stage breakdown: the same frames again with per-stage events (launch shapes already tuned)
    inst = run_gpu(pkg, sc, faces, frames_dev, warmup, steps, pkg.SCHED_LEARNED if learned is not None
                   else pkg.SCHED_PLAIN, torch, weights=weights, autotune=False, instrument=True)
    log(f"instrumented pass: {steps / inst['wall_s']:.1f} frames/s ({inst['restructured']} restructure frames, "
        f"{inst['updates']} UPDATE frames)")
    roof, stages = roofline(inst)
    roof["copy_ceiling_GBps"] = copy_ceiling(torch)
    traffic = profiled_traffic(roof["kernel"], args.workload, "learned" if learned is not None else "plain")
    if traffic is not None:  # measured HBM bytes per launch; its rate over the same event time
        roof["traffic"] = traffic
        roof["traffic_GBps"] = round(traffic / (roof["avg_ms"] * 1e-3) / 1e9, 1)
        roof["traffic_frac"] = round(roof["traffic_GBps"] / PEAK_HBM_GBS, 4)
        roof["traffic_source"] = os.path.relpath(TRAFFIC_PROFILE, ROOT)

    host_fps = None
    if args.host_upload:  # same frames, vertices handed over from (pinned) host memory every frame
        pinned = torch.empty((warmup + steps,) + tuple(frames_dev.shape[1:]), dtype=torch.float32, pin_memory=True)
        pinned.copy_(frames_dev[: warmup + steps].cpu())
        host_frames = [pinned[f].numpy() for f in range(warmup + steps)]
        ctx = make_ctx(pkg, sc, faces, pkg.SCHED_LEARNED if learned is not None else pkg.SCHED_PLAIN, weights)
        for f in range(warmup):
            ctx.update_vertices(host_frames[f])
            ctx.step_scheduled()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(warmup, warmup + steps):
            ctx.update_vertices(host_frames[f])
            ctx.step_scheduled()
        torch.cuda.synchronize()
        host_fps = round(steps / (time.perf_counter() - t0), 3)
        ctx.close()
        log(f"host-upload (PCIe-inclusive) rate: {host_fps} frames/s")

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        done, t_cpu, match, checked = cpu_baseline(pkg, sc, faces, args.cpu_budget, threads, plain["digests"])
        cpu = {"value": round(done / t_cpu, 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"frames 0..{done - 1} of the same sequence, plain schedule, host oracle "
                         f"(C++ restatement, std::thread x{threads})",
               "pairs_match_gpu_frames": checked, "pairs_match": bool(match)}
        log(f"cpu oracle: {cpu['value']} frames/s over {done} frames, GPU digests match: {match} ({checked} frames)")

    out = {
        "metric": "CD frames/sec @2M-tri cloth, learned vs plain sched, exact pairs; HBM GB/s",
        "value": round(value, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(1e3 * rep["wall_s"] / steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: seeded folding cloth, {sc.num_tris} triangles, "
                               f"{sc.num_verts} vertices, frames {warmup}..{warmup + steps - 1} of {sc.frames}",
                   "scheduler": "learned" if learned is not None else "plain",
                   "narrow_phase": True},
        "plain_fps": round(plain_fps, 3),
        "learned_fps": round(steps / learned["wall_s"], 3) if learned is not None else None,
        "learned_vs_plain": round((steps / learned["wall_s"]) / plain_fps, 4) if learned is not None else None,
        "learned_pairs_identical_all_frames": identical,
        "host_upload_fps": host_fps,
        "restructure_frames": inst["restructured"],
        "instrumented_fps": round(steps / inst["wall_s"], 3),
        "pairs_last_frame": rep["last_digest"][1],
        "roofline": roof,
        "stages": stages,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.calibrate_copy:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration: 3 copies of 1 GiB done")


if __name__ == "__main__":
    main()
