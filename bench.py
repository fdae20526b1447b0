"""Benchmark: batched 84x84 phong_reflection_shadow rendering of 1024 uneven
Brax-like scenes (256x256 shadow maps) on one GPU, against the plain
per-scene launch and the CPU oracle.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  `value` is batched forward frames/s with every input
already resident in HBM (one HIP-graph replay per step renders all scenes:
shadow pass then main pass).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_times(n_calls: int):
    import numpy as np

    from jaxrenderer_amd import _hip

    L = _hip.lib()
    buf = (ctypes.c_float * (4 * n_calls))()
    n = L.jr_stage_times(buf, n_calls)
    return np.array(buf[: 4 * n], dtype=np.float64).reshape(n, 4)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scenes", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--tile", default="auto",
                    help="'auto' (time every launch shape on this batch) or WxH, e.g. 8x8")
    ap.add_argument("--shadow-tile", default=None, help="WxH for the shadow pass (with --tile WxH)")
    ap.add_argument("--no-autotune-cache", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-per-scene", action="store_true")
    ap.add_argument("--no-backward", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy, the reference "
                         "kernel for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build renders on one GPU; --gpus must be 1 "
                                   "(scenes shard across GPUs with no exchange, see DESIGN.md)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from jaxrenderer_amd import _hip, batch, workloads

    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 268435456 bytes read, 268435456 bytes written")
        del src, dst

    W = H = 84
    B = args.scenes
    t0 = time.time()
    scenes, smspec = workloads.config3(n_scenes=B, seed=args.seed, width=W, height=H)
    tris = workloads.triangle_counts(scenes)
    calls = batch.plan_batch(scenes, W, H, smspec)
    log(f"generated {B} scenes ({tris.sum()} triangles, {tris.min()}..{tris.max()} per scene) "
        f"in {time.time() - t0:.1f}s")

    def wh(s):
        w, h = (int(v) for v in s.lower().split("x"))
        return (w, h)

    auto = args.tile == "auto"
    tile = (8, 8) if auto else wh(args.tile)
    shadow_tile = wh(args.shadow_tile) if args.shadow_tile else tile
    br = batch.BatchRenderer(calls, tile=tile, shadow_tile=shadow_tile, persistent=True)
    tuned = None
    if auto:
        t0 = time.time()
        tuned = br.autotune(use_cache=not args.no_autotune_cache)
        for tag, c in tuned.items():
            log(f"autotune {tag}: tile {c['tile']} persistent={c['persistent']} "
                f"{c['ms']:.3f} ms ({'cached' if c['cached'] else 'timed'})")
        log(f"autotune took {time.time() - t0:.1f}s")
    pairs = br.calibrate()
    log(f"binned triangle-tile pairs: {pairs}")
    br.capture()
    for _ in range(args.warmup):
        br.replay()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        br.replay()
    torch.cuda.synchronize()
    t_batched = (time.perf_counter() - t_start) / args.steps
    fps = B / t_batched
    log(f"batched: {t_batched * 1e3:.3f} ms per step -> {fps:.0f} frames/s")
    out = br.render()
    torch.cuda.synchronize()
    canvas_b = out["canvas"].clone()
    zbuf_b = out["zbuffer"].clone()

    # per-stage durations, events recorded on the rendering stream
    lib = _hip.lib()
    lib.jr_stage_timing(1)
    n_prof = max(3, min(args.steps, 10))
    for _ in range(n_prof):
        br.render()
    torch.cuda.synchronize()
    lib.jr_stage_timing(0)
    st = stage_times(2 * n_prof)
    shadow_ms = st[0::2].mean(axis=0)
    main_ms = st[1::2].mean(axis=0)
    names = ["setup", "scan", "bin", "raster"]
    stages = {f"shadow.{n}": float(v) for n, v in zip(names, shadow_ms)}
    stages.update({f"main.{n}": float(v) for n, v in zip(names, main_ms)})
    dom = max(stages, key=stages.get)
    log("stage ms: " + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))

    # kept pixels of the main pass (for the algorithmic byte count)
    kept = int((zbuf_b < 1.0).sum())

    # forward + backward: L2 image loss against a fixed synthetic target,
    # gradients to vertices, normals, diffuse/specular maps and scene params
    fwd_bwd_fps = None
    grads_finite = None
    if not args.no_backward:
        br.enable_backward()
        gen = torch.Generator(device="cuda").manual_seed(1)
        target = torch.rand(canvas_b.shape, generator=gen, device="cuda")
        br.capture_train(target)
        for _ in range(args.warmup):
            br.train_graph.replay()
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            br.train_graph.replay()
        torch.cuda.synchronize()
        t_fb = (time.perf_counter() - t_start) / args.steps
        fwd_bwd_fps = B / t_fb
        grads_finite = bool(all(torch.isfinite(g).all() for g in br.grads.values()))
        log(f"fwd+bwd: {t_fb * 1e3:.3f} ms per step -> {fwd_bwd_fps:.0f} frames/s "
            f"({fwd_bwd_fps / fps:.2f}x forward), gradients finite: {grads_finite}")

    # plain per-scene forward: the same pass sequence with a batch of one
    per_scene_fps = None
    mismatches = None
    if not args.no_per_scene:
        singles = []
        for i in range(B):
            r = batch.BatchRenderer(batch.plan_batch([scenes[i]], W, H, smspec),
                                    tile=(8, 8), shadow_tile=(8, 8), persistent=False)
            singles.append(r)
        for r in singles[: min(8, B)]:
            r.render()
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for r in singles:
            r.render()
        torch.cuda.synchronize()
        t_single = time.perf_counter() - t_start
        per_scene_fps = B / t_single
        log(f"per-scene launches: {t_single * 1e3:.1f} ms for {B} scenes -> "
            f"{per_scene_fps:.0f} frames/s")
        bad = 0
        for i, r in enumerate(singles):
            zs = r.passes["main"].zbuffer[0]
            cs = r.passes["main"].canvas[0]
            same_z = torch.equal(zs.view(torch.int32), zbuf_b[i].view(torch.int32))
            close_c = bool((cs - canvas_b[i]).abs().max() <= 1e-5)
            bad += int(not (same_z and close_c))
        mismatches = bad
        log(f"batched vs per-scene: {bad} of {B} scenes differ")
        del singles

    # CPU oracle on a bounded sample of the same scenes
    cpu = None
    if not args.no_cpu_baseline:
        from oracle.oracle import run_oracle_plan

        n_done, t_cpu = 0, 0.0
        while t_cpu < args.cpu_seconds and n_done < B:
            plan = batch.plan_batch([scenes[n_done]], W, H, smspec)
            t1 = time.perf_counter()
            run_oracle_plan(plan, prim_id=False, bary=False)
            t_cpu += time.perf_counter() - t1
            n_done += 1
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cores = min(cores, os.cpu_count() or cores)
        cpu = {"value": n_done / t_cpu, "unit": "frames/s", "cores": cores, "kind": "port",
               "sample": f"first {n_done} of the {B} scenes (same seed), brute-force C oracle, "
                         f"OpenMP over rows, {t_cpu:.1f}s"}
        log(f"cpu oracle: {n_done} scenes in {t_cpu:.1f}s")

    # roofline of the dominant stage: algorithmic bytes of its launch
    # (definitions in DESIGN.md) over its event-timed duration
    n_tri = int(tris.sum())
    n_pix_main = B * W * H
    n_pix_shadow = B * smspec.width * smspec.height
    alg = {
        # read 3 slot ids + 3 position gathers (+3 attribute indices for the
        # main pass), write one 64-byte record and a 16-byte tile range
        "setup": lambda pass_: n_tri * ((12 + 36 + (12 if pass_ == "main" else 0)) + 80),
        "raster": lambda pass_: (
            (pairs[pass_] * 52 + n_pix_shadow * 4) if pass_ == "shadow" else
            (pairs[pass_] * 52 + n_pix_main * 16
             + kept * (64 + 3 * (4 + 12 + 12 + 8 + 12) + 4 + 8 + 12 + 4 + 4))),
        "bin": lambda pass_: n_tri * 16 + pairs[pass_] * 4,
        "scan": lambda pass_: (n_pix_main if pass_ == "main" else n_pix_shadow) // 64 * 12,
    }
    dpass, dstage = dom.split(".")
    dom_bytes = alg[dstage](dpass)
    achieved = dom_bytes / (stages[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom, "kernel_ms": round(stages[dom], 4), "alg_bytes": int(dom_bytes)}

    res = {
        "metric": "frames/s at 84x84 batch 1024 (fwd, fwd+bwd) vs per-scene launch, 1 GPU",
        "value": round(fps, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batched * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "config 3: 1024 Brax-like scenes (ground plane + 4..30 boxes/capsules/"
                        "spheres, 650..10k triangles), 84x84, phong_reflection_shadow, "
                        "256x256 shadow map, batched forward in one HIP-graph replay",
            "scenes": B,
            "triangles": n_tri,
            "width": W,
            "height": H,
            "shadow_map": [smspec.width, smspec.height],
            "tile": {tag: {"tile": list(p.tile), "persistent": p.persistent}
                     for tag, p in br.passes.items()},
            "autotuned": auto,
        },
        "mpix_per_s": round(fps * W * H / 1e6, 2),
        "per_scene_fps": None if per_scene_fps is None else round(per_scene_fps, 1),
        "batched_over_per_scene": None if per_scene_fps is None else round(fps / per_scene_fps, 2),
        "batched_vs_per_scene_mismatched_scenes": mismatches,
        "fwd_bwd_fps": None if fwd_bwd_fps is None else round(fwd_bwd_fps, 1),
        "fwd_bwd_over_fwd": None if fwd_bwd_fps is None else round(fwd_bwd_fps / fps, 3),
        "fwd_bwd_grads_finite": grads_finite,
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "binned_pairs": pairs,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
