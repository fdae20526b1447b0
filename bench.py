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
# fp32 VALU issue rate of the chip without packed or fused operations:
# 256 CUs x 64 lanes x 2.4 GHz, in Gops/s (the operations of the work model
# below are unfused fp32 operations, the code being built with contraction off)
VALU_PEAK_GOPS = 256 * 64 * 2.4


def work_model(scenes, width: int, height: int, smspec, sample: dict) -> dict:
    """Algorithmic work of one frame of a batch, per stage and in total.

    Bytes are the compulsory input and output bytes: 12 B per unique
    position, 12 B per unique normal, 8 B per unique uv, 36 B per triangle
    (three int32 index triples), 16 B per texel of every object's diffuse +
    specular map plus 76 B of per-object pose, 16 B per pixel of colour +
    depth.  The shadow map is an intermediate and counts nothing.  Ops are
    unfused fp32 VALU operations: 120 per corner for the main vertex stage,
    80 per triangle for each pass's setup (M, det, inverse), 28 per unique
    vertex for the shadow vertex stage, 30 per (pixel, covering primitive)
    pair and 220 per covered pixel in the main pass, 25 per (texel, covering
    primitive) pair in the shadow pass.  "Covering" means passing keep,
    inside and front-facing (the depth complexity); it is counted by the CPU
    oracle on `sample` (a prefix of the batch) and scaled by the batch's
    triangle count over the sample's.

    A stage is charged the bytes it must read or write itself, so the two
    setups each count the positions and faces, which the frame total counts
    once."""
    F = Vu = Nu = UVu = maps = 0
    for s in scenes:
        m = s.model
        F += len(m.faces)
        Vu += len(m.verts)
        Nu += len(m.norms)
        UVu += len(m.uvs)
        ts = m.texture_shape
        maps += sum(16 * int(w) * int(h) + 76 for w, h in ts)
    B = len(scenes)
    scale = F / max(sample["triangles"], 1)
    P_main = sample["P_main"] * scale
    C_main = sample["C_main"] * scale
    P_shadow = sample["P_shadow"] * scale if smspec is not None else 0.0
    out_b = B * width * height * 16
    geo_b = 12 * Vu + 36 * F
    stage_bytes = {"main.setup": geo_b, "main.raster": 12 * Nu + 8 * UVu + maps + out_b}
    stage_ops = {"main.setup": 120 * 3 * F + 80 * F, "main.raster": 30 * P_main + 220 * C_main}
    if smspec is not None:
        stage_bytes.update({"shadow.setup": geo_b, "shadow.raster": 0})
        stage_ops.update({"shadow.setup": 28 * Vu + 80 * F, "shadow.raster": 25 * P_shadow})
    total_b = 12 * Vu + 12 * Nu + 8 * UVu + 36 * F + maps + out_b
    total_ops = sum(stage_ops.values())
    return {"bytes": stage_bytes, "ops": stage_ops, "bytes_total": total_b,
            "ops_total": total_ops,
            "counts": {"scenes": B, "triangles": F, "unique_positions": Vu,
                       "unique_normals": Nu, "unique_uvs": UVu,
                       "P_main": round(P_main), "C_main": round(C_main),
                       "P_shadow": round(P_shadow)}}


def fractions(nbytes: float, ops: float, ms: float) -> dict:
    """Achieved rates of `nbytes` and `ops` done in `ms`, against the HBM and
    VALU peaks; `bound` names the larger fraction."""
    t = ms * 1e-3
    gbs = nbytes / t / 1e9
    gops = ops / t / 1e9
    hbm, valu = gbs / HBM_PEAK_GBS, gops / VALU_PEAK_GOPS
    return {"bound": "valu" if valu >= hbm else "hbm",
            "hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(hbm, 4), "bytes": int(nbytes)},
            "valu": {"achieved": round(gops, 1), "peak": round(VALU_PEAK_GOPS, 1),
                     "unit": "Gops/s", "frac": round(valu, 4), "ops": int(ops)}}


def complexity_sample(scenes, width: int, height: int, smspec, n: int) -> dict:
    """Depth complexity of the first n scenes, counted by the CPU oracle."""
    from jaxrenderer_amd import batch
    from oracle.oracle import run_oracle_plan

    P_main = C_main = P_shadow = tris = 0
    for s in scenes[:n]:
        res = run_oracle_plan(batch.plan_batch([s], width, height, smspec), prim_id=False,
                              bary=False, complexity=True)
        cx = res["main"]["complexity"]
        P_main += int(cx.sum())
        C_main += int((cx > 0).sum())
        if "shadow" in res:
            P_shadow += int(res["shadow"]["complexity"].sum())
        tris += len(s.model.faces)
    return {"scenes": min(n, len(scenes)), "triangles": tris, "P_main": P_main,
            "C_main": C_main, "P_shadow": P_shadow}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stage_times(n_calls: int):
    import numpy as np

    from jaxrenderer_amd import _hip

    L = _hip.lib()
    k = _hip.STAGES
    buf = (ctypes.c_float * (k * n_calls))()
    n = L.jr_stage_times(buf, n_calls)
    return np.array(buf[: k * n], dtype=np.float64).reshape(n, k)


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
    ap.add_argument("--no-pose", action="store_true", help="skip timing the pose update")
    ap.add_argument("--full-shadow-map", action="store_true",
                    help="rasterise every shadow-map texel (default: the main pass computes "
                         "the texels it reads from the shadow pass's binning)")
    ap.add_argument("--workload", choices=("c3", "c4", "c5"), default="c3",
                    help="c3 (default): 1024 Brax-like scenes at 84x84; c4: 64-view "
                         "differentiation of a 5000-triangle sphere at 256x256; c5: 512 "
                         "scenes of 1k..20k triangles at 480x270, forward + backward")
    ap.add_argument("--no-extra-configs", action="store_true",
                    help="skip the summaries of configs 1, 2, 4 and 5 in the default line")
    ap.add_argument("--c5-per-scene", type=int, default=64,
                    help="config 5: scenes timed one by one for the per-scene comparison")
    ap.add_argument("--complexity-scenes", type=int, default=32,
                    help="scenes whose depth complexity the oracle counts for the work model")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 256 MiB device-to-device copy, the reference "
                         "kernel for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py renders on one GPU; --gpus must be 1 (a batch "
                                   "shards by scene across GPUs with no exchange: one process "
                                   "per GPU, each rendering its own scenes)"}))
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

    if args.workload == "c4":
        print(json.dumps(run_views(args)), flush=True)
        return
    if args.workload == "c5":
        print(json.dumps(run_large(args)), flush=True)
        return

    W = H = 84
    B = args.scenes
    t0 = time.time()
    objects: list = []
    scenes, smspec = workloads.config3(n_scenes=B, seed=args.seed, width=W, height=H,
                                       objects=objects)
    tris = workloads.triangle_counts(scenes)
    log(f"generated {B} scenes ({tris.sum()} triangles, {tris.min()}..{tris.max()} per scene) "
        f"in {time.time() - t0:.1f}s")
    # planning the batch: the per-scene host planner, and the device planner
    # (jr_pack plus batched records) whose calls are rendered below; both
    # give the same calls bit for bit
    from jaxrenderer_amd import packing

    t0 = time.perf_counter()
    batch.plan_batch(scenes, W, H, smspec)
    plan_host_ms = (time.perf_counter() - t0) * 1e3
    packing.plan_batch_device(scenes[:8], W, H, smspec)  # first-call set-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    calls = packing.plan_batch_device(scenes, W, H, smspec)
    torch.cuda.synchronize()
    plan_dev_ms = (time.perf_counter() - t0) * 1e3
    log(f"planning {B} scenes: host {plan_host_ms:.0f} ms, device {plan_dev_ms:.1f} ms")

    def wh(s):
        w, h = (int(v) for v in s.lower().split("x"))
        return (w, h)

    auto = args.tile == "auto"
    tile = (8, 8) if auto else wh(args.tile)
    shadow_tile = wh(args.shadow_tile) if args.shadow_tile else tile
    br = batch.BatchRenderer(calls, tile=tile, shadow_tile=shadow_tile, persistent=True,
                             shadow_on_demand="auto" if not args.full_shadow_map else False)
    tuned = None
    if auto:
        t0 = time.time()
        tuned = br.autotune(use_cache=not args.no_autotune_cache)
        for tag, c in tuned.items():
            log(f"autotune {tag}: tile {c['tile']} form {c['form']} x{c['ppl']} "
                f"{c['ms']:.3f} ms ({'cached' if c['cached'] else 'timed'})")
            if not c["cached"]:
                log("  candidates: " + ", ".join(
                    f"{t['tile'][0]}x{t['tile'][1]}/{t['form']}/x{t['ppl']}={t['ms']:.3f}"
                    for t in sorted(c["timings"], key=lambda t: t["ms"])))
        log(f"autotune took {time.time() - t0:.1f}s")
    pairs = br.calibrate()
    br_on_demand = br.shadow_on_demand
    log(f"binned triangle-tile pairs: {pairs}; shadow texels on demand: {br_on_demand}")
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
    names = list(_hip.STAGE_NAMES)
    stages = {f"shadow.{n}": float(v) for n, v in zip(names, shadow_ms)}
    stages.update({f"main.{n}": float(v) for n, v in zip(names, main_ms)})
    dom = max(stages, key=stages.get)
    log("stage ms: " + ", ".join(f"{k}={v:.3f}" for k, v in stages.items()))

    # moving every object of the batch on the device (jr_pose): the world
    # positions and normals of all scenes for a new [n_objects, 4, 4] stack of
    # poses, beside the host merge_objects it replaces (timed on a prefix)
    pose = None
    if not args.no_pose:
        from jaxrenderer_amd.model import merge_objects
        from jaxrenderer_amd.pose import ScenePoser

        poser = ScenePoser(objects)
        T = torch.from_numpy(np.ascontiguousarray(poser.initial_transforms)).cuda()
        for _ in range(3):
            poser.update(br, T)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            poser.update(br, T)
        e1.record()
        torch.cuda.synchronize()
        pose_ms = e0.elapsed_time(e1) / args.steps
        n_host = min(B, 32)
        t0 = time.perf_counter()
        for s in objects[:n_host]:
            merge_objects(s)
        host_ms = (time.perf_counter() - t0) * 1e3 * B / n_host
        out = br.render()
        torch.cuda.synchronize()
        same = bool(torch.equal(out["canvas"], canvas_b))
        pose = {"objects": poser.n_objects, "vertices": poser.n_verts, "normals": poser.n_norms,
                "device_ms": round(pose_ms, 4), "host_merge_ms": round(host_ms, 1),
                "render_unchanged": same}
        log(f"pose update of {poser.n_objects} objects ({poser.n_verts} vertices): "
            f"{pose_ms:.3f} ms on the device vs {host_ms:.0f} ms of host merge_objects "
            f"(extrapolated from {n_host} scenes); render unchanged: {same}")


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

    # CPU leg: the oracle timed on a bounded sample of the same scenes (the
    # CPU baseline), and, untimed, the depth complexity the work model needs
    cpu = sample = None
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
        t1 = time.perf_counter()
        sample = complexity_sample(scenes, W, H, smspec, args.complexity_scenes)
        log(f"depth complexity of {sample['scenes']} scenes counted by the oracle in "
            f"{time.perf_counter() - t1:.1f}s: {sample}")

    # triangles that won at least one pixel of the main pass (their shading
    # inputs are read by the raster stage)
    main_pass = br.passes["main"]
    had_ids = main_pass.prim_id is not None
    if not had_ids:
        main_pass.prim_id = torch.empty(zbuf_b.shape, dtype=torch.int32, device=zbuf_b.device)
    br.render()
    pid = main_pass.prim_id
    n_visible = int(torch.unique(pid[pid >= 0]).numel())
    if not had_ids:
        main_pass.prim_id = None

    n_tri = int(tris.sum())
    px = {"main": B * W * H, "shadow": B * smspec.width * smspec.height}
    tiles = {tag: p.n_tiles for tag, p in br.passes.items()}

    # the implementation's own byte model per stage (what the kernels must
    # move as built: key buffers, records, lists), beside the algorithmic
    # bytes.  The shadow pass keeps 32-bit depth keys in its output buffer
    # (a depth pass with an offset and no id outputs); the main pass 64-bit
    # keys.
    key_b = {"shadow": 4, "main": 8}

    def impl_bytes(pass_: str, stage: str) -> int:
        if stage == "clear":  # key per pixel, tile counters
            return px[pass_] * key_b[pass_] + tiles[pass_] * 8
        if stage == "setup":  # 3 slot ids (+3 position indices in the main
            # pass) and 3 positions read, one 4 B binning word written per
            # triangle (records of the triangles that survive culling not
            # counted: a lower bound)
            return n_tri * (12 + (12 if pass_ == "main" else 0) + 36 + 4)
        if stage == "wide":  # the wide list and ranges of wide triangles
            return 0
        if stage == "scan":  # count read, offset written, per tile
            return tiles[pass_] * 8
        if stage == "bin":  # binning word per triangle, 4 B per binned pair
            return n_tri * 4 + pairs[pass_] * 4
        # raster: keys read, depth (+ colour) written, 52 B per binned pair
        # (id + 48 B record), and per visible triangle of the main pass its
        # record (64 B), face row (12 B) and three corners' indices and
        # attributes (4 x 4 B indices + position, normal, uv: 48 B each)
        out = px[pass_] * (4 if pass_ == "shadow" else 16)
        shade = n_visible * (64 + 12 + 3 * 48) if pass_ == "main" else 0
        return px[pass_] * key_b[pass_] + out + pairs[pass_] * 52 + shade

    # roofline: the algorithmic work of the work model over event-timed
    # durations, for the whole frame and for the dominant stage, against
    # both the HBM and the VALU peak
    wm = work_model(scenes, W, H, smspec, sample) if sample is not None else None
    dom_fr = frame_fr = None
    if wm is not None:
        dom_fr = fractions(wm["bytes"].get(dom, 0), wm["ops"].get(dom, 0), stages[dom])
        frame_fr = fractions(wm["bytes_total"], wm["ops_total"], t_batched * 1e3)
    traffic = valu_busy = prof_tile = None
    tpath = os.path.join(ROOT, "profiles", "config3_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            prof = json.load(f)
        measured = prof.get("kernels", {})
        if dom in measured:
            traffic = measured[dom]["hbm_bytes"]
            valu_busy = measured[dom].get("valu_busy")
            prof_tile = measured[dom].get("tile", prof.get("tile"))
            run_p = br.passes.get(dom.split(".")[0])
            if isinstance(prof_tile, dict) and run_p is not None:
                same = (list(prof_tile.get("tile", [])) == list(run_p.tile)
                        and prof_tile.get("persistent") == run_p.persistent
                        and prof_tile.get("tiles_per_group") == run_p.tiles_per_group
                        and prof_tile.get("pixels_per_lane") == run_p.pixels_per_lane)
                prof_tile = (f"profiled launch {prof_tile}: "
                             f"{'the same as' if same else 'differs from'} this run's")
    if dom_fr is not None:
        best = dom_fr[dom_fr["bound"]]
        roofline = {"bound": dom_fr["bound"], "achieved": best["achieved"], "peak": best["peak"],
                    "unit": best["unit"], "frac": best["frac"], "traffic": traffic}
    else:
        roofline = {"bound": None, "achieved": None, "peak": None, "unit": None, "frac": None,
                    "traffic": traffic}
    alg_b = wm["bytes"].get(dom, 0) if wm is not None else None
    roofline.update({
        "kernel": dom, "kernel_ms": round(stages[dom], 4),
        "kernel_fractions": dom_fr,
        "frame_fractions": frame_fr,
        "work": None if wm is None else {
            "bytes_by_stage": wm["bytes"], "ops_by_stage": {k: int(v) for k, v in wm["ops"].items()},
            "bytes_total": int(wm["bytes_total"]), "ops_total": int(wm["ops_total"]),
            "counts": wm["counts"], "complexity_sample": sample},
        # HBM bytes the dominant stage's kernel moved per launch, from a
        # rocprofv3 FETCH_SIZE / WRITE_SIZE run committed under profiles/
        # (not measured in this run); its launch shape is named beside it
        "traffic_source": None if traffic is None else
        f"profiles/config3_traffic.json ({prof_tile or 'launch shape not recorded'})",
        "traffic_over_alg_bytes": (round(traffic / alg_b, 2) if traffic and alg_b else None),
        "valu_busy_from_profile": valu_busy,
        "traffic_model": {k: impl_bytes(*k.split(".")) for k in stages},
        "visible_triangles": n_visible,
    })

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
            "workload": f"config 3: {B} Brax-like scenes (ground plane + 4..30 boxes/"
                        f"capsules/spheres, {int(tris.min())}..{int(tris.max())} triangles), "
                        "84x84, phong_reflection_shadow, 256x256 shadow map, batched forward "
                        "in one HIP-graph replay",
            "scenes": B,
            "triangles": n_tri,
            "width": W,
            "height": H,
            "shadow_map": [smspec.width, smspec.height],
            "tile": {tag: {"tile": list(p.tile), "persistent": p.persistent,
                           "tiles_per_group": p.tiles_per_group,
                           "pixels_per_lane": p.pixels_per_lane}
                     for tag, p in br.passes.items()},
            "autotuned": auto,
            "shadow_texels": "on demand (main raster computes the texels it reads from the "
                             "shadow binning)" if br_on_demand else "full map",
        },
        "mpix_per_s": round(fps * W * H / 1e6, 2),
        "per_scene_fps": None if per_scene_fps is None else round(per_scene_fps, 1),
        "batched_over_per_scene": None if per_scene_fps is None else round(fps / per_scene_fps, 2),
        "batched_vs_per_scene_mismatched_scenes": mismatches,
        "fwd_bwd_fps": None if fwd_bwd_fps is None else round(fwd_bwd_fps, 1),
        "fwd_bwd_over_fwd": None if fwd_bwd_fps is None else round(fwd_bwd_fps / fps, 3),
        "fwd_bwd_grads_finite": grads_finite,
        "stage_ms": {k: round(v, 4) for k, v in stages.items()},
        "pose_update": pose,
        "planning_ms": {"host_per_scene": round(plan_host_ms, 1),
                        "device": round(plan_dev_ms, 2)},
        "binned_pairs": pairs,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "other_configs": None,
    }
    if not args.no_extra_configs:
        import gc

        del br, out, main_pass, pid
        gc.collect()
        torch.cuda.empty_cache()
        sub = argparse.Namespace(**{**vars(args), "c5_per_scene": min(args.c5_per_scene, 16)})
        res["other_configs"] = other_configs(sub, cpu_sample=not args.no_cpu_baseline)
    print(json.dumps(res), flush=True)


def _timed(fn, steps: int, warmup: int) -> float:
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def single_scene(args, scene, width: int, height: int, cpu_sample: bool) -> dict:
    """One scene's plain forward (phong_reflection, no shadow) replayed from a
    HIP graph, and the CPU oracle on the same frame."""
    from jaxrenderer_amd import batch

    calls = batch.plan_batch([scene], width, height)
    br = batch.BatchRenderer(calls, tile=(8, 8), persistent=False)
    if args.tile == "auto":
        br.autotune(use_cache=not args.no_autotune_cache)
    br.calibrate()
    br.capture()
    t = _timed(br.graph.replay, args.steps, args.warmup)
    d = {"triangles": int(len(scene.model.faces)), "width": width, "height": height,
         "fwd_ms": round(t * 1e3, 4), "frames_per_s": round(1.0 / t, 1),
         "mpix_per_s": round(width * height / t / 1e6, 2),
         "tile": list(br.passes["main"].tile)}
    if cpu_sample:
        from oracle.oracle import run_oracle_plan

        t1 = time.perf_counter()
        run_oracle_plan(calls, prim_id=False, bary=False)
        tc = time.perf_counter() - t1
        d["cpu_oracle_frames_per_s"] = round(1.0 / tc, 3)
    return d


def ant_leg(args, n: int = 200) -> dict:
    """The reference's published batched figure is for its Brax ant scene
    (18 objects, 3,276 triangles, get_camera_image with a canvas-sized shadow
    map, 84x84, N=200 frames per call: 888.6 frames/s on an A100,
    notebooks/84x84/A100.ipynb).  Its rollout ships only as pickles, so the
    frames here are synthetic ant-like walkers of the same composition
    (workloads.ant_like), rendered batched in one HIP-graph replay."""
    from jaxrenderer_amd import batch, workloads

    scenes, spec = workloads.ant_like(n_scenes=n, seed=args.seed)
    br = batch.BatchRenderer(batch.plan_batch(scenes, 84, 84, spec), persistent=False,
                             shadow_on_demand="auto")
    if args.tile == "auto":
        br.autotune(use_cache=not args.no_autotune_cache)
    br.calibrate()
    br.capture()
    t = _timed(br.graph.replay, args.steps, args.warmup)
    return {"scenes": n, "triangles_per_scene": 3276, "width": 84, "height": 84,
            "shadow_map": [84, 84], "ms_per_call": round(t * 1e3, 4),
            "frames_per_s": round(n / t, 1),
            "published_reference": {"frames_per_s": 888.6, "hardware": "NVIDIA A100-SXM 40GB",
                                    "source": "notebooks/84x84/A100.ipynb:1308",
                                    "scene": "Brax ant rollout (pickled inputs)"},
            "data": "synthetic ant-like frames of the same composition (not the reference's "
                    "rollout)"}


def other_configs(args, cpu_sample: bool) -> dict:
    """Short summaries of the other workloads, measured in the same run:
    config 1 (the cube, 320x240) and config 2 (the 9k-triangle sphere,
    960x540) single-scene forwards; config 4 (64 views, forward and forward +
    backward per shader); config 5 (512 scenes at 480x270, forward, forward +
    backward, peak HBM, per-scene forward + backward on a prefix)."""
    import gc

    import torch

    from jaxrenderer_amd import workloads

    out = {}
    t0 = time.time()
    out["config1"] = single_scene(args, workloads.config1(), 320, 240, cpu_sample)
    out["config2"] = single_scene(args, workloads.config2(seed=args.seed), 960, 540, cpu_sample)
    log(f"configs 1 and 2: {out['config1']['fwd_ms']:.4f} / {out['config2']['fwd_ms']:.4f} ms "
        f"per frame ({time.time() - t0:.1f}s)")
    t0 = time.time()
    out["ant_like"] = ant_leg(args)
    log(f"ant-like frames: {out['ant_like']['frames_per_s']:.0f} frames/s "
        f"({time.time() - t0:.1f}s)")
    t0 = time.time()
    c4 = run_views(args)
    out["config4"] = {k: {kk: v[kk] for kk in ("fwd_ms", "fwd_bwd_ms", "fwd_bwd_over_fwd",
                                                "fwd_views_per_s", "fwd_bwd_views_per_s",
                                                "grads_finite")}
                      for k, v in c4["shaders"].items()}
    log(f"config 4 summary in {time.time() - t0:.1f}s")
    gc.collect()
    torch.cuda.empty_cache()
    t0 = time.time()
    c5 = run_large(args)
    out["config5"] = {k: c5[k] for k in ("value", "unit", "ms_per_step", "fwd_fps", "fwd_ms",
                                         "fwd_bwd_over_fwd", "peak_hbm_bytes",
                                         "per_scene_fwd_bwd_fps", "batched_over_per_scene",
                                         "grads_finite", "shadow_texels_on_demand")}
    out["config5"]["triangles"] = c5["config"]["triangles"]
    log(f"config 5 summary in {time.time() - t0:.1f}s")
    return out


def run_views(args) -> dict:
    """64 ring views of a deformed 5000-triangle sphere at 256x256 (512x512
    texture), gouraud_texture and phong: batched forward, and forward + L2
    loss gradient + backward to vertices, normals, texture and the per-view
    parameters, each in one HIP graph."""
    import numpy as np
    import torch

    from jaxrenderer_amd import batch, workloads
    from jaxrenderer_amd.autograd import build_views_call
    from jaxrenderer_amd.shaders import (GouraudTextureExtraInput, GouraudTextureShader,
                                         PhongTextureExtraInput, PhongTextureShader)
    from jaxrenderer_amd.types import LightSource

    f32 = np.float32
    V, S = 64, 256
    cams, faces, v, n, uv, col, tex = workloads.config4(seed=args.seed, n_views=V, size=S)
    _, _, vt, _, _, _, text = workloads.config4(seed=args.seed, n_views=V, size=S,
                                                perturbed=True)
    light = LightSource(direction=np.array((0.3, -0.5, 0.8), f32), colour=np.ones(3, f32))
    lines = {}
    for shader, X in ((GouraudTextureShader, GouraudTextureExtraInput),
                      (PhongTextureShader, PhongTextureExtraInput)):
        extra = X(position=v, normal=n, uv=uv, light=light, texture=tex)
        target_extra = X(position=vt, normal=n, uv=uv, light=light, texture=text)
        call = build_views_call(cams, shader, faces, extra, S, S, init_buffers=True)
        tcall = build_views_call(cams, shader, faces, target_extra, S, S, init_buffers=True)
        br = batch.BatchRenderer([call], tile=(8, 8), persistent=False)
        if args.tile == "auto":
            br.autotune(use_cache=not args.no_autotune_cache)
        br.calibrate()
        target = batch.BatchRenderer([tcall], tile=(8, 8), persistent=False).render()["canvas"]
        target = target.clone()
        br.capture()
        t_f = _timed(br.graph.replay, args.steps, args.warmup)
        br.enable_backward()
        br.capture_train(target)
        t_fb = _timed(br.train_graph.replay, args.steps, args.warmup)
        finite = bool(all(torch.isfinite(g).all() for g in br.grads.values()))
        name = shader.__name__
        lines[name] = {"fwd_views_per_s": round(V / t_f, 1), "fwd_ms": round(t_f * 1e3, 4),
                       "fwd_bwd_views_per_s": round(V / t_fb, 1),
                       "fwd_bwd_ms": round(t_fb * 1e3, 4),
                       "fwd_bwd_over_fwd": round(t_f / t_fb, 3), "grads_finite": finite,
                       "tile": list(br.passes["main"].tile),
                       "tiles_per_group": br.passes["main"].tiles_per_group}
        log(f"{name}: forward {t_f * 1e3:.3f} ms, forward+backward {t_fb * 1e3:.3f} ms "
            f"for {V} views")
    first = lines["GouraudTextureShader"]
    return ({
        "metric": "views/s, 64 views 256x256 forward+backward (differentiation workload), 1 GPU",
        "value": first["fwd_bwd_views_per_s"], "unit": "views/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": first["fwd_bwd_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "config 4: deformed UV sphere (5000 triangles), 64 ring views, "
                               "256x256, 512x512x3 texture, L2 loss against a perturbed "
                               "target (vertex sigma 0.02, texture sigma 0.1)",
                   "views": V, "width": S, "height": S, "triangles": int(len(faces))},
        "shaders": lines})


def run_large(args) -> dict:
    """512 scenes of 1k..20k triangles at 480x270, phong_reflection_shadow with
    512x512 shadow maps: batched forward, batched forward + backward (only the
    primitive ids are kept; fragments are recomputed), peak HBM, and the plain
    per-scene forward + backward."""
    import torch

    from jaxrenderer_amd import batch, workloads

    B = min(args.scenes, 512)
    W, H = 480, 270
    t0 = time.time()
    scenes, smspec = workloads.config5(n_scenes=B, seed=args.seed, width=W, height=H)
    tris = workloads.triangle_counts(scenes)
    log(f"generated {B} scenes ({tris.sum()} triangles, {tris.min()}..{tris.max()}) in "
        f"{time.time() - t0:.1f}s")
    calls = batch.plan_batch(scenes, W, H, smspec)
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    br = batch.BatchRenderer(calls, persistent=False,
                             shadow_on_demand="auto" if not args.full_shadow_map else False)
    if args.tile == "auto":
        for tag, c in br.autotune(use_cache=not args.no_autotune_cache).items():
            log(f"autotune {tag}: tile {c['tile']} form {c['form']} x{c['ppl']} "
                f"{c['ms']:.3f} ms")
            if not c["cached"]:
                log("  candidates: " + ", ".join(
                    f"{t['tile'][0]}x{t['tile'][1]}/{t['form']}/x{t['ppl']}={t['ms']:.3f}"
                    for t in sorted(c["timings"], key=lambda t: t["ms"])))
    pairs = br.calibrate()
    br.capture()
    t_f = _timed(br.graph.replay, args.steps, args.warmup)
    out = br.render()
    gen = torch.Generator(device="cuda").manual_seed(1)
    target = torch.rand(out["canvas"].shape, generator=gen, device="cuda")
    br.enable_backward()
    br.capture_train(target)
    t_fb = _timed(br.train_graph.replay, args.steps, args.warmup)
    # device memory this workload added at its peak (what was allocated
    # before it, e.g. by an earlier leg of the same run, is not counted)
    peak = torch.cuda.max_memory_allocated() - base
    finite = bool(all(torch.isfinite(g).all() for g in br.grads.values()))
    log(f"forward {t_f * 1e3:.2f} ms, forward+backward {t_fb * 1e3:.2f} ms, peak HBM "
        f"{peak / 2**30:.2f} GiB")
    per_scene = None
    if not args.no_per_scene:
        n_single = min(B, args.c5_per_scene)
        singles = []
        for i in range(n_single):
            r = batch.BatchRenderer(batch.plan_batch([scenes[i]], W, H, smspec),
                                    persistent=False)
            r.calibrate()
            r.enable_backward()
            singles.append(r)
        tg = [target[i:i + 1] for i in range(n_single)]

        def step():
            for r, t in zip(singles, tg):
                r.train_step(t)

        t_s = _timed(step, 1, 1)
        per_scene = n_single / t_s
        log(f"per-scene forward+backward: {per_scene:.1f} frames/s over {n_single} scenes")
    return ({
        "metric": "frames/s at 480x270 batch 512 (fwd+bwd) vs per-scene launch, 1 GPU",
        "value": round(B / t_fb, 1), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_fb * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "config 5: 512 Brax-like scenes of 1k..20k triangles, 480x270, "
                               "phong_reflection_shadow, 512x512 shadow maps, forward + L2 "
                               "backward in one HIP graph",
                   "scenes": B, "triangles": int(tris.sum()), "width": W, "height": H},
        "fwd_fps": round(B / t_f, 1), "fwd_ms": round(t_f * 1e3, 4),
        "fwd_bwd_over_fwd": round(t_f / t_fb, 3), "peak_hbm_bytes": int(peak),
        "per_scene_fwd_bwd_fps": None if per_scene is None else round(per_scene, 1),
        "batched_over_per_scene": None if per_scene is None else round(B / t_fb / per_scene, 2),
        "grads_finite": finite, "binned_pairs": pairs,
        "shadow_texels_on_demand": br.shadow_on_demand})


if __name__ == "__main__":
    main()
