"""Launch-shape autotuner for batched rendering.

For every pass of a BatchRenderer (shadow, main) it times the candidate
launch shapes of `jr_forward` on the live data: the screen-tile shape (which
also sets the workgroup: one lane per pixel, or per 4 pixels for the
256-pixel tiles), the launch form of the raster stage (a persistent
work-queue grid, or a grid of workgroups taking 1, 2 or 4 consecutive tiles
each: one after another, or, at 4 pixels per lane, one wavefront per tile side
by side), and, for each, the bin capacity sized from the binned
triangle-tile pairs that shape produces.
Every candidate must reproduce the first candidate's image bit for bit (the
binning is exact, so the tile shape cannot change it): the pass's depth
buffer, or, tuned inside a renderer, the main pass's depth and colour; a
mismatch raises.  With shadows on demand the main raster's cost depends on
the shadow pass's tiles, so candidates are timed by the whole render.

Results are cached in a JSON file keyed by device, pass, shader, frame size,
scene count and the triangle count rounded to a power of two; the path is
`$JR_AUTOTUNE_CACHE`, else `~/.cache/jaxrenderer_amd/autotune.json`.  The
bin capacity is not cached: it is recalibrated from the data on every tune.
"""

from __future__ import annotations

import json
import math
import os
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import torch

from . import _raster

TILES: Tuple[Tuple[int, int], ...] = ((8, 8), (4, 16), (16, 16), (8, 32))
# launch forms: 0 = persistent work queue, n > 0 = n tiles per workgroup
FORMS: Tuple[int, ...] = (0, 1, 2, 4)
# pixels per raster thread: 4 only for the 256-pixel tiles (one wavefront each)
PIXELS_PER_LANE: Tuple[int, ...] = (1, 4)
DEFAULT_CANDIDATES: Tuple[Tuple[Tuple[int, int], int, int], ...] = tuple(
    (t, m, ppl) for t in TILES for m in FORMS for ppl in PIXELS_PER_LANE
    if ppl == 1 or t[0] * t[1] == 256)
SLACK = 1.25


def cache_path() -> str:
    return os.environ.get("JR_AUTOTUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "jaxrenderer_amd", "autotune.json")


def _load(path: str) -> Dict[str, dict]:
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path: str, d: Dict[str, dict]) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def pass_key(tag: str, call, device: torch.device, on_demand: bool = False) -> str:
    # the architecture and CU count identify the chip even where the marketing
    # name is unavailable (it reads empty under some profilers)
    props = torch.cuda.get_device_properties(device)
    arch = getattr(props, "gcnArchName", "") or torch.cuda.get_device_name(device)
    name = f"{arch.split(':')[0]}/{props.multi_processor_count}cu"
    tris = 1 << max(0, math.ceil(math.log2(max(call.n_tris, 1))))
    return (f"{name}|{tag}|shader{call.shader}|{call.width}x{call.height}|"
            f"scenes{call.n_scenes}|tris<={tris}" + ("|shadow-on-demand" if on_demand else ""))


def _time_launches(p, reps: int, renderer=None) -> float:
    """Milliseconds per launch of pass `p`: the sum over its stages of the
    median stage time over `reps` launches, from the HIP events the library
    records on the launching stream (medians per stage keep one slow launch
    of one stage from deciding the choice).  With `renderer`, every launch is
    a whole render() -- all passes in order, as a replay runs them -- and only
    pass `p`'s stages are counted, or every pass's with shadows on demand."""
    import ctypes

    import numpy as np

    from . import _hip

    L = _hip.lib()
    passes = list(renderer.passes.values()) if renderer is not None else [p]
    idx = next(i for i, q in enumerate(passes) if q is p)
    run = renderer.render if renderer is not None else p.launch
    run()
    torch.cuda.synchronize()
    L.jr_stage_times(None, 0)  # forget earlier records
    L.jr_stage_timing(1)
    try:
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
    finally:
        L.jr_stage_timing(0)
    k, P = _hip.STAGES, len(passes)
    buf = (ctypes.c_float * (k * reps * P))()
    n = L.jr_stage_times(buf, reps * P)
    st = np.array(buf[: k * n], dtype=np.float64).reshape(n // P, P, k)
    if renderer is not None and getattr(renderer, "shadow_on_demand", False):
        return float(np.median(st, axis=0).sum())
    return float(np.median(st[:, idx, :], axis=0).sum())


def _apply(p, tile, form: int, ppl: int) -> None:
    p.tile = tuple(tile)
    p.persistent = form == 0
    p.tiles_per_group = max(1, int(form))
    p.pixels_per_lane = int(ppl)


def tune_pass(p, candidates: Sequence = DEFAULT_CANDIDATES, reps: int = 9,
              renderer=None) -> Dict[str, object]:
    """Times each (tile, form, pixels per lane) candidate on pass `p` (a
    batch pass whose inputs are ready, including any shadow map it reads; with
    `renderer`, inside whole renders) and leaves `p` set to the fastest.
    Returns the per-candidate timings and the choice."""
    timings: List[dict] = []
    ref: Optional[List[torch.Tensor]] = None
    # a pass inside a renderer is launched by whole renders: a main pass that
    # reads the shadow pass's binning needs that binning to run first
    run = renderer.render if renderer is not None else p.launch
    for tile, form, ppl in candidates:
        _apply(p, tile, form, ppl)
        p.set_capacity(_raster.default_capacity(p.call.n_tris, p.n_tiles))
        run()
        torch.cuda.synchronize()
        pairs = p.pairs()
        p.set_capacity(int(pairs * SLACK) + 1024)
        run()
        torch.cuda.synchronize()
        if renderer is not None:
            main = renderer.passes["main"]
            img = [main.zbuffer.view(torch.int32)]
            if main.canvas is not None:
                img.append(main.canvas.view(torch.int32))
        else:
            img = [p.zbuffer.view(torch.int32)]
        if ref is None:
            ref = [x.clone() for x in img]
        elif not all(torch.equal(a, b) for a, b in zip(ref, img)):
            raise RuntimeError(f"tile {tile} form {form} x{ppl} changed the image")
        ms = _time_launches(p, reps, renderer)
        timings.append({"tile": list(tile), "form": int(form), "ppl": int(ppl), "ms": ms,
                        "pairs": int(pairs)})
    best = min(timings, key=lambda t: t["ms"])
    _apply(p, best["tile"], best["form"], best["ppl"])
    p.set_capacity(int(best["pairs"] * SLACK) + 1024)
    return {"tile": best["tile"], "form": best["form"], "ppl": best["ppl"],
            "persistent": best["form"] == 0, "tiles_per_group": max(1, best["form"]),
            "ms": best["ms"], "timings": timings}


def tune(renderer, candidates: Optional[Iterable] = None, reps: int = 9,
         use_cache: bool = True, path: Optional[str] = None) -> Dict[str, dict]:
    """Picks tile shape, launch form and bin capacity for every pass of a
    BatchRenderer.  Cached choices are applied without timing (the capacity is
    still calibrated).  Returns {pass tag: choice}."""
    cands = tuple(candidates) if candidates is not None else DEFAULT_CANDIDATES
    path = path or cache_path()
    cache = _load(path) if use_cache else {}
    renderer.render()  # the main pass may read the shadow pass's output
    torch.cuda.synchronize()
    chosen: Dict[str, dict] = {}
    dirty = False
    for tag, p in renderer.passes.items():
        key = pass_key(tag, p.call, p.zbuffer.device,
                       getattr(renderer, "shadow_on_demand", False))
        hit = cache.get(key)
        entry = None if hit is None or "ppl" not in hit else (
            tuple(hit["tile"]), hit["form"], hit["ppl"])
        if entry is not None and entry in [(tuple(t), m, q) for t, m, q in cands]:
            got = tune_pass(p, [entry], reps=1, renderer=renderer)
            got["cached"] = True
        else:
            got = tune_pass(p, cands, reps=reps, renderer=renderer)
            got["cached"] = False
            cache[key] = {"tile": got["tile"], "form": got["form"], "ppl": got["ppl"],
                          "ms": got["ms"]}
            dirty = True
        chosen[tag] = got
    if dirty and use_cache:
        _store(path, cache)
    renderer.graph = None
    return chosen


__all__ = ["DEFAULT_CANDIDATES", "FORMS", "PIXELS_PER_LANE", "TILES", "cache_path",
           "pass_key", "tune", "tune_pass"]
