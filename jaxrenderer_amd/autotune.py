"""Launch-shape autotuner for batched rendering.

For every pass of a BatchRenderer (shadow, main) it times the candidate
launch shapes of `jr_forward` on the live data: the screen-tile shape (which
is also the workgroup shape: one lane per pixel of the tile), the persistent
work-queue grid against one workgroup per tile, and, for each, the bin
capacity sized from the binned triangle-tile pairs that shape produces.
Every candidate must reproduce the first candidate's depth buffer bit for bit
(the binning is exact, so the tile shape cannot change the image); a
mismatch raises.

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
DEFAULT_CANDIDATES: Tuple[Tuple[Tuple[int, int], bool], ...] = tuple(
    (t, p) for t in TILES for p in (True, False))
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


def pass_key(tag: str, call, device: torch.device) -> str:
    # the architecture and CU count identify the chip even where the marketing
    # name is unavailable (it reads empty under some profilers)
    props = torch.cuda.get_device_properties(device)
    arch = getattr(props, "gcnArchName", "") or torch.cuda.get_device_name(device)
    name = f"{arch.split(':')[0]}/{props.multi_processor_count}cu"
    tris = 1 << max(0, math.ceil(math.log2(max(call.n_tris, 1))))
    return (f"{name}|{tag}|shader{call.shader}|{call.width}x{call.height}|"
            f"scenes{call.n_scenes}|tris<={tris}")


def _time_launches(p, reps: int) -> float:
    """Milliseconds per launch of pass `p`, events on the launching stream."""
    stream = torch.cuda.current_stream()
    p.launch()
    best = float("inf")
    for _ in range(3):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(reps):
            p.launch()
        b.record(stream)
        b.synchronize()
        best = min(best, a.elapsed_time(b) / reps)
    return best


def tune_pass(p, candidates: Sequence[Tuple[Tuple[int, int], bool]] = DEFAULT_CANDIDATES,
              reps: int = 5) -> Dict[str, object]:
    """Times each (tile, persistent) candidate on pass `p` (a batch pass whose
    inputs are ready, including any shadow map it reads) and leaves `p` set to
    the fastest.  Returns the per-candidate timings and the choice."""
    timings: List[dict] = []
    ref: Optional[torch.Tensor] = None
    for tile, persistent in candidates:
        p.tile, p.persistent = tuple(tile), bool(persistent)
        p.set_capacity(_raster.default_capacity(p.call.n_tris, p.n_tiles))
        p.launch()
        torch.cuda.synchronize()
        pairs = p.pairs()
        p.set_capacity(int(pairs * SLACK) + 1024)
        ms = _time_launches(p, reps)
        z = p.zbuffer.view(torch.int32)
        if ref is None:
            ref = z.clone()
        elif not torch.equal(ref, z):
            raise RuntimeError(f"tile {tile} persistent={persistent} changed the depth buffer")
        timings.append({"tile": list(tile), "persistent": bool(persistent), "ms": ms,
                        "pairs": int(pairs)})
    best = min(timings, key=lambda t: t["ms"])
    p.tile, p.persistent = tuple(best["tile"]), best["persistent"]
    p.set_capacity(int(best["pairs"] * SLACK) + 1024)
    return {"tile": best["tile"], "persistent": best["persistent"], "ms": best["ms"],
            "timings": timings}


def tune(renderer, candidates: Optional[Iterable] = None, reps: int = 5,
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
        key = pass_key(tag, p.call, p.zbuffer.device)
        hit = cache.get(key)
        if hit is not None and [tuple(hit["tile"]), hit["persistent"]] in \
                [[tuple(t), q] for t, q in cands]:
            got = tune_pass(p, [(tuple(hit["tile"]), hit["persistent"])], reps=1)
            got["cached"] = True
        else:
            got = tune_pass(p, cands, reps=reps)
            got["cached"] = False
            cache[key] = {"tile": got["tile"], "persistent": got["persistent"],
                          "ms": got["ms"]}
            dirty = True
        chosen[tag] = got
    if dirty and use_cache:
        _store(path, cache)
    renderer.graph = None
    return chosen


__all__ = ["DEFAULT_CANDIDATES", "TILES", "cache_path", "pass_key", "tune", "tune_pass"]
