"""A rasterisation call described independently of where it runs.

`RasterCall` carries everything jr_forward needs (include/jrhip.h): shader,
frame size, per-scene uniforms and the geometry/map arrays.  The renderer
builds calls on the host and `run` executes them on the GPU through
libjrhip.so.  Keeping the description separate lets the test suite hand the
very same call to the CPU oracle.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _hip, _raster

_FLOAT_INPUTS = {"positions", "normals", "uvs", "colours", "textures", "specular_maps",
                 "normal_maps", "shadow_maps", "clip"}


def check_index(idx, n: int, name: str) -> None:
    """Raises ValueError unless every entry of the index array `idx` lies in
    [0, n).  The kernels gather attr[idx] without a clamp, so an index out of
    range would read past the array (or another scene's rows); the reference's
    Model.value_checks (renderer/model.py:62-78) rejects the same inputs.
    Device tensors are reduced on the device (one synchronisation)."""
    if idx is None:
        return
    if isinstance(idx, torch.Tensor):
        if idx.numel() == 0:
            return
        lo, hi = int(idx.min()), int(idx.max())
    else:
        a = np.asarray(idx)
        if a.size == 0:
            return
        lo, hi = int(a.min()), int(a.max())
    if lo < 0 or hi >= n:
        raise ValueError(f"{name} out of bound, expected [0, {n}), got [{lo}, {hi}]")


def check_rows(arr, n: int, name: str) -> None:
    """Raises ValueError when the per-slot array `arr` has fewer than n rows."""
    if arr is None:
        return
    rows = int(arr.shape[0]) if getattr(arr, "ndim", 1) > 0 else 0
    if rows < n:
        raise ValueError(f"{name} has {rows} rows; the faces index {n} vertex slots")


@dataclass
class RasterCall:
    shader: int
    width: int
    height: int
    scenes: List[_hip.JrScene]
    inputs: Dict[str, object]  # numpy arrays or device tensors
    n_slots: int
    n_tris: int
    shadow_shape: Tuple[int, int] = (0, 0)
    init_buffers: bool = True
    zbuffer_default: float = 1.0
    colour_default: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    apply_offset: bool = False
    depth_offset: float = 0.0
    want_canvas: bool = True
    # name of an output of an earlier call that feeds inputs["shadow_maps"]
    shadow_from: Optional[str] = None
    tag: str = "main"
    extras: Dict[str, object] = field(default_factory=dict)

    @property
    def n_scenes(self) -> int:
        return len(self.scenes)


def to_device(x, dev: torch.device, name: str) -> Optional[torch.Tensor]:
    if x is None:
        return None
    dtype = torch.float32 if name in _FLOAT_INPUTS else torch.int32
    if isinstance(x, torch.Tensor):
        t = x.to(device=dev, dtype=dtype)
    else:
        t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32 if dtype == torch.float32
                                                  else np.int32)).to(dev)
    return t.contiguous()


def device_inputs(call: RasterCall, dev: torch.device) -> Dict[str, Optional[torch.Tensor]]:
    """The call's inputs as device tensors.  A call whose face table is the
    identity of a per-corner flattening (extras["identity_faces"]) passes no
    table: the library then takes slots 3t, 3t+1, 3t+2 for row t without
    reading them."""
    inputs = {k: to_device(v, dev, k) for k, v in call.inputs.items()}
    if call.extras.get("identity_faces"):
        inputs["faces"] = None
    return inputs


def run(call: RasterCall, *, zbuffer: Optional[torch.Tensor] = None,
        canvas: Optional[torch.Tensor] = None, prim_id: bool = False, bary: bool = False,
        display: bool = False, tile=_raster.DEFAULT_TILE, persistent: bool = False,
        tiles_per_group: int = 1, pixels_per_lane: int = 1,
        capacity: Optional[int] = None, noalloc: bool = False) -> Dict[str, torch.Tensor]:
    """Executes `call` on the GPU.  zbuffer/canvas given here are updated in
    place (merge semantics when call.init_buffers is False).  `noalloc`: run
    through jr_forward_noalloc (the library's scratch as workspace)."""
    dev = _raster.require_device()
    B, W, H = call.n_scenes, call.width, call.height
    if zbuffer is None:
        zbuffer = torch.empty((B, W, H), dtype=torch.float32, device=dev)
        if not call.init_buffers:
            zbuffer.fill_(call.zbuffer_default)
    if canvas is None and call.want_canvas:
        canvas = torch.empty((B, W, H, 3), dtype=torch.float32, device=dev)
        if not call.init_buffers:
            canvas[:] = torch.tensor(call.colour_default, dtype=torch.float32, device=dev)
    out = {"zbuffer": zbuffer}
    if canvas is not None:
        out["canvas"] = canvas
    pid = torch.empty((B, W, H), dtype=torch.int32, device=dev) if prim_id else None
    bar = torch.empty((B, W, H, 3), dtype=torch.float32, device=dev) if bary else None
    if pid is not None:
        out["prim_id"] = pid
    if bar is not None:
        out["bary"] = bar
    disp = (torch.empty((B, H, W, 3), dtype=torch.uint8, device=dev)
            if display and canvas is not None else None)
    if disp is not None:
        out["display"] = disp
    inputs = device_inputs(call, dev)
    scenes = _raster.scenes_to_device(call.scenes, dev)
    _raster.forward(
        shader=call.shader, width=W, height=H, scenes=scenes, n_scenes=B, inputs=inputs,
        n_slots=call.n_slots, n_tris=call.n_tris, zbuffer=zbuffer, canvas=canvas,
        prim_id=pid, bary=bar, display=disp, shadow_shape=call.shadow_shape,
        init_buffers=call.init_buffers, zbuffer_default=call.zbuffer_default,
        colour_default=call.colour_default, apply_offset=call.apply_offset,
        depth_offset=call.depth_offset, tile=tile, persistent=persistent,
        tiles_per_group=tiles_per_group, pixels_per_lane=pixels_per_lane, capacity=capacity,
        noalloc=noalloc,
    )
    return out


def run_plan(calls: List[RasterCall], **kw) -> Dict[str, Dict[str, torch.Tensor]]:
    """Runs dependent calls in order (a shadow pass feeding a main pass)."""
    results: Dict[str, Dict[str, torch.Tensor]] = {}
    for c in calls:
        if c.shadow_from is not None:
            c.inputs["shadow_maps"] = results[c.shadow_from]["zbuffer"]
        results[c.tag] = run(c, **kw)
    return results
