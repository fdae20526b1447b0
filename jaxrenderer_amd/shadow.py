"""Shadow maps (renderer/shadow.py)."""

from __future__ import annotations

from typing import Any, NamedTuple

import numpy as np
import torch

from . import _hip
from .geometry import Camera, asf32


class Shadow(NamedTuple):
    """renderer/shadow.py:27-38."""

    shadow_map: Any  # [W, H] depth from the light
    strength: Any  # [3]
    camera: Camera

    @staticmethod
    def camera_for(light_direction, viewport_matrix, centre, up,
                   distance: float = 10.0) -> Camera:
        """Orthographic light camera of renderer/shadow.py:84-103: eye at
        centre + light_direction * distance (direction not normalised)."""
        eye = (asf32(centre) + (asf32(light_direction) * np.float32(distance)).astype(np.float32)).astype(np.float32)
        view = Camera.view_matrix(eye=eye, centre=centre, up=up)
        proj = Camera.orthographic_projection_matrix(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
        return Camera.create(view=view, projection=proj, viewport=viewport_matrix)

    @staticmethod
    def render_shadow_map(shadow_map, verts, faces, light_direction, viewport_matrix, centre,
                          up, strength, offset: float = 0.001, distance: float = 10.0,
                          loop_unroll: int = 1) -> "Shadow":
        """renderer/shadow.py:40-125: DepthShader pass from the light, merged
        into `shadow_map`, then `+ offset` everywhere."""
        from .pipeline import build_call, _device_copy
        from . import _plan, _raster
        from .shaders.depth import DepthExtraInput, DepthShader

        dev = _raster.require_device()
        cam = Shadow.camera_for(light_direction, viewport_matrix, centre, up, distance)
        W, H = int(shadow_map.shape[0]), int(shadow_map.shape[1])
        call = build_call(cam, DepthShader, faces, DepthExtraInput(position=verts), W, H,
                          apply_offset=True, depth_offset=float(offset))
        sm = _device_copy(shadow_map, dev)
        _plan.run(call, zbuffer=sm.view(1, W, H))
        return Shadow(shadow_map=sm, strength=strength, camera=cam)

    def get(self, position) -> torch.Tensor:
        """renderer/shadow.py:127-153: depth at round(position) (half away from
        zero); negative indices count from the end; out of range -> +inf."""
        sm = self.shadow_map
        if not isinstance(sm, torch.Tensor):
            sm = torch.as_tensor(np.asarray(sm))
        pos = torch.as_tensor(position, device=sm.device)
        r = torch.sign(pos[..., :2]) * torch.floor(torch.abs(pos[..., :2]) + 0.5)
        # XLA's float -> int32 convert: NaN -> 0, saturating at the int32 range
        r = torch.nan_to_num(r, nan=0.0).clamp(-2147483648.0, 2147483647.0)
        ix = r[..., 0].to(torch.int64)
        iy = r[..., 1].to(torch.int64)
        W, H = sm.shape[0], sm.shape[1]
        ix = torch.where(ix < 0, ix + W, ix)
        iy = torch.where(iy < 0, iy + H, iy)
        ok = (ix >= 0) & (ix < W) & (iy >= 0) & (iy < H)
        val = sm[ix.clamp(0, W - 1), iy.clamp(0, H - 1)]
        return torch.where(ok, val, torch.full_like(val, float("inf")))


__all__ = ["Shadow"]
