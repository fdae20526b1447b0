"""`render`: the pipeline entry point (renderer/pipeline.py:462-537).

For the built-in shaders the whole pipeline -- vertex stage, per-primitive
setup, visibility, interpolation, fragment, mix and merge -- runs as the fused
HIP pass of libjrhip.so.  Other Shader subclasses run their own hooks in torch
around the HIP visibility pass (see `_custom`).
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from . import _hip, _plan, _raster
from ._plan import RasterCall
from .geometry import Camera
from .shader import Shader
from .shaders import BUILTIN
from .types import Buffers


def _arr(x):
    """Device tensors pass through; everything else becomes a numpy array."""
    if isinstance(x, torch.Tensor):
        return x
    return np.asarray(x)


def _host(x) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _vec3(x) -> np.ndarray:
    return np.asarray(_host(x), dtype=np.float32).reshape(-1)


def _map_shape(m) -> Tuple[int, int, int]:
    shp = tuple(m.shape)
    if len(shp) == 2:
        return (shp[0], shp[1], 1)
    if shp[2] not in (1, 3):
        raise NotImplementedError("texture maps must have 1 or 3 channels")
    return shp  # type: ignore[return-value]


def build_call(camera: Camera, shader: type, face_indices, extra, width: int, height: int, *,
               init_buffers: bool = False, zbuffer_default: float = 1.0,
               colour_default=(1.0, 1.0, 1.0), apply_offset: bool = False,
               depth_offset: float = 0.0) -> RasterCall:
    """Describes `render(camera, shader, ..., face_indices, extra)` for a
    built-in shader as one jr_forward call."""
    sid = shader._jr_shader_id
    faces = _arr(face_indices)
    n_tris = int(faces.shape[0])
    n_slots = int(extra[0].shape[0])
    # every slot a face names must exist in every per-vertex array
    _plan.check_index(faces, n_slots, "face_indices")
    for name in ("normal", "uv", "colour", "texture_index", "id_to_face"):
        if name in getattr(extra, "_fields", ()):
            _plan.check_rows(getattr(extra, name), n_slots, f"extra.{name}")
    if sid == _hip.SHADER_PHONG_DARBOUX:
        fi = _arr(extra.faces_indices)
        _plan.check_index(fi, n_slots, "extra.faces_indices")
        _plan.check_index(_arr(extra.id_to_face), int(fi.shape[0]), "extra.id_to_face")
    inputs = {"positions": _arr(extra.position), "faces": faces}
    sc = dict(
        world_to_clip=camera.world_to_clip,
        viewport=camera.viewport,
        world_to_eye_norm=camera.world_to_eye_norm,
        tri_begin=0,
        tri_end=n_tris,
    )
    shadow_shape = (0, 0)
    if sid != _hip.SHADER_DEPTH:
        sc["light_direction"] = _vec3(extra.light.direction)
        sc["light_colour"] = _vec3(extra.light.colour)
        inputs["normals"] = _arr(extra.normal)
    if sid == _hip.SHADER_GOURAUD:
        inputs["colours"] = _arr(extra.colour)
    if sid in (_hip.SHADER_GOURAUD_TEXTURE, _hip.SHADER_PHONG, _hip.SHADER_PHONG_DARBOUX,
               _hip.SHADER_PHONG_REFLECTION, _hip.SHADER_PHONG_REFLECTION_SHADOW):
        inputs["uvs"] = _arr(extra.uv)
        inputs["textures"] = _arr(extra.texture)
        sc["tex_shape"] = _map_shape(extra.texture)
    if sid == _hip.SHADER_PHONG_DARBOUX:
        inputs["normal_maps"] = _arr(extra.normal_map)
        inputs["id_to_face"] = _arr(extra.id_to_face)
        inputs["face_slots"] = _arr(extra.faces_indices)
        sc["nmap_shape"] = tuple(extra.normal_map.shape[:2])
    if sid in (_hip.SHADER_PHONG_REFLECTION, _hip.SHADER_PHONG_REFLECTION_SHADOW):
        inputs["tex_index"] = _arr(extra.texture_index)
        inputs["specular_maps"] = _arr(extra.specular_map)
        ts = _arr(extra.texture_shape)
        inputs["texture_shapes"] = ts
        sc.update(
            light_dir_eye=_vec3(extra.light_dir_eye),
            ambient=_vec3(extra.ambient),
            diffuse=_vec3(extra.diffuse),
            specular=_vec3(extra.specular),
            spec_shape=tuple(extra.specular_map.shape[:2]),
            texshape_count=int(ts.shape[0]),
            tex_offset=int(_host(extra.texture_offset)),
        )
    if sid == _hip.SHADER_PHONG_REFLECTION_SHADOW:
        shadow = extra.shadow
        inputs["shadow_maps"] = _arr(shadow.shadow_map)
        shadow_shape = tuple(int(v) for v in shadow.shadow_map.shape[:2])
        sc.update(
            shadow_world_to_clip=shadow.camera.world_to_clip,
            shadow_viewport=shadow.camera.viewport,
            shadow_strength=_vec3(shadow.strength),
        )
    return RasterCall(
        shader=sid,
        width=int(width),
        height=int(height),
        scenes=[_raster.scene_struct(**sc)],
        inputs=inputs,
        n_slots=n_slots,
        n_tris=n_tris,
        shadow_shape=shadow_shape,
        init_buffers=init_buffers,
        zbuffer_default=float(zbuffer_default),
        colour_default=tuple(float(c) for c in np.asarray(colour_default, np.float32).reshape(-1)[:3]),
        apply_offset=apply_offset,
        depth_offset=float(depth_offset),
        want_canvas=sid not in (_hip.SHADER_DEPTH, _hip.SHADER_VISIBILITY),
    )


def _device_copy(x, dev) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.detach().to(device=dev, dtype=torch.float32).clone().contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float32), device=dev).clone()


def is_builtin(shader: type) -> bool:
    return any(shader is b for b in BUILTIN)


def render(camera: Camera, shader: type, buffers: Buffers, face_indices, extra,
           loop_unroll: int = 1) -> Buffers:
    """Render one scene with `shader` into (copies of) `buffers`.

    Pixels are merged with `keep ? new : old`; like the reference there is no
    depth test against the incoming zbuffer (renderer/pipeline.py:401-440).
    `loop_unroll` is accepted for signature compatibility; it tunes JAX's row
    scan and has no meaning for the tiled kernels.
    """
    dev = _raster.require_device()
    zbuffer = buffers.zbuffer
    W, H = int(zbuffer.shape[0]), int(zbuffer.shape[1])
    if not is_builtin(shader):
        from ._custom import render_custom

        return render_custom(camera, shader, buffers, face_indices, extra)
    call = build_call(camera, shader, face_indices, extra, W, H)
    zb = _device_copy(zbuffer, dev)
    targets = tuple(buffers.targets)
    cv = None
    if call.want_canvas:
        if len(targets) != 1:
            raise ValueError(f"{shader.__name__} writes one target (canvas); got {len(targets)}")
        cv = _device_copy(targets[0], dev)
        if tuple(cv.shape) != (W, H, 3):
            raise ValueError(f"canvas must be [{W}, {H}, 3], got {tuple(cv.shape)}")
    _plan.run(call, zbuffer=zb.view(1, W, H), canvas=None if cv is None else cv.view(1, W, H, 3))
    return Buffers(zbuffer=zb, targets=(cv,) if cv is not None else ())


__all__ = ["build_call", "render"]
