"""The public facade (renderer/renderer.py): Renderer, CameraParameters,
LightParameters, ShadowParameters.

`Renderer.plan` turns a merged model, light, camera and optional shadow into
the one or two jr_forward calls that render it (shadow-map pass, then the
phong_reflection(_shadow) pass); `Renderer.render` and `get_camera_image`
execute that plan on the GPU.
"""

from __future__ import annotations

from typing import Any, List, NamedTuple, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _hip, _plan, _raster
from ._plan import RasterCall
from .geometry import Camera, asf32, normalise
from .model import MergedModel, ModelObject, merge_objects
from .shadow import Shadow
from .types import Buffers

f32 = np.float32


class CameraParameters(NamedTuple):
    """renderer/renderer.py:46-73 (defaults from PyTinyrenderer)."""

    viewWidth: int = 640
    viewHeight: int = 480
    viewDepth: float = 1.0
    near: float = 0.01
    far: float = 1000.0
    hfov: float = 58.0
    vfov: float = 45.0
    position: Any = (1.0, 1.0, 1.0)
    target: Any = (0.0, 0.0, 0.0)
    up: Any = (0.0, 0.0, 1.0)


class LightParameters(NamedTuple):
    """renderer/renderer.py:76-109."""

    direction: Any = normalise(np.array((0.57735, 0.57735, 0.57735), dtype=f32))
    colour: Any = np.ones(3, dtype=f32)
    ambient: Any = np.array((0.6, 0.6, 0.6), dtype=f32)
    diffuse: Any = np.array((0.35, 0.35, 0.35), dtype=f32)
    specular: Any = np.array((0.05, 0.05, 0.05), dtype=f32)


class ShadowParameters(NamedTuple):
    """renderer/renderer.py:112-133."""

    centre: Any = np.zeros(3, dtype=f32)
    up: Any = np.array((0.0, 0.0, 1.0), dtype=f32)
    strength: Any = (f32(1.0) - np.array((0.4, 0.4, 0.4), dtype=f32)).astype(f32)
    offset: float = 0.05


def _tan_half(deg) -> np.float32:
    return f32(np.tan(f32(f32(f32(deg) * f32(np.pi / 180)) / f32(2.0))))


class Renderer:
    @staticmethod
    def create_camera_from_parameters(camera: CameraParameters) -> Camera:
        """renderer/renderer.py:137-196."""
        eye = asf32(camera.position)
        centre = asf32(camera.target)
        up = asf32(camera.up)
        view = Camera.view_matrix(eye=eye, centre=centre, up=up)
        view_inv = Camera.view_matrix_inv(eye=eye, centre=centre, up=up)
        aspect = f32(_tan_half(camera.hfov) / _tan_half(camera.vfov))
        projection = Camera.perspective_projection_matrix(
            fovy=camera.vfov, aspect=aspect, z_near=camera.near, z_far=camera.far
        )
        viewport = Camera.viewport_matrix(
            lowerbound=np.zeros(2, dtype=np.int32),
            dimension=np.array((camera.viewWidth, camera.viewHeight), dtype=np.int32),
            depth=f32(camera.viewDepth),
        )
        return Camera.create(view=view, projection=projection, viewport=viewport,
                             view_inv=view_inv)

    @staticmethod
    def create_buffers(width: int, height: int, batch: Optional[int] = None,
                       colour_default=(1.0, 1.0, 1.0), zbuffer_default=1.0) -> Buffers:
        """renderer/renderer.py:198-243 (GPU tensors, x-major)."""
        dev = _raster.require_device()
        lead = (batch,) if batch is not None else ()
        zbuffer = torch.full(lead + (width, height), float(np.asarray(zbuffer_default)),
                             dtype=torch.float32, device=dev)
        cd = torch.as_tensor(np.asarray(colour_default, dtype=f32), device=dev)
        canvas = cd.expand(lead + (width, height, cd.numel())).contiguous()
        return Buffers(zbuffer=zbuffer, targets=(canvas,))

    @staticmethod
    def plan(model: MergedModel, light: LightParameters, camera: Camera, width: int,
             height: int, shadow_param: Optional[ShadowParameters] = None, *,
             init_buffers: bool = True, zbuffer_default: float = 1.0,
             colour_default=(1.0, 1.0, 1.0)) -> List[RasterCall]:
        """Renderer.render (renderer/renderer.py:254-385) as jr_forward calls."""
        faces = np.asarray(model.faces, dtype=np.int32)
        F = int(faces.shape[0])
        faces_uv = np.asarray(model.faces_uv, dtype=np.int32).reshape(-1)
        tex_index = np.asarray(model.texture_index, dtype=np.int32)
        # texture_index is per vertex but the reference gathers it with faces_uv
        # (renderer/renderer.py:287-289); out-of-range rows clamp like a JAX gather
        slot_tex = tex_index[np.clip(faces_uv, 0, max(len(tex_index) - 1, 0))]
        for name, f, n in (("faces", model.faces, len(model.verts)),
                           ("faces_norm", model.faces_norm, len(model.norms)),
                           ("faces_uv", model.faces_uv, len(model.uvs))):
            f = np.asarray(f)
            if f.size and (f.min() < 0 or f.max() >= n):
                raise ValueError(f"{name} out of bound, expected [0, {n})")
        light_dir = normalise(asf32(light.direction))
        light_dir_eye = Camera.apply_vec(light_dir, camera.view)
        tex = asf32(model.diffuse_map)
        spec = asf32(model.specular_map)
        texshape = np.asarray(model.texture_shape, dtype=np.int32)
        sc = dict(
            world_to_clip=camera.world_to_clip,
            viewport=camera.viewport,
            world_to_eye_norm=camera.world_to_eye_norm,
            light_direction=light_dir,
            light_colour=asf32(light.colour),
            light_dir_eye=light_dir_eye,
            ambient=asf32(light.ambient),
            diffuse=asf32(light.diffuse),
            specular=asf32(light.specular),
            tri_begin=0,
            tri_end=F,
            tex_shape=(tex.shape[0], tex.shape[1], tex.shape[2] if tex.ndim == 3 else 1),
            spec_shape=spec.shape[:2],
            texshape_count=int(texshape.shape[0]),
            tex_offset=int(model.offset),
        )
        inputs = {
            "positions": asf32(model.verts),
            "pos_index": faces.reshape(-1),
            "normals": asf32(model.norms),
            "nrm_index": np.asarray(model.faces_norm, dtype=np.int32).reshape(-1),
            "uvs": asf32(model.uvs),
            "uv_index": faces_uv,
            "tex_index": slot_tex,
            "faces": np.arange(3 * F, dtype=np.int32).reshape(F, 3),
            "textures": tex,
            "specular_maps": spec,
            "texture_shapes": texshape,
        }
        calls: List[RasterCall] = []
        shader = _hip.SHADER_PHONG_REFLECTION
        shadow_shape = (0, 0)
        shadow_from = None
        if shadow_param is not None:
            scam = Shadow.camera_for(light.direction, camera.viewport, shadow_param.centre,
                                     shadow_param.up)
            calls.append(RasterCall(
                shader=_hip.SHADER_DEPTH, width=width, height=height,
                scenes=[_raster.scene_struct(world_to_clip=scam.world_to_clip,
                                             viewport=scam.viewport, tri_begin=0, tri_end=F)],
                inputs={"positions": asf32(model.verts), "faces": faces},
                n_slots=int(len(model.verts)), n_tris=F, init_buffers=True,
                zbuffer_default=float(np.finfo(f32).max), apply_offset=True,
                depth_offset=float(shadow_param.offset), want_canvas=False, tag="shadow",
            ))
            sc.update(shadow_world_to_clip=scam.world_to_clip, shadow_viewport=scam.viewport,
                      shadow_strength=asf32(shadow_param.strength))
            shader = _hip.SHADER_PHONG_REFLECTION_SHADOW
            shadow_shape = (width, height)
            shadow_from = "shadow"
        calls.append(RasterCall(
            shader=shader, width=width, height=height, scenes=[_raster.scene_struct(**sc)],
            inputs=inputs, n_slots=3 * F, n_tris=F, shadow_shape=shadow_shape,
            init_buffers=init_buffers, zbuffer_default=float(zbuffer_default),
            colour_default=tuple(float(c) for c in np.asarray(colour_default, f32).reshape(-1)[:3]),
            shadow_from=shadow_from, tag="main", extras={"identity_faces": True},
        ))
        return calls

    @classmethod
    def render(cls, model: MergedModel, light: LightParameters, camera: Camera,
               buffers: Buffers, shadow_param: Optional[ShadowParameters] = None,
               loop_unroll: int = 1) -> Buffers:
        """renderer/renderer.py:245-385; merges into copies of `buffers`."""
        dev = _raster.require_device()
        zb = torch.as_tensor(buffers.zbuffer).to(device=dev, dtype=torch.float32).clone()
        cv = torch.as_tensor(buffers.targets[0]).to(device=dev, dtype=torch.float32).clone()
        W, H = int(zb.shape[-2]), int(zb.shape[-1])
        calls = cls.plan(model, light, camera, W, H, shadow_param, init_buffers=False)
        if shadow_param is not None:
            res = _plan.run(calls[0])
            calls[1].inputs["shadow_maps"] = res["zbuffer"]
        _plan.run(calls[-1], zbuffer=zb.view(1, W, H), canvas=cv.view(1, W, H, 3))
        return Buffers(zbuffer=zb, targets=(cv,))

    @classmethod
    def get_camera_image(cls, objects: Sequence[ModelObject], light: LightParameters,
                         camera: Union[Camera, CameraParameters], width: int, height: int,
                         colour_default=(1.0, 1.0, 1.0), zbuffer_default=1.0,
                         shadow_param: Optional[ShadowParameters] = None,
                         loop_unroll: int = 1) -> torch.Tensor:
        """renderer/renderer.py:387-476: the [width, height, 3] canvas."""
        buffers = cls.get_camera_buffers(objects, light, camera, width, height, colour_default,
                                         zbuffer_default, shadow_param)
        return buffers.targets[0]

    @classmethod
    def get_camera_buffers(cls, objects, light, camera, width, height,
                           colour_default=(1.0, 1.0, 1.0), zbuffer_default=1.0,
                           shadow_param=None) -> Buffers:
        """get_camera_image that also returns the zbuffer."""
        if isinstance(camera, CameraParameters):
            camera = cls.create_camera_from_parameters(camera)
        model = merge_objects(objects)
        calls = cls.plan(model, light, camera, width, height, shadow_param, init_buffers=True,
                         zbuffer_default=float(np.asarray(zbuffer_default)),
                         colour_default=colour_default)
        res = _plan.run_plan(calls)["main"]
        return Buffers(zbuffer=res["zbuffer"][0], targets=(res["canvas"][0],))


__all__ = ["CameraParameters", "LightParameters", "Renderer", "ShadowParameters"]
