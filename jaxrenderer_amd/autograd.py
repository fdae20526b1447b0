"""Differentiable rendering: jr_forward / jr_backward behind
torch.autograd.Function.

The reference differentiates its pipeline with jax.grad
(tests/smoke_test_grad.py:92-128).  Here `render_differentiable` runs the HIP
forward, keeps only the per-pixel primitive ids, and its backward calls the
hand-written jr_backward kernels, which recompute every fragment from those
ids.  Gradients reach vertex positions, normals, per-vertex colours, texture
and specular maps (and for phong_darboux, whose tangent frame is built from
each face's corners and uvs, the uvs and the normal map), and -- through one per-scene parameter vector laid out like
include/jrhip.h's JR_DSCENE_* -- the light, the material coefficients, the
shadow strength and the camera matrices (world_to_clip, viewport,
world_to_eye_norm).  Visibility is piecewise constant, as under jax.grad.
"""

from __future__ import annotations

import ctypes
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _hip, _plan, _raster
from ._plan import RasterCall
from .pipeline import build_call, is_builtin

_D = _hip.DSCENE
_SLICES = {
    "light_direction": (_D["light_direction"], 3),
    "light_colour": (_D["light_colour"], 3),
    "light_dir_eye": (_D["light_dir_eye"], 3),
    "ambient": (_D["ambient"], 3),
    "diffuse": (_D["diffuse"], 3),
    "specular": (_D["specular"], 3),
    "world_to_clip": (_D["world_to_clip"], 16),
    "viewport": (_D["viewport"], 16),
    "world_to_eye_norm": (_D["world_to_eye_norm"], 16),
    "shadow_strength": (_D["shadow_strength"], 3),
}


def scene_rows(scenes: Sequence[_hip.JrScene]) -> np.ndarray:
    """The differentiable fields of jr_scene records as [n_scenes, 69]."""
    out = np.zeros((len(scenes), _hip.DSCENE_FLOATS), dtype=np.float32)
    for i, s in enumerate(scenes):
        for name, (o, n) in _SLICES.items():
            out[i, o:o + n] = np.frombuffer(getattr(s, name), dtype=np.float32)[:n]
    return out


def scenes_with_rows(scenes: Sequence[_hip.JrScene], rows: np.ndarray) -> list:
    """Copies of the jr_scene records with their differentiable fields
    replaced by `rows`."""
    out = []
    for s, r in zip(scenes, rows):
        c = _hip.JrScene.from_buffer_copy(bytes(s))
        for name, (o, n) in _SLICES.items():
            setattr(c, name, (ctypes.c_float * n)(*r[o:o + n]))
        out.append(c)
    return out


_TENSOR_INPUTS = ("positions", "normals", "colours", "textures", "specular_maps", "uvs",
                  "normal_maps")
_GRAD_NAMES = ("d_positions", "d_normals", "d_colours", "d_textures", "d_specular_maps",
               "d_uvs", "d_normal_maps")


class RenderFunction(torch.autograd.Function):
    """(params [B, 69], positions, normals, colours, textures, specular_maps,
    uvs, normal_maps) -> (zbuffer [B, W, H], canvas [B, W, H, 3]).  uvs and
    normal_maps receive gradients under phong_darboux only (elsewhere uvs
    reach the image through floor alone)."""

    @staticmethod
    def forward(ctx, call: RasterCall, params, positions, normals, colours, textures,
                specular_maps, uvs=None, normal_maps=None):
        dev = params.device
        rows = params.detach().float().cpu().numpy()
        scenes = _raster.scenes_to_device(scenes_with_rows(call.scenes, rows), dev)
        inputs = {k: _plan.to_device(v, dev, k) for k, v in call.inputs.items()
                  if k not in _TENSOR_INPUTS}
        tensors = (positions, normals, colours, textures, specular_maps, uvs, normal_maps)
        for name, t in zip(_TENSOR_INPUTS, tensors):
            if t is not None:
                inputs[name] = t.detach().float().contiguous()
            elif name in ("uvs", "normal_maps") and name in call.inputs:
                inputs[name] = _plan.to_device(call.inputs[name], dev, name)
            else:
                inputs[name] = None
        B, W, H = call.n_scenes, call.width, call.height
        zbuffer = torch.empty((B, W, H), dtype=torch.float32, device=dev)
        canvas = (torch.empty((B, W, H, 3), dtype=torch.float32, device=dev)
                  if call.want_canvas else None)
        prim_id = torch.empty((B, W, H), dtype=torch.int32, device=dev)
        _raster.forward(
            shader=call.shader, width=W, height=H, scenes=scenes, n_scenes=B, inputs=inputs,
            n_slots=call.n_slots, n_tris=call.n_tris, zbuffer=zbuffer, canvas=canvas,
            prim_id=prim_id, shadow_shape=call.shadow_shape, init_buffers=True,
            zbuffer_default=call.zbuffer_default, colour_default=call.colour_default,
            apply_offset=call.apply_offset, depth_offset=call.depth_offset)
        ctx.call = call
        ctx.scenes = scenes
        ctx.inputs = inputs
        ctx.prim_id = prim_id
        ctx.shapes = [None if t is None else tuple(t.shape) for t in tensors]
        ctx.needs = ctx.needs_input_grad
        if canvas is None:
            canvas = torch.zeros((B, W, H, 3), dtype=torch.float32, device=dev)
        return zbuffer, canvas

    @staticmethod
    def backward(ctx, d_zbuffer, d_canvas):
        call = ctx.call
        dev = ctx.prim_id.device
        grads: Dict[str, Optional[torch.Tensor]] = {}
        for i, (name, shape) in enumerate(zip(_GRAD_NAMES, ctx.shapes)):
            need = ctx.needs[2 + i] and shape is not None
            grads[name] = torch.zeros(shape, dtype=torch.float32, device=dev) if need else None
        grads["d_scene"] = (torch.zeros((call.n_scenes, _hip.DSCENE_FLOATS), dtype=torch.float32,
                                        device=dev) if ctx.needs[1] else None)
        dz = None if d_zbuffer is None else d_zbuffer.float().contiguous()
        dc = None if (d_canvas is None or not call.want_canvas) else d_canvas.float().contiguous()
        if dz is not None or dc is not None:
            _raster.backward(
                shader=call.shader, width=call.width, height=call.height, scenes=ctx.scenes,
                n_scenes=call.n_scenes, inputs=ctx.inputs, n_slots=call.n_slots,
                n_tris=call.n_tris, prim_id=ctx.prim_id, d_canvas=dc, d_zbuffer=dz,
                grads=grads, shadow_shape=call.shadow_shape)
        return (None, grads["d_scene"]) + tuple(grads[n] for n in _GRAD_NAMES)


def darboux_leaves(shader, leaf) -> tuple:
    """(uvs, normal_maps) tensors for phong_darboux, whose image depends on
    them continuously; (None, None) for the other shaders."""
    if shader._jr_shader_id != _hip.SHADER_PHONG_DARBOUX:
        return (None, None)
    return (leaf("uv"), leaf("normal_map"))


def _as_tensor(x, dev) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(dev)
    return torch.as_tensor(np.asarray(x, dtype=np.float32), device=dev)


def _row_tensor(dev, camera, extra, call: RasterCall) -> torch.Tensor:
    """[1, 69] parameter row assembled (differentiably) from the camera and the
    extra input's tensors, defaults from the call's scene record."""
    base = torch.as_tensor(scene_rows(call.scenes)[0], device=dev)
    parts = []
    pos = 0
    vals = {}
    light = getattr(extra, "light", None)
    if light is not None:
        vals["light_direction"] = light.direction
        vals["light_colour"] = light.colour
    for name in ("light_dir_eye", "ambient", "diffuse", "specular"):
        if hasattr(extra, name):
            vals[name] = getattr(extra, name)
    vals["world_to_clip"] = camera.world_to_clip
    vals["viewport"] = camera.viewport
    vals["world_to_eye_norm"] = camera.world_to_eye_norm
    shadow = getattr(extra, "shadow", None)
    if shadow is not None:
        vals["shadow_strength"] = shadow.strength
    for name, (o, n) in sorted(_SLICES.items(), key=lambda kv: kv[1][0]):
        if o > pos:
            parts.append(base[pos:o])
        v = vals.get(name)
        if v is None:
            parts.append(base[o:o + n])
        else:
            t = _as_tensor(v, dev).float().reshape(-1)
            if t.numel() == 1:
                t = t.expand(n)
            parts.append(t[:n])
        pos = o + n
    if pos < _hip.DSCENE_FLOATS:
        parts.append(base[pos:])
    return torch.cat(parts)[None, :]


def build_views_call(cameras: Sequence, shader, face_indices, extra, width: int, height: int,
                     **kw) -> RasterCall:
    """One jr_forward call rendering the same scene through several cameras:
    scene b is view b, every view shares the vertex, attribute and texture
    arrays (its face rows are the same slots), so gradients of shared inputs
    sum over the views in the backward pass.  A shadow map is repeated per
    view (each scene reads its own)."""
    cameras = list(cameras)
    if not cameras:
        raise ValueError("at least one camera is needed")
    base = build_call(cameras[0], shader, face_indices, extra, width, height, **kw)
    B, F = len(cameras), base.n_tris
    scenes = []
    for b, cam in enumerate(cameras):
        c = _hip.JrScene.from_buffer_copy(bytes(base.scenes[0]))
        for name in ("world_to_clip", "viewport", "world_to_eye_norm"):
            m = np.asarray(getattr(cam, name) if not isinstance(getattr(cam, name), torch.Tensor)
                           else getattr(cam, name).detach().cpu().numpy(), np.float32)
            setattr(c, name, (ctypes.c_float * 16)(*m.reshape(-1)))
        c.tri_begin, c.tri_end = b * F, (b + 1) * F
        scenes.append(c)
    inputs = dict(base.inputs)
    faces = inputs["faces"]
    inputs["faces"] = (faces.repeat(B, 1) if isinstance(faces, torch.Tensor)
                       else np.tile(np.asarray(faces), (B, 1)))
    if "shadow_maps" in inputs:
        sm = inputs["shadow_maps"]
        inputs["shadow_maps"] = (sm[None].repeat(B, 1, 1) if isinstance(sm, torch.Tensor)
                                 else np.tile(np.asarray(sm)[None], (B, 1, 1)))
    call = RasterCall(**{**base.__dict__, "scenes": scenes, "inputs": inputs, "n_tris": B * F})
    return call


def render_views_differentiable(cameras: Sequence, shader, face_indices, extra, width: int,
                                height: int, *, zbuffer_default: float = 1.0,
                                colour_default=(1.0, 1.0, 1.0)
                                ) -> Tuple[torch.Tensor, torch.Tensor]:
    """Like render_differentiable for a list of cameras, in one batched
    forward and one backward launch.  Returns (zbuffer [B, W, H], canvas
    [B, W, H, 3]); gradients reach every camera's tensors and the shared
    tensor leaves of `extra`."""
    if not is_builtin(shader):
        raise NotImplementedError("render_views_differentiable supports the built-in shaders")
    dev = _raster.require_device()

    def host(x):
        return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x

    def host_tuple(nt):
        return type(nt)(*(host(v) if not hasattr(v, "_fields") else host_tuple(v) for v in nt))

    cams_h = [type(c)(*(host(v) for v in c)) for c in cameras]
    extra_h = host_tuple(extra)
    call = build_views_call(cams_h, shader, host(face_indices), extra_h, width, height,
                            init_buffers=True, zbuffer_default=zbuffer_default,
                            colour_default=colour_default)
    one = RasterCall(**{**call.__dict__, "scenes": call.scenes[:1]})
    params = torch.cat([_row_tensor(dev, cam, extra, one) for cam in cameras])

    def leaf(name):
        v = getattr(extra, name, None)
        return None if v is None else _as_tensor(v, dev).float()

    return RenderFunction.apply(call, params, leaf("position"), leaf("normal"), leaf("colour"),
                                leaf("texture"), leaf("specular_map"), *darboux_leaves(shader, leaf))


def render_differentiable(camera, shader, face_indices, extra, width: int, height: int, *,
                          zbuffer_default: float = 1.0,
                          colour_default=(1.0, 1.0, 1.0)) -> Tuple[torch.Tensor, torch.Tensor]:
    """`render` into fresh buffers, differentiable in every tensor leaf of
    `extra` and `camera` (see module docstring).  Returns (zbuffer [W, H],
    canvas [W, H, 3])."""
    if not is_builtin(shader):
        raise NotImplementedError("render_differentiable supports the built-in shaders")
    dev = _raster.require_device()

    def host(x):
        return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x

    def host_tuple(nt):
        return type(nt)(*(host(v) if not hasattr(v, "_fields") else host_tuple(v) for v in nt))

    cam_h = type(camera)(*(host(v) for v in camera))
    extra_h = host_tuple(extra)
    call = build_call(cam_h, shader, host(face_indices), extra_h, width, height,
                      init_buffers=True, zbuffer_default=zbuffer_default,
                      colour_default=colour_default)
    params = _row_tensor(dev, camera, extra, call)

    def leaf(name):
        v = getattr(extra, name, None)
        return None if v is None else _as_tensor(v, dev).float()

    zb, cv = RenderFunction.apply(call, params, leaf("position"), leaf("normal"), leaf("colour"),
                                  leaf("texture"), leaf("specular_map"),
                                  *darboux_leaves(shader, leaf))
    return zb[0], cv[0]


__all__ = ["RenderFunction", "build_views_call", "render_differentiable",
           "render_views_differentiable", "scene_rows", "scenes_with_rows"]
