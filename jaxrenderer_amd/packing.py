"""Batch planning with the per-corner index work on the GPU.

`plan_batch` (batch.py) walks the scenes in Python: per scene it offsets the
face tables into global index arrays (the per-corner flattening of
Renderer.render, renderer/renderer.py:277-327) and builds the scene's light
camera and records.  `plan_batch_device` produces the same calls, bit for
bit, for a whole batch at once:

* the per-scene uniforms (camera matrices, the normalised light, its
  eye-space direction, the light's orthographic camera of
  renderer/shadow.py:84-103, triangle and map offsets) are computed for all
  scenes together in numpy, with the same float32 operations in the same
  order as the per-scene code (`scene_records`);
* the scenes' rows are concatenated and uploaded once, and `jr_pack` writes
  the global position / normal / uv / texture-index arrays on the GPU, with
  the bounds check of every local index done there (an index outside its
  scene's rows raises ValueError).
"""

from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _hip, _raster
from ._plan import RasterCall
from .geometry import Camera, asf32

f32 = np.float32


def _v3(x) -> np.ndarray:
    a = np.asarray(x.detach().cpu() if hasattr(x, "detach") else x, dtype=f32).reshape(-1)
    return np.repeat(a, 3) if a.size == 1 else a[:3]


def _stack3(values) -> np.ndarray:
    return np.stack([_v3(v) for v in values]).astype(f32)


def _norm3(v: np.ndarray) -> np.ndarray:
    """geometry.normalise of each row of [B, 3]: the sum of squares left to
    right in float32, then one division per component."""
    ss = (v[:, 0] * v[:, 0]).astype(f32)
    ss = (ss + (v[:, 1] * v[:, 1]).astype(f32)).astype(f32)
    ss = (ss + (v[:, 2] * v[:, 2]).astype(f32)).astype(f32)
    return (v / np.sqrt(ss)[:, None]).astype(f32)


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.stack([
        (a[:, 1] * b[:, 2]).astype(f32) - (a[:, 2] * b[:, 1]).astype(f32),
        (a[:, 2] * b[:, 0]).astype(f32) - (a[:, 0] * b[:, 2]).astype(f32),
        (a[:, 0] * b[:, 1]).astype(f32) - (a[:, 1] * b[:, 0]).astype(f32),
    ], axis=1).astype(f32)


def _matmul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """[B, n, m] @ [B, m, p] in float32, each entry summed left to right."""
    acc = (a[:, :, 0:1] * b[:, 0:1, :]).astype(f32)
    for k in range(1, a.shape[2]):
        acc = (acc + (a[:, :, k:k + 1] * b[:, k:k + 1, :]).astype(f32)).astype(f32)
    return acc


def _view_matrix(eye: np.ndarray, centre: np.ndarray, up: np.ndarray) -> np.ndarray:
    """Camera.view_matrix (renderer/geometry.py:577-632) of B cameras."""
    B = eye.shape[0]
    forward = _norm3((centre - eye).astype(f32))
    up = _norm3(up)
    side = _norm3(_cross(forward, up))
    up = _cross(side, forward)
    m = np.tile(np.identity(4, dtype=f32), (B, 1, 1))
    m[:, 0, :3] = side
    m[:, 1, :3] = up
    m[:, 2, :3] = -forward
    t = np.tile(np.identity(4, dtype=f32), (B, 1, 1))
    t[:, :3, 3] = -eye
    return _matmul(m, t)


def _apply_vec(v: np.ndarray, M: np.ndarray) -> np.ndarray:
    """Camera.apply_vec of one vector per scene: normalise, w = 0, apply (sums
    left to right), drop w, normalise."""
    n = _norm3(v)
    h = [(n[:, 0] * M[:, i, 0]).astype(f32) for i in range(4)]
    for j in range(1, 3):
        h = [(h[i] + (n[:, j] * M[:, i, j]).astype(f32)).astype(f32) for i in range(4)]
    h = [(h[i] + (f32(0.0) * M[:, i, 3]).astype(f32)).astype(f32) for i in range(4)]
    return _norm3(np.stack(h[:3], axis=1))


def _records(n: int) -> np.ndarray:
    """n jr_scene records holding scene_struct's defaults."""
    base = np.frombuffer(bytes(_raster.scene_struct(world_to_clip=None, viewport=None)),
                         dtype=np.dtype(_hip.JrScene))
    return np.repeat(base, n)


def scene_records(scenes: Sequence, width: int, height: int, shadow=None
                  ) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """The main pass's and (with `shadow`, a ShadowMapSpec) the shadow pass's
    jr_scene records of a batch, as numpy structured arrays: the bytes
    plan_batch's calls carry, computed for all scenes at once."""
    B = len(scenes)
    F = np.array([len(s.model.faces) for s in scenes], dtype=np.int64)
    tri = np.concatenate([[0], np.cumsum(F)])
    dshape = np.array([np.shape(s.model.diffuse_map) for s in scenes], dtype=np.int64)
    sshape = np.array([np.shape(s.model.specular_map)[:2] for s in scenes], dtype=np.int64)
    nts = np.array([len(s.model.texture_shape) for s in scenes], dtype=np.int64)
    rec = _records(B)
    rec["world_to_clip"] = np.stack([asf32(s.camera.world_to_clip).reshape(16) for s in scenes])
    rec["viewport"] = np.stack([asf32(s.camera.viewport).reshape(16) for s in scenes])
    rec["world_to_eye_norm"] = np.stack([asf32(s.camera.world_to_eye_norm).reshape(16)
                                         for s in scenes])
    light_dir = _norm3(_stack3(s.light.direction for s in scenes))
    rec["light_direction"] = light_dir
    rec["light_colour"] = _stack3(s.light.colour for s in scenes)
    views = np.stack([asf32(s.camera.view) for s in scenes])
    rec["light_dir_eye"] = _apply_vec(light_dir, views)
    rec["ambient"] = _stack3(s.light.ambient for s in scenes)
    rec["diffuse"] = _stack3(s.light.diffuse for s in scenes)
    rec["specular"] = _stack3(s.light.specular for s in scenes)
    rec["tri_begin"] = tri[:-1]
    rec["tri_end"] = tri[1:]
    rec["tex_base"] = np.concatenate([[0], np.cumsum(np.prod(dshape, axis=1))])[:-1]
    rec["tex_w"], rec["tex_h"], rec["tex_c"] = dshape[:, 0], dshape[:, 1], dshape[:, 2]
    rec["spec_base"] = np.concatenate([[0], np.cumsum(np.prod(sshape, axis=1))])[:-1]
    rec["spec_w"], rec["spec_h"] = sshape[:, 0], sshape[:, 1]
    rec["texshape_base"] = np.concatenate([[0], np.cumsum(nts)])[:-1]
    rec["texshape_count"] = nts
    rec["tex_offset"] = np.array([int(s.model.offset) for s in scenes], dtype=np.int64)
    if shadow is None:
        return rec, None
    svp = Camera.viewport_matrix(np.zeros(2, np.int32), np.array((shadow.width, shadow.height)),
                                 1.0)
    centre = np.stack([asf32(shadow.centre if getattr(s, "shadow_centre", None) is None
                             else s.shadow_centre) for s in scenes])
    raw_dir = _stack3(s.light.direction for s in scenes)
    eye = (centre + (raw_dir * f32(shadow.distance)).astype(f32)).astype(f32)
    up = np.tile(asf32(shadow.up), (B, 1))
    P = Camera.orthographic_projection_matrix(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
    w2c = _matmul(np.tile(P, (B, 1, 1)), _view_matrix(eye, centre, up))
    rec["shadow_world_to_clip"] = w2c.reshape(B, 16)
    rec["shadow_viewport"] = np.tile(svp.reshape(16), (B, 1))
    rec["shadow_strength"] = np.tile(_v3(shadow.strength), (B, 1))
    srec = _records(B)
    srec["world_to_clip"] = w2c.reshape(B, 16)
    srec["viewport"] = np.tile(svp.reshape(16), (B, 1))
    srec["tri_begin"] = tri[:-1]
    srec["tri_end"] = tri[1:]
    return rec, srec


def _structs(rec: np.ndarray) -> List[_hip.JrScene]:
    """The records as jr_scene ctypes structs (what RasterCall.scenes holds)."""
    raw = memoryview(np.ascontiguousarray(rec).view(np.uint8))
    n = rec.itemsize
    return [_hip.JrScene.from_buffer_copy(raw[i * n:(i + 1) * n]) for i in range(len(rec))]


def plan_batch_device(scenes: Sequence, width: int, height: int, shadow=None, *,
                      zbuffer_default: float = 1.0, colour_default=(1.0, 1.0, 1.0),
                      device: Optional[torch.device] = None) -> List[RasterCall]:
    """batch.plan_batch with the index arrays written by jr_pack on `device`
    and the records built for all scenes at once.  Returns the same calls
    (bit for bit), holding device tensors."""
    dev = device or _raster.require_device()
    B = len(scenes)
    if B == 0:
        raise ValueError("a batch needs at least one scene")
    models = [s.model for s in scenes]

    def cat(arrs, dtype):
        return np.ascontiguousarray(np.concatenate([np.asarray(a, dtype=dtype).reshape(-1)
                                                    for a in arrs]))

    def offsets(counts):
        return np.concatenate([[0], np.cumsum(np.asarray(counts, dtype=np.int64))])

    for m in models:
        if np.shape(m.faces_norm) != np.shape(m.faces) or np.size(m.faces_uv) != np.size(m.faces):
            raise ValueError("faces, faces_norm and faces_uv must all be [F, 3]")
    F = np.array([len(m.faces) for m in models], dtype=np.int64)
    n_faces = int(F.sum())
    host = {
        "positions": cat([m.verts for m in models], f32).reshape(-1, 3),
        "normals": cat([m.norms for m in models], f32).reshape(-1, 3),
        "uvs": cat([m.uvs for m in models], f32).reshape(-1, 2),
        "textures": cat([m.diffuse_map for m in models], f32),
        "specular_maps": cat([m.specular_map for m in models], f32),
        "texture_shapes": cat([m.texture_shape for m in models], np.int32).reshape(-1, 2),
    }
    local = {
        "faces": cat([m.faces for m in models], np.int32),
        "faces_norm": cat([m.faces_norm for m in models], np.int32),
        "faces_uv": cat([m.faces_uv for m in models], np.int32),
        "texture_index": cat([m.texture_index for m in models], np.int32),
    }
    offs = {
        "tri_off": offsets(F),
        "vert_off": offsets([len(m.verts) for m in models]),
        "norm_off": offsets([len(m.norms) for m in models]),
        "uv_off": offsets([len(m.uvs) for m in models]),
        "texidx_off": offsets([len(m.texture_index) for m in models]),
    }
    d = {k: torch.from_numpy(v).to(dev) for k, v in {**host, **local, **offs}.items()}
    out = {k: torch.empty(3 * n_faces, dtype=torch.int32, device=dev)
           for k in ("pos_index", "nrm_index", "uv_index", "tex_index")}
    err = torch.zeros(1, dtype=torch.int32, device=dev)
    a = _hip.JrPackArgs()
    a.n_scenes, a.n_faces = B, n_faces
    for k in offs:
        setattr(a, k, d[k].data_ptr())
    for k in local:
        setattr(a, k, d[k].data_ptr())
    for k in out:
        setattr(a, k, out[k].data_ptr())
    a.error = err.data_ptr()
    stream = torch.cuda.current_stream(dev)
    _hip.check(_hip.lib().jr_pack(ctypes.byref(a), ctypes.c_void_p(stream.cuda_stream)),
               "jr_pack")
    if int(err.item()) != 0:
        raise ValueError("a face, faces_norm or faces_uv index lies outside its scene's "
                         "vertices, normals or uvs")
    rec, srec = scene_records(scenes, width, height, shadow)
    inputs = {
        "positions": d["positions"], "pos_index": out["pos_index"], "normals": d["normals"],
        "nrm_index": out["nrm_index"], "uvs": d["uvs"], "uv_index": out["uv_index"],
        "tex_index": out["tex_index"],
        "faces": torch.arange(3 * n_faces, dtype=torch.int32, device=dev).view(n_faces, 3),
        "textures": d["textures"], "specular_maps": d["specular_maps"],
        "texture_shapes": d["texture_shapes"],
    }
    calls: List[RasterCall] = []
    shader = _hip.SHADER_PHONG_REFLECTION
    shadow_shape = (0, 0)
    if shadow is not None:
        calls.append(RasterCall(
            shader=_hip.SHADER_DEPTH, width=shadow.width, height=shadow.height,
            scenes=_structs(srec),
            inputs={"positions": d["positions"], "faces": out["pos_index"].view(n_faces, 3)},
            n_slots=int(d["positions"].shape[0]), n_tris=n_faces, init_buffers=True,
            zbuffer_default=float(np.finfo(f32).max), apply_offset=True,
            depth_offset=float(shadow.offset), want_canvas=False, tag="shadow"))
        shader = _hip.SHADER_PHONG_REFLECTION_SHADOW
        shadow_shape = (shadow.width, shadow.height)
    calls.append(RasterCall(
        shader=shader, width=width, height=height, scenes=_structs(rec), inputs=inputs,
        n_slots=3 * n_faces, n_tris=n_faces, shadow_shape=shadow_shape, init_buffers=True,
        zbuffer_default=float(zbuffer_default),
        colour_default=tuple(float(c) for c in np.asarray(colour_default, f32).reshape(-1)[:3]),
        shadow_from="shadow" if shadow is not None else None, tag="main",
        extras={"identity_faces": True,
                "scene_verts": [len(m.verts) for m in models],
                "scene_norms": [len(m.norms) for m in models]}))
    return calls


__all__ = ["plan_batch_device", "scene_records"]
