"""Rendering with user-defined Shader subclasses.

The reference traces the user's hooks into one XLA program
(renderer/pipeline.py:470-537).  Here the user's `vertex` hook runs under
torch.func.vmap over vertex ids (as jax.vmap does in the reference), its
gl_Position goes to the HIP visibility pass (the default primitive_chooser,
JR_SHADER_VISIBILITY with caller-side clip coordinates), and `interpolate`,
`fragment` and `mix` run under torch.func.vmap over the covered pixels before
the `keep ? new : old` merge.

Hooks receive torch tensors: one vertex / one fragment at a time, with the
reference's NamedTuple structure.  A custom `primitive_chooser` is not
supported (the visibility decision is the HIP kernel's); overriding it
raises NotImplementedError.
"""

from __future__ import annotations

from typing import Callable, Optional

import numpy as np
import torch

from . import _hip, _plan, _raster
from ._plan import RasterCall
from .shader import PerFragment, Shader
from .types import Buffers


def _t(x, dev, dtype=None):
    if isinstance(x, torch.Tensor):
        return x.to(dev) if dtype is None else x.to(dev, dtype)
    return torch.as_tensor(np.asarray(x), device=dev, dtype=dtype)


def _rebuild(like, items):
    items = tuple(items)
    return type(like)(*items) if hasattr(like, "_fields") else items


def _select(t: torch.Tensor, key):
    if isinstance(key, torch.Tensor) and key.dim() == 0 and not key.is_floating_point() \
            and key.dtype != torch.bool:
        return torch.index_select(t, 0, key.reshape(1)).squeeze(0)
    return None


class _VertexArray(torch.Tensor):
    """A shader input array that can be indexed by a vmapped scalar id.

    Inside torch.func.vmap, `array[gl_VertexID]` with a 0-d integer tensor
    would go through Tensor.item(); this view routes such keys (alone or
    first in a tuple) to index_select.  Results are plain tensors."""

    __torch_function__ = torch._C._disabled_torch_function_impl

    def __getitem__(self, key):
        plain = self.as_subclass(torch.Tensor)
        got = _select(plain, key)
        if got is not None:
            return got
        if isinstance(key, tuple) and key:
            got = _select(plain, key[0])
            if got is not None:
                return got[key[1:]] if len(key) > 1 else got
        return plain[key]


def _to_tensors(obj, dev, indexable: bool = False):
    """Arrays inside (nested) NamedTuples -> torch tensors on `dev`."""
    if hasattr(obj, "_fields"):
        return type(obj)(*(_to_tensors(v, dev, indexable) for v in obj))
    if isinstance(obj, (np.ndarray, torch.Tensor)):
        t = _t(obj, dev)
        t = t.float() if t.is_floating_point() else t
        return t.as_subclass(_VertexArray) if indexable and t.dim() > 0 else t
    return obj


def visibility_call(clip: torch.Tensor, faces, viewport, width: int, height: int) -> RasterCall:
    """The default primitive chooser over caller-computed clip coordinates."""
    F = int(faces.shape[0])
    return RasterCall(
        shader=_hip.SHADER_VISIBILITY, width=width, height=height,
        scenes=[_raster.scene_struct(world_to_clip=None, viewport=viewport, tri_begin=0,
                                     tri_end=F)],
        inputs={"clip": clip.detach().float().contiguous(), "faces": faces},
        n_slots=int(clip.shape[0]), n_tris=F, init_buffers=True, want_canvas=False,
        tag="visibility")


def gpu_visibility(call: RasterCall):
    out = _plan.run(call, prim_id=True)
    return out["prim_id"][0], out["zbuffer"][0]


def render_custom(camera, shader: type, buffers: Buffers, face_indices, extra,
                  visibility: Optional[Callable] = None) -> Buffers:
    if shader.primitive_chooser is not Shader.primitive_chooser:
        raise NotImplementedError(
            "custom primitive_chooser hooks are not supported; the chooser is the HIP "
            "visibility kernel (argmin of depth over kept front-facing primitives)")
    dev = (buffers.zbuffer.device if isinstance(buffers.zbuffer, torch.Tensor)
           and visibility is not None else _raster.require_device())
    vmap = torch.func.vmap
    ex = _to_tensors(extra, dev, indexable=True)
    cam = _to_tensors(camera, dev)
    faces = _t(face_indices, dev, torch.int64)
    V = int(ex[0].shape[0])
    _plan.check_index(faces, V, "face_indices")
    inst = torch.zeros((), dtype=torch.int64, device=dev)
    per_vertex, varyings = vmap(lambda vid: shader.vertex(vid, inst, cam, ex))(
        torch.arange(V, device=dev))
    clip = per_vertex.gl_Position.float()
    zb0 = _t(buffers.zbuffer, dev, torch.float32)
    W, H = int(zb0.shape[0]), int(zb0.shape[1])
    vp_np = np.asarray(camera.viewport.detach().cpu() if isinstance(camera.viewport, torch.Tensor)
                       else camera.viewport, dtype=np.float32)
    call = visibility_call(clip, faces.to(torch.int32), vp_np, W, H)
    pid, zwin = (visibility or gpu_visibility)(call)
    pid = _t(pid, dev, torch.int64)
    zwin = _t(zwin, dev, torch.float32)

    ii, jj = torch.nonzero(pid >= 0, as_tuple=True)
    targets = [_t(t, dev, torch.float32).clone() for t in buffers.targets]
    zb = zb0.clone()
    if ii.numel() == 0:
        return Buffers(zbuffer=zb, targets=tuple(targets))
    t = pid[ii, jj]
    fv = faces[t]  # [P, 3] slots
    tri = clip[fv]  # [P, 3, 4]
    M = torch.stack([tri[..., 0], tri[..., 1], tri[..., 3]], dim=-1)
    vp = torch.as_tensor(vp_np, device=dev)
    x = (ii.float() - vp[0, 3]) / vp[0, 0]
    y = (jj.float() - vp[1, 3]) / vp[1, 1]
    e = torch.einsum("pr,prc->pc", torch.stack([x, y, torch.ones_like(x)], -1),
                     torch.linalg.inv(M))
    w = e.sum(-1)
    b = e / w[:, None]
    frag_coord = torch.stack([ii.float(), jj.float(), zwin[ii, jj], w], dim=-1)
    front = torch.linalg.det(M) >= 0
    values = _rebuild(varyings, (v[fv] for v in varyings))
    point = torch.zeros(2, device=dev)

    def per_fragment(fc, ff, vals, bary):
        varying = shader.interpolate(vals, bary, bary)
        pf, out = shader.fragment(fc, ff, point, varying, ex)
        pf = PerFragment(*(torch.as_tensor(v, device=dev) for v in pf))
        depth = torch.where(pf.use_default_depth.bool(), fc[2], pf.gl_FragDepth.float())
        mixed, tgt = shader.mix(depth[None], pf.keeps.bool()[None],
                                _rebuild(out, (torch.as_tensor(v, device=dev)[None] for v in out)))
        return mixed.keep, mixed.zbuffer, tuple(tgt)

    keep, zval, tg = vmap(per_fragment)(frag_coord, front, values, b)
    keep = keep.bool()
    zb[ii, jj] = torch.where(keep, zval.float(), zb[ii, jj])
    for k, tv in enumerate(tg[: len(targets)]):
        old = targets[k][ii, jj]
        new = tv.to(old.dtype).reshape(old.shape)
        targets[k][ii, jj] = torch.where(keep.reshape((-1,) + (1,) * (old.ndim - 1)), new, old)
    return Buffers(zbuffer=zb, targets=tuple(targets))
