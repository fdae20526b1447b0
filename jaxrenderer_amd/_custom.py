"""Rendering with Shader classes other than the seven built-ins themselves
(user-defined shaders, and subclasses of the built-ins that override hooks).

The reference traces the user's hooks into one XLA program
(renderer/pipeline.py:470-537).  Here the `vertex` hook runs under
torch.func.vmap over vertex ids (as jax.vmap does in the reference).  Then:

* with the default `primitive_chooser`, its gl_Position goes to the HIP
  visibility pass (JR_SHADER_VISIBILITY with caller-side clip coordinates),
  which returns per pixel the chosen primitive, its window depth and its
  perspective-correct barycentrics -- the fused kernels' own visibility, bit
  for bit;
* with an overridden `primitive_chooser`, every pixel evaluates every
  primitive in torch (renderer/pipeline.py:163-279: edge values
  (x, y, 1) @ M^-1, w = sum, inside test, window depth, front facing) and the
  hook picks the primitives to shade: the reference's brute force, chunked
  over pixels.

`interpolate`, `fragment` and `mix` then run under torch.func.vmap over every
pixel (renderer/pipeline.py:283-328), and the merge writes `keep ? new :
old`.  A pixel that no kept front-facing primitive covers receives primitive
0 with keeps False, as the reference's argmin gives it, so a custom `mix` sees
the reference's inputs there too.
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
    """(primitive id, window depth, barycentrics) per pixel from the HIP
    visibility pass."""
    out = _plan.run(call, prim_id=True, bary=True)
    return out["prim_id"][0], out["zbuffer"][0], out["bary"][0]


def _det3(M: torch.Tensor) -> torch.Tensor:
    """jnp.linalg.det of [..., 3, 3] by JAX's closed form _det_3x3, left to
    right (renderer/pipeline.py:93), as the kernels evaluate it."""
    a = [[M[..., r, c] for c in range(3)] for r in range(3)]
    d = (a[0][0] * a[1][1]) * a[2][2]
    d = d + (a[0][1] * a[1][2]) * a[2][0]
    d = d + (a[0][2] * a[1][0]) * a[2][1]
    d = d - (a[0][2] * a[1][1]) * a[2][0]
    d = d - (a[0][0] * a[1][2]) * a[2][1]
    d = d - (a[0][1] * a[1][0]) * a[2][2]
    return d


def _edges(M: torch.Tensor, x: torch.Tensor, y: torch.Tensor):
    """Edge values e = (x, y, 1) @ M^-1 of primitives M [..., 3, 3] at NDC
    (x, y) (renderer/pipeline.py:170-195)."""
    Mi = torch.linalg.inv(M)
    e = (x[..., None] * Mi[..., 0, :] + y[..., None] * Mi[..., 1, :]) + Mi[..., 2, :]
    return e


def render_custom(camera, shader: type, buffers: Buffers, face_indices, extra,
                  visibility: Optional[Callable] = None) -> Buffers:
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
    vp = torch.as_tensor(vp_np, device=dev)
    targets = [_t(t, dev, torch.float32).clone() for t in buffers.targets]
    zb = zb0.clone()
    F = int(faces.shape[0])
    if F == 0:
        return Buffers(zbuffer=zb, targets=tuple(targets))

    tri = clip[faces]  # [F, 3, 4]
    M = torch.stack([tri[..., 0], tri[..., 1], tri[..., 3]], dim=-1)  # rows (x, y, w)
    det = _det3(M)
    prim_keep = det.abs() > 1e-6
    front_all = det >= 0
    z_clip = tri[..., 2]  # [F, 3]
    values_all = _rebuild(varyings, (v[faces] for v in varyings))  # [F, 3, ...]
    ii, jj = torch.meshgrid(torch.arange(W, device=dev), torch.arange(H, device=dev),
                            indexing="ij")
    ii, jj = ii.reshape(-1), jj.reshape(-1)
    x = (ii.float() - vp[0, 3]) / vp[0, 0]
    y = (jj.float() - vp[1, 3]) / vp[1, 1]
    point = torch.zeros(2, device=dev)

    def shade(frag_coord, front, keeps, values, bary):
        # renderer/pipeline.py:283-328 for the chosen primitives [K, ...],
        # then the mixer over them
        def one(fc, ff, kk, vals, b):
            varying = shader.interpolate(vals, b, b)
            pf, out = shader.fragment(fc, ff, point, varying, ex)
            pf = PerFragment(*(torch.as_tensor(v, device=dev) for v in pf))
            depth = torch.where(pf.use_default_depth.bool(), fc[2], pf.gl_FragDepth.float())
            out = _rebuild(out, (torch.as_tensor(v, device=dev) for v in out))
            return depth, kk & pf.keeps.bool(), out

        depths, kept, outs = vmap(one)(frag_coord, front, keeps, values, bary)
        mixed, tgt = shader.mix(depths, kept, outs)
        return mixed.keep, mixed.zbuffer, tuple(tgt)

    custom_chooser = shader.primitive_chooser is not Shader.primitive_chooser
    if not custom_chooser:
        # the HIP visibility pass decides; primitive 0 (keeps False) where it
        # found nothing
        call = visibility_call(clip, faces.to(torch.int32), vp_np, W, H)
        pid, zwin, bary = (visibility or gpu_visibility)(call)
        pid = _t(pid, dev, torch.int64).reshape(-1)
        zwin = _t(zwin, dev, torch.float32).reshape(-1)
        bary = _t(bary, dev, torch.float32).reshape(-1, 3)

    def default_chunk(sl):
        covered = pid[sl] >= 0
        t = torch.where(covered, pid[sl], torch.zeros_like(pid[sl]))
        e = _edges(M[t], x[sl], y[sl])  # w for gl_FragCoord; primitive 0's data
        w = e.sum(-1)
        z0 = ((e[:, 0] * z_clip[t, 0] + e[:, 1] * z_clip[t, 1]) + e[:, 2] * z_clip[t, 2]) \
            * vp[2, 2] + vp[2, 3]
        z = torch.where(covered, zwin[sl], z0)
        b = torch.where(covered[:, None], bary[sl], e / w[:, None])
        fc = torch.stack([ii[sl].float(), jj[sl].float(), z, w], dim=-1)[:, None]
        values = _rebuild(values_all, (v[t][:, None] for v in values_all))
        return vmap(shade)(fc, front_all[t][:, None], covered[:, None], values, b[:, None])

    def chooser_chunk(sl):
        # every primitive at every pixel of the chunk; the hook chooses
        e = _edges(M[None], x[sl, None], y[sl, None])  # [P, F, 3]
        w = e.sum(-1)
        inside = (e >= 0).all(-1)
        z = ((e[..., 0] * z_clip[None, :, 0] + e[..., 1] * z_clip[None, :, 1])
             + e[..., 2] * z_clip[None, :, 2]) * vp[2, 2] + vp[2, 3]
        P = e.shape[0]
        fc = torch.stack([ii[sl, None].float().expand(P, F), jj[sl, None].float().expand(P, F),
                          z, w], dim=-1)
        b = e / w[..., None]
        keeps = prim_keep[None] & inside
        pc = torch.zeros((P, F, 2), device=dev)

        def choose_and_shade(fc_, ff_, pc_, kk_, b_):
            c_fc, c_ff, _, c_kk, c_vals, _, c_bc = shader.primitive_chooser(
                fc_, ff_, pc_, kk_, values_all, b_, b_)
            return shade(c_fc, c_ff, c_kk, c_vals, c_bc)

        return vmap(choose_and_shade)(fc, front_all[None].expand(P, F), pc, keeps, b)

    chunk = max(1, (1 << 22) // max(F, 1)) if custom_chooser else 1 << 16
    parts = [(chooser_chunk if custom_chooser else default_chunk)(slice(p0, p0 + chunk))
             for p0 in range(0, W * H, chunk)]
    keep = torch.cat([q[0].reshape(-1) for q in parts])
    zval = torch.cat([q[1].reshape(-1) for q in parts])
    tg = tuple(torch.cat([q[2][k] for q in parts]) for k in range(len(parts[0][2])))

    keep = keep.bool().reshape(-1)
    zflat = zb.reshape(-1)
    zb = torch.where(keep, zval.float().reshape(-1), zflat).reshape(W, H)
    for k, tv in enumerate(tg[: len(targets)]):
        old = targets[k].reshape((W * H,) + tuple(targets[k].shape[2:]))
        new = tv.to(old.dtype).reshape(old.shape)
        mask = keep.reshape((-1,) + (1,) * (old.ndim - 1))
        targets[k] = torch.where(mask, new, old).reshape(targets[k].shape)
    return Buffers(zbuffer=zb, targets=tuple(targets))
