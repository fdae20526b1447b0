"""Torch building blocks of the built-in shaders' hooks.

The hooks run one vertex or one fragment at a time under torch.func.vmap (as
the reference's run under jax.vmap), so everything here is written with
batchable tensor operations: gathers by a (vmapped) 0-d index go through
index_select / advanced indexing instead of Python indexing, and branches are
torch.where.  Sums and products follow the fused kernels' order (left to
right, one rounding per operation), so a hook's output equals the kernel's
intermediate bit for bit where the operations are the same.

Index semantics are JAX's (the reference's gathers): a negative index counts
from the end, anything still out of range is clamped; float -> int32
conversion (astype(int)) maps NaN to 0 and saturates.
"""

from __future__ import annotations

import numpy as np
import torch


def as_tensor(x, like: torch.Tensor, dtype=None) -> torch.Tensor:
    """x on like's device (float32 unless it is already a tensor or dtype is
    given)."""
    if isinstance(x, torch.Tensor):
        t = x.to(like.device)
        return t if dtype is None else t.to(dtype)
    a = np.asarray(x)
    if dtype is None:
        dtype = torch.float32 if a.dtype.kind == "f" else torch.int64
    return torch.as_tensor(a, device=like.device, dtype=dtype)


def index_clamp(i: torch.Tensor, n: int) -> torch.Tensor:
    """JAX gather index rule: negative counts from the end, then clamp."""
    i = torch.where(i < 0, i + n, i)
    return i.clamp(0, n - 1)


def take(a, i) -> torch.Tensor:
    """a[i] along axis 0 for an integer index (a Python int or a 0-d tensor,
    vmapped or not), with JAX's index rule."""
    if not isinstance(a, torch.Tensor):
        a = torch.as_tensor(np.asarray(a))
    if not isinstance(i, torch.Tensor):
        i = torch.as_tensor(int(i), device=a.device)
    a = a.as_subclass(torch.Tensor)
    idx = index_clamp(i.to(torch.int64).reshape(1), a.shape[0])
    return torch.index_select(a, 0, idx).squeeze(0)


def take2(a: torch.Tensor, i: torch.Tensor, j: torch.Tensor) -> torch.Tensor:
    """a[i, j] for 0-d integer indices, each with JAX's index rule."""
    a = a.as_subclass(torch.Tensor)
    ii = index_clamp(i.to(torch.int64).reshape(1), a.shape[0])
    jj = index_clamp(j.to(torch.int64).reshape(1), a.shape[1])
    return a[ii, jj].squeeze(0)


def to_int(x: torch.Tensor) -> torch.Tensor:
    """XLA's float -> int32 convert: NaN -> 0, saturating (as int64 tensors)."""
    x = torch.nan_to_num(x, nan=0.0, posinf=2147483647.0, neginf=-2147483648.0)
    return x.clamp(-2147483648.0, 2147483647.0).to(torch.int64)


def dot3(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """jnp.dot of two 3-vectors, left to right."""
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]


def normalise3(v: torch.Tensor) -> torch.Tensor:
    """geometry.normalise of a 3-vector: v / sqrt(v0^2 + v1^2 + v2^2)."""
    return v / torch.sqrt(dot3(v, v))


def apply(M: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
    """Camera.apply of one homogeneous point: out_i = sum_j p_j M[i, j], the
    sum taken left to right (renderer/geometry.py:280-315)."""
    rows = []
    for i in range(M.shape[0]):
        s = p[0] * M[i, 0]
        for j in range(1, p.shape[-1]):
            s = s + p[j] * M[i, j]
        rows.append(s)
    return torch.stack(rows)


def homogeneous(v: torch.Tensor, w: float = 1.0) -> torch.Tensor:
    return torch.cat([v, torch.full_like(v[..., :1], float(w))], dim=-1)


def to_clip(camera, position: torch.Tensor) -> torch.Tensor:
    """camera.to_clip(to_homogeneous(position))."""
    M = as_tensor(camera.world_to_clip, position)
    return apply(M, homogeneous(position))


def apply_vec(v: torch.Tensor, M) -> torch.Tensor:
    """Camera.apply_vec (renderer/geometry.py:346-386): normalise, w = 0,
    apply, drop w, normalise."""
    M = as_tensor(M, v)
    return normalise3(apply(M, homogeneous(normalise3(v), 0.0))[:3])


def normalise_homogeneous(h: torch.Tensor) -> torch.Tensor:
    return h / h[..., -1:]


def to_cartesian(h: torch.Tensor) -> torch.Tensor:
    """renderer/geometry.py to_cartesian: divide by w unless w is 0."""
    return torch.where(h[..., -1:] == 0.0, h[..., :-1], (h / h[..., -1:])[..., :-1])


def floor_mod_index(uv: torch.Tensor, shape) -> tuple:
    """lax.floor(uv).astype(int) % shape[:2], per axis (always in range)."""
    u = to_int(torch.floor(uv[0]))
    v = to_int(torch.floor(uv[1]))
    return torch.remainder(u, int(shape[0])), torch.remainder(v, int(shape[1]))


def uv_repeat_index(uv: torch.Tensor, shape: torch.Tensor, map_index: torch.Tensor,
                    offset) -> tuple:
    """MergedModel.uv_repeat (renderer/model.py:303-339) then floor to int:
    the fractional part of uv (in [0, 1)) scaled by the object's map shape,
    axis 0 shifted by map_index * offset rows of the merged map."""
    whole = torch.where(uv >= 0, torch.floor(uv), torch.ceil(uv))
    frac = uv - whole
    frac = torch.where(frac < 0, frac + 1, frac)
    shape = shape.to(torch.int64)
    shift = (map_index.to(torch.int64) * int(offset)).to(torch.int32).to(frac.dtype)
    r0 = frac[0] * shape[0].to(frac.dtype) + shift
    r1 = frac[1] * shape[1].to(frac.dtype)
    return to_int(torch.floor(r0)), to_int(torch.floor(r1))


def shadow_get(shadow_map: torch.Tensor, xy: torch.Tensor) -> torch.Tensor:
    """Shadow.get (renderer/shadow.py:127-153): lax.round (half away from
    zero), XLA's convert, negative indices from the end, +inf out of range."""
    sm = shadow_map.as_subclass(torch.Tensor)
    r = torch.sign(xy[..., :2]) * torch.floor(torch.abs(xy[..., :2]) + 0.5)
    ix, iy = to_int(r[..., 0]), to_int(r[..., 1])
    W, H = sm.shape[0], sm.shape[1]
    ix = torch.where(ix < 0, ix + W, ix)
    iy = torch.where(iy < 0, iy + H, iy)
    ok = (ix >= 0) & (ix < W) & (iy >= 0) & (iy < H)
    val = sm[ix.clamp(0, W - 1).reshape(-1), iy.clamp(0, H - 1).reshape(-1)].reshape(ix.shape)
    return torch.where(ok, val, torch.full_like(val, float("inf")))


def max_nan(a: torch.Tensor, b) -> torch.Tensor:
    """lax.max / jnp.maximum: NaN propagates."""
    b = torch.as_tensor(b, dtype=a.dtype, device=a.device)
    return torch.where(torch.isnan(a), a, torch.maximum(a, b))


__all__ = ["apply", "apply_vec", "as_tensor", "dot3", "floor_mod_index", "homogeneous",
           "index_clamp", "max_nan", "normalise3", "normalise_homogeneous", "shadow_get",
           "take", "take2", "to_cartesian", "to_clip", "to_int", "uv_repeat_index"]
