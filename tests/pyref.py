"""Independent float64 restatement of the reference pipeline for tiny scenes.

Written from renderer/pipeline.py and renderer/shader.py directly (vmap over
pixels x primitives, argmin chooser, keep/merge rules), in numpy float64 with
library linear algebra (np.linalg.det / inv).  It shares no code with the C
oracle; tests compare the two where float32 rounding cannot flip a decision.
"""

from __future__ import annotations

import numpy as np


def primitives(clip: np.ndarray, faces: np.ndarray):
    """PerPrimitive.create (renderer/pipeline.py:76-113) in float64."""
    tri = clip[faces]  # [F, 3, 4]
    M = tri[:, :, [0, 1, 3]]
    det = np.linalg.det(M)
    keep = np.abs(det) > 1e-6
    with np.errstate(all="ignore"):
        inv = np.linalg.inv(np.where(keep[:, None, None], M, np.eye(3)))
    return tri, det, keep, inv


def visibility(clip: np.ndarray, faces: np.ndarray, W: int, H: int, vp: np.ndarray):
    """Per pixel: chosen primitive index, its keep, depth, barycentrics and a
    margin telling how far the decision is from a float32-sensitive tie."""
    tri, det, keep, inv = primitives(clip.astype(np.float64), faces)
    vp = vp.astype(np.float64)
    i = np.arange(W, dtype=np.float64)[:, None]
    j = np.arange(H, dtype=np.float64)[None, :]
    x = ((i - vp[0, 3]) / vp[0, 0])[..., None, None] * np.ones((1, H, 1, 1))
    y = ((j - vp[1, 3]) / vp[1, 1])[..., None, None] * np.ones((W, 1, 1, 1))
    # clip_coef = (x, y, 1) @ Minv  (renderer/pipeline.py:177-193)
    e = x * inv[None, None, :, 0, :] + y * inv[None, None, :, 1, :] + inv[None, None, :, 2, :]
    inside = np.all(e >= 0, axis=-1)
    z = (e * tri[None, None, :, :, 2]).sum(-1) * vp[2, 2] + vp[2, 3]
    front = det >= 0
    cand = keep[None, None, :] & inside & front[None, None, :]
    depth = np.where(cand, z, np.inf)
    idx = np.argmin(depth, axis=-1)  # first minimum; index 0 when all inf
    take = lambda a: np.take_along_axis(a, idx[..., None], axis=2)[..., 0]  # noqa: E731
    chosen_keep = keep[idx] & take(inside)
    ech = np.take_along_axis(e, idx[..., None, None], axis=2)[:, :, 0, :]
    zch = take(z)
    bary = ech / ech.sum(-1, keepdims=True)
    # margin: distance of any edge value from 0 and of the best depth from the
    # runner-up (relative), the quantities float32 rounding could flip
    scale = np.abs(inv).sum(axis=1).max(axis=-1)  # |x a| + |y b| + |c| bound
    edge_margin = np.min(np.abs(e) / (scale[None, None, :, None] + 1e-30), axis=(2, 3))
    sd = np.sort(np.where(cand, z, np.inf), axis=-1)
    if sd.shape[-1] > 1:
        both = np.isfinite(sd[..., 1]) & np.isfinite(sd[..., 0])
        gap = np.full(sd.shape[:2], np.inf)
        gap[both] = np.abs(sd[..., 1] - sd[..., 0])[both]
    else:
        gap = np.full(sd.shape[:2], np.inf)
    return dict(idx=idx, keep=chosen_keep, depth=zch, bary=bary, front=front[idx],
                edge_margin=edge_margin, depth_gap=gap)


def normalise(v: np.ndarray) -> np.ndarray:
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def gouraud_colour(pos, normal, colour, light_dir, light_col, faces, vis):
    """GouraudShader vertex + fragment (shaders/gouraud.py:54-124), float64."""
    inten = (normalise(normal) * normalise(light_dir)).sum(-1)
    vcol = colour * light_col[None, :] * inten[:, None]
    tri = vcol[faces][vis["idx"]]  # [W, H, 3 verts, 3]
    col = (vis["bary"][..., None] * tri).sum(-2)
    keep = vis["keep"] & vis["front"] & np.all(col >= 0, axis=-1)
    return col, keep


def phong_reflection_colour(normal_eye, uv, tex_index, texture, spec_map, tex_shapes, tex_offset,
                            light_dir_eye, light_col, ambient, diffuse, specular, faces, vis,
                            shadow=None):
    """PhongReflection(Shadow)TextureShader fragment (shaders/phong_reflection.py:154-235,
    phong_reflection_shadow.py:173-271) in float64 from per-slot varyings."""
    tri_n = normal_eye[faces][vis["idx"]]
    tri_uv = uv[faces][vis["idx"]]
    b = vis["bary"][..., None]
    n = normalise((b * tri_n).sum(-2))
    uvi = (b * tri_uv).sum(-2)
    ti = tex_index[faces[vis["idx"], 0]]
    shape = tex_shapes[np.clip(ti, 0, len(tex_shapes) - 1)]
    frac = uvi - np.where(uvi >= 0, np.floor(uvi), np.ceil(uvi))
    frac = np.where(frac < 0, frac + 1, frac)
    r0 = frac[..., 0] * shape[..., 0] + ti * tex_offset
    r1 = frac[..., 1] * shape[..., 1]
    u = np.clip(np.floor(r0).astype(np.int64), 0, texture.shape[0] - 1)
    v = np.clip(np.floor(r1).astype(np.int64), 0, texture.shape[1] - 1)
    tex = texture[u, v]
    sp = spec_map[np.clip(u, 0, spec_map.shape[0] - 1), np.clip(v, 0, spec_map.shape[1] - 1)]
    l = normalise(light_dir_eye)
    d = (n * l).sum(-1)
    diff = np.maximum(d, 0)
    r = normalise(2 * d[..., None] * n - l)
    s = np.power(np.maximum(r[..., 2], 0), sp)
    k = diffuse * diff[..., None] + specular * s[..., None]
    if shadow is None:
        col = ambient * tex + k * light_col * tex
    else:
        col = ambient * tex + shadow * k * tex * light_col
    return col, vis["keep"] & vis["front"]
