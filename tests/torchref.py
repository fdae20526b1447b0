"""Float64 torch restatement of the reference's per-pixel forward with the
chosen primitive per pixel given (frozen visibility), for gradient checks.

Written from renderer/pipeline.py:163-328 and the shader files; torch autograd
on it gives reference gradients with visibility held piecewise constant, the
same semantics jax.grad has through the reference (comparisons, argmin and
floor carry no gradient).  Pixels whose float32 decisions (texel index,
branch, shadow test) sit within a hair of a boundary are reported in
`ambiguous` so tests can give them zero weight.
"""

from __future__ import annotations

import numpy as np
import torch

from jaxrenderer_amd import _hip


def N(v):
    return v / torch.linalg.norm(v, dim=-1, keepdim=True)


def apply(M, p, w):
    """Camera.apply on [..., 3] points with homogeneous coordinate w."""
    ph = torch.cat([p, torch.full_like(p[..., :1], w)], dim=-1)
    return ph @ M.T


def near_int(v, tol=1e-3):
    f = v - torch.floor(v)
    return (f < tol) | (f > 1 - tol)


def forward(shader: int, W: int, H: int, prim_id: np.ndarray, faces: np.ndarray, P: dict):
    """P holds float64 tensors: pos, normal, colour, uv, texture, spec, w2c, vp,
    eye, light_dir, light_col, light_dir_eye, ambient, diffuse, specular,
    strength, tex_shapes, tex_index, tex_offset, shadow_map, s_w2c, s_vp.
    Returns canvas [W, H, 3], zbuffer [W, H] (kept pixels only, zeros
    elsewhere) and an `ambiguous` mask."""
    pid = torch.as_tensor(prim_id, dtype=torch.int64)
    ii, jj = torch.nonzero(pid >= 0, as_tuple=True)
    t = pid[ii, jj]
    fc = torch.as_tensor(faces, dtype=torch.int64)[t]  # [K, 3] slots
    vp = P["vp"]
    dt = P["pos"].dtype  # float64 for reference gradients; float32 to measure rounding
    x = (ii.to(dt) - vp[0, 3]) / vp[0, 0]
    y = (jj.to(dt) - vp[1, 3]) / vp[1, 1]
    clip = apply(P["w2c"], P["pos"][fc], 1.0)  # [K, 3, 4]
    M = torch.stack([clip[..., 0], clip[..., 1], clip[..., 3]], dim=-1)
    Mi = torch.linalg.inv(M)
    xy1 = torch.stack([x, y, torch.ones_like(x)], dim=-1)
    e = torch.einsum("kr,krc->kc", xy1, Mi)
    b = e / e.sum(-1, keepdim=True)
    z = (e * clip[..., 2]).sum(-1) * vp[2, 2] + vp[2, 3]
    amb = torch.zeros_like(x, dtype=torch.bool)
    K = len(t)
    col = torch.zeros((K, 3), dtype=dt)
    if shader in (_hip.SHADER_GOURAUD, _hip.SHADER_GOURAUD_TEXTURE):
        L = N(P["light_dir"])
        inten = (N(P["normal"][fc]) * L).sum(-1, keepdim=True)  # [K, 3, 1]
        if shader == _hip.SHADER_GOURAUD:
            vcol = P["colour"][fc] * P["light_col"] * inten
            col = (b[..., None] * vcol).sum(1)
            amb |= (col.abs() < 1e-4).any(-1)
        else:
            lc = P["light_col"] * inten
            light = (b[..., None] * lc).sum(1)
            uv = (b[..., None] * P["uv"][fc]).sum(1)
            tw, th = P["texture"].shape[:2]
            amb |= near_int(uv).any(-1) | (light.abs() < 1e-4).any(-1)
            u = torch.remainder(torch.floor(uv[:, 0]).long(), tw)
            v = torch.remainder(torch.floor(uv[:, 1]).long(), th)
            col = P["texture"][u, v] * light
    elif shader == _hip.SHADER_PHONG:
        ne = N(apply(P["eye"], N(N(P["normal"][fc])), 0.0)[..., :3])
        n = (b[..., None] * ne).sum(1)
        uv = (b[..., None] * P["uv"][fc]).sum(1)
        tw, th = P["texture"].shape[:2]
        u = torch.remainder(torch.floor(uv[:, 0]).long(), tw)
        v = torch.remainder(torch.floor(uv[:, 1]).long(), th)
        d = (N(n) * N(P["light_dir"])).sum(-1, keepdim=True)
        lc = P["light_col"] * d
        pos = (lc >= 0).all(-1, keepdim=True)
        amb |= near_int(uv).any(-1) | (lc.abs() < 1e-4).any(-1)
        col = torch.where(pos, P["texture"][u, v] * lc, torch.zeros_like(lc))
    elif shader == _hip.SHADER_PHONG_DARBOUX:
        # renderer/shaders/phong_darboux.py:116-293: the triangle of the face
        # of the primitive's first vertex, in NDC, and its uvs; a tangent frame
        # from inv([p1 - p0; p2 - p0; n]) and the normal map in that frame
        ne = N(apply(P["eye"], N(N(P["normal"][fc])), 0.0)[..., :3])
        n = (b[..., None] * ne).sum(1)
        uv = (b[..., None] * P["uv"][fc]).sum(1)
        tw, th = P["texture"].shape[:2]
        u = torch.remainder(torch.floor(uv[:, 0]).long(), tw)
        v = torch.remainder(torch.floor(uv[:, 1]).long(), th)
        fs = P["faces_indices"][P["id_to_face"][fc[:, 0]]]  # [K, 3] slots
        tc = apply(P["w2c"], P["pos"][fs], 1.0)
        w = tc[..., 3:]
        tri = torch.where(w == 0, tc[..., :3], tc[..., :3] / torch.where(w == 0, 1.0, w))
        tuv = P["uv"][fs]
        nh = N(n)
        A = torch.stack([tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0], nh], dim=1)
        AI = torch.linalg.inv(A)
        zero = torch.zeros_like(uv[:, 0])
        du = torch.stack([tuv[:, 1, 0] - tuv[:, 0, 0], tuv[:, 2, 0] - tuv[:, 0, 0], zero], -1)
        dv = torch.stack([tuv[:, 1, 1] - tuv[:, 0, 1], tuv[:, 2, 1] - tuv[:, 0, 1], zero], -1)
        ti = (AI @ du[..., None])[..., 0]
        tj = (AI @ dv[..., None])[..., 0]
        B = torch.stack([N(ti), N(tj), nh], dim=-1)  # columns
        nm = P["normal_map"]
        nmv = nm[u.clamp(0, nm.shape[0] - 1), v.clamp(0, nm.shape[1] - 1)]
        n2 = N((B @ nmv[..., None])[..., 0])
        d = (n2 * N(P["light_dir"])).sum(-1, keepdim=True)
        lc = P["light_col"] * d
        pos = (lc >= 0).all(-1, keepdim=True)
        # float32 rounding grows with the condition number of the frame matrix
        cond = torch.linalg.cond(A)
        amb |= near_int(uv).any(-1) | (lc.abs() < 1e-4).any(-1) | ~(cond < 1e3)
        col = torch.where(pos, P["texture"][u, v] * lc, torch.zeros_like(lc))
    elif shader in (_hip.SHADER_PHONG_REFLECTION, _hip.SHADER_PHONG_REFLECTION_SHADOW):
        ne = N(apply(P["eye"], N(N(P["normal"][fc])), 0.0)[..., :3])
        n = (b[..., None] * ne).sum(1)
        uv = (b[..., None] * P["uv"][fc]).sum(1)
        ti = P["tex_index"][fc[:, 0]]
        shp = P["tex_shapes"][ti.clamp(0, len(P["tex_shapes"]) - 1)]
        fr = uv - torch.where(uv >= 0, torch.floor(uv), torch.ceil(uv))
        fr = torch.where(fr < 0, fr + 1, fr)
        shp = shp.to(dt)
        r0 = fr[:, 0] * shp[:, 0] + (ti * int(P["tex_offset"])).to(dt)
        r1 = fr[:, 1] * shp[:, 1]
        amb |= near_int(r0) | near_int(r1)
        tw, th = P["texture"].shape[:2]
        u = torch.floor(r0).long().clamp(0, tw - 1)
        v = torch.floor(r1).long().clamp(0, th - 1)
        tex = P["texture"][u, v]
        S = P["spec"][u.clamp(0, P["spec"].shape[0] - 1), v.clamp(0, P["spec"].shape[1] - 1)]
        nh = N(n)
        lh = N(P["light_dir_eye"])
        d = (nh * lh).sum(-1)
        diff = torch.clamp(d, min=0.0)
        r = N(2 * d[:, None] * nh - lh)
        m = torch.clamp(r[:, 2], min=0.0)
        spec = m ** S
        amb |= (d.abs() < 1e-4) | (r[:, 2].abs() < 1e-3)
        k = P["diffuse"] * diff[:, None] + P["specular"] * spec[:, None]
        if shader == _hip.SHADER_PHONG_REFLECTION:
            col = P["ambient"] * tex + k * P["light_col"] * tex
        else:
            sh4 = apply(P["s_w2c"], P["pos"][fc], 1.0)
            sh4 = sh4 / sh4[..., 3:]
            s4 = (b[..., None] * sh4).sum(1)
            sv = s4 @ P["s_vp"].T
            sv = sv / sv[:, 3:]
            smap = P["shadow_map"]
            # lax.round rounds half away from zero
            ix = (torch.sign(sv[:, 0]) * torch.floor(sv[:, 0].abs() + 0.5)).long()
            iy = (torch.sign(sv[:, 1]) * torch.floor(sv[:, 1].abs() + 0.5)).long()
            ix = torch.where(ix < 0, ix + smap.shape[0], ix)
            iy = torch.where(iy < 0, iy + smap.shape[1], iy)
            ok = (ix >= 0) & (ix < smap.shape[0]) & (iy >= 0) & (iy < smap.shape[1])
            depth = torch.where(ok, smap[ix.clamp(0, smap.shape[0] - 1),
                                         iy.clamp(0, smap.shape[1] - 1)],
                                torch.full_like(sv[:, 0], float("inf")))
            lit = sv[:, 2] <= depth
            amb |= near_int(sv[:, 0] + 0.5) | near_int(sv[:, 1] + 0.5) | \
                ((sv[:, 2] - depth).abs() < 1e-4)
            shade = torch.where(lit[:, None], torch.ones_like(P["strength"]).expand(K, 3),
                                (1.0 - P["strength"]).expand(K, 3))
            col = P["ambient"] * tex + shade * k * tex * P["light_col"]
    canvas = torch.zeros((W, H, 3), dtype=dt)
    zbuf = torch.zeros((W, H), dtype=dt)
    canvas = canvas.index_put((ii, jj), col)
    zbuf = zbuf.index_put((ii, jj), z)
    ambiguous = torch.zeros((W, H), dtype=torch.bool)
    ambiguous[ii, jj] = amb
    return canvas, zbuf, ambiguous
