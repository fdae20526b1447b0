"""Multi-view differentiation (the 64-camera workload, at test sizes): one
batched forward/backward over all views sharing the mesh and texture.

* Batched views reproduce per-view renders bit for bit (depth) and within
  1e-5 (colour).
* Gradients of the shared inputs equal the sum of the per-view gradients
  (1e-5 relative: only the order of the atomic sums differs); per-view
  camera gradients equal the single-view ones.
* Central differences (h = 1e-3) of the L2 image loss for randomly chosen
  texel values, per-vertex colours, light direction and (untextured shader)
  vertex coordinates.  Parameters whose +-h moves any primitive id are
  excluded (the gradient is that of the piecewise-constant visibility, as
  under jax.grad) and counted; the rest must agree with the analytic
  gradient.
"""
import numpy as np
import pytest
import torch

from jaxrenderer_amd import workloads
from jaxrenderer_amd.autograd import (build_views_call, render_differentiable,
                                      render_views_differentiable)
from jaxrenderer_amd import _plan
from jaxrenderer_amd.shaders import (GouraudExtraInput, GouraudShader, GouraudTextureExtraInput,
                                     GouraudTextureShader, PhongTextureExtraInput,
                                     PhongTextureShader)
from jaxrenderer_amd.types import LightSource

pytestmark = pytest.mark.gpu
f32 = np.float32


def _scene(shader, n_views=4, size=64, tex=64, perturbed=False):
    cams, faces, v, n, uv, col, texture = workloads.config4(
        seed=5, n_views=n_views, size=size, tex_size=tex, n_lat=20, n_lon=24, perturbed=perturbed)
    light = LightSource(direction=np.array((0.3, -0.5, 0.8), f32), colour=np.ones(3, f32))
    if shader is GouraudShader:
        extra = GouraudExtraInput(position=v, colour=col, normal=n, light=light)
    elif shader is GouraudTextureShader:
        extra = GouraudTextureExtraInput(position=v, normal=n, uv=uv, light=light, texture=texture)
    else:
        extra = PhongTextureExtraInput(position=v, normal=n, uv=uv, light=light, texture=texture)
    return cams, faces, extra


def _leaves(cams, extra, dev):
    def t(x):
        return torch.tensor(np.asarray(x, f32), device=dev, requires_grad=True)

    light = LightSource(direction=t(extra.light.direction), colour=t(extra.light.colour))
    ex = extra._replace(**{k: t(getattr(extra, k)) for k in ("position", "normal", "uv",
                                                              "texture", "colour")
                           if hasattr(extra, k)}, light=light)
    cams_t = [type(c)(*(t(x) for x in c)) for c in cams]
    return cams_t, ex


SHADERS = [GouraudShader, GouraudTextureShader, PhongTextureShader]


@pytest.mark.parametrize("shader", SHADERS, ids=lambda s: s.__name__)
def test_views_match_single_renders(shader):
    dev = torch.device("cuda")
    cams, faces, extra = _scene(shader)
    W = H = 64
    w = torch.as_tensor(np.random.default_rng(0).normal(size=(len(cams), W, H, 3)), device=dev).float()

    cams_t, ex = _leaves(cams, extra, dev)
    zb, cv = render_views_differentiable(cams_t, shader, faces, ex, W, H)
    loss = (w * cv).sum()
    names = [k for k in ("position", "normal", "texture", "colour") if hasattr(ex, k)]
    leaves = [getattr(ex, k) for k in names] + [ex.light.direction]
    g_batch = torch.autograd.grad(loss, leaves + [c.world_to_clip for c in cams_t])

    total = [torch.zeros_like(x) for x in leaves]
    for b in range(len(cams)):
        cams_s, ex_s = _leaves(cams[b:b + 1], extra, dev)
        zs, cs = render_differentiable(cams_s[0], shader, faces, ex_s, W, H)
        assert torch.equal(zs.view(torch.int32), zb[b].view(torch.int32))
        assert float((cs - cv[b]).abs().max()) <= 1e-5
        ls = (w[b] * cs).sum()
        ll = [getattr(ex_s, k) for k in names] + [ex_s.light.direction]
        gs = torch.autograd.grad(ls, ll + [cams_s[0].world_to_clip])
        for i in range(len(leaves)):
            total[i] += gs[i]
        gb_cam, gs_cam = g_batch[len(leaves) + b], gs[-1]
        assert float((gb_cam - gs_cam).norm()) <= 1e-5 * max(float(gs_cam.norm()), 1e-3)
    for name, a, b in zip(names + ["light_direction"], g_batch[:len(leaves)], total):
        scale = max(float(b.norm()), 1e-6)
        assert float((a - b).norm()) <= 1e-5 * scale, name


def _loss_np(cams, shader, faces, extra, target, W, H):
    call = build_views_call(cams, shader, faces, extra, W, H, init_buffers=True)
    out = _plan.run(call, prim_id=True)
    cv = out["canvas"].double().cpu().numpy()
    return float(((cv - target) ** 2).sum()), out["prim_id"].cpu().numpy()


@pytest.mark.parametrize("shader", SHADERS, ids=lambda s: s.__name__)
def test_views_finite_differences(shader):
    dev = torch.device("cuda")
    W = H = 48
    cams, faces, extra = _scene(shader, n_views=6, size=W, tex=32)
    _, _, target_extra = _scene(shader, n_views=6, size=W, tex=32, perturbed=True)
    tgt_call = build_views_call(cams, shader, faces, target_extra, W, H, init_buffers=True)
    target = _plan.run(tgt_call)["canvas"].double().cpu().numpy()

    cams_t, ex = _leaves(cams, extra, dev)
    zb, cv = render_views_differentiable(cams_t, shader, faces, ex, W, H)
    loss = ((cv - torch.as_tensor(target, device=dev).float()) ** 2).sum()
    # vertex coordinates only where no texture is read: moving a vertex also
    # moves the interpolated uv, and where that crosses a texel edge the loss
    # jumps (floor has no derivative, as under jax.grad), which central
    # differences would pick up
    names = [k for k in ("texture", "colour") if hasattr(ex, k)]
    if not hasattr(ex, "texture"):
        names.append("position")
    grads = dict(zip(names + ["light_direction"],
                     torch.autograd.grad(loss, [getattr(ex, k) for k in names]
                                         + [ex.light.direction])))
    _, pid0 = _loss_np(cams, shader, faces, extra, target, W, H)

    rng = np.random.default_rng(1)
    h = 1e-3
    checked = excluded = 0
    for name in grads:
        g = grads[name].detach().cpu().numpy().reshape(-1)
        base = (np.asarray(extra.light.direction if name == "light_direction"
                           else getattr(extra, name), f32))
        # only parameters that touch the image: nonzero gradient or a seen texel
        cand = np.flatnonzero(np.abs(g) > 0) if name != "light_direction" else np.arange(3)
        if cand.size == 0:
            continue
        for k in rng.choice(cand, size=min(40, cand.size), replace=False):
            vals = []
            moved = False
            for d in (h, -h):
                arr = base.copy()
                arr.reshape(-1)[k] += d
                e = (extra._replace(light=LightSource(arr, extra.light.colour))
                     if name == "light_direction" else extra._replace(**{name: arr}))
                lv, pid = _loss_np(cams, shader, faces, e, target, W, H)
                moved |= not np.array_equal(pid, pid0)
                vals.append(lv)
            if moved:
                excluded += 1
                continue
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[k]) <= 2e-2 * max(1.0, abs(fd)), f"{name}[{k}]: fd {fd} vs {g[k]}"
            checked += 1
    assert checked >= 25, (checked, excluded)
