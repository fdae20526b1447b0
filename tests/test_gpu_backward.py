"""jr_backward (hand-written HIP backward) against reference gradients.

Reference 1: torch autograd on a float64 restatement of the reference's
per-pixel forward with the GPU's chosen primitives frozen (tests/torchref.py),
the semantics jax.grad has through renderer/pipeline.py.  Pixels whose float32
decisions sit next to a boundary (texel index, branch, shadow test) get zero
weight in the loss.  Tolerance: 1e-4 relative (vector 2-norm per parameter).

Reference 2: central finite differences (h = 1e-3) of the GPU forward itself
for parameters that cannot move visibility (texture values, light colour,
material coefficients, shadow strength)."""
import numpy as np
import pytest
import torch

from jaxrenderer_amd import _hip, _plan
from jaxrenderer_amd.autograd import render_differentiable
from jaxrenderer_amd.geometry import Camera
from jaxrenderer_amd.pipeline import build_call
from jaxrenderer_amd.shaders import (
    DepthShader,
    GouraudShader,
    GouraudTextureShader,
    PhongReflectionShadowTextureShader,
    PhongReflectionTextureShader,
    PhongTextureDarbouxShader,
    PhongTextureShader,
)
from jaxrenderer_amd.types import LightSource
from tests import scenes, torchref

pytestmark = pytest.mark.gpu

SHADERS = [DepthShader, GouraudShader, GouraudTextureShader, PhongTextureShader,
           PhongTextureDarbouxShader, PhongReflectionTextureShader,
           PhongReflectionShadowTextureShader]
W, H = 48, 40


def _shadow_map(shadow_call):
    return _plan.run(shadow_call)["zbuffer"][0].cpu().numpy()


def _setup(shader, seed, w=W, h=H):
    cam, faces, extra, shadow_call = scenes.soup_inputs(shader, seed, w, h)
    if hasattr(extra, "specular_map"):
        # exponents below 1 make pow's derivative at 0 infinite and jax.grad's
        # (and this backward's) gradient NaN wherever the specular term is 0
        extra = extra._replace(specular_map=(1.0 + 0.75 * extra.specular_map).astype(np.float32))
    if shadow_call is not None:
        extra = extra._replace(shadow=extra.shadow._replace(shadow_map=_shadow_map(shadow_call)))
    return cam, faces, extra


def _to(x, dtype, dev, grad=True):
    t = torch.as_tensor(np.asarray(x), dtype=dtype, device=dev).clone()
    return t.requires_grad_(grad)


GRAD_FIELDS = ("position", "normal", "colour", "texture", "specular_map", "light_dir_eye",
               "ambient", "diffuse", "specular", "uv", "normal_map")


def _leaves(cam, extra, dtype, dev):
    """Copies of every differentiable input as leaves (dict name -> tensor),
    plus camera / extra NamedTuples built from them."""
    L = {}
    for f in GRAD_FIELDS:
        if hasattr(extra, f):
            L[f] = _to(getattr(extra, f), dtype, dev)
    if hasattr(extra, "light"):
        L["light_direction"] = _to(extra.light.direction, dtype, dev)
        L["light_colour"] = _to(extra.light.colour, dtype, dev)
    for f in ("world_to_clip", "viewport", "world_to_eye_norm"):
        L[f] = _to(getattr(cam, f), dtype, dev)
    if hasattr(extra, "shadow"):
        L["shadow_strength"] = _to(extra.shadow.strength, dtype, dev)
    ex = {f: (L[f] if f in L else getattr(extra, f)) for f in extra._fields}
    if "light" in ex:
        ex["light"] = LightSource(direction=L["light_direction"], colour=L["light_colour"])
    if "shadow" in ex:
        ex["shadow"] = extra.shadow._replace(strength=L["shadow_strength"])
    cam2 = cam._replace(world_to_clip=L["world_to_clip"], viewport=L["viewport"],
                        world_to_eye_norm=L["world_to_eye_norm"])
    return L, cam2, type(extra)(**ex)


def _ref_params(L, extra, shadow_cam):
    P = {"pos": L["position"], "w2c": L["world_to_clip"], "vp": L["viewport"],
         "eye": L["world_to_eye_norm"]}
    for k, f in (("normal", "normal"), ("colour", "colour"), ("texture", "texture"),
                 ("spec", "specular_map"), ("light_dir_eye", "light_dir_eye"),
                 ("ambient", "ambient"), ("diffuse", "diffuse"), ("specular", "specular"),
                 ("light_dir", "light_direction"), ("light_col", "light_colour"),
                 ("strength", "shadow_strength")):
        if f in L:
            P[k] = L[f]
    if hasattr(extra, "uv"):
        P["uv"] = L["uv"]
    if hasattr(extra, "normal_map"):
        P["normal_map"] = L["normal_map"]
        P["id_to_face"] = torch.as_tensor(np.asarray(extra.id_to_face), dtype=torch.int64)
        P["faces_indices"] = torch.as_tensor(np.asarray(extra.faces_indices), dtype=torch.int64)
    if hasattr(extra, "texture_shape"):
        P["tex_shapes"] = torch.as_tensor(np.asarray(extra.texture_shape), dtype=torch.int64)
        P["tex_index"] = torch.as_tensor(np.asarray(extra.texture_index), dtype=torch.int64)
        P["tex_offset"] = int(extra.texture_offset)
    if shadow_cam is not None:
        P["shadow_map"] = torch.as_tensor(np.asarray(extra.shadow.shadow_map), dtype=torch.float64)
        P["s_w2c"] = torch.as_tensor(shadow_cam.world_to_clip, dtype=torch.float64)
        P["s_vp"] = torch.as_tensor(shadow_cam.viewport, dtype=torch.float64)
    return P


@pytest.mark.parametrize("shader", SHADERS, ids=lambda s: s.__name__)
@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_float64_autograd(shader, seed):
    _check_against_float64(shader, seed, W, H)


# The backward's wavefronts take the full 16 x 16 tiles first, then the 2x2
# pixel blocks below them and right of them packed together: odd frames have
# both kinds of leftover blocks and blocks cut by the frame's edge.
@pytest.mark.parametrize("shader", [GouraudShader, PhongReflectionShadowTextureShader],
                         ids=lambda s: s.__name__)
@pytest.mark.parametrize("size", [(37, 29), (51, 39)], ids=lambda s: f"{s[0]}x{s[1]}")
def test_backward_on_odd_frame_sizes(shader, size):
    _check_against_float64(shader, 0, *size)


def _check_against_float64(shader, seed, W, H):
    dev = torch.device("cuda")
    cam, faces, extra = _setup(shader, seed, W, H)
    call = build_call(cam, shader, faces, extra, W, H, init_buffers=True)
    pid = _plan.run(call, prim_id=True)["prim_id"][0].cpu().numpy()
    assert (pid >= 0).sum() > 50
    L64, _, _ = _leaves(cam, extra, torch.float64, torch.device("cpu"))
    shadow_cam = extra.shadow.camera if hasattr(extra, "shadow") else None
    P = _ref_params(L64, extra, shadow_cam)
    canvas64, z64, amb = torchref.forward(shader._jr_shader_id, W, H, pid, faces, P)
    rng = np.random.default_rng(100 + seed)
    wc = torch.as_tensor(rng.normal(size=(W, H, 3)))
    wz = torch.as_tensor(rng.normal(size=(W, H)))
    keep = torch.as_tensor(pid >= 0) & ~amb
    wc = wc * keep[..., None]
    wz = wz * keep
    if shader is DepthShader:
        wc = wc * 0
    loss64 = (wc * canvas64).sum() + (wz * z64).sum()
    names = list(L64)
    g64 = torch.autograd.grad(loss64, [L64[n] for n in names], allow_unused=True)

    L32, cam32, extra32 = _leaves(cam, extra, torch.float32, dev)
    zb, cv = render_differentiable(cam32, shader, faces, extra32, W, H)
    loss32 = (wc.to(dev).float() * cv).sum() + (wz.to(dev).float() * zb).sum()
    g32 = torch.autograd.grad(loss32, [L32[n] for n in names], allow_unused=True)
    checked = 0
    for n, a, b in zip(names, g32, g64):
        ref = torch.zeros_like(L64[n]) if b is None else b
        got = torch.zeros_like(L64[n]) if a is None else a.detach().double().cpu()
        scale = float(ref.norm())
        err = float((got - ref).norm())
        if scale < 1e-8:
            assert err < 1e-6, f"{n}: expected ~0 gradient, got norm {err}"
            continue
        checked += 1
        assert err <= 1e-4 * scale, f"{n}: relative error {err / scale:.3e} (|ref|={scale:.3e})"
    assert checked >= 3


def _loss_of(shader, cam, faces, extra, w):
    call = build_call(cam, shader, faces, extra, W, H, init_buffers=True)
    cv = _plan.run(call)["canvas"][0].double().cpu().numpy()
    return float((w * cv).sum())


@pytest.mark.parametrize("shader", [GouraudTextureShader, PhongReflectionShadowTextureShader],
                         ids=lambda s: s.__name__)
def test_backward_matches_finite_differences(shader):
    """Central differences (h=1e-3) of the GPU forward for parameters that do
    not move visibility."""
    dev = torch.device("cuda")
    cam, faces, extra = _setup(shader, 3)
    w = np.random.default_rng(10).normal(size=(W, H, 3))
    L32, cam32, extra32 = _leaves(cam, extra, torch.float32, dev)
    zb, cv = render_differentiable(cam32, shader, faces, extra32, W, H)
    loss = (torch.as_tensor(w, device=dev).float() * cv).sum()
    names = [n for n in ("texture", "light_colour", "ambient", "diffuse", "specular",
                         "shadow_strength") if n in L32]
    grads = dict(zip(names, torch.autograd.grad(loss, [L32[n] for n in names])))
    rng = np.random.default_rng(11)
    h = 1e-3
    n_checked = 0
    for n in names:
        g = grads[n].detach().cpu().numpy().reshape(-1)
        flat_idx = rng.choice(g.size, size=min(16, g.size), replace=False)
        for k in flat_idx:
            def perturbed(delta):
                if n == "texture":
                    t = np.array(extra.texture, dtype=np.float32)
                    t.reshape(-1)[k] += delta
                    return cam, extra._replace(texture=t)
                if n == "light_colour":
                    c = np.array(extra.light.colour, dtype=np.float32)
                    c[k] += delta
                    return cam, extra._replace(light=LightSource(extra.light.direction, c))
                if n == "shadow_strength":
                    s = np.array(extra.shadow.strength, dtype=np.float32)
                    s[k] += delta
                    return cam, extra._replace(shadow=extra.shadow._replace(strength=s))
                v = np.array(getattr(extra, n), dtype=np.float32)
                v[k] += delta
                return cam, extra._replace(**{n: v})
            lp = _loss_of(shader, cam, faces, perturbed(h)[1], w)
            lm = _loss_of(shader, cam, faces, perturbed(-h)[1], w)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[k]) <= 2e-3 * max(1.0, abs(fd)), f"{n}[{k}]: fd {fd} vs {g[k]}"
            n_checked += 1
    assert n_checked >= 15
