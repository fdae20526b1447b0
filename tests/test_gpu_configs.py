"""The benchmark workloads at their full stated sizes, on the GPU, against the
CPU oracle (primitive ids, depths and barycentrics bit for bit, colours within
1e-5) and, for the differentiation workload, against float64 autograd with
the GPU's primitives frozen (1e-4 relative per parameter tensor on pixels of
well-conditioned triangles; on all pixels 1e-3, or within 1.5 times the
distance of a float32 evaluation of the same formulas).

* Config 2: the 48 x 96 UV sphere (9,216 triangles, pole triangles
  degenerate) with 256x256 maps, phong_reflection, 960x540.
* Config 4: 64 ring views of the deformed 5,000-triangle sphere at 256x256
  with a 512x512x3 texture, gouraud_texture and phong: four views against
  the oracle, all 64 against their own single-view renders, and the backward
  of two full-size views against float64 autograd.
* Config 5: scenes at both ends of the 1k..20k-triangle range at 480x270
  with 512x512 shadow maps, phong_reflection_shadow, batched in one call.
"""
import numpy as np
import pytest
import torch

from jaxrenderer_amd import _plan, batch, workloads
from jaxrenderer_amd.autograd import build_views_call, render_differentiable
from jaxrenderer_amd.shaders import (GouraudShader, GouraudExtraInput, GouraudTextureExtraInput,
                                     GouraudTextureShader, PhongTextureExtraInput,
                                     PhongTextureShader)
from jaxrenderer_amd.types import LightSource
from oracle.oracle import run_oracle, run_oracle_plan
from tests import torchref
from tests.parity import as_np, assert_parity

pytestmark = pytest.mark.gpu
f32 = np.float32


def _check_plan(calls_fn, label, **kw):
    print(f"configs: {label}", flush=True)
    ref = run_oracle_plan(calls_fn())
    got = _plan.run_plan(calls_fn(), prim_id=True, bary=True, **kw)
    torch.cuda.synchronize()
    for tag in ref:
        assert_parity(got[tag], ref[tag], label=f"{label}/{tag}")
    return got, ref


@pytest.mark.parametrize("tile,ppl", [((8, 8), 1), ((16, 16), 4)], ids=str)
def test_config2_sphere_960x540(tile, ppl):
    scene = workloads.config2(seed=0)
    assert len(scene.model.faces) == 9216
    got, _ = _check_plan(lambda: batch.plan_batch([scene], 960, 540), f"c2/{tile}",
                         tile=tile, pixels_per_lane=ppl)
    assert int((as_np(got["main"]["prim_id"]) >= 0).sum()) > 50_000


C4_LIGHT = LightSource(direction=np.array((0.3, -0.5, 0.8), f32), colour=np.ones(3, f32))


def _c4(shader):
    cams, faces, v, n, uv, col, tex = workloads.config4(seed=0)
    assert len(cams) == 64 and len(faces) == 5000 and tex.shape == (512, 512, 3)
    if shader is GouraudShader:
        extra = GouraudExtraInput(position=v, colour=col, normal=n, light=C4_LIGHT)
    elif shader is GouraudTextureShader:
        extra = GouraudTextureExtraInput(position=v, normal=n, uv=uv, light=C4_LIGHT, texture=tex)
    else:
        extra = PhongTextureExtraInput(position=v, normal=n, uv=uv, light=C4_LIGHT, texture=tex)
    return cams, faces, extra


C4_SHADERS = [GouraudTextureShader, PhongTextureShader]


@pytest.mark.parametrize("shader", C4_SHADERS, ids=lambda s: s.__name__)
def test_config4_64_views_forward(shader):
    cams, faces, extra = _c4(shader)
    S = 256
    print(f"configs: c4 forward {shader.__name__}", flush=True)
    allv = _plan.run(build_views_call(cams, shader, faces, extra, S, S, init_buffers=True),
                     prim_id=True, bary=True)
    torch.cuda.synchronize()
    sel = [0, 17, 33, 50]
    ref = run_oracle(build_views_call([cams[b] for b in sel], shader, faces, extra, S, S,
                                      init_buffers=True))
    got = {k: as_np(v[sel]).copy() for k, v in allv.items()}
    # primitive ids count the faces of every view before this one: compare the
    # face row within the view
    F = len(faces)
    for i, b in enumerate(sel):
        g, r = got["prim_id"][i], ref["prim_id"][i]
        got["prim_id"][i] = np.where(g >= 0, g - b * F, -1)
        ref["prim_id"][i] = np.where(r >= 0, r - i * F, -1)
    assert_parity(got, ref, label=f"c4/{shader.__name__}/oracle")
    # every view against its own single-view render
    for b in range(len(cams)):
        one = _plan.run(build_views_call([cams[b]], shader, faces, extra, S, S,
                                         init_buffers=True), prim_id=True)
        pb = allv["prim_id"][b]
        assert torch.equal(one["prim_id"][0], torch.where(pb >= 0, pb - b * F, pb)), b
        assert torch.equal(one["zbuffer"][0].view(torch.int32),
                           allv["zbuffer"][b].view(torch.int32)), b
        assert float((one["canvas"][0] - allv["canvas"][b]).abs().max()) <= 1e-5, b
    assert int((as_np(allv["prim_id"]) >= 0).sum()) > 64 * 10_000


def _leaves(cam, extra, dtype, dev):
    def t(x):
        return torch.as_tensor(np.asarray(x), dtype=dtype, device=dev).clone().requires_grad_(True)

    # phong_darboux's image depends on the uvs and the normal map continuously
    keys = ("position", "normal", "colour", "texture") + (
        ("uv", "normal_map") if hasattr(extra, "normal_map") else ())
    L = {k: t(getattr(extra, k)) for k in keys if hasattr(extra, k)}
    L["light_direction"] = t(extra.light.direction)
    L["light_colour"] = t(extra.light.colour)
    for k in ("world_to_clip", "viewport", "world_to_eye_norm"):
        L[k] = t(getattr(cam, k))
    ex = extra._replace(**{k: L[k] for k in keys if k in L},
                        light=LightSource(direction=L["light_direction"],
                                          colour=L["light_colour"]))
    cam2 = cam._replace(world_to_clip=L["world_to_clip"], viewport=L["viewport"],
                        world_to_eye_norm=L["world_to_eye_norm"])
    return L, cam2, ex


def _ref_forward(shader, cam, faces, extra, pid, dtype):
    """torchref's forward with leaves of `dtype` on the CPU."""
    L, _, _ = _leaves(cam, extra, dtype, torch.device("cpu"))
    P = {"pos": L["position"], "w2c": L["world_to_clip"], "vp": L["viewport"],
         "eye": L["world_to_eye_norm"], "normal": L["normal"],
         "light_dir": L["light_direction"], "light_col": L["light_colour"]}
    if "texture" in L:
        P["texture"] = L["texture"]
        P["uv"] = L["uv"] if "uv" in L else torch.as_tensor(np.asarray(extra.uv), dtype=dtype)
    if "normal_map" in L:
        P["normal_map"] = L["normal_map"]
        P["id_to_face"] = torch.as_tensor(np.asarray(extra.id_to_face), dtype=torch.int64)
        P["faces_indices"] = torch.as_tensor(np.asarray(extra.faces_indices), dtype=torch.int64)
    if "colour" in L:
        P["colour"] = L["colour"]
    S = pid.shape[0]
    return (L,) + torchref.forward(shader._jr_shader_id, S, S, pid, np.asarray(faces), P)


def _grads(L, canvas, z, wc, wz):
    names = list(L)
    loss = (wc.to(canvas) * canvas).sum() + (wz.to(z) * z).sum()
    g = torch.autograd.grad(loss, [L[n] for n in names], allow_unused=True, retain_graph=True)
    return {n: (torch.zeros_like(L[n]) if x is None else x).detach().double().cpu()
            for n, x in zip(names, g)}


@pytest.mark.parametrize("shader", C4_SHADERS + [GouraudShader], ids=lambda s: s.__name__)
@pytest.mark.parametrize("view", [5, 40])
def test_config4_backward_full_size_against_float64(shader, view):
    """One full-size view (256x256, 5,000 triangles, 512x512 texture) of the
    differentiation workload: jr_backward's gradients of vertex positions,
    normals, per-vertex colours (gouraud), texture, light and camera against
    torch autograd of the float64 restatement with the GPU's primitive ids
    frozen.  Pixels whose float32 decisions sit at a boundary carry no loss
    weight.

    Two bars.  (1) 1e-4 relative per tensor where the loss weighs only pixels
    of triangles whose M = [x y w] has a float64 condition number below 300
    (97-98 % of the covered pixels).  (2) Over every pixel, 1e-3 relative, or
    no further from float64 than 1.5 times what the same formulas evaluated
    in float32 (torch, the reference's arithmetic type) land: through
    d(M^-1) = -M^-T G M^-T the rounding grows with cond(M)^2, and a few
    grazing triangles (condition numbers up to 8e4) put the float32 position
    gradient of a view 5e-4 to 4e-3 away from float64 in a float32
    implementation, by amounts that depend on the order of its operations."""
    cams, faces, extra = _c4(shader)
    cam = cams[view]
    S = 256
    print(f"configs: c4 backward {shader.__name__} view {view}", flush=True)
    dev = torch.device("cuda")
    call = build_views_call([cam], shader, faces, extra, S, S, init_buffers=True)
    pid = _plan.run(call, prim_id=True)["prim_id"][0].cpu().numpy()
    assert (pid >= 0).sum() > 10_000
    L64, canvas64, z64, amb = _ref_forward(shader, cam, faces, extra, pid, torch.float64)
    L32c, canvas32, z32, _ = _ref_forward(shader, cam, faces, extra, pid, torch.float32)
    w2c = np.asarray(cam.world_to_clip, np.float64)
    clip = np.concatenate([np.asarray(extra.position, np.float64),
                           np.ones((len(extra.position), 1))], axis=1) @ w2c.T
    cond = np.linalg.cond(clip[np.asarray(faces)][..., [0, 1, 3]])
    pix_cond = np.where(pid >= 0, cond[np.maximum(pid, 0)], np.inf)
    rng = np.random.default_rng(300 + view)
    wc_all = torch.as_tensor(rng.normal(size=(S, S, 3)))
    wz_all = torch.as_tensor(rng.normal(size=(S, S)))
    checked = 0
    for label, mask in (("well-conditioned", (pid >= 0) & (pix_cond < 300)),
                        ("all", pid >= 0)):
        keep = torch.as_tensor(mask) & ~amb
        wc, wz = wc_all * keep[..., None], wz_all * keep
        g64 = _grads(L64, canvas64, z64, wc, wz)
        g32c = _grads(L32c, canvas32, z32, wc, wz)
        L32, cam32, ex32 = _leaves(cam, extra, torch.float32, dev)
        zb, cv = render_differentiable(cam32, shader, faces, ex32, S, S)
        gpu = _grads(L32, cv, zb, wc, wz)
        for n in g64:
            ref, got = g64[n], gpu[n]
            scale = float(ref.norm())
            err = float((got - ref).norm())
            if scale < 1e-8:
                assert err < 1e-6, f"{label} {n}: expected ~0 gradient, got norm {err}"
                continue
            checked += 1
            if label == "all":
                f32_err = float((g32c[n] - ref).norm())
                assert err <= max(1e-3 * scale, 1.5 * f32_err), \
                    f"{label} {n}: relative error {err / scale:.3e}, float32 restatement " \
                    f"{f32_err / scale:.3e}"
            else:
                assert err <= 1e-4 * scale, \
                    f"{label} {n}: relative error {err / scale:.3e} (|ref|={scale:.3e})"
    assert checked >= 10


def test_config5_range_ends_480x270():
    """Scenes of about 1,000 and 20,000 triangles (the ends of the config-5
    range) in one batched call at 480x270 with 512x512 shadow maps."""
    scs, sm = workloads.config5(seed=5, targets=(1000, 20000))
    tris = workloads.triangle_counts(scs)
    assert tris.min() < 1100 and tris.max() >= 20000
    assert (sm.width, sm.height) == (512, 512)
    got, _ = _check_plan(lambda: batch.plan_batch(scs, 480, 270, sm), "c5", tile=(16, 16),
                         pixels_per_lane=4)
    pid = as_np(got["main"]["prim_id"])
    assert (pid[0] >= 0).sum() > 1000 and (pid[1] >= 0).sum() > 1000


def test_ant_like_frames_against_oracle():
    """The ant-like benchmark frames (18 objects, 3,276 triangles, shadow map
    the size of the 84x84 canvas centred on each camera's target) batched,
    against the oracle."""
    scs, sm = workloads.ant_like(n_scenes=3, seed=7)
    assert set(workloads.triangle_counts(scs)) == {3276}
    assert all(len(s.model.texture_shape) == 18 for s in scs)
    assert (sm.width, sm.height) == (84, 84)
    got, _ = _check_plan(lambda: batch.plan_batch(scs, 84, 84, sm), "ant")
    pid = as_np(got["main"]["prim_id"])
    for s in range(3):
        # the walker itself (faces after the ground box's 12) is on screen
        assert ((pid[s] >= s * 3276 + 12) & (pid[s] < (s + 1) * 3276)).sum() > 200


@pytest.mark.parametrize("tile,ppl", [((8, 8), 1), ((16, 16), 4)], ids=str)
def test_darboux_head_512(tile, ppl):
    """phong_darboux at scale: the reference examples' african head (2,492
    triangles, tests/golden/african_head.obj) with a 512x512 diffuse map and a
    tangent-space normal map, 512x512 pixels, against the oracle."""
    from jaxrenderer_amd.pipeline import build_call
    from jaxrenderer_amd.shaders import PhongTextureDarbouxShader
    from tests import scenes

    cam, faces, extra = scenes.darboux_head()
    got, _ = _check_plan(lambda: [build_call(cam, PhongTextureDarbouxShader, faces, extra, 512,
                                             512, init_buffers=True)],
                         f"darboux-head/{tile}", tile=tile, pixels_per_lane=ppl)
    assert int((as_np(got["main"]["prim_id"]) >= 0).sum()) > 40_000


def test_darboux_head_backward_against_float64():
    """phong_darboux's backward at scale: the african head at 512x512 with its
    512x512 diffuse and normal maps.  jr_backward's gradients of positions
    (through visibility and the tangent frame's face corners), normals, uvs
    (the tangent frame), texture, normal map, light and camera against torch
    autograd of the float64 restatement with the GPU's primitive ids frozen.
    Pixels at a float32 decision boundary (texel index, lit/unlit branch) or
    whose tangent-frame matrix has a float64 condition number of 1e3 or more
    carry no loss weight.

    Two bars, as for the view workload: 1e-4 relative per tensor over pixels
    of triangles whose M = [x y w] has a condition number below 300, and over
    every pixel 1e-3 relative or no further from float64 than 1.5 times the
    same formulas evaluated in float32."""
    from jaxrenderer_amd.shaders import PhongTextureDarbouxShader as Sh
    from tests import scenes

    S = 512
    print("configs: darboux head backward", flush=True)
    cam, faces, extra = scenes.darboux_head(size=S)
    dev = torch.device("cuda")
    call = build_views_call([cam], Sh, faces, extra, S, S, init_buffers=True)
    pid = _plan.run(call, prim_id=True)["prim_id"][0].cpu().numpy()
    assert (pid >= 0).sum() > 40_000
    L64, canvas64, z64, amb = _ref_forward(Sh, cam, faces, extra, pid, torch.float64)
    L32c, canvas32, z32, _ = _ref_forward(Sh, cam, faces, extra, pid, torch.float32)
    w2c = np.asarray(cam.world_to_clip, np.float64)
    clip = np.concatenate([np.asarray(extra.position, np.float64),
                           np.ones((len(extra.position), 1))], axis=1) @ w2c.T
    cond = np.linalg.cond(clip[np.asarray(faces)][..., [0, 1, 3]])
    pix_cond = np.where(pid >= 0, cond[np.maximum(pid, 0)], np.inf)
    rng = np.random.default_rng(500)
    wc_all = torch.as_tensor(rng.normal(size=(S, S, 3)))
    wz_all = torch.as_tensor(rng.normal(size=(S, S)))
    checked = 0
    for label, mask in (("well-conditioned", (pid >= 0) & (pix_cond < 300)),
                        ("all", pid >= 0)):
        keep = torch.as_tensor(mask) & ~amb
        wc, wz = wc_all * keep[..., None], wz_all * keep
        g64 = _grads(L64, canvas64, z64, wc, wz)
        g32c = _grads(L32c, canvas32, z32, wc, wz)
        L32, cam32, ex32 = _leaves(cam, extra, torch.float32, dev)
        zb, cv = render_differentiable(cam32, Sh, faces, ex32, S, S)
        gpu = _grads(L32, cv, zb, wc, wz)
        for n in g64:
            ref, got = g64[n], gpu[n]
            scale = float(ref.norm())
            err = float((got - ref).norm())
            if scale < 1e-8:
                assert err < 1e-6, f"{label} {n}: expected ~0 gradient, got norm {err}"
                continue
            checked += 1
            if label == "all":
                f32_err = float((g32c[n] - ref).norm())
                assert err <= max(1e-3 * scale, 1.5 * f32_err), \
                    f"{label} {n}: relative error {err / scale:.3e}, float32 restatement " \
                    f"{f32_err / scale:.3e}"
            else:
                assert err <= 1e-4 * scale, \
                    f"{label} {n}: relative error {err / scale:.3e} (|ref|={scale:.3e})"
    assert checked >= 14
