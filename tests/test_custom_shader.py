"""User-defined Shader subclasses (renderer/pipeline.py:470-537 with hooks from
the caller): the hooks run as torch code around the visibility pass.

The CPU tests feed the visibility pass from the oracle; the GPU tests use the
HIP visibility kernel.  `test_perspective_interpolation_*` carries the
assertions of the reference's tests/smoke_test.py:235-330 (its checkerboard
shader, restated here with torch hooks).  A Gouraud shader written as hooks
must give the built-in kernel's image (same visibility, colours within 1e-5).

The built-in shaders' own hooks (renderer/shaders/*.py restated in torch) run
whenever a subclass is rendered: with nothing overridden they must give the
oracle's image (CPU) and the fused kernel's (GPU), depth bit for bit and
colour within 1e-5.  A user primitive_chooser runs the reference's brute
force in torch: one restating the default must give the default path's
image, and a "farthest primitive" chooser the float64 brute force's depths.
"""

from __future__ import annotations

from typing import Any, NamedTuple

import numpy as np
import pytest
import torch

from jaxrenderer_amd import pipeline
from jaxrenderer_amd._custom import render_custom
from jaxrenderer_amd.geometry import normalise, to_homogeneous
from jaxrenderer_amd.shader import MixerOutput, PerFragment, PerVertex, Shader
from jaxrenderer_amd.geometry import Camera
from jaxrenderer_amd.shaders import (BUILTIN, DepthShader, GouraudExtraInput, GouraudShader,
                                     PhongReflectionTextureShader)
from jaxrenderer_amd.types import Buffers, LightSource
from oracle.oracle import run_oracle
from tests import scenes
from tests.scenes import SMOKE_COLOURS, SMOKE_FACES, SMOKE_POSITIONS, smoke_camera

f32 = np.float32


class CheckerInput(NamedTuple):
    position: Any
    normal: Any
    uv: Any
    light: LightSource
    colour: Any


class CheckerVarying(NamedTuple):
    colour: Any
    uv: Any


class CheckerOutput(NamedTuple):
    canvas: Any


class CheckerShader(Shader):
    """tests/smoke_test.py:160-232: Gouraud colour, halved on alternate uv squares."""

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        gl_Position = camera.to_clip(to_homogeneous(extra.position[gl_VertexID]))
        intensity = torch.dot(normalise(extra.normal[gl_VertexID]),
                              normalise(extra.light.direction))
        return PerVertex(gl_Position=gl_Position), CheckerVarying(
            colour=extra.light.colour * intensity * extra.colour[gl_VertexID],
            uv=extra.uv[gl_VertexID])

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        built_in = Shader.fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra)[0]
        a = torch.frac(varying.uv) < 0.5
        colour = torch.where(a[0] != a[1], varying.colour, varying.colour * 0.5)
        return built_in, CheckerVarying(colour=colour, uv=varying.uv)

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        out, ex = Shader.mix(gl_FragDepth, keeps, extra)
        return out, CheckerOutput(canvas=ex.colour)


class HookGouraud(Shader):
    """renderer/shaders/gouraud.py:49-124 written as torch hooks."""

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        gl_Position = camera.to_clip(to_homogeneous(extra.position[gl_VertexID]))
        intensity = torch.dot(normalise(extra.normal[gl_VertexID]),
                              normalise(extra.light.direction))
        colour = extra.colour[gl_VertexID] * extra.light.colour * intensity
        return PerVertex(gl_Position=gl_Position), (colour,)

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        keeps = gl_FrontFacing & torch.all(varying[0] >= 0)
        return PerFragment(keeps=keeps, use_default_depth=True), varying

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        out, ex = Shader.mix(gl_FragDepth, keeps, extra)
        return out, ex


def _pick(idx, fc, ff, pc, keeps, values, bs, bc):
    take = lambda a: torch.index_select(a, 0, idx)  # noqa: E731
    return (take(fc), take(ff), take(pc), take(keeps), type(values)(*(take(v) for v in values)),
            take(bs), take(bc))


class DefaultChooserGouraud(GouraudShader):
    """renderer/shader.py:207-251's default chooser, restated as a hook."""

    @staticmethod
    def primitive_chooser(fc, ff, pc, keeps, values, bs, bc):
        depths = torch.where(keeps & ff, fc[:, 2], torch.full_like(fc[:, 2], float("inf")))
        return _pick(torch.argmin(depths).reshape(1), fc, ff, pc, keeps, values, bs, bc)


class FarthestDepth(DepthShader):
    """The kept front-facing primitive of LARGEST depth (index 0 if none)."""

    @staticmethod
    def primitive_chooser(fc, ff, pc, keeps, values, bs, bc):
        depths = torch.where(keeps & ff, fc[:, 2], torch.full_like(fc[:, 2], -float("inf")))
        return _pick(torch.argmax(depths).reshape(1), fc, ff, pc, keeps, values, bs, bc)


def _soup(shader, seed, W, H, vis):
    """A soup scene's inputs with the shadow map (from `vis`'s producer, the
    oracle) filled in for the shadow shader."""
    cam, faces, extra, shadow_call = scenes.soup_inputs(shader, seed, W, H)
    if shadow_call is not None:
        smap = run_oracle(shadow_call, prim_id=False, bary=False)["zbuffer"][0]
        extra = extra._replace(shadow=extra.shadow._replace(shadow_map=smap))
    return cam, faces, extra


def _checker_scene(W=1920, H=1080):
    camera = smoke_camera(W, H, eye=(0.0, 0.0, 1.0))
    position = np.array(((-1.0, -1.0, -2.0), (1.0, -1.0, -1.0), (0.0, 1.0, -1.0)), f32)
    extra = CheckerInput(
        position=position,
        normal=np.tile(LightSource().direction, (3, 1)).astype(f32),
        uv=np.array(((0.0, 0.0), (10.0, 0.0), (0.0, 10.0)), f32),
        light=LightSource(),
        colour=np.eye(3, dtype=f32),
    )
    faces = np.array(((0, 1, 2),), np.int32)
    return camera, faces, extra


def _gouraud_scene(W, H):
    camera = smoke_camera(W, H)
    extra = GouraudExtraInput(position=SMOKE_POSITIONS, colour=SMOKE_COLOURS,
                              normal=np.tile(LightSource().direction, (6, 1)).astype(f32),
                              light=LightSource())
    return camera, SMOKE_FACES, extra


def _buffers(W, H, dev, zdef=1.0, cdef=0.0):
    return Buffers(zbuffer=torch.full((W, H), zdef, device=dev),
                   targets=(torch.full((W, H, 3), cdef, device=dev),))


def oracle_visibility(call):
    out = run_oracle(call, prim_id=True, bary=True)
    return (torch.from_numpy(out["prim_id"][0]), torch.from_numpy(out["zbuffer"][0]),
            torch.from_numpy(out["bary"][0]))


def _checker_assertions(zbuffer, canvas, W, H):
    """tests/smoke_test.py:310-330 (display orientation: [H, W], row 0 on top)."""
    z = torch.flip(zbuffer.cpu().T, (0,))
    c = torch.flip(canvas.cpu().transpose(0, 1), (0,))
    assert z.shape == (H, W) and c.shape == (H, W, 3)
    assert int((z == 1.0).sum()) > W * H // 2
    assert z[679, 701] < z[779, 1388]
    empty = int((c == 0.0).all(dim=2).sum())
    assert W * H // 2 < empty < W * H
    # the checkerboard: both the full and the halved colour occur
    lit = c[(c != 0.0).any(dim=2)]
    ratio = lit.sum(-1)
    assert lit.shape[0] > 1000
    assert float(ratio.max()) > 1.5 * float(ratio.min())


def test_perspective_interpolation_cpu():
    W, H = 1920, 1080
    camera, faces, extra = _checker_scene(W, H)
    out = render_custom(camera, CheckerShader, _buffers(W, H, "cpu"), faces, extra,
                        visibility=oracle_visibility)
    _checker_assertions(out.zbuffer, out.targets[0], W, H)


def test_hook_gouraud_matches_oracle_cpu():
    W, H = 192, 108
    camera, faces, extra = _gouraud_scene(W, H)
    out = render_custom(camera, HookGouraud, _buffers(W, H, "cpu"), faces, extra,
                        visibility=oracle_visibility)
    ref = run_oracle(pipeline.build_call(camera, GouraudShader, faces, extra, W, H),
                     zbuffer=np.full((1, W, H), 1.0, f32), canvas=np.zeros((1, W, H, 3), f32),
                     prim_id=False, bary=False)
    assert torch.equal(out.zbuffer, torch.from_numpy(ref["zbuffer"][0]))
    np.testing.assert_allclose(out.targets[0].numpy(), ref["canvas"][0], atol=1e-5, rtol=0)
    assert int((out.zbuffer != 1.0).sum()) > 1000


@pytest.mark.parametrize("shader", BUILTIN, ids=lambda s: s.__name__)
@pytest.mark.parametrize("seed", [0, 1])
def test_builtin_hooks_match_oracle_cpu(shader, seed):
    """A subclass overriding nothing runs the built-in's own torch hooks."""
    W, H = 48, 40
    cam, faces, extra = _soup(shader, seed, W, H, oracle_visibility)
    sub = type("Sub" + shader.__name__, (shader,), {})
    assert not pipeline.is_builtin(sub)
    want_canvas = shader is not DepthShader
    bufs = Buffers(zbuffer=torch.full((W, H), 1.0),
                   targets=(torch.full((W, H, 3), 0.25),) if want_canvas else ())
    out = render_custom(cam, sub, bufs, faces, extra, visibility=oracle_visibility)
    call = pipeline.build_call(cam, shader, faces, extra, W, H)
    ref = run_oracle(call, zbuffer=np.full((1, W, H), 1.0, f32),
                     canvas=np.full((1, W, H, 3), 0.25, f32) if want_canvas else None,
                     prim_id=False, bary=False)
    assert np.array_equal(out.zbuffer.numpy().view(np.int32), ref["zbuffer"][0].view(np.int32))
    if want_canvas:
        np.testing.assert_allclose(out.targets[0].numpy(), ref["canvas"][0], atol=1e-5, rtol=0)
    assert int((out.zbuffer != 1.0).sum()) > 50


def test_builtin_vertex_hooks_called_directly():
    """GouraudShader.vertex / DepthShader.vertex on plain arrays: the clip
    position of Camera.apply and the Gouraud colour colour * light * dot(n, l)
    in float32, left to right."""
    cam, faces, extra, _ = scenes.soup_inputs(GouraudShader, 3, 32, 24)
    for vid in (0, 7, 31):
        pv, data = GouraudShader.vertex(vid, 0, cam, extra)
        p = np.concatenate([extra.position[vid], [1.0]]).astype(f32)
        assert np.array_equal(pv.gl_Position.numpy(), Camera.apply(p, cam.world_to_clip))
        n = normalise(extra.normal[vid])
        ld = normalise(extra.light.direction)
        it = f32(f32(f32(n[0] * ld[0]) + f32(n[1] * ld[1])) + f32(n[2] * ld[2]))
        want = ((extra.colour[vid] * extra.light.colour).astype(f32) * it).astype(f32)
        assert np.array_equal(data.colour.numpy(), want)
        pd, _ = DepthShader.vertex(vid, 0, cam, extra)
        assert np.array_equal(pd.gl_Position.numpy(), pv.gl_Position.numpy())


def test_default_restating_chooser_matches_default_path_cpu():
    W, H = 48, 40
    cam, faces, extra = _soup(GouraudShader, 0, W, H, oracle_visibility)
    a = render_custom(cam, DefaultChooserGouraud, _buffers(W, H, "cpu", cdef=0.25), faces,
                      extra, visibility=oracle_visibility)
    b = render_custom(cam, type("G", (GouraudShader,), {}), _buffers(W, H, "cpu", cdef=0.25),
                      faces, extra, visibility=oracle_visibility)
    # M^-1 from torch (chooser path) and from the kernel's LU may round apart
    assert float((a.zbuffer - b.zbuffer).abs().max()) <= 1e-6
    assert float((a.targets[0] - b.targets[0]).abs().max()) <= 1e-5
    assert int((a.zbuffer != 1.0).sum()) > 50


def test_farthest_chooser_against_float64_cpu():
    W, H = 48, 40
    cam, faces, extra, _ = scenes.soup_inputs(DepthShader, 2, W, H)
    out = render_custom(cam, FarthestDepth, Buffers(zbuffer=torch.full((W, H), 1.0), targets=()),
                        faces, extra, visibility=oracle_visibility)
    P = np.concatenate([extra.position.astype(np.float64),
                        np.ones((len(extra.position), 1))], 1) @ cam.world_to_clip.astype(
        np.float64).T
    vp = cam.viewport.astype(np.float64)
    want = np.full((W, H), np.nan)
    ambiguous = np.zeros((W, H), bool)
    first = None
    for k, t in enumerate(faces):
        M = P[t][:, [0, 1, 3]]
        d = np.linalg.det(M)
        Mi = np.linalg.inv(M)
        i, j = np.meshgrid(np.arange(W), np.arange(H), indexing="ij")
        xy1 = np.stack([(i - vp[0, 3]) / vp[0, 0], (j - vp[1, 3]) / vp[1, 1], np.ones_like(
            i, dtype=np.float64)], -1)
        e = xy1 @ Mi
        inside = (e >= 0).all(-1)
        ambiguous |= (e > -1e-5).all(-1) & (np.abs(e) < 1e-5).any(-1)
        z = (e @ P[t][:, 2]) * vp[2, 2] + vp[2, 3]
        if k == 0:
            first = np.where(inside & (abs(d) > 1e-6), z, np.nan)
        if abs(d) > 1e-6 and d >= 0:
            want = np.where(inside, np.fmax(want, z), want)
    # nothing kept and front-facing: argmax's index 0, drawn if it covers
    want = np.where(np.isnan(want), first, want)
    want = np.where(np.isnan(want), 1.0, want)
    got = out.zbuffer.numpy()
    ok = ~ambiguous
    assert int((want != 1.0).sum()) > 500
    assert float(np.abs(got - want)[ok].max()) <= 1e-5


@pytest.mark.gpu
def test_perspective_interpolation_gpu():
    W, H = 1920, 1080
    camera, faces, extra = _checker_scene(W, H)
    out = pipeline.render(camera, CheckerShader, _buffers(W, H, "cuda"), faces, extra)
    _checker_assertions(out.zbuffer, out.targets[0], W, H)
    ref = render_custom(camera, CheckerShader, _buffers(W, H, "cpu"), faces, extra,
                        visibility=oracle_visibility)
    assert torch.equal(out.zbuffer.cpu(), ref.zbuffer)
    # the checkerboard is discontinuous in uv: a uv that torch's GPU and CPU
    # kernels round to either side of a square's edge may flip a pixel there
    diff = (out.targets[0].cpu() - ref.targets[0]).abs().amax(-1)
    assert float((diff > 1e-5).float().mean()) < 1e-3


@pytest.mark.gpu
@pytest.mark.parametrize("shader", BUILTIN, ids=lambda s: s.__name__)
def test_builtin_subclasses_match_fused_kernels_gpu(shader):
    """Rendering a subclass runs the built-in's torch hooks around the HIP
    visibility pass; the fused kernel of the class itself must agree."""
    W, H = 64, 48
    cam, faces, extra = _soup(shader, 4, W, H, oracle_visibility)
    want_canvas = shader is not DepthShader

    def bufs():
        return Buffers(zbuffer=torch.full((W, H), 1.0, device="cuda"),
                       targets=(torch.full((W, H, 3), 0.25, device="cuda"),)
                       if want_canvas else ())

    fused = pipeline.render(cam, shader, bufs(), faces, extra)
    hooks = pipeline.render(cam, type("Sub", (shader,), {}), bufs(), faces, extra)
    assert torch.equal(hooks.zbuffer.view(torch.int32), fused.zbuffer.view(torch.int32))
    if want_canvas:
        assert float((hooks.targets[0] - fused.targets[0]).abs().max()) <= 1e-5


class MixOnlyReflection(PhongReflectionTextureShader):
    """Overrides only `mix` (with the reference's own mixing)."""

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        mixed, out = Shader.mix(gl_FragDepth, keeps, extra)
        return mixed, (out.colour,)


@pytest.mark.gpu
def test_reflection_subclass_overriding_mix_matches_fused_gpu():
    W, H = 320, 240
    calls = scenes.c1_plan()
    from jaxrenderer_amd import _plan
    fused = _plan.run_plan(scenes.c1_plan())["main"]
    c = calls[-1]
    cam = Camera.create(view=np.identity(4, f32), projection=np.identity(4, f32),
                        viewport=np.ctypeslib.as_array(c.scenes[0].viewport).reshape(4, 4))
    cam = cam._replace(world_to_clip=np.ctypeslib.as_array(c.scenes[0].world_to_clip)
                       .reshape(4, 4).copy(),
                       world_to_eye_norm=np.ctypeslib.as_array(c.scenes[0].world_to_eye_norm)
                       .reshape(4, 4).copy())
    from jaxrenderer_amd.shaders import PhongReflectionTextureExtraInput
    ins = c.inputs
    pos = ins["positions"][ins["pos_index"]]
    extra = PhongReflectionTextureExtraInput(
        position=pos, normal=ins["normals"][ins["nrm_index"]], uv=ins["uvs"][ins["uv_index"]],
        light=LightSource(direction=np.ctypeslib.as_array(c.scenes[0].light_direction).copy(),
                          colour=np.ctypeslib.as_array(c.scenes[0].light_colour).copy()),
        light_dir_eye=np.ctypeslib.as_array(c.scenes[0].light_dir_eye).copy(),
        texture_shape=ins["texture_shapes"], texture_index=ins["tex_index"],
        texture_offset=int(c.scenes[0].tex_offset), texture=ins["textures"],
        specular_map=ins["specular_maps"],
        ambient=np.ctypeslib.as_array(c.scenes[0].ambient).copy(),
        diffuse=np.ctypeslib.as_array(c.scenes[0].diffuse).copy(),
        specular=np.ctypeslib.as_array(c.scenes[0].specular).copy())
    faces = np.arange(3 * c.n_tris, dtype=np.int32).reshape(c.n_tris, 3)
    bufs = Buffers(zbuffer=torch.full((W, H), 1.0, device="cuda"),
                   targets=(torch.full((W, H, 3), 1.0, device="cuda"),))
    out = pipeline.render(cam, MixOnlyReflection, bufs, faces, extra)
    assert torch.equal(out.zbuffer.view(torch.int32), fused["zbuffer"][0].view(torch.int32))
    assert float((out.targets[0] - fused["canvas"][0]).abs().max()) <= 1e-5


@pytest.mark.gpu
def test_custom_chooser_gpu():
    W, H = 48, 40
    cam, faces, extra = _soup(GouraudShader, 0, W, H, oracle_visibility)
    a = pipeline.render(cam, DefaultChooserGouraud, _buffers(W, H, "cuda", cdef=0.25), faces,
                        extra)
    b = pipeline.render(cam, GouraudShader, _buffers(W, H, "cuda", cdef=0.25), faces, extra)
    assert float((a.zbuffer - b.zbuffer).abs().max()) <= 1e-6
    assert float((a.targets[0] - b.targets[0]).abs().max()) <= 1e-5


@pytest.mark.gpu
def test_hook_gouraud_matches_builtin_kernel_gpu():
    W, H = 1920, 1080
    camera, faces, extra = _gouraud_scene(W, H)
    hook = pipeline.render(camera, HookGouraud, _buffers(W, H, "cuda"), faces, extra)
    fused = pipeline.render(camera, GouraudShader, _buffers(W, H, "cuda"), faces, extra)
    assert torch.equal(hook.zbuffer, fused.zbuffer)
    assert float((hook.targets[0] - fused.targets[0]).abs().max()) <= 1e-5
