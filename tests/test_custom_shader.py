"""User-defined Shader subclasses (renderer/pipeline.py:470-537 with hooks from
the caller): the hooks run as torch code around the visibility pass.

The CPU tests feed the visibility pass from the oracle; the GPU tests use the
HIP visibility kernel.  `test_perspective_interpolation_*` carries the
assertions of the reference's tests/smoke_test.py:235-330 (its checkerboard
shader, restated here with torch hooks).  A Gouraud shader written as hooks
must give the built-in kernel's image (same visibility, colours within 1e-5).
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
from jaxrenderer_amd.shaders import GouraudExtraInput, GouraudShader
from jaxrenderer_amd.types import Buffers, LightSource
from oracle.oracle import run_oracle
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


class ChooserOverride(HookGouraud):
    @staticmethod
    def primitive_chooser(*args):
        return args


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
    out = run_oracle(call, prim_id=True, bary=False)
    return torch.from_numpy(out["prim_id"][0]), torch.from_numpy(out["zbuffer"][0])


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


def test_custom_chooser_is_rejected():
    W, H = 16, 16
    camera, faces, extra = _gouraud_scene(W, H)
    with pytest.raises(NotImplementedError):
        render_custom(camera, ChooserOverride, _buffers(W, H, "cpu"), faces, extra,
                      visibility=oracle_visibility)


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
def test_hook_gouraud_matches_builtin_kernel_gpu():
    W, H = 1920, 1080
    camera, faces, extra = _gouraud_scene(W, H)
    hook = pipeline.render(camera, HookGouraud, _buffers(W, H, "cuda"), faces, extra)
    fused = pipeline.render(camera, GouraudShader, _buffers(W, H, "cuda"), faces, extra)
    assert torch.equal(hook.zbuffer, fused.zbuffer)
    assert float((hook.targets[0] - fused.targets[0]).abs().max()) <= 1e-5
