"""PhongTextureShader (renderer/shaders/phong.py:65-187)."""
from typing import Any, NamedTuple

import numpy as np
import torch

from .._hip import SHADER_PHONG
from ..shader import PerFragment, PerVertex, Shader
from ..types import LightSource
from . import _ops


class PhongTextureExtraInput(NamedTuple):
    position: Any  # [vertices, 3] world space
    normal: Any  # [vertices, 3] world space
    uv: Any  # [vertices, 2]
    light: LightSource
    texture: Any  # [W, H, C]


class PhongTextureExtraFragmentData(NamedTuple):
    normal: Any = np.zeros(3, dtype=np.float32)
    uv: Any = np.zeros(2, dtype=np.float32)
    colour: Any = np.zeros(3, dtype=np.float32)


class PhongTextureExtraMixerOutput(NamedTuple):
    canvas: Any


class PhongTextureShader(Shader):
    """Per-fragment normal (eye space) lighting times texture."""

    _jr_shader_id = SHADER_PHONG

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        """renderer/shaders/phong.py:74-104: eye-space normal, uv."""
        position = _ops.take(extra.position, gl_VertexID)
        gl_Position = _ops.to_clip(camera, position)
        normal = _ops.apply_vec(_ops.normalise3(_ops.take(extra.normal, gl_VertexID)),
                                camera.world_to_eye_norm)
        return PerVertex(gl_Position=gl_Position), PhongTextureExtraFragmentData(
            normal=normal, uv=_ops.take(extra.uv, gl_VertexID),
            colour=torch.zeros(3, dtype=position.dtype, device=position.device))

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        """renderer/shaders/phong.py:106-164: texture times light colour times
        dot(normal, light direction), black where any channel is negative."""
        built_in = Shader.fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra)[0]
        tex = _ops.as_tensor(extra.texture, gl_FragCoord)
        u, v = _ops.floor_mod_index(varying.uv, tex.shape[:2])
        texture_colour = _ops.take2(tex, u, v)
        light_dir = _ops.normalise3(_ops.as_tensor(extra.light.direction, gl_FragCoord))
        light_colour = (_ops.as_tensor(extra.light.colour, gl_FragCoord)
                        * _ops.dot3(_ops.normalise3(varying.normal), light_dir))
        colour = torch.where((light_colour >= 0).all(), texture_colour * light_colour,
                             torch.zeros_like(light_colour))
        keeps = torch.as_tensor(built_in.keeps, device=gl_FragCoord.device) & torch.as_tensor(
            gl_FrontFacing)
        return (PerFragment(keeps=keeps, use_default_depth=built_in.use_default_depth),
                PhongTextureExtraFragmentData(normal=varying.normal, uv=varying.uv,
                                              colour=colour))

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        mixed, out = Shader.mix(gl_FragDepth, keeps, extra)
        return mixed, PhongTextureExtraMixerOutput(canvas=out.colour)
