"""GouraudTextureShader (renderer/shaders/gouraud_texture.py:55-167)."""
from typing import Any, NamedTuple

import numpy as np
import torch

from .._hip import SHADER_GOURAUD_TEXTURE
from ..shader import PerFragment, PerVertex, Shader
from ..types import LightSource
from . import _ops


class GouraudTextureExtraInput(NamedTuple):
    position: Any  # [vertices, 3] world space
    normal: Any  # [vertices, 3] world space
    uv: Any  # [vertices, 2] texture space
    light: LightSource
    texture: Any  # [W, H, C]


class GouraudTextureExtraFragmentData(NamedTuple):
    colour: Any = np.zeros(3, dtype=np.float32)
    uv: Any = np.zeros(2, dtype=np.float32)


class GouraudTextureExtraMixerOutput(NamedTuple):
    canvas: Any


class GouraudTextureShader(Shader):
    """Gouraud lighting times texture[floor(uv) % shape]."""

    _jr_shader_id = SHADER_GOURAUD_TEXTURE

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        """renderer/shaders/gouraud_texture.py:64-101."""
        position = _ops.take(extra.position, gl_VertexID)
        gl_Position = _ops.to_clip(camera, position)
        normal = _ops.normalise3(_ops.take(extra.normal, gl_VertexID))
        light_dir = _ops.normalise3(_ops.as_tensor(extra.light.direction, position))
        intensity = _ops.dot3(normal, light_dir)
        light_colour = _ops.as_tensor(extra.light.colour, position) * intensity
        return PerVertex(gl_Position=gl_Position), GouraudTextureExtraFragmentData(
            colour=light_colour, uv=_ops.take(extra.uv, gl_VertexID))

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        """renderer/shaders/gouraud_texture.py:103-144: texture[floor(uv) %
        shape] times the interpolated light colour; discard back faces and
        negative light."""
        built_in = Shader.fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra)[0]
        tex = _ops.as_tensor(extra.texture, gl_FragCoord)
        u, v = _ops.floor_mod_index(varying.uv, tex.shape[:2])
        texture_colour = _ops.take2(tex, u, v)
        light_colour = varying.colour
        keeps = (torch.as_tensor(built_in.keeps, device=gl_FragCoord.device)
                 & torch.as_tensor(gl_FrontFacing) & (light_colour >= 0).all())
        return (PerFragment(keeps=keeps, use_default_depth=built_in.use_default_depth),
                GouraudTextureExtraFragmentData(colour=texture_colour * light_colour,
                                                uv=varying.uv))

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        mixed, out = Shader.mix(gl_FragDepth, keeps, extra)
        return mixed, GouraudTextureExtraMixerOutput(canvas=out.colour)
