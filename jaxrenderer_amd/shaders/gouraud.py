"""GouraudShader (renderer/shaders/gouraud.py:49-124)."""
from typing import Any, NamedTuple

import numpy as np
import torch

from .._hip import SHADER_GOURAUD
from ..shader import PerFragment, PerVertex, Shader
from ..types import LightSource
from . import _ops


class GouraudExtraInput(NamedTuple):
    position: Any  # [vertices, 3] world space
    colour: Any  # [vertices, 3]
    normal: Any  # [vertices, 3] world space
    light: LightSource


class GouraudExtraFragmentData(NamedTuple):
    colour: Any = np.zeros(3, dtype=np.float32)


class GouraudExtraMixerOutput(NamedTuple):
    canvas: Any


class GouraudShader(Shader):
    """Per-vertex colour * light colour * dot(normal, light dir); fragments
    with a negative colour channel or facing away are discarded."""

    _jr_shader_id = SHADER_GOURAUD

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        """renderer/shaders/gouraud.py:54-88."""
        position = _ops.take(extra.position, gl_VertexID)
        gl_Position = _ops.to_clip(camera, position)
        normal = _ops.normalise3(_ops.take(extra.normal, gl_VertexID))
        light_dir = _ops.normalise3(_ops.as_tensor(extra.light.direction, position))
        intensity = _ops.dot3(normal, light_dir)
        colour = (_ops.take(extra.colour, gl_VertexID)
                  * _ops.as_tensor(extra.light.colour, position)) * intensity
        return PerVertex(gl_Position=gl_Position), GouraudExtraFragmentData(colour=colour)

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        """renderer/shaders/gouraud.py:90-124: discard back faces and negative
        colours."""
        built_in = Shader.fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra)[0]
        keeps = (torch.as_tensor(built_in.keeps, device=gl_FragCoord.device)
                 & torch.as_tensor(gl_FrontFacing) & (varying.colour >= 0).all())
        return PerFragment(keeps=keeps, use_default_depth=built_in.use_default_depth), varying

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        mixed, out = Shader.mix(gl_FragDepth, keeps, extra)
        return mixed, GouraudExtraMixerOutput(canvas=out.colour)
