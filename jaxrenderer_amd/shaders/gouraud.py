"""GouraudShader (renderer/shaders/gouraud.py:49-124)."""
from typing import Any, NamedTuple

import numpy as np

from .._hip import SHADER_GOURAUD
from ..shader import Shader
from ..types import LightSource


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
