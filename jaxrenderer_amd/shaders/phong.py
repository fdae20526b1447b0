"""PhongTextureShader (renderer/shaders/phong.py:65-187)."""
from typing import Any, NamedTuple

import numpy as np

from .._hip import SHADER_PHONG
from ..shader import Shader
from ..types import LightSource


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
