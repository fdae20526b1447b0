"""PhongReflectionTextureShader (renderer/shaders/phong_reflection.py:91-258)."""
from typing import Any, NamedTuple

import numpy as np

from .._hip import SHADER_PHONG_REFLECTION
from ..shader import Shader
from ..types import LightSource


class PhongReflectionTextureExtraInput(NamedTuple):
    position: Any  # [vertices, 3] world space
    normal: Any  # [vertices, 3] world space
    uv: Any  # [vertices, 2]
    light: LightSource
    light_dir_eye: Any  # [3] eye space
    texture_shape: Any  # [objects, 2]
    texture_index: Any  # [vertices]
    texture_offset: Any  # int
    texture: Any  # merged diffuse maps [W, H, C]
    specular_map: Any  # merged specular maps [W, H]
    ambient: Any
    diffuse: Any
    specular: Any


class PhongReflectionTextureExtraFragmentData(NamedTuple):
    normal: Any = np.zeros(3, dtype=np.float32)
    uv: Any = np.zeros(2, dtype=np.float32)
    texture_index: Any = 0
    colour: Any = np.zeros(3, dtype=np.float32)


class PhongReflectionTextureExtraMixerOutput(NamedTuple):
    canvas: Any


class PhongReflectionTextureShader(Shader):
    """Phong reflection model (ambient + diffuse + specular) with merged maps."""

    _jr_shader_id = SHADER_PHONG_REFLECTION
