"""PhongReflectionShadowTextureShader
(renderer/shaders/phong_reflection_shadow.py:99-295)."""
from typing import Any, NamedTuple

import numpy as np

from .._hip import SHADER_PHONG_REFLECTION_SHADOW
from ..shader import Shader
from ..types import LightSource


class PhongReflectionShadowTextureExtraInput(NamedTuple):
    position: Any
    normal: Any
    uv: Any
    light: LightSource
    light_dir_eye: Any
    texture_shape: Any
    texture_index: Any
    texture_offset: Any
    texture: Any
    specular_map: Any
    shadow: Any  # jaxrenderer_amd.shadow.Shadow
    camera: Any  # jaxrenderer_amd.geometry.Camera
    ambient: Any
    diffuse: Any
    specular: Any


class PhongReflectionShadowTextureExtraFragmentData(NamedTuple):
    normal: Any = np.zeros(3, dtype=np.float32)
    uv: Any = np.zeros(2, dtype=np.float32)
    texture_index: Any = 0
    shadow_coord: Any = np.zeros(4, dtype=np.float32)
    colour: Any = np.zeros(3, dtype=np.float32)


class PhongReflectionShadowTextureExtraMixerOutput(NamedTuple):
    canvas: Any


class PhongReflectionShadowTextureShader(Shader):
    """Phong reflection attenuated by a shadow-map lookup."""

    _jr_shader_id = SHADER_PHONG_REFLECTION_SHADOW
