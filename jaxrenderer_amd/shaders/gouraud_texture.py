"""GouraudTextureShader (renderer/shaders/gouraud_texture.py:55-167)."""
from typing import Any, NamedTuple

import numpy as np

from .._hip import SHADER_GOURAUD_TEXTURE
from ..shader import Shader
from ..types import LightSource


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
