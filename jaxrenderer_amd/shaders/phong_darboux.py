"""PhongTextureDarbouxShader (renderer/shaders/phong_darboux.py:106-316)."""
from typing import Any, NamedTuple

import numpy as np

from .._hip import SHADER_PHONG_DARBOUX
from ..shader import Shader
from ..types import LightSource


class PhongTextureDarbouxExtraInput(NamedTuple):
    position: Any  # [vertices, 3] world space
    normal: Any  # [vertices, 3] world space
    uv: Any  # [vertices, 2]
    light: LightSource
    texture: Any  # [W, H, C]
    normal_map: Any  # [W, H, 3] in the Darboux frame
    id_to_face: Any  # [vertices] face of each vertex
    faces_indices: Any  # [faces, 3] vertices of each face


class PhongTextureDarbouxExtraFragmentData(NamedTuple):
    normal: Any = np.zeros(3, dtype=np.float32)
    uv: Any = np.zeros(2, dtype=np.float32)
    triangle: Any = np.zeros((3, 3), dtype=np.float32)
    triangle_uv: Any = np.zeros((3, 2), dtype=np.float32)
    colour: Any = np.zeros(3, dtype=np.float32)


class PhongTextureDarbouxExtraMixerOutput(NamedTuple):
    canvas: Any


class PhongTextureDarbouxShader(Shader):
    """Phong with normals from a tangent-space (Darboux frame) normal map."""

    _jr_shader_id = SHADER_PHONG_DARBOUX
