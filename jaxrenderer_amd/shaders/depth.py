"""DepthShader (renderer/shaders/depth.py:38-61): position only; writes depth."""
from typing import Any, NamedTuple

from .._hip import SHADER_DEPTH
from ..shader import Shader


class DepthExtraInput(NamedTuple):
    position: Any  # [vertices, 3] world space


class DepthExtraFragmentData(NamedTuple):
    pass


class DepthExtraMixerOutput(NamedTuple):
    pass


class DepthShader(Shader):
    """Depth only (used by the shadow-map pass)."""

    _jr_shader_id = SHADER_DEPTH
