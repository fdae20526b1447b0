"""DepthShader (renderer/shaders/depth.py:38-61): position only; writes depth."""
from typing import Any, NamedTuple

from .._hip import SHADER_DEPTH
from ..shader import PerVertex, Shader
from . import _ops


class DepthExtraInput(NamedTuple):
    position: Any  # [vertices, 3] world space


class DepthExtraFragmentData(NamedTuple):
    pass


class DepthExtraMixerOutput(NamedTuple):
    pass


class DepthShader(Shader):
    """Depth only (used by the shadow-map pass).  The default interpolate,
    fragment and mix hooks; the class itself runs as the fused depth kernel."""

    _jr_shader_id = SHADER_DEPTH

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        """renderer/shaders/depth.py:43-61: gl_Position = camera.to_clip(position)."""
        position = _ops.take(extra.position, gl_VertexID)
        return PerVertex(gl_Position=_ops.to_clip(camera, position)), DepthExtraFragmentData()
