"""Built-in shaders (renderer/shaders).  Each class is executed by the fused
HIP kernel named by its `_jr_shader_id`; the NamedTuples are the extra inputs
and fragment/mixer payloads of the reference, field for field."""

from .depth import DepthExtraInput, DepthShader
from .gouraud import GouraudExtraInput, GouraudShader
from .gouraud_texture import GouraudTextureExtraInput, GouraudTextureShader
from .phong import PhongTextureExtraInput, PhongTextureShader
from .phong_darboux import PhongTextureDarbouxExtraInput, PhongTextureDarbouxShader
from .phong_reflection import PhongReflectionTextureExtraInput, PhongReflectionTextureShader
from .phong_reflection_shadow import (
    PhongReflectionShadowTextureExtraInput,
    PhongReflectionShadowTextureShader,
)

BUILTIN = (
    DepthShader,
    GouraudShader,
    GouraudTextureShader,
    PhongTextureShader,
    PhongTextureDarbouxShader,
    PhongReflectionTextureShader,
    PhongReflectionShadowTextureShader,
)

__all__ = [c.__name__ for c in BUILTIN] + [
    "DepthExtraInput",
    "GouraudExtraInput",
    "GouraudTextureExtraInput",
    "PhongTextureExtraInput",
    "PhongTextureDarbouxExtraInput",
    "PhongReflectionTextureExtraInput",
    "PhongReflectionShadowTextureExtraInput",
    "BUILTIN",
]
