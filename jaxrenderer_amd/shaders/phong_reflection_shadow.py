"""PhongReflectionShadowTextureShader
(renderer/shaders/phong_reflection_shadow.py:99-295)."""
from typing import Any, NamedTuple

import numpy as np
import torch

from .._hip import SHADER_PHONG_REFLECTION_SHADOW
from ..shader import PerFragment, PerVertex, Shader
from ..types import LightSource
from . import _ops
from .phong_reflection import reflection_colour


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

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        """renderer/shaders/phong_reflection_shadow.py:110-150: as the
        reflection shader, plus the vertex in the shadow camera's clip space,
        normalised by w."""
        position = _ops.take(extra.position, gl_VertexID)
        gl_Position = _ops.to_clip(camera, position)
        normal = _ops.apply_vec(_ops.normalise3(_ops.take(extra.normal, gl_VertexID)),
                                camera.world_to_eye_norm)
        shadow_coord = _ops.normalise_homogeneous(_ops.to_clip(extra.shadow.camera, position))
        return PerVertex(gl_Position=gl_Position), PhongReflectionShadowTextureExtraFragmentData(
            normal=normal, uv=_ops.take(extra.uv, gl_VertexID),
            texture_index=_ops.take(extra.texture_index, gl_VertexID),
            shadow_coord=shadow_coord,
            colour=torch.zeros(3, dtype=position.dtype, device=position.device))

    @staticmethod
    def interpolate(values, barycentric_screen, barycentric_clip):
        """renderer/shaders/phong_reflection_shadow.py:152-171."""
        varying = Shader.interpolate(values, barycentric_screen, barycentric_clip)
        return varying._replace(texture_index=values.texture_index[0])

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        """renderer/shaders/phong_reflection_shadow.py:173-271: the shadow
        test s.z <= shadow_map[round(s.xy)] of s = normalise_homogeneous(
        viewport @ shadow_coord); in shadow only 1 - strength of the diffuse
        and specular light is kept."""
        built_in = Shader.fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra)[0]
        like = gl_FragCoord
        vp = _ops.as_tensor(extra.shadow.camera.viewport, like)
        s = _ops.normalise_homogeneous(_ops.apply(vp, varying.shadow_coord))
        depth = _ops.shadow_get(_ops.as_tensor(extra.shadow.shadow_map, like), s[:2])
        strength = _ops.as_tensor(extra.shadow.strength, like)
        shadow = torch.where(s[2] <= depth, torch.ones_like(strength), 1.0 - strength)
        colour = reflection_colour(varying, extra, like, shadow=shadow)
        keeps = torch.as_tensor(built_in.keeps, device=like.device) & torch.as_tensor(
            gl_FrontFacing)
        return (PerFragment(keeps=keeps, use_default_depth=built_in.use_default_depth),
                PhongReflectionShadowTextureExtraFragmentData(
                    normal=varying.normal, uv=varying.uv,
                    texture_index=torch.zeros_like(varying.texture_index),
                    shadow_coord=varying.shadow_coord, colour=colour))

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        mixed, out = Shader.mix(gl_FragDepth, keeps, extra)
        return mixed, PhongReflectionShadowTextureExtraMixerOutput(canvas=out.colour)
