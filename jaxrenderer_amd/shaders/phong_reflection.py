"""PhongReflectionTextureShader (renderer/shaders/phong_reflection.py:91-258)."""
from typing import Any, NamedTuple

import numpy as np
import torch

from .._hip import SHADER_PHONG_REFLECTION
from ..shader import PerFragment, PerVertex, Shader
from ..types import LightSource
from . import _ops


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


def reflection_colour(varying, extra, like: torch.Tensor, shadow=None) -> torch.Tensor:
    """The phong reflection model of renderer/shaders/phong_reflection.py:
    175-220 (and, with `shadow`, the attenuated form of
    phong_reflection_shadow.py:228-257): the merged map texel of
    uv_repeat(uv), diffuse max(n . l, 0), specular
    pow(max(normalise(2 (n . l) n - l)_z, 0), specular_map texel)."""
    ti = varying.texture_index.to(torch.int64)
    shape = _ops.take(_ops.as_tensor(extra.texture_shape, like, torch.int64), ti)
    u, v = _ops.uv_repeat_index(varying.uv, shape, ti, int(np.asarray(
        extra.texture_offset.cpu() if isinstance(extra.texture_offset, torch.Tensor)
        else extra.texture_offset)))
    tex = _ops.as_tensor(extra.texture, like)
    spec_map = _ops.as_tensor(extra.specular_map, like)
    texture_colour = _ops.take2(tex, u, v)
    normal = _ops.normalise3(varying.normal)
    light_dir = _ops.normalise3(_ops.as_tensor(extra.light_dir_eye, like))
    d = _ops.dot3(normal, light_dir)
    diffuse = _ops.max_nan(d, 0.0)
    reflected = _ops.normalise3((2.0 * d) * normal - light_dir)
    specular = torch.pow(_ops.max_nan(reflected[2], 0.0), _ops.take2(spec_map, u, v))
    k = (_ops.as_tensor(extra.diffuse, like) * diffuse
         + _ops.as_tensor(extra.specular, like) * specular)
    ambient = _ops.as_tensor(extra.ambient, like) * texture_colour
    light = _ops.as_tensor(extra.light.colour, like)
    if shadow is None:
        return ambient + (k * light) * texture_colour
    return ambient + ((shadow * k) * texture_colour) * light


class PhongReflectionTextureShader(Shader):
    """Phong reflection model (ambient + diffuse + specular) with merged maps."""

    _jr_shader_id = SHADER_PHONG_REFLECTION

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        """renderer/shaders/phong_reflection.py:100-132."""
        position = _ops.take(extra.position, gl_VertexID)
        gl_Position = _ops.to_clip(camera, position)
        normal = _ops.apply_vec(_ops.normalise3(_ops.take(extra.normal, gl_VertexID)),
                                camera.world_to_eye_norm)
        return PerVertex(gl_Position=gl_Position), PhongReflectionTextureExtraFragmentData(
            normal=normal, uv=_ops.take(extra.uv, gl_VertexID),
            texture_index=_ops.take(extra.texture_index, gl_VertexID),
            colour=torch.zeros(3, dtype=position.dtype, device=position.device))

    @staticmethod
    def interpolate(values, barycentric_screen, barycentric_clip):
        """renderer/shaders/phong_reflection.py:134-152: smooth, but the
        texture index of the first vertex."""
        varying = Shader.interpolate(values, barycentric_screen, barycentric_clip)
        return varying._replace(texture_index=values.texture_index[0])

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        """renderer/shaders/phong_reflection.py:154-235."""
        built_in = Shader.fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra)[0]
        colour = reflection_colour(varying, extra, gl_FragCoord)
        keeps = torch.as_tensor(built_in.keeps, device=gl_FragCoord.device) & torch.as_tensor(
            gl_FrontFacing)
        return (PerFragment(keeps=keeps, use_default_depth=built_in.use_default_depth),
                PhongReflectionTextureExtraFragmentData(
                    normal=varying.normal, uv=varying.uv,
                    texture_index=torch.zeros_like(varying.texture_index), colour=colour))

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        mixed, out = Shader.mix(gl_FragDepth, keeps, extra)
        return mixed, PhongReflectionTextureExtraMixerOutput(canvas=out.colour)
