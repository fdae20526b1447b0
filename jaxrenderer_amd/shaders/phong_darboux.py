"""PhongTextureDarbouxShader (renderer/shaders/phong_darboux.py:106-316)."""
from typing import Any, NamedTuple

import numpy as np
import torch

from .._hip import SHADER_PHONG_DARBOUX
from ..geometry import Interpolation, interpolate
from ..shader import PerFragment, PerVertex, Shader
from ..types import LightSource
from . import _ops


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


def _normalise(v: torch.Tensor) -> torch.Tensor:
    return v / torch.linalg.vector_norm(v)


class PhongTextureDarbouxShader(Shader):
    """Phong with normals from a tangent-space (Darboux frame) normal map."""

    _jr_shader_id = SHADER_PHONG_DARBOUX

    @staticmethod
    def vertex(gl_VertexID, gl_InstanceID, camera, extra):
        """renderer/shaders/phong_darboux.py:116-170: besides the eye-space
        normal and uv, the NDC corners and uvs of the vertex's own face."""
        position = _ops.take(extra.position, gl_VertexID)
        gl_Position = _ops.to_clip(camera, position)
        face = _ops.take(extra.faces_indices, _ops.take(extra.id_to_face, gl_VertexID))
        corners = [_ops.take(extra.position, face[k]) for k in range(3)]
        triangle = torch.stack([_ops.to_cartesian(_ops.to_clip(camera, c)) for c in corners])
        triangle_uv = torch.stack([_ops.take(extra.uv, face[k]) for k in range(3)])
        normal = _ops.apply_vec(_ops.normalise3(_ops.take(extra.normal, gl_VertexID)),
                                camera.world_to_eye_norm)
        return PerVertex(gl_Position=gl_Position), PhongTextureDarbouxExtraFragmentData(
            normal=normal, uv=_ops.take(extra.uv, gl_VertexID), triangle=triangle,
            triangle_uv=triangle_uv,
            colour=torch.zeros(3, dtype=position.dtype, device=position.device))

    @staticmethod
    def interpolate(values, barycentric_screen, barycentric_clip):
        """renderer/shaders/phong_darboux.py:172-206: normal and uv smooth, the
        face's corners taken from the first vertex un-interpolated."""
        smooth = Interpolation.SMOOTH
        return PhongTextureDarbouxExtraFragmentData(
            normal=interpolate(values.normal, barycentric_screen, barycentric_clip, smooth),
            uv=interpolate(values.uv, barycentric_screen, barycentric_clip, smooth),
            triangle=values.triangle[0], triangle_uv=values.triangle_uv[0],
            colour=values.colour[0])

    @staticmethod
    def fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra):
        """renderer/shaders/phong_darboux.py:208-293: the tangent frame from
        inv([p1 - p0; p2 - p0; n]) and the uv differences, the normal map
        looked up in that frame, then phong lighting times texture."""
        built_in = Shader.fragment(gl_FragCoord, gl_FrontFacing, gl_PointCoord, varying, extra)[0]
        tex = _ops.as_tensor(extra.texture, gl_FragCoord)
        nmap = _ops.as_tensor(extra.normal_map, gl_FragCoord)
        u, v = _ops.floor_mod_index(varying.uv, tex.shape[:2])
        normal = _normalise(varying.normal)
        tri, tuv = varying.triangle, varying.triangle_uv
        A = torch.stack([tri[1] - tri[0], tri[2] - tri[0], normal])
        AI = torch.linalg.inv(A)
        zero = torch.zeros_like(tuv[0, 0])
        i = AI @ torch.stack([tuv[1, 0] - tuv[0, 0], tuv[2, 0] - tuv[0, 0], zero])
        j = AI @ torch.stack([tuv[1, 1] - tuv[0, 1], tuv[2, 1] - tuv[0, 1], zero])
        B = torch.stack([_normalise(i), _normalise(j), normal], dim=1)
        normal = _normalise(B @ _ops.take2(nmap, u, v))
        texture_colour = _ops.take2(tex, u, v)
        light_dir = _normalise(_ops.as_tensor(extra.light.direction, gl_FragCoord))
        light_colour = _ops.as_tensor(extra.light.colour, gl_FragCoord) * torch.dot(normal,
                                                                                     light_dir)
        colour = torch.where((light_colour >= 0).all(), texture_colour * light_colour,
                             torch.zeros_like(light_colour))
        keeps = torch.as_tensor(built_in.keeps, device=gl_FragCoord.device) & torch.as_tensor(
            gl_FrontFacing)
        return (PerFragment(keeps=keeps, use_default_depth=built_in.use_default_depth),
                varying._replace(colour=colour))

    @staticmethod
    def mix(gl_FragDepth, keeps, extra):
        mixed, out = Shader.mix(gl_FragDepth, keeps, extra)
        return mixed, PhongTextureDarbouxExtraMixerOutput(canvas=out.colour)
