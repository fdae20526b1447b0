"""create_cube (renderer/shapes/cube.py:143-159)."""
import numpy as np

from ..geometry import asf32
from ..model import Model
from ._tables import TABLES

_verts = TABLES["cube_verts"]
_normals = TABLES["cube_normals"]
_uvs = TABLES["cube_uvs"]
_faces = TABLES["cube_faces"]


def create_cube(half_extents, texture_scaling, diffuse_map, specular_map) -> Model:
    """Unit cube (24 vertices, 12 triangles) scaled by half_extents, uvs by
    texture_scaling."""
    return Model(
        verts=(_verts * asf32(half_extents)).astype(np.float32),
        norms=_normals.copy(),
        uvs=(_uvs * asf32(texture_scaling)).astype(np.float32),
        faces=_faces.copy(),
        faces_norm=_faces.copy(),
        faces_uv=_faces.copy(),
        diffuse_map=asf32(diffuse_map),
        specular_map=asf32(specular_map),
    )
