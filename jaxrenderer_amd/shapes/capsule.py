"""create_capsule (renderer/shapes/capsule.py:1958-2015)."""
import enum

import numpy as np

from ..geometry import asf32
from ..model import Model
from ._tables import TABLES

_verts = TABLES["capsule_verts"]
_normals = TABLES["capsule_normals"]
_uvs = TABLES["capsule_uvs"]
_faces = TABLES["capsule_faces"]


class UpAxis(enum.IntEnum):
    X = 0
    Y = 1
    Z = 2


_SHUFFLE = np.array(((1, 2, 0), (0, 1, 2), (2, 0, 1)), dtype=np.int64)


def create_capsule(radius, half_height, up_axis, diffuse_map, specular_map) -> Model:
    """Capsule of 576 vertices / 192 triangles: the sphere table with its axes
    shuffled so `up_axis` points up, scaled by radius and each hemisphere
    moved by +-half_height along the up axis."""
    shuffled = _SHUFFLE[int(up_axis)]
    verts = (_verts[:, shuffled] * np.float32(radius)).astype(np.float32)
    ax = int(up_axis)
    hh = np.float32(half_height)
    verts[:, ax] = (verts[:, ax] + np.where(verts[:, ax] > 0, hh, -hh)).astype(np.float32)
    return Model(
        verts=verts,
        norms=_normals[:, shuffled].copy(),
        uvs=_uvs.copy(),
        faces=_faces.copy(),
        faces_norm=_faces.copy(),
        faces_uv=_faces.copy(),
        diffuse_map=asf32(diffuse_map),
        specular_map=asf32(specular_map),
    )
