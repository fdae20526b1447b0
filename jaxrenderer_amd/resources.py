"""Loaders for Wavefront OBJ meshes and TGA textures, as the reference's test
resources read them (test_resources/utils.py:47-114): the african-head mesh
and diffuse map its examples render with the Darboux-frame shader.

Plain file parsing on the host; nothing here runs on the GPU.
"""

from __future__ import annotations

import re
from typing import Iterable, NamedTuple, Union

import numpy as np

f32 = np.float32

_FLOAT = re.compile(r"(-?\d+\.?\d*(?:e[+-]\d+)?)")
_INTEGER = re.compile(r"\d+")
_ONE_VERTEX = re.compile(r"\d+/\d*/\d*")


class ObjModel(NamedTuple):
    """test_resources/utils.py:20-44: the arrays of one OBJ file (indices from
    0).  verts/norms/uv are float32, the face tables int32."""

    verts: np.ndarray
    norms: np.ndarray
    uv: np.ndarray
    faces: np.ndarray
    faces_norm: np.ndarray
    faces_uv: np.ndarray

    @property
    def nverts(self) -> int:
        return int(self.verts.shape[0])

    @property
    def nfaces(self) -> int:
        return int(self.faces.shape[0])


def make_model(lines: Iterable[str]) -> ObjModel:
    """test_resources/utils.py:47-96: `v`, `vn`, `vt` lines (the first three,
    three and two numbers) and triangular `f v/vt/vn` lines (1-based)."""
    verts, norms, uv = [], [], []
    faces, faces_norm, faces_uv = [], [], []
    for line in lines:
        if line.startswith("v "):
            verts.append(tuple(map(float, _FLOAT.findall(line, 2)[:3])))
        elif line.startswith("vn "):
            norms.append(tuple(map(float, _FLOAT.findall(line, 2)[:3])))
        elif line.startswith("vt "):
            uv.append(tuple(map(float, _FLOAT.findall(line, 2)[:2])))
        elif line.startswith("f "):
            corners = _ONE_VERTEX.findall(line)
            if len(corners) != 3:
                raise ValueError(f"expected 3 vertices in a face, got {len(corners)}")
            face, face_norm, face_uv = [], [], []
            for c in corners:
                idx = list(map(int, _INTEGER.findall(c)))
                if len(idx) != 3:
                    raise ValueError(f"expected 3 indices (v/vt/vn), got {len(idx)}")
                v, vt, vn = idx
                face.append(v - 1)
                face_norm.append(vn - 1)
                face_uv.append(vt - 1)
            faces.append(face)
            faces_norm.append(face_norm)
            faces_uv.append(face_uv)
    return ObjModel(
        verts=np.array(verts, dtype=f32).reshape(-1, 3),
        norms=np.array(norms, dtype=f32).reshape(-1, 3),
        uv=np.array(uv, dtype=f32).reshape(-1, 2),
        faces=np.array(faces, dtype=np.int32).reshape(-1, 3),
        faces_norm=np.array(faces_norm, dtype=np.int32).reshape(-1, 3),
        faces_uv=np.array(faces_uv, dtype=np.int32).reshape(-1, 3),
    )


def load_obj(path: str) -> ObjModel:
    with open(path) as f:
        return make_model(f.readlines())


def load_tga(path_or_image: Union[str, "object"]) -> np.ndarray:
    """test_resources/utils.py:99-114: a [width, height, 3] float32 texture
    whose entry [y, x] is the image's pixel at (x, y), channel values 0..255
    (not rescaled).  As in the reference, the layout needs a square image."""
    from PIL import Image

    image = Image.open(path_or_image) if isinstance(path_or_image, str) else path_or_image
    width, height = image.size
    if width != height:
        raise ValueError(f"load_tga needs a square image, got {width}x{height}")
    rgb = np.asarray(image.convert("RGB"), dtype=f32)  # [row y][column x]
    return np.ascontiguousarray(rgb[:, :, :3])


__all__ = ["ObjModel", "load_obj", "load_tga", "make_model"]
