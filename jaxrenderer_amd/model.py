"""Models, objects and scene merging (renderer/model.py), host side.

`merge_objects` flattens a list of `ModelObject` into one mesh exactly as the
reference does: world-space vertices, inverse-transpose normals normalised by
the norm of the whole array (renderer/geometry.py:36-47 applied to [N, 3]),
maps zero-padded and concatenated along axis 0, and a per-vertex texture
index.  The merged float32 arrays are what the GPU kernels and the CPU oracle
both consume.
"""

from __future__ import annotations

from typing import List, NamedTuple, Optional, Sequence, Tuple

import numpy as np

from .geometry import Camera, asf32, inv, transform_matrix_from_rotation

f32 = np.float32


def _i32(x) -> np.ndarray:
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x), dtype=np.int32)


class Model(NamedTuple):
    """renderer/model.py:37-143."""

    verts: np.ndarray
    norms: np.ndarray
    uvs: np.ndarray
    faces: np.ndarray
    faces_norm: np.ndarray
    faces_uv: np.ndarray
    diffuse_map: np.ndarray
    specular_map: np.ndarray

    @classmethod
    def create(cls, verts, norms, uvs, faces, diffuse_map, specular_map=None) -> "Model":
        diffuse_map = asf32(diffuse_map)
        if specular_map is None:
            specular_map = np.full(diffuse_map.shape[:2], 2.0, dtype=f32)
        faces = _i32(faces)
        return cls(
            verts=asf32(verts),
            norms=asf32(norms),
            uvs=asf32(uvs),
            faces=faces,
            faces_norm=faces,
            faces_uv=faces,
            diffuse_map=diffuse_map,
            specular_map=asf32(specular_map),
        )

    def asserts(self) -> None:
        assert self.verts.ndim == 2 and self.verts.shape[1] == 3
        assert self.norms.ndim == 2 and self.norms.shape[1] == 3
        assert self.uvs.ndim == 2 and self.uvs.shape[1] == 2
        for f in (self.faces, self.faces_norm, self.faces_uv):
            assert f.ndim == 2 and f.shape[1] == 3
        assert self.diffuse_map.ndim == 3
        assert self.specular_map.ndim == 2

    def value_checks(self) -> Tuple[Optional[str], None]:
        """renderer/model.py:105-143: index bounds.  Returns (error or None, None)."""
        for name, f, n in (
            ("faces", self.faces, len(self.verts)),
            ("faces_norm", self.faces_norm, len(self.norms)),
            ("faces_uv", self.faces_uv, len(self.uvs)),
        ):
            f = np.asarray(f)
            if f.size and (f.min() < 0 or f.max() >= n):
                return f"{name} out of bound, expected [0, {n}), got {f.max()}.", None
        return None, None


class MergedModel(NamedTuple):
    """renderer/model.py:155-339."""

    verts: np.ndarray
    norms: np.ndarray
    uvs: np.ndarray
    faces: np.ndarray
    faces_norm: np.ndarray
    faces_uv: np.ndarray
    texture_index: np.ndarray
    double_sided: np.ndarray
    texture_shape: np.ndarray
    offset: int
    diffuse_map: np.ndarray
    specular_map: np.ndarray

    @staticmethod
    def generate_object_vert_info(counts: Sequence[int], values: Sequence) -> np.ndarray:
        parts = [np.full((c,) + np.asarray(v).shape, v) for c, v in zip(counts, values)]
        return np.concatenate(parts, axis=0)

    @staticmethod
    def merge_verts(vs: Sequence[np.ndarray], fs: Sequence[np.ndarray]):
        cumsum = [0]
        for v in vs[:-1]:
            cumsum.append(cumsum[-1] + v.shape[0])
        dtype = np.result_type(*vs)
        verts = np.concatenate([np.asarray(v, dtype=dtype) for v in vs], axis=0)
        faces = np.concatenate([_i32(f) + np.int32(c) for f, c in zip(fs, cumsum)], axis=0)
        return verts, faces.astype(np.int32)

    @staticmethod
    def merge_maps(maps: Sequence[np.ndarray]) -> Tuple[np.ndarray, Tuple[int, int]]:
        """Pad every map with zeros to the largest extent of each axis and
        concatenate along axis 0 (renderer/model.py:249-301)."""
        dims = maps[0].ndim
        single = tuple(max(m.shape[i] for m in maps) for i in range(dims))
        dtype = np.result_type(*maps)
        padded = [
            np.pad(np.asarray(m), [(0, single[i] - m.shape[i]) for i in range(dims)]).astype(dtype)
            for m in maps
        ]
        return np.concatenate(padded, axis=0), (single[0], single[1])

    @staticmethod
    def uv_repeat(uv, shape, map_index, offset):
        """renderer/model.py:303-339 (torch or numpy)."""
        import torch

        if isinstance(uv, torch.Tensor):
            whole = torch.where(uv >= 0, torch.floor(uv), torch.ceil(uv))
            frac = uv - whole
            frac = torch.where(frac < 0, frac + 1, frac)
            r = frac * shape.to(frac.dtype)
            add = (map_index * offset).to(frac.dtype)
            return torch.stack([r[..., 0] + add, r[..., 1]], dim=-1)
        uv = asf32(uv)
        whole = np.where(uv >= 0, np.floor(uv), np.ceil(uv))
        frac = (uv - whole).astype(f32)
        frac = np.where(frac < 0, frac + f32(1), frac).astype(f32)
        r = (frac * np.asarray(shape).astype(f32)).astype(f32)
        r[..., 0] = r[..., 0] + f32(np.int32(map_index) * np.int32(offset))
        return r


class ModelObject(NamedTuple):
    """renderer/model.py:342-421."""

    model: Model
    local_scaling: np.ndarray = np.ones(3, dtype=f32)
    transform: np.ndarray = np.identity(4, dtype=f32)
    double_sided: bool = False

    def replace_with_position(self, position) -> "ModelObject":
        t = asf32(self.transform).copy()
        t[:3, 3] = asf32(position)
        return self._replace(transform=t)

    def replace_with_orientation(self, orientation=None, rotation_matrix=None) -> "ModelObject":
        if rotation_matrix is None:
            if orientation is None:
                orientation = np.array((0.0, 0.0, 0.0, 1.0), dtype=f32)
            rotation_matrix = transform_matrix_from_rotation(orientation)
        t = asf32(self.transform).copy()
        t[:3, :3] = asf32(rotation_matrix)
        return self._replace(transform=t)

    def replace_with_local_scaling(self, local_scaling) -> "ModelObject":
        return self._replace(local_scaling=asf32(local_scaling))

    def replace_with_double_sided(self, double_sided) -> "ModelObject":
        return self._replace(double_sided=bool(double_sided))


def batch_models(models: Sequence[MergedModel]) -> MergedModel:
    """renderer/model.py:424-442: stack equally-shaped merged models."""
    return MergedModel._make(
        np.stack([np.asarray(m[i]) for m in models], axis=0) for i in range(len(models[0]))
    )


def merge_objects(objects: Sequence[ModelObject]) -> MergedModel:
    """renderer/model.py:445-555."""
    models = [o.model for o in objects]
    counts: List[int] = [len(m.verts) for m in models]
    map_indices = MergedModel.generate_object_vert_info(counts, list(range(len(models)))).astype(
        np.int32
    )
    map_wh = np.asarray([m.diffuse_map.shape[:2] for m in models], dtype=np.int32)
    double_sided = MergedModel.generate_object_vert_info(
        counts, [bool(o.double_sided) for o in objects]
    ).astype(bool)
    diffuse_map, single_shape = MergedModel.merge_maps([asf32(m.diffuse_map) for m in models])
    specular_map = MergedModel.merge_maps([asf32(m.specular_map) for m in models])[0]

    world = [
        Camera.apply_pos((asf32(o.model.verts) * asf32(o.local_scaling)).astype(f32), o.transform)
        for o in objects
    ]
    verts, faces = MergedModel.merge_verts(world, [m.faces for m in models])
    normals = [
        Camera.apply_vec(asf32(o.model.norms), np.ascontiguousarray(inv(o.transform).T))
        for o in objects
    ]
    norms, faces_norm = MergedModel.merge_verts(normals, [m.faces_norm for m in models])
    uvs, faces_uv = MergedModel.merge_verts([asf32(m.uvs) for m in models],
                                            [m.faces_uv for m in models])
    return MergedModel(
        verts=verts.astype(f32),
        norms=norms.astype(f32),
        uvs=uvs.astype(f32),
        faces=faces,
        faces_norm=faces_norm,
        faces_uv=faces_uv,
        texture_index=map_indices,
        double_sided=double_sided,
        texture_shape=map_wh,
        offset=int(single_shape[0]),
        diffuse_map=diffuse_map.astype(f32),
        specular_map=specular_map.astype(f32),
    )


__all__ = ["Model", "MergedModel", "ModelObject", "batch_models", "merge_objects"]
