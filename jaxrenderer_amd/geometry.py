"""Camera and small-matrix math on the host (numpy float32).

Restates renderer/geometry.py.  These run once per call on 4x4 matrices and
3-vectors, so they stay on the host; their float32 results are the inputs the
HIP kernels (and the CPU oracle) receive.  Each product is evaluated one IEEE
float32 operation at a time, sums left to right, in the order the reference
writes them; where JAX leaves the order to XLA/LAPACK (4x4 matmul, `inv`) the
canonical order is left-to-right sums and LU with partial pivoting.
"""

from __future__ import annotations

import enum
from typing import NamedTuple, Optional, Sequence, Union

import numpy as np

f32 = np.float32

ArrayLike = Union[np.ndarray, Sequence[float], float]


def asf32(x) -> np.ndarray:
    """float32 numpy copy of an array-like (torch tensors included)."""
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.array(x, dtype=f32, copy=True)


def _sum_sq(v: np.ndarray) -> np.float32:
    flat = np.ascontiguousarray(v, dtype=f32).reshape(-1)
    if flat.size <= 4:
        acc = f32(0.0)
        first = True
        for e in flat:
            acc = f32(e * e) if first else f32(acc + f32(e * e))
            first = False
        return acc
    # arrays (merge_objects normals): numpy's float32 sum of squares
    return f32(np.sum(flat * flat, dtype=f32))


def _is_torch(x) -> bool:
    return hasattr(x, "detach") and hasattr(x, "device")


def _torch_matrix(m, like):
    import torch

    if _is_torch(m):
        return m.to(device=like.device, dtype=like.dtype)
    return torch.as_tensor(np.asarray(m), device=like.device, dtype=like.dtype)


def normalise(vector: ArrayLike):
    """renderer/geometry.py:36-47: divide by the 2-norm of the WHOLE array.
    Torch tensors stay torch (and differentiable); anything else becomes a
    float32 numpy array."""
    if _is_torch(vector):
        import torch

        if vector.numel() <= 4:
            # small vectors: the sum of squares left to right, as the kernels
            flat = vector.reshape(-1)
            ss = flat[0] * flat[0]
            for k in range(1, flat.shape[0]):
                ss = ss + flat[k] * flat[k]
            return vector / torch.sqrt(ss)
        return vector / torch.linalg.norm(vector)
    v = asf32(vector)
    n = f32(np.sqrt(_sum_sq(v)))
    return (v / n).astype(f32)


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = asf32(a)
    b = asf32(b)
    return np.array(
        [
            f32(a[1] * b[2]) - f32(a[2] * b[1]),
            f32(a[2] * b[0]) - f32(a[0] * b[2]),
            f32(a[0] * b[1]) - f32(a[1] * b[0]),
        ],
        dtype=f32,
    )


def matmul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """float32 matrix product, each entry summed left to right."""
    a = asf32(a)
    b = asf32(b)
    squeeze = b.ndim == 1
    if squeeze:
        b = b[:, None]
    acc = (a[:, 0:1] * b[0:1, :]).astype(f32)
    for k in range(1, a.shape[1]):
        acc = (acc + (a[:, k : k + 1] * b[k : k + 1, :]).astype(f32)).astype(f32)
    return acc[:, 0] if squeeze else acc


def inv(a: ArrayLike) -> np.ndarray:
    """jnp.linalg.inv in the canonical order: LU with partial pivoting (first
    maximal pivot, reciprocal scaling) and reference-BLAS triangular solves."""
    A = asf32(a).copy()
    n = A.shape[0]
    perm = list(range(n))
    for j in range(n):
        p = j
        amax = abs(A[j, j])
        for i in range(j + 1, n):
            if abs(A[i, j]) > amax:
                amax = abs(A[i, j])
                p = i
        if A[p, j] != 0:
            if p != j:
                A[[p, j], :] = A[[j, p], :]
                perm[p], perm[j] = perm[j], perm[p]
            if abs(A[j, j]) >= np.finfo(f32).tiny:
                r = f32(f32(1.0) / A[j, j])
                for i in range(j + 1, n):
                    A[i, j] = f32(A[i, j] * r)
            else:
                for i in range(j + 1, n):
                    A[i, j] = f32(A[i, j] / A[j, j])
        for i in range(j + 1, n):
            for k in range(j + 1, n):
                A[i, k] = f32(A[i, k] - f32(A[i, j] * A[j, k]))
    X = np.zeros((n, n), dtype=f32)
    for r in range(n):
        X[r, perm[r]] = 1.0
    with np.errstate(all="ignore"):
        for c in range(n):
            for k in range(n):
                if X[k, c] != 0:
                    for i in range(k + 1, n):
                        X[i, c] = f32(X[i, c] - f32(X[k, c] * A[i, k]))
            for k in range(n - 1, -1, -1):
                if X[k, c] != 0:
                    X[k, c] = f32(X[k, c] / A[k, k])
                    for i in range(k):
                        X[i, c] = f32(X[i, c] - f32(X[k, c] * A[i, k]))
    return X


class Interpolation(enum.Enum):
    """renderer/geometry.py:50-110."""

    FLAT = 0
    NOPERSPECTIVE = 1
    SMOOTH = 2

    def __call__(self, values, barycentric_screen, barycentric_clip):
        import torch

        if self == Interpolation.FLAT:
            coef = torch.tensor([1.0, 0.0, 0.0], dtype=barycentric_clip.dtype,
                                device=barycentric_clip.device)
        elif self == Interpolation.NOPERSPECTIVE:
            coef = barycentric_screen
        else:
            coef = barycentric_clip
        v = values.to(coef.dtype)
        shape = (3,) + (1,) * (v.ndim - 1)
        c = coef.reshape(shape)
        return (c[0] * v[0] + c[1] * v[1]) + c[2] * v[2]


def interpolate(values, barycentric_screen, barycentric_clip,
                mode: Interpolation = Interpolation.SMOOTH):
    """renderer/geometry.py:113-138."""
    return mode(values, barycentric_screen, barycentric_clip)


def to_homogeneous(coordinates: ArrayLike, value: float = 1.0):
    if _is_torch(coordinates):
        import torch

        pad = torch.full_like(coordinates[..., :1], float(value))
        return torch.cat([coordinates, pad], dim=-1)
    c = asf32(coordinates)
    pad = np.full(c.shape[:-1] + (1,), value, dtype=f32)
    return np.concatenate([c, pad], axis=-1)


def normalise_homogeneous(coordinates: ArrayLike):
    if _is_torch(coordinates):
        return coordinates / coordinates[..., -1:]
    c = asf32(coordinates)
    with np.errstate(all="ignore"):
        return (c / c[..., -1:]).astype(f32)


def to_cartesian(coordinates: ArrayLike):
    if _is_torch(coordinates):
        import torch

        return torch.where(coordinates[..., -1:] == 0.0, coordinates[..., :-1],
                           normalise_homogeneous(coordinates)[..., :-1])
    c = asf32(coordinates)
    with np.errstate(all="ignore"):
        return np.where(c[..., -1:] == 0.0, c[..., :-1],
                        normalise_homogeneous(c)[..., :-1]).astype(f32)


def _radians(deg) -> np.float32:
    return f32(f32(deg) * f32(np.pi / 180))


class Camera(NamedTuple):
    """renderer/geometry.py:205-886 (float32 numpy matrices)."""

    view: np.ndarray
    projection: np.ndarray
    viewport: np.ndarray
    world_to_clip: np.ndarray
    world_to_eye_norm: np.ndarray
    world_to_screen: np.ndarray
    view_inv: np.ndarray
    screen_to_world: np.ndarray

    @classmethod
    def create(cls, view, projection, viewport, view_inv=None) -> "Camera":
        view = asf32(view)
        projection = asf32(projection)
        viewport = asf32(viewport)
        view_inv = inv(view) if view_inv is None else asf32(view_inv)
        if np.isclose(projection[3, 3], 0):
            projection_inv = cls.perspective_projection_matrix_inv(projection)
        else:
            projection_inv = cls.orthographic_projection_matrix_inv(projection)
        viewport_inv = cls.viewport_matrix_inv(viewport)
        return cls(
            view=view,
            projection=projection,
            viewport=viewport,
            world_to_clip=matmul(projection, view),
            world_to_eye_norm=np.ascontiguousarray(view_inv.T),
            world_to_screen=matmul(matmul(viewport, projection), view),
            view_inv=view_inv,
            screen_to_world=matmul(matmul(view_inv, projection_inv), viewport_inv),
        )

    @staticmethod
    def apply(points: ArrayLike, matrix: ArrayLike) -> np.ndarray:
        """out_i = sum_j p_j M[i, j] (renderer/geometry.py:280-315)."""
        if _is_torch(points):
            import torch

            M = _torch_matrix(matrix, points)
            # sums left to right, one rounding per operation, as the kernels
            rows = []
            for i in range(M.shape[0]):
                s = points[..., 0] * M[i, 0]
                for j in range(1, points.shape[-1]):
                    s = s + points[..., j] * M[i, j]
                rows.append(s)
            return torch.stack(rows, dim=-1)
        p = asf32(points)
        M = asf32(matrix)
        single = p.ndim == 1
        p2 = p[None, :] if single else p
        acc = (p2[:, 0:1] * M[None, :, 0]).astype(f32)
        for j in range(1, 4):
            acc = (acc + (p2[:, j : j + 1] * M[None, :, j]).astype(f32)).astype(f32)
        return acc[0] if single else acc

    @classmethod
    def apply_pos(cls, points, matrix) -> np.ndarray:
        return to_cartesian(cls.apply(to_homogeneous(points), matrix))

    @classmethod
    def apply_vec(cls, vectors, matrix) -> np.ndarray:
        v = normalise(vectors)
        h = cls.apply(to_homogeneous(v, 0.0), matrix)
        return normalise(h[..., :3])

    def to_screen(self, points) -> np.ndarray:
        return normalise_homogeneous(self.apply(points, self.world_to_screen))

    def to_clip(self, points) -> np.ndarray:
        return self.apply(points, self.world_to_clip)

    @staticmethod
    def inv_scale_translation_matrix(m: np.ndarray) -> np.ndarray:
        m = asf32(m)
        scale_inv = np.diag((f32(1.0) / np.diag(m)).astype(f32)).astype(f32)
        translation = matmul(scale_inv, m)
        t_inv = np.identity(4, dtype=f32)
        t_inv[:3, 3] = -translation[:3, 3]
        return matmul(t_inv, scale_inv)

    @staticmethod
    def view_matrix(eye, centre, up) -> np.ndarray:
        eye, centre, up = asf32(eye), asf32(centre), asf32(up)
        forward = normalise((centre - eye).astype(f32))
        up = normalise(up)
        side = normalise(_cross(forward, up))
        up = _cross(side, forward)
        m = np.identity(4, dtype=f32)
        m[0, :3] = side
        m[1, :3] = up
        m[2, :3] = -forward
        translation = np.identity(4, dtype=f32)
        translation[:3, 3] = -eye
        return matmul(m, translation)

    @staticmethod
    def view_matrix_inv(eye, centre, up) -> np.ndarray:
        eye, centre, up = asf32(eye), asf32(centre), asf32(up)
        forward = normalise((centre - eye).astype(f32))
        up = normalise(up)
        side = normalise(_cross(forward, up))
        up = _cross(side, forward)
        m = np.identity(4, dtype=f32)
        m[0, :3] = side
        m[1, :3] = up
        m[2, :3] = -forward
        t_inv = np.identity(4, dtype=f32)
        t_inv[:3, 3] = eye
        return matmul(t_inv, np.ascontiguousarray(m.T))

    @staticmethod
    def perspective_projection_matrix(fovy, aspect, z_near, z_far) -> np.ndarray:
        deg = f32(fovy)
        f = f32(f32(1.0) / f32(np.tan(f32(_radians(deg) / f32(2.0)))))
        aspect, z_near, z_far = f32(aspect), f32(z_near), f32(z_far)
        p = np.zeros((4, 4), dtype=f32)
        p[0, 0] = f32(f / aspect)
        p[1, 1] = f
        p[2, 2] = f32(f32(z_far + z_near) / f32(z_near - z_far))
        p[2, 3] = f32(f32(f32(f32(2.0) * z_far) * z_near) / f32(z_near - z_far))
        p[3, 2] = -1.0
        return p

    @classmethod
    def perspective_projection_matrix_inv(cls, mat) -> np.ndarray:
        shuffle = [0, 1, 3, 2]
        m = asf32(mat)[:, shuffle]
        return cls.inv_scale_translation_matrix(m)[shuffle, :]

    @staticmethod
    def orthographic_projection_matrix(left, right, bottom, top, z_near, z_far) -> np.ndarray:
        left, right, bottom, top = f32(left), f32(right), f32(bottom), f32(top)
        z_near, z_far = f32(z_near), f32(z_far)
        l_op = np.array([right, top, z_far], dtype=f32)
        r_op = np.array([left, bottom, z_near], dtype=f32)
        p = np.zeros((4, 4), dtype=f32)
        p[0, 0] = f32(f32(2.0) / f32(right - left))
        p[1, 1] = f32(f32(2.0) / f32(top - bottom))
        p[2, 2] = f32(f32(-2.0) / f32(z_far - z_near))
        p[3, 3] = 1.0
        p[:3, 3] = (-(l_op + r_op).astype(f32) / (l_op - r_op).astype(f32)).astype(f32)
        return p

    @classmethod
    def orthographic_projection_matrix_inv(cls, mat) -> np.ndarray:
        return cls.inv_scale_translation_matrix(mat)

    @staticmethod
    def viewport_matrix(lowerbound, dimension, depth) -> np.ndarray:
        lb = np.asarray(lowerbound)
        dim = np.asarray(dimension)
        half = (dim.astype(f32) / f32(2.0)).astype(f32)
        v = np.identity(4, dtype=f32)
        v[:2, 3] = (lb.astype(f32) + half).astype(f32)
        v[0, 0] = half[0]
        v[1, 1] = half[1]
        v[2, 2:] = f32(f32(depth) / f32(2.0))
        return v

    @classmethod
    def viewport_matrix_inv(cls, viewport) -> np.ndarray:
        return cls.inv_scale_translation_matrix(viewport)

    @staticmethod
    def world_to_screen_matrix(width: int, height: int) -> np.ndarray:
        a = np.identity(4, dtype=f32)
        a[0, 0] = 0.5
        a[1, 1] = 0.5
        b = np.eye(4, dtype=f32)
        b[0, 0] = width
        b[1, 1] = height
        c = np.identity(4, dtype=f32)
        c[:2, -1] = 1
        return matmul(matmul(a, b), c)


def quaternion(rotation_axis, rotation_angle) -> np.ndarray:
    """(w, x, y, z); angle in degrees (renderer/geometry.py:909-935)."""
    axis = normalise(asf32(rotation_axis))
    angle = _radians(rotation_angle)
    half = f32(angle / f32(2.0))
    q = np.zeros(4, dtype=f32)
    q[0] = f32(np.cos(half))
    q[1:] = (axis * f32(np.sin(half))).astype(f32)
    return q


def quaternion_mul(quatA, quatB) -> np.ndarray:
    """renderer/geometry.py:938-961."""
    a, b = asf32(quatA), asf32(quatB)

    def d(u, v):
        s = f32(u[0] * v[0])
        for k in range(1, len(u)):
            s = f32(s + f32(u[k] * v[k]))
        return s

    return np.array(
        [
            f32(f32(a[0] * b[0]) - d(a[1:], b[1:])),
            f32(d(a[:3], b[[1, 0, 3]]) - f32(a[3] * b[2])),
            f32(d(a[[2, 3, 0]], b[:3]) - f32(a[1] * b[3])),
            f32(d(a[[0, 1, 3]], b[[3, 2, 0]]) - f32(a[2] * b[1])),
        ],
        dtype=f32,
    )


def rotation_matrix(rotation_axis, rotation_angle) -> np.ndarray:
    """renderer/geometry.py:964-997 (glRotate convention)."""
    axis = normalise(asf32(rotation_axis))
    angle = _radians(rotation_angle)
    c = f32(np.cos(angle))
    s = f32(np.sin(angle))
    eye = np.identity(3, dtype=f32)
    cross = np.stack([_cross(axis, eye[i]) for i in range(3)]).astype(f32)
    outer = (axis[:, None] * axis[None, :]).astype(f32)
    return ((eye * c).astype(f32) - (s * cross).astype(f32)
            + (f32(1 - c) * outer).astype(f32)).astype(f32)


def transform_matrix_from_rotation(rotation) -> np.ndarray:
    """Quaternion (w, x, y, z) -> 3x3 (renderer/geometry.py:1000-1031)."""
    q = asf32(rotation)
    d = f32(f32(f32(f32(q[0] * q[0]) + f32(q[1] * q[1])) + f32(q[2] * q[2])) + f32(q[3] * q[3]))
    s = f32(f32(2.0) / d)
    rs = (q[1:] * s).astype(f32)
    outer = (q[:3, None] * rs[None, :]).astype(f32)
    (wx, wy, wz), (xx, xy, xz), (_, yy, yz) = outer
    zz = f32(q[3] * rs[2])
    one = f32(1.0)
    return np.array(
        [
            [one - f32(yy + zz), f32(xy - wz), f32(xz + wy)],
            [f32(xy + wz), one - f32(xx + zz), f32(yz - wx)],
            [f32(xz - wy), f32(yz + wx), one - f32(xx + yy)],
        ],
        dtype=f32,
    )


def compute_normal(triangle_verts) -> np.ndarray:
    t = asf32(triangle_verts)
    n = _cross((t[2] - t[0]).astype(f32), (t[1] - t[0]).astype(f32))
    return normalise(n)


__all__ = [
    "Camera",
    "Interpolation",
    "interpolate",
    "inv",
    "matmul",
    "normalise",
    "normalise_homogeneous",
    "quaternion",
    "quaternion_mul",
    "rotation_matrix",
    "to_cartesian",
    "to_homogeneous",
    "transform_matrix_from_rotation",
    "compute_normal",
]
