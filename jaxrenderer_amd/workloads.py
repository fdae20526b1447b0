"""Seeded synthetic workloads used by the benchmark and the tests.

* config 1: the cube of the reference's examples (renderer/shapes/cube.py),
  phong_reflection, 320x240.
* config 2: one UV sphere of 48 x 96 segments with per-vertex radial noise,
  256x256 diffuse / specular maps, phong_reflection, 960x540.
* config 3: Brax-like scenes -- a ground plane plus 4..30 boxes, capsules
  (8 x 16 segments) and spheres (12 x 24 segments) at poses U[-3,3]^3 with
  uniform random orientations and 8x8 textures -- rendered at 84x84 with
  phong_reflection_shadow and a 256x256 shadow map.
* config 4: a deformed 5000-triangle sphere seen by 64 ring cameras at
  256x256 with a 512x512 texture, for differentiation.
* config 5: config-3-like scenes of 1k..20k triangles at 480x270 with a
  512x512 shadow map.

Everything is drawn from numpy.random.Generator(PCG64(seed)).
"""

from __future__ import annotations

import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .batch import BatchScene, ShadowMapSpec
from .geometry import Camera, transform_matrix_from_rotation
from .model import Model, ModelObject, merge_objects
from .renderer import CameraParameters, LightParameters, Renderer
from .shapes.cube import create_cube

f32 = np.float32


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(seed))


def uv_sphere(n_lat: int, n_lon: int, radius=1.0, half_height: float = 0.0,
              radial: np.ndarray = None) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """(verts, normals, uvs, faces) of a latitude/longitude sphere.  Rings of
    the upper hemisphere move up by half_height and the lower ones down (a
    capsule).  Quads touching a pole yield one degenerate triangle each."""
    lat = np.linspace(0.0, math.pi, n_lat + 1)
    lon = np.linspace(0.0, 2.0 * math.pi, n_lon + 1)
    th, ph = np.meshgrid(lat, lon, indexing="ij")
    n = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=-1)
    r = radius if radial is None else radial.reshape(n_lat + 1, n_lon + 1, 1)
    v = n * r
    v[..., 2] += np.where(n[..., 2] > 0, half_height, np.where(n[..., 2] < 0, -half_height, 0.0))
    uv = np.stack([ph / (2 * math.pi), 1.0 - th / math.pi], axis=-1)
    idx = np.arange((n_lat + 1) * (n_lon + 1)).reshape(n_lat + 1, n_lon + 1)
    a, b = idx[:-1, :-1], idx[1:, :-1]
    c, d = idx[1:, 1:], idx[:-1, 1:]
    faces = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3),
                            np.stack([a, c, d], -1).reshape(-1, 3)])
    return (v.reshape(-1, 3).astype(f32), n.reshape(-1, 3).astype(f32),
            uv.reshape(-1, 2).astype(f32), faces.astype(np.int32))


def mesh_model(verts, normals, uvs, faces, diffuse, specular) -> Model:
    return Model(verts=verts, norms=normals, uvs=uvs, faces=faces, faces_norm=faces,
                 faces_uv=faces, diffuse_map=diffuse.astype(f32), specular_map=specular.astype(f32))


def plane_model(half: float, diffuse, specular, uv_repeat: float = 4.0) -> Model:
    v = np.array([[-half, -half, 0.0], [half, -half, 0.0], [half, half, 0.0], [-half, half, 0.0]], f32)
    n = np.tile(np.array([0.0, 0.0, 1.0], f32), (4, 1))
    uv = np.array([[0, 0], [uv_repeat, 0], [uv_repeat, uv_repeat], [0, uv_repeat]], f32)
    faces = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return mesh_model(v, n, uv, faces, diffuse, specular)


def random_quaternion(rng) -> np.ndarray:
    """Uniform unit quaternion (w, x, y, z) (Shoemake's method)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([math.sqrt(u1) * math.cos(2 * math.pi * u3),
                  math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
                  math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
                  math.sqrt(u1) * math.sin(2 * math.pi * u3)])
    return q.astype(f32)


NOTEBOOK_LIGHT = LightParameters(
    direction=np.array((0.57735, -0.57735, 0.57735), f32),
    colour=np.ones(3, f32),
    ambient=np.full(3, 0.8, f32),
    diffuse=np.full(3, 0.8, f32),
    specular=np.full(3, 0.6, f32),
)


def brax_like_objects(rng) -> List[ModelObject]:
    spec8 = np.full((8, 8), 2.0, f32)
    objs = [ModelObject(model=plane_model(10.0, rng.random((8, 8, 3)).astype(f32), spec8),
                        transform=np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, -3.5],
                                            [0, 0, 0, 1]], f32))]
    for _ in range(int(rng.integers(4, 31))):
        kind = int(rng.integers(0, 3))
        tex = rng.random((8, 8, 3)).astype(f32)
        if kind == 0:
            he = rng.uniform(0.2, 0.6, size=3).astype(f32)
            model = create_cube(he, np.ones(2, f32), tex, spec8)
        elif kind == 1:
            r = float(rng.uniform(0.1, 0.3))
            model = mesh_model(*uv_sphere(8, 16, r, float(rng.uniform(0.2, 0.5))), tex, spec8)
        else:
            model = mesh_model(*uv_sphere(12, 24, float(rng.uniform(0.2, 0.6))), tex, spec8)
        t = np.identity(4, dtype=f32)
        t[:3, :3] = transform_matrix_from_rotation(random_quaternion(rng))
        t[:3, 3] = rng.uniform(-3.0, 3.0, size=3).astype(f32)
        objs.append(ModelObject(model=model, transform=t))
    return objs


def ring_camera(i: int, n: int, width: int, height: int, radius: float = 8.0,
                height_z: float = 4.0, fov: float = 60.0) -> Camera:
    a = 2.0 * math.pi * i / max(n, 1)
    cp = CameraParameters(viewWidth=width, viewHeight=height, hfov=fov, vfov=fov,
                          position=(radius * math.cos(a), radius * math.sin(a), height_z),
                          target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0))
    return Renderer.create_camera_from_parameters(cp)


def config3(n_scenes: int = 1024, seed: int = 0, width: int = 84, height: int = 84,
            shadow_size: int = 256, objects: Optional[list] = None
            ) -> Tuple[List[BatchScene], ShadowMapSpec]:
    """`objects`: a list that receives each scene's ModelObject list (for
    a ScenePoser)."""
    rng = _rng(seed)
    scenes = []
    for i in range(n_scenes):
        objs = brax_like_objects(rng)
        if objects is not None:
            objects.append(objs)
        model = merge_objects(objs)
        scenes.append(BatchScene(model=model, camera=ring_camera(i, n_scenes, width, height),
                                 light=NOTEBOOK_LIGHT))
    return scenes, ShadowMapSpec(width=shadow_size, height=shadow_size)


def config1(width: int = 320, height: int = 240) -> BatchScene:
    """The cube of create_cube(half_extents=1, texture_scaling=1) with a pure
    blue 2x2 diffuse map and specular 2.0, seen from (2, 4, 1) looking at the
    origin (up +z, vfov 60, square pixels), default light, phong_reflection."""
    hfov = float(np.degrees(2 * math.atan(math.tan(math.radians(30)) * width / height)))
    cp = CameraParameters(viewWidth=width, viewHeight=height, position=(2.0, 4.0, 1.0),
                          target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), vfov=60.0, hfov=hfov)
    diffuse = np.zeros((2, 2, 3), f32)
    diffuse[..., 2] = 1.0
    model = create_cube(np.ones(3, f32), np.ones(2, f32), diffuse, np.full((2, 2), 2.0, f32))
    return BatchScene(model=merge_objects([ModelObject(model=model)]),
                      camera=Renderer.create_camera_from_parameters(cp), light=LightParameters())


def config2(seed: int = 0, width: int = 960, height: int = 540) -> BatchScene:
    rng = _rng(seed)
    n_lat, n_lon = 48, 96
    noise = rng.normal(size=(n_lat + 1, n_lon + 1))
    # smooth along both grid axes (wrapping in longitude)
    for _ in range(4):
        noise = (noise + np.roll(noise, 1, 1) + np.roll(noise, -1, 1)) / 3.0
        noise[1:-1] = (noise[1:-1] + noise[:-2] + noise[2:]) / 3.0
    radial = 1.0 + 0.05 * noise / max(noise.std(), 1e-6)
    v, n, uv, f = uv_sphere(n_lat, n_lon, 1.0, 0.0, radial)
    tex = rng.random((256, 256, 3)).astype(f32)
    spec = rng.uniform(0.0, 4.0, size=(256, 256)).astype(f32)
    model = merge_objects([ModelObject(model=mesh_model(v, n, uv, f, tex, spec))])
    az = float(rng.uniform(0.0, 2.0 * math.pi))
    hfov = float(np.degrees(2 * math.atan(math.tan(math.radians(30)) * width / height)))
    cp = CameraParameters(viewWidth=width, viewHeight=height, vfov=60.0, hfov=hfov,
                          position=(3.0 * math.cos(az), 3.0 * math.sin(az), 0.5),
                          target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0))
    return BatchScene(model=model, camera=Renderer.create_camera_from_parameters(cp),
                      light=LightParameters())


def dense_brax_like_objects(rng, target_tris: int) -> List[ModelObject]:
    """A ground plane plus boxes, capsules and spheres of finer tessellation,
    added until the scene holds at least `target_tris` triangles."""
    spec8 = np.full((8, 8), 2.0, f32)
    objs = [ModelObject(model=plane_model(10.0, rng.random((8, 8, 3)).astype(f32), spec8),
                        transform=np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, -3.5],
                                            [0, 0, 0, 1]], f32))]
    total = 2
    while total < target_tris:
        kind = int(rng.integers(0, 3))
        tex = rng.random((8, 8, 3)).astype(f32)
        if kind == 0:
            he = rng.uniform(0.2, 0.6, size=3).astype(f32)
            model = create_cube(he, np.ones(2, f32), tex, spec8)
        else:
            n_lat = int(rng.choice([12, 16, 24, 32]))
            r = float(rng.uniform(0.15, 0.6))
            hh = float(rng.uniform(0.2, 0.5)) if kind == 1 else 0.0
            model = mesh_model(*uv_sphere(n_lat, 2 * n_lat, r, hh), tex, spec8)
        t = np.identity(4, dtype=f32)
        t[:3, :3] = transform_matrix_from_rotation(random_quaternion(rng))
        t[:3, 3] = rng.uniform(-3.0, 3.0, size=3).astype(f32)
        objs.append(ModelObject(model=model, transform=t))
        total += len(model.faces)
    return objs


def config5(n_scenes: int = 512, seed: int = 0, width: int = 480, height: int = 270,
            shadow_size: int = 512, min_tris: int = 1000, max_tris: int = 20000,
            targets: Optional[Sequence[int]] = None
            ) -> Tuple[List[BatchScene], ShadowMapSpec]:
    """Batched forward + backward workload: Brax-like scenes of 1k..20k
    triangles (target drawn log-uniformly per scene) at 480x270 with
    phong_reflection_shadow and a 512x512 shadow map.  `targets`: the
    triangle targets of the scenes instead (one scene each)."""
    rng = _rng(seed)
    scenes = []
    if targets is not None:
        n_scenes = len(targets)
    for i in range(n_scenes):
        target = int(math.exp(rng.uniform(math.log(min_tris), math.log(max_tris))))
        if targets is not None:
            target = int(targets[i])
        model = merge_objects(dense_brax_like_objects(rng, target))
        scenes.append(BatchScene(model=model, camera=ring_camera(i, n_scenes, width, height),
                                 light=NOTEBOOK_LIGHT))
    return scenes, ShadowMapSpec(width=shadow_size, height=shadow_size)


def config4(seed: int = 0, n_views: int = 64, size: int = 256, tex_size: int = 512,
            n_lat: int = 50, n_lon: int = 50, perturbed: bool = False):
    """Differentiation workload: a deformed UV sphere (n_lat x n_lon quads,
    5000 triangles by default) seen by `n_views` cameras on a ring of radius 3,
    size x size pixels, with a tex_size^2 x 3 diffuse texture and per-vertex
    colours.  `perturbed` gives the target scene: vertex noise sigma 0.02 and
    texture noise sigma 0.1 drawn from the same seed stream.  Returns
    (cameras, faces, positions, normals, uvs, colours, texture)."""
    rng = _rng(seed)
    noise = rng.normal(size=(n_lat + 1, n_lon + 1))
    for _ in range(4):
        noise = (noise + np.roll(noise, 1, 1) + np.roll(noise, -1, 1)) / 3.0
        noise[1:-1] = (noise[1:-1] + noise[:-2] + noise[2:]) / 3.0
    radial = 1.0 + 0.05 * noise / max(noise.std(), 1e-6)
    v, n, uv, f = uv_sphere(n_lat, n_lon, 1.0, 0.0, radial)
    # gouraud_texture and phong index texture[floor(uv) % shape]: texel units
    uv = (uv * f32(tex_size)).astype(f32)
    tex = rng.random((tex_size, tex_size, 3)).astype(f32)
    colours = rng.uniform(0.3, 1.0, size=(len(v), 3)).astype(f32)
    if perturbed:
        v = (v + rng.normal(0.0, 0.02, size=v.shape)).astype(f32)
        tex = np.clip(tex + rng.normal(0.0, 0.1, size=tex.shape), 0.0, 1.0).astype(f32)
    cams = []
    for i in range(n_views):
        a = 2.0 * math.pi * i / n_views
        cp = CameraParameters(viewWidth=size, viewHeight=size, hfov=45.0, vfov=45.0,
                              position=(3.0 * math.cos(a), 3.0 * math.sin(a), 0.8),
                              target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0))
        cams.append(Renderer.create_camera_from_parameters(cp))
    return cams, f, v, n, uv, colours, tex


def _quat_z_to(d: np.ndarray) -> np.ndarray:
    """Unit quaternion (w, x, y, z) rotating +z onto the unit vector d."""
    z = np.array((0.0, 0.0, 1.0))
    c = float(np.dot(z, d))
    if c < -1.0 + 1e-9:
        return np.array((0.0, 1.0, 0.0, 0.0), f32)
    axis = np.cross(z, d)
    q = np.array((1.0 + c, axis[0], axis[1], axis[2]))
    return (q / np.linalg.norm(q)).astype(f32)


def ant_like_objects(rng, torso_xy=(0.0, 0.0)) -> Tuple[List[ModelObject], np.ndarray]:
    """One frame of an ant-like walker, the composition of the reference's
    Brax ant benchmark scene (notebooks/84x84/A100.ipynb; 18 objects, 3,276
    triangles): a ground box (12 triangles) and 17 create_capsule capsules
    (192 triangles each) -- a spherical torso and four legs of four segments
    (hip, upper, thigh, shin) at joint angles drawn from `rng`.  Returns the
    objects and the torso centre."""
    from .shapes.capsule import UpAxis, create_capsule

    spec8 = np.full((8, 8), 2.0, f32)
    ground = create_cube(np.array((10.0, 10.0, 0.05), f32), np.full(2, 4.0, f32),
                         rng.random((8, 8, 3)).astype(f32), spec8)
    gt = np.identity(4, dtype=f32)
    gt[2, 3] = -0.05
    objs = [ModelObject(model=ground, transform=gt)]
    torso = np.array((torso_xy[0], torso_xy[1], 0.75 + 0.1 * rng.uniform(-1, 1)), f32)
    tex = rng.random((8, 8, 3)).astype(f32)

    def capsule(a: np.ndarray, b: np.ndarray, radius: float) -> ModelObject:
        d = b - a
        half = float(np.linalg.norm(d)) / 2.0
        t = np.identity(4, dtype=f32)
        t[:3, :3] = transform_matrix_from_rotation(_quat_z_to(d / (2.0 * half)))
        t[:3, 3] = ((a + b) / 2.0).astype(f32)
        model = create_capsule(f32(radius), f32(max(half - radius, 0.0)), UpAxis.Z, tex, spec8)
        return ModelObject(model=model, transform=t)

    sphere = create_capsule(f32(0.25), f32(0.0), UpAxis.Z, tex, spec8)
    tt = np.identity(4, dtype=f32)
    tt[:3, 3] = torso
    objs.append(ModelObject(model=sphere, transform=tt))
    yaw = rng.uniform(0.0, 2.0 * math.pi)
    for k in range(4):
        a = yaw + math.pi / 4 + k * math.pi / 2 + 0.3 * rng.uniform(-1, 1)
        out = np.array((math.cos(a), math.sin(a), 0.0))
        p0 = torso + 0.2 * out
        p1 = p0 + 0.15 * out
        p2 = p1 + 0.2 * out + np.array((0.0, 0.0, 0.05 * rng.uniform(-1, 1)))
        lift = rng.uniform(0.2, 0.6)
        p3 = p2 + 0.28 * (math.cos(lift) * out + np.array((0.0, 0.0, -math.sin(lift))))
        p4 = np.array((p3[0] + 0.15 * out[0], p3[1] + 0.15 * out[1], 0.08))
        for s0, s1 in ((p0, p1), (p1, p2), (p2, p3), (p3, p4)):
            objs.append(capsule(s0.astype(f32), s1.astype(f32), 0.08))
    return objs, torso


def ant_like(n_scenes: int = 200, seed: int = 0, width: int = 84, height: int = 84
             ) -> Tuple[List[BatchScene], ShadowMapSpec]:
    """Frames of ant-like walkers rendered as the reference's Brax benchmark
    renders its ant (notebooks/84x84/A100.ipynb cells 4 and 7): 3,276
    triangles per frame, a camera tracking the torso, the notebook's light,
    a shadow map the size of the canvas centred on the camera target
    (get_camera_image with ShadowParameters(centre=target)),
    phong_reflection_shadow."""
    rng = _rng(seed)
    scenes = []
    for i in range(n_scenes):
        xy = rng.uniform(-2.0, 2.0, size=2)
        objs, torso = ant_like_objects(rng, tuple(xy))
        az = rng.uniform(0.0, 2.0 * math.pi)
        target = tuple(float(v) for v in torso)
        cp = CameraParameters(viewWidth=width, viewHeight=height, hfov=60.0, vfov=60.0,
                              position=(torso[0] + 2.5 * math.cos(az),
                                        torso[1] + 2.5 * math.sin(az), 1.8),
                              target=target, up=(0.0, 0.0, 1.0))
        scenes.append(BatchScene(model=merge_objects(objs),
                                 camera=Renderer.create_camera_from_parameters(cp),
                                 light=NOTEBOOK_LIGHT, shadow_centre=target))
    return scenes, ShadowMapSpec(width=width, height=height)


def triangle_counts(scenes: Sequence[BatchScene]) -> np.ndarray:
    return np.array([len(s.model.faces) for s in scenes])


__all__ = ["ant_like", "ant_like_objects", "config1", "config2", "config3", "config4", "config5", "brax_like_objects",
           "dense_brax_like_objects", "uv_sphere", "NOTEBOOK_LIGHT"]
