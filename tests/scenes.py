"""Seeded scene builders shared by the CPU and GPU tests.

Every builder returns host-side descriptions (jaxrenderer_amd._plan.RasterCall
lists) that can be executed both by libjrhip.so and by the CPU oracle.
"""

from __future__ import annotations

import math
from typing import List

import numpy as np

import jaxrenderer_amd as jr
from jaxrenderer_amd import _hip
from jaxrenderer_amd._plan import RasterCall
from jaxrenderer_amd.geometry import Camera
from jaxrenderer_amd.pipeline import build_call
from jaxrenderer_amd.renderer import CameraParameters, LightParameters, Renderer
from jaxrenderer_amd.shaders import (
    DepthExtraInput,
    DepthShader,
    GouraudExtraInput,
    GouraudShader,
    GouraudTextureExtraInput,
    GouraudTextureShader,
    PhongReflectionShadowTextureExtraInput,
    PhongReflectionShadowTextureShader,
    PhongReflectionTextureExtraInput,
    PhongReflectionTextureShader,
    PhongTextureDarbouxExtraInput,
    PhongTextureDarbouxShader,
    PhongTextureExtraInput,
    PhongTextureShader,
)
from jaxrenderer_amd.shadow import Shadow
from jaxrenderer_amd.types import LightSource

f32 = np.float32


def c1_plan() -> List[RasterCall]:
    """Config 1: create_cube(1, 1), blue 2x2 diffuse, specular 2.0, camera at
    (2,4,1) looking at the origin (up +z, vfov 60, square pixels), default light,
    phong_reflection, 320x240."""
    W, H = 320, 240
    hfov = float(np.degrees(2 * math.atan(math.tan(math.radians(30)) * 4 / 3)))
    cam = CameraParameters(viewWidth=W, viewHeight=H, position=(2.0, 4.0, 1.0),
                           target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), vfov=60.0, hfov=hfov)
    diffuse = np.zeros((2, 2, 3), f32)
    diffuse[..., 2] = 1.0
    spec = np.full((2, 2), 2.0, f32)
    obj = jr.ModelObject(model=jr.create_cube(np.ones(3, f32), np.ones(2, f32), diffuse, spec))
    camera = Renderer.create_camera_from_parameters(cam)
    return Renderer.plan(jr.merge_objects([obj]), LightParameters(), camera, W, H)


def smoke_camera(width: int, height: int, eye=(0.0, 0.0, 2.0), depth: float = 255.0) -> Camera:
    """The camera of the reference's tests/smoke_test.py:27-51."""
    return Camera.create(
        view=Camera.view_matrix(eye=np.array(eye, f32), centre=np.zeros(3, f32),
                                up=np.array((0.0, 1.0, 0.0), f32)),
        projection=Camera.perspective_projection_matrix(fovy=90.0, aspect=1.0, z_near=-1.0,
                                                        z_far=1.0),
        viewport=Camera.viewport_matrix(lowerbound=np.zeros(2, np.int32),
                                        dimension=np.array((width, height)), depth=depth),
    )


SMOKE_FACES = np.array(((0, 1, 2), (1, 3, 2), (0, 2, 4), (0, 4, 3), (2, 5, 1)), np.int32)
SMOKE_POSITIONS = np.array(((0.0, 0.0, 0.0), (2.0, 0.0, 0.0), (0.0, 1.0, 0.0), (1.0, 1.0, 0.0),
                            (-1, -1, 1.0), (-2, 0.0, 0.0)), f32)
SMOKE_COLOURS = np.array(((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (0.0, 0.0, 0.0),
                          (1.0, 1.0, 1.0), (1.0, 1.0, 0.0)), f32)


def smoke_gouraud_call(width: int = 1920, height: int = 1080, zdef: float = 1.0,
                       cdef: float = 0.0, depth: float = 255.0) -> RasterCall:
    """tests/smoke_test.py:27-132: five triangles, GouraudShader, merged into
    buffers of zbuffer=zdef, canvas=cdef."""
    light = LightSource()
    extra = GouraudExtraInput(
        position=SMOKE_POSITIONS,
        colour=SMOKE_COLOURS,
        normal=np.tile(np.asarray(light.direction, f32), (6, 1)),
        light=light,
    )
    call = build_call(smoke_camera(width, height, depth=depth), GouraudShader, SMOKE_FACES, extra,
                      width, height, init_buffers=True, zbuffer_default=zdef,
                      colour_default=(cdef, cdef, cdef))
    return call


def _soup(rng, n_tris: int, behind: float = 0.1, degenerate: float = 0.05):
    """Random triangles around the origin, some reaching behind the camera and
    some degenerate (repeated vertex)."""
    centres = rng.uniform(-1.5, 1.5, size=(n_tris, 1, 3)).astype(f32)
    pos = (centres + rng.normal(0.0, 0.5, size=(n_tris, 3, 3))).astype(f32)
    k = rng.random(n_tris) < behind
    pos[k, 0, 2] = rng.uniform(4.0, 9.0, size=k.sum()).astype(f32)  # past the eye at z=3
    d = rng.random(n_tris) < degenerate
    pos[d, 2] = pos[d, 1]
    return pos.reshape(-1, 3)


def soup_call(shader: type, seed: int, width: int = 64, height: int = 48, n_tris: int = 60,
              init: bool = True) -> List[RasterCall]:
    """A seeded triangle soup rendered with one of the built-in shaders.
    Returns [shadow pass, main pass] for the shadow shader, else [main]."""
    rng = np.random.default_rng(seed)
    V = 3 * n_tris
    pos = _soup(rng, n_tris)
    faces = rng.permutation(V).astype(np.int32).reshape(n_tris, 3)
    normal = rng.normal(size=(V, 3)).astype(f32)
    uv = rng.uniform(-2.0, 3.0, size=(V, 2)).astype(f32)
    colour = rng.uniform(-0.2, 1.0, size=(V, 3)).astype(f32)
    tex = rng.uniform(0.0, 1.0, size=(7, 5, 3)).astype(f32)
    camera = Camera.create(
        view=Camera.view_matrix(eye=np.array((0.3, 0.2, 3.0), f32), centre=np.zeros(3, f32),
                                up=np.array((0.0, 1.0, 0.0), f32)),
        projection=Camera.perspective_projection_matrix(60.0, width / height, 0.1, 20.0),
        viewport=Camera.viewport_matrix(np.zeros(2, np.int32), np.array((width, height)), 1.0),
    )
    light = LightSource(direction=np.array((0.3, 0.5, 0.8), f32),
                        colour=np.array((1.0, 0.9, 0.8), f32))
    kw = dict(init_buffers=init, zbuffer_default=1.0, colour_default=(0.1, 0.2, 0.3))
    if shader is DepthShader:
        return [build_call(camera, shader, faces, DepthExtraInput(position=pos), width, height, **kw)]
    if shader is GouraudShader:
        ex = GouraudExtraInput(position=pos, colour=colour, normal=normal, light=light)
    elif shader is GouraudTextureShader:
        ex = GouraudTextureExtraInput(position=pos, normal=normal, uv=uv, light=light, texture=tex)
    elif shader is PhongTextureShader:
        ex = PhongTextureExtraInput(position=pos, normal=normal, uv=uv, light=light, texture=tex)
    elif shader is PhongTextureDarbouxShader:
        nmap = rng.normal(size=(6, 4, 3)).astype(f32)
        id_to_face = np.repeat(np.arange(n_tris, dtype=np.int32), 3)
        # faces_indices[id_to_face[v]] must list the slots of v's face
        slot_faces = np.arange(V, dtype=np.int32).reshape(n_tris, 3)
        ex = PhongTextureDarbouxExtraInput(position=pos, normal=normal, uv=uv, light=light,
                                           texture=tex, normal_map=nmap, id_to_face=id_to_face,
                                           faces_indices=slot_faces)
        faces = slot_faces
    else:
        n_obj = 3
        texm = rng.uniform(0.0, 1.0, size=(4 * n_obj, 6, 3)).astype(f32)
        spec = rng.uniform(0.0, 4.0, size=(4 * n_obj, 6)).astype(f32)
        tshape = np.array([[4, 6], [3, 5], [4, 2]], np.int32)
        tindex = rng.integers(0, n_obj, size=V).astype(np.int32)
        ldir = np.array((0.3, 0.5, 0.8), f32)
        common = dict(position=pos, normal=normal, uv=uv, light=light,
                      light_dir_eye=Camera.apply_vec(ldir, camera.view), texture_shape=tshape,
                      texture_index=tindex, texture_offset=4, texture=texm, specular_map=spec,
                      ambient=np.array((0.3, 0.3, 0.3), f32),
                      diffuse=np.array((0.5, 0.6, 0.7), f32),
                      specular=np.array((0.4, 0.3, 0.2), f32))
        if shader is PhongReflectionTextureShader:
            ex = PhongReflectionTextureExtraInput(**common)
        else:
            SW, SH = 40, 36
            vp = Camera.viewport_matrix(np.zeros(2, np.int32), np.array((SW, SH)), 1.0)
            scam = Shadow.camera_for(np.array((0.2, 0.3, 1.0), f32), vp, np.zeros(3, f32),
                                     np.array((0.0, 1.0, 0.0), f32), distance=4.0)
            scam = scam._replace(world_to_clip=(scam.world_to_clip * np.array(
                [[0.5], [0.5], [0.2], [1.0]], f32)).astype(f32))
            shadow_call = build_call(scam, DepthShader, faces, DepthExtraInput(position=pos), SW,
                                     SH, init_buffers=True,
                                     zbuffer_default=float(np.finfo(f32).max),
                                     apply_offset=True, depth_offset=0.05)
            shadow_call.tag = "shadow"
            sh = Shadow(shadow_map=np.zeros((SW, SH), f32), strength=np.array((0.6, 0.5, 0.4), f32),
                        camera=scam)
            ex = PhongReflectionShadowTextureExtraInput(**common, shadow=sh, camera=camera)
            main = build_call(camera, shader, faces, ex, width, height, **kw)
            main.shadow_from = "shadow"
            return [shadow_call, main]
    return [build_call(camera, shader, faces, ex, width, height, **kw)]


def soup_inputs(shader: type, seed: int, width: int = 48, height: int = 40, n_tris: int = 40):
    """Raw inputs of a seeded soup for the gradient tests: (camera, faces,
    extra, shadow_call) with numpy arrays; shadow_call is the depth pass whose
    zbuffer feeds extra.shadow.shadow_map (None for other shaders)."""
    calls = soup_call(shader, seed, width, height, n_tris)
    rng = np.random.default_rng(seed)
    V = 3 * n_tris
    pos = _soup(rng, n_tris)
    faces = rng.permutation(V).astype(np.int32).reshape(n_tris, 3)
    normal = rng.normal(size=(V, 3)).astype(f32)
    uv = rng.uniform(-2.0, 3.0, size=(V, 2)).astype(f32)
    colour = rng.uniform(-0.2, 1.0, size=(V, 3)).astype(f32)
    tex = rng.uniform(0.0, 1.0, size=(7, 5, 3)).astype(f32)
    camera = Camera.create(
        view=Camera.view_matrix(eye=np.array((0.3, 0.2, 3.0), f32), centre=np.zeros(3, f32),
                                up=np.array((0.0, 1.0, 0.0), f32)),
        projection=Camera.perspective_projection_matrix(60.0, width / height, 0.1, 20.0),
        viewport=Camera.viewport_matrix(np.zeros(2, np.int32), np.array((width, height)), 1.0),
    )
    light = LightSource(direction=np.array((0.3, 0.5, 0.8), f32),
                        colour=np.array((1.0, 0.9, 0.8), f32))
    shadow_call = calls[0] if len(calls) == 2 else None
    if shader is DepthShader:
        return camera, faces, DepthExtraInput(position=pos), None
    if shader is GouraudShader:
        return camera, faces, GouraudExtraInput(position=pos, colour=colour, normal=normal,
                                                light=light), None
    if shader is GouraudTextureShader:
        return camera, faces, GouraudTextureExtraInput(position=pos, normal=normal, uv=uv,
                                                       light=light, texture=tex), None
    if shader is PhongTextureShader:
        return camera, faces, PhongTextureExtraInput(position=pos, normal=normal, uv=uv,
                                                     light=light, texture=tex), None
    if shader is PhongTextureDarbouxShader:
        nmap = rng.normal(size=(6, 4, 3)).astype(f32)
        id_to_face = np.repeat(np.arange(n_tris, dtype=np.int32), 3)
        slot_faces = np.arange(V, dtype=np.int32).reshape(n_tris, 3)
        return camera, slot_faces, PhongTextureDarbouxExtraInput(
            position=pos, normal=normal, uv=uv, light=light, texture=tex, normal_map=nmap,
            id_to_face=id_to_face, faces_indices=slot_faces), None
    main = calls[-1]
    # the reflection shaders' soup draws its extra arrays after the ones above
    n_obj = 3
    texm = rng.uniform(0.0, 1.0, size=(4 * n_obj, 6, 3)).astype(f32)
    spec = rng.uniform(0.0, 4.0, size=(4 * n_obj, 6)).astype(f32)
    tshape = np.array([[4, 6], [3, 5], [4, 2]], np.int32)
    tindex = rng.integers(0, n_obj, size=V).astype(np.int32)
    ldir = np.array((0.3, 0.5, 0.8), f32)
    common = dict(position=pos, normal=normal, uv=uv, light=light,
                  light_dir_eye=Camera.apply_vec(ldir, camera.view), texture_shape=tshape,
                  texture_index=tindex, texture_offset=4, texture=texm, specular_map=spec,
                  ambient=np.array((0.3, 0.3, 0.3), f32), diffuse=np.array((0.5, 0.6, 0.7), f32),
                  specular=np.array((0.4, 0.3, 0.2), f32))
    assert np.array_equal(main.inputs["textures"], texm)
    if shader is PhongReflectionTextureShader:
        return camera, faces, PhongReflectionTextureExtraInput(**common), None
    scam = None
    SW, SH = shadow_call.width, shadow_call.height
    vp = Camera.viewport_matrix(np.zeros(2, np.int32), np.array((SW, SH)), 1.0)
    scam = Shadow.camera_for(np.array((0.2, 0.3, 1.0), f32), vp, np.zeros(3, f32),
                             np.array((0.0, 1.0, 0.0), f32), distance=4.0)
    scam = scam._replace(world_to_clip=(scam.world_to_clip * np.array(
        [[0.5], [0.5], [0.2], [1.0]], f32)).astype(f32))
    sh = Shadow(shadow_map=None, strength=np.array((0.6, 0.5, 0.4), f32), camera=scam)
    return camera, faces, PhongReflectionShadowTextureExtraInput(**common, shadow=sh,
                                                                 camera=camera), shadow_call


def tiny_soup_call(seed: int, width: int = 96, height: int = 64, n_tris: int = 3000,
                   shader: type = DepthShader) -> List[RasterCall]:
    """Edge cases of the pixel-range culling: sub-pixel and few-pixel
    triangles, slivers, and (through an orthographic camera whose clip x, y are
    the world x, y) triangles whose vertices and edges sit exactly on pixel
    sample positions, so that edge values of exactly 0 decide coverage."""
    rng = np.random.default_rng(seed)
    W, H = width, height
    # NDC sample position of pixel i is (i - W/2) / (W/2), exactly representable
    sx = lambda i: (np.asarray(i, f32) - f32(W / 2)) / f32(W / 2)  # noqa: E731
    sy = lambda j: (np.asarray(j, f32) - f32(H / 2)) / f32(H / 2)  # noqa: E731
    n_grid = n_tris // 2
    i0 = rng.integers(-2, W + 2, size=n_grid)
    j0 = rng.integers(-2, H + 2, size=n_grid)
    di = rng.integers(-3, 4, size=(n_grid, 2))
    dj = rng.integers(-3, 4, size=(n_grid, 2))
    grid = np.zeros((n_grid, 3, 3), f32)
    grid[:, 0, 0], grid[:, 0, 1] = sx(i0), sy(j0)
    grid[:, 1, 0], grid[:, 1, 1] = sx(i0 + di[:, 0]), sy(j0 + dj[:, 0])
    grid[:, 2, 0], grid[:, 2, 1] = sx(i0 + di[:, 1]), sy(j0 + dj[:, 1])
    # half of those moved by a fraction of a pixel, some by one float ulp
    frac = rng.random(n_grid) < 0.5
    grid[frac, :, :2] += rng.choice([-0.5, 0.25, 1e-7, -1e-7], size=(frac.sum(), 3, 2)).astype(
        f32) * f32(2.0 / W)
    grid[:, :, 2] = rng.uniform(-0.9, 0.9, size=(n_grid, 3)).astype(f32)
    n_free = n_tris - n_grid
    c = rng.uniform(-1.05, 1.05, size=(n_free, 1, 3)).astype(f32)
    size = np.exp(rng.uniform(np.log(0.002), np.log(0.1), size=(n_free, 1, 1))).astype(f32)
    free = (c + size * rng.normal(size=(n_free, 3, 3))).astype(f32)
    sliver = rng.random(n_free) < 0.2
    t = rng.random((sliver.sum(), 1)).astype(f32)
    free[sliver, 2] = (free[sliver, 0] * t + free[sliver, 1] * (1 - t)).astype(f32)
    free[:, :, 2] = np.clip(free[:, :, 2], -0.95, 0.95)
    pos = np.concatenate([grid, free]).reshape(-1, 3)
    # random winding: about half the triangles face away
    faces = np.arange(3 * n_tris, dtype=np.int32).reshape(n_tris, 3)
    flip = rng.random(n_tris) < 0.5
    faces[flip] = faces[flip][:, ::-1]
    camera = Camera.create(
        view=np.identity(4, dtype=f32),
        projection=np.identity(4, dtype=f32),
        viewport=Camera.viewport_matrix(np.zeros(2, np.int32), np.array((W, H)), 1.0),
    )
    kw = dict(init_buffers=True, zbuffer_default=1.0, colour_default=(0.1, 0.2, 0.3))
    if shader is DepthShader:
        return [build_call(camera, shader, faces, DepthExtraInput(position=pos), W, H, **kw)]
    V = 3 * n_tris
    ex = GouraudExtraInput(position=pos, colour=rng.uniform(0.0, 1.0, size=(V, 3)).astype(f32),
                           normal=np.tile(np.array((0.0, 0.0, 1.0), f32), (V, 1)),
                           light=LightSource(direction=np.array((0.0, 0.0, 1.0), f32)))
    return [build_call(camera, shader, faces, ex, W, H, **kw)]


def darboux_head(size: int = 512, map_size: int = 512, seed: int = 0):
    """The reference examples' african head (tests/golden/african_head.obj)
    flattened per corner as Renderer.render does, for the Darboux-frame
    shader: uvs in texel units, a seeded diffuse map and a tangent-space
    normal map of smooth bumps around (0, 0, 1).  Returns (camera, faces,
    extra)."""
    import os

    from jaxrenderer_amd.resources import load_obj

    m = load_obj(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                              "african_head.obj"))
    rng = np.random.default_rng(seed)
    F = m.nfaces
    pos = m.verts[m.faces.reshape(-1)]
    nrm = m.norms[m.faces_norm.reshape(-1)]
    uv = (m.uv[m.faces_uv.reshape(-1)] * f32(map_size)).astype(f32)
    tex = rng.uniform(0.2, 1.0, size=(map_size, map_size, 3)).astype(f32)
    g = np.linspace(0.0, 8.0 * math.pi, map_size, dtype=np.float64)
    bump = np.stack([0.3 * np.sin(g)[:, None] * np.ones((1, map_size)),
                     0.3 * np.cos(g)[None, :] * np.ones((map_size, 1)),
                     np.ones((map_size, map_size))], axis=-1)
    nmap = (bump / np.linalg.norm(bump, axis=-1, keepdims=True)).astype(f32)
    slots = np.arange(3 * F, dtype=np.int32).reshape(F, 3)
    extra = PhongTextureDarbouxExtraInput(
        position=pos, normal=nrm, uv=uv,
        light=LightSource(direction=np.array((0.4, 0.5, 0.9), f32), colour=np.ones(3, f32)),
        texture=tex, normal_map=nmap, id_to_face=np.repeat(np.arange(F, dtype=np.int32), 3),
        faces_indices=slots)
    camera = Camera.create(
        view=Camera.view_matrix(eye=np.array((0.8, 0.3, 2.5), f32), centre=np.zeros(3, f32),
                                up=np.array((0.0, 1.0, 0.0), f32)),
        projection=Camera.perspective_projection_matrix(45.0, 1.0, 0.1, 10.0),
        viewport=Camera.viewport_matrix(np.zeros(2, np.int32), np.array((size, size)), 1.0),
    )
    return camera, slots, extra
