"""Pins the CPU oracle (oracle/jr_oracle.c) before it is trusted as the GPU's
checker: against library float64 linear algebra, an independent float64
restatement of the reference pipeline (tests/pyref.py), closed forms, and the
assertions of the reference's own tests (tests/smoke_test.py)."""

import math

import numpy as np
import pytest

import jaxrenderer_amd as jr
from jaxrenderer_amd import _hip
from jaxrenderer_amd.geometry import Camera
from jaxrenderer_amd.pipeline import build_call
from jaxrenderer_amd.renderer import LightParameters
from jaxrenderer_amd.shaders import (
    DepthExtraInput,
    DepthShader,
    GouraudExtraInput,
    GouraudShader,
    PhongReflectionShadowTextureShader,
    PhongReflectionTextureShader,
)
from jaxrenderer_amd.types import LightSource
from jaxrenderer_amd.utils import transpose_for_display
from oracle import oracle
from oracle.oracle import run_oracle, run_oracle_plan
from tests import pyref, scenes

f32 = np.float32


# ---- 3x3 determinant and inverse ------------------------------------------

def test_inv3_against_float64_library():
    rng = np.random.default_rng(0)
    for _ in range(300):
        a = rng.normal(size=(3, 3)).astype(f32)
        if np.linalg.cond(a.astype(np.float64)) > 1e3:
            continue
        x = oracle.inv3(a).astype(np.float64)
        ref = np.linalg.inv(a.astype(np.float64))
        assert np.allclose(x, ref, rtol=1e-4, atol=1e-5 * np.abs(ref).max())
        assert np.allclose(a.astype(np.float64) @ x, np.eye(3), atol=1e-4)


def test_inv3_exact_cases():
    # diagonal: exact reciprocals; permutation: exact transpose (pins the
    # pivoting and the row permutation of the solve)
    d = np.diag([2.0, -4.0, 0.5]).astype(f32)
    assert np.array_equal(oracle.inv3(d), np.diag([0.5, -0.25, 2.0]).astype(f32))
    for perm in ([1, 2, 0], [2, 0, 1], [1, 0, 2], [0, 2, 1]):
        p = np.eye(3, dtype=f32)[perm]
        assert np.array_equal(oracle.inv3(p), p.T)
    # needs a row swap in the first column
    a = np.array([[1e-3, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]], f32)
    assert np.allclose(oracle.inv3(a).astype(np.float64) @ a, np.eye(3), atol=1e-6)


def test_det3_against_float64_library():
    rng = np.random.default_rng(1)
    for _ in range(300):
        a = rng.normal(size=(3, 3)).astype(f32)
        ref = np.linalg.det(a.astype(np.float64))
        assert abs(float(oracle.det3(a)) - ref) <= 1e-5 * (1 + np.abs(a).max() ** 3)
    p = np.eye(3, dtype=f32)[[1, 0, 2]]
    assert float(oracle.det3(p)) == -1.0


# ---- pipeline against the float64 restatement -----------------------------

def _tiny_gouraud(seed, W=40, H=30, n_tris=40):
    rng = np.random.default_rng(seed)
    pos = scenes._soup(rng, n_tris)
    faces = rng.permutation(3 * n_tris).astype(np.int32).reshape(n_tris, 3)
    normal = rng.normal(size=(3 * n_tris, 3)).astype(f32)
    colour = rng.uniform(-0.2, 1.0, size=(3 * n_tris, 3)).astype(f32)
    cam = Camera.create(
        view=Camera.view_matrix(np.array((0.3, 0.2, 3.0), f32), np.zeros(3, f32),
                                np.array((0.0, 1.0, 0.0), f32)),
        projection=Camera.perspective_projection_matrix(60.0, W / H, 0.1, 20.0),
        viewport=Camera.viewport_matrix(np.zeros(2, np.int32), np.array((W, H)), 1.0))
    light = LightSource(direction=np.array((0.3, 0.5, 0.8), f32), colour=np.array((1.0, 0.9, 0.8), f32))
    ex = GouraudExtraInput(position=pos, colour=colour, normal=normal, light=light)
    call = build_call(cam, GouraudShader, faces, ex, W, H, init_buffers=True)
    clip64 = np.concatenate([pos, np.ones((len(pos), 1), f32)], 1).astype(np.float64) @ \
        cam.world_to_clip.astype(np.float64).T
    return call, clip64, faces, cam, (pos, normal, colour, light)


@pytest.mark.parametrize("seed", range(6))
def test_gouraud_matches_float64_restatement(seed):
    call, clip64, faces, cam, (pos, normal, colour, light) = _tiny_gouraud(seed)
    got = run_oracle(call)
    vis = pyref.visibility(clip64, faces, call.width, call.height, cam.viewport)
    col, keep = pyref.gouraud_colour(pos.astype(np.float64), normal.astype(np.float64),
                                     colour.astype(np.float64),
                                     np.asarray(light.direction, np.float64),
                                     np.asarray(light.colour, np.float64), faces, vis)
    confident = (vis["edge_margin"] > 1e-4) & (vis["depth_gap"] > 1e-4) & \
        ((np.abs(col).min(-1) > 1e-4) | ~vis["keep"])
    assert confident.mean() > 0.7
    pid = np.where(keep, vis["idx"], -1)
    o_pid = got["prim_id"][0]
    assert np.array_equal(o_pid[confident], pid[confident])
    k = confident & keep
    assert k.sum() > 15
    z = got["zbuffer"][0]
    assert np.allclose(z[k], vis["depth"][k], rtol=2e-5, atol=2e-5)
    assert np.allclose(got["canvas"][0][k], col[k], atol=1e-4)
    assert np.allclose(got["bary"][0][k], vis["bary"][k], atol=1e-3)


def test_backfacing_primitive_zero_fallback():
    """When no kept front-facing primitive covers a pixel, the chooser's
    argmin returns index 0 and primitive 0's own keep & inside decides
    (renderer/shader.py:207-217, renderer/pipeline.py:275); the depth shader
    does not discard back faces, so a back-facing primitive 0 is drawn."""
    W = H = 16
    cam = scenes.smoke_camera(W, H, depth=1.0)
    # primitive 0 is clockwise on screen (back-facing), primitive 1 small and
    # front-facing in one corner
    pos = np.array([[-0.8, -0.8, 0.0], [-0.8, 0.8, 0.0], [0.8, -0.8, 0.0],
                    [0.4, 0.4, 0.5], [0.7, 0.4, 0.5], [0.4, 0.7, 0.5]], f32)
    faces = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    call = build_call(cam, DepthShader, faces, DepthExtraInput(position=pos), W, H,
                      init_buffers=True)
    got = run_oracle(call)
    clip64 = np.concatenate([pos, np.ones((6, 1), f32)], 1).astype(np.float64) @ \
        cam.world_to_clip.astype(np.float64).T
    vis = pyref.visibility(clip64, faces, W, H, cam.viewport)
    assert not vis["front"][vis["idx"] == 0].any()
    pid = got["prim_id"][0]
    assert (pid == 0).sum() > 20
    assert np.array_equal(pid, np.where(vis["keep"], vis["idx"], -1))


# ---- closed forms -----------------------------------------------------------

def test_config1_face_colours_closed_form():
    """Flat cube faces with a constant blue texture: every pixel of a face has
    ambient + diffuse*cos + specular*r_z^2, computed here in float64."""
    calls = scenes.c1_plan()
    res = run_oracle_plan(calls)["main"]
    pid, col = res["prim_id"][0], res["canvas"][0]
    cam = jr.Renderer.create_camera_from_parameters(jr.CameraParameters(
        viewWidth=320, viewHeight=240, position=(2.0, 4.0, 1.0), target=(0.0, 0.0, 0.0),
        up=(0.0, 0.0, 1.0), vfov=60.0,
        hfov=float(np.degrees(2 * math.atan(math.tan(math.radians(30)) * 4 / 3)))))
    light = LightParameters()
    R = cam.view[:3, :3].astype(np.float64)
    l_eye = R @ (np.asarray(light.direction, np.float64))
    l_eye /= np.linalg.norm(l_eye)
    from jaxrenderer_amd.shapes.cube import _faces, _normals
    seen = set()
    for f in np.unique(pid):
        if f < 0:
            continue
        n = R @ _normals[_faces[f, 0]].astype(np.float64)
        d = float(n @ l_eye)
        r = 2 * d * n - l_eye
        r /= np.linalg.norm(r)
        spec = max(r[2], 0.0) ** 2.0
        blue = 0.6 + (0.35 * max(d, 0.0) + 0.05 * spec)
        m = pid == f
        assert np.allclose(col[m][:, 2], blue, atol=1e-5)
        assert np.all(col[m][:, :2] == 0.0)
        seen.add(int(f) // 2)
    assert seen == {3, 5}  # +x and +y faces face the camera at (2, 4, 1)


# ---- the reference's own test assertions -----------------------------------

def test_reference_smoke_test_assertions():
    """tests/smoke_test.py:27-132 (1920x1080, GouraudShader, depth 255)."""
    res = run_oracle(scenes.smoke_gouraud_call(1920, 1080, zdef=1.0, cdef=0.0), prim_id=False,
                     bary=False)
    zbuffer = transpose_for_display(res["zbuffer"][0])
    canvas = transpose_for_display(res["canvas"][0])
    assert zbuffer.shape == (1080, 1920)
    assert canvas.shape == (1080, 1920, 3)
    assert np.unique(zbuffer[293:528, 964:1423].astype(np.uint8)).shape == (1,)
    assert np.all(zbuffer[590:1049, 964:1423] == 1.0)
    assert zbuffer[551, 914] < zbuffer[1026, 92]
    empty = (canvas == 0.0).all(axis=2).sum()
    assert 1920 * 1080 // 2 < empty < 1920 * 1080


def test_reference_perspective_test_depth_assertions():
    """tests/smoke_test.py:235-323: the zbuffer assertions (the custom
    checkerboard fragment keeps every fragment at its default depth, so the
    visibility pass decides the zbuffer)."""
    W, H = 1920, 1080
    cam = scenes.smoke_camera(W, H, eye=(0.0, 0.0, 1.0))
    pos = np.array([[-1.0, -1.0, -2.0], [1.0, -1.0, -1.0], [0.0, 1.0, -1.0]], f32)
    call = build_call(cam, DepthShader, np.array([[0, 1, 2]], np.int32),
                      DepthExtraInput(position=pos), W, H, init_buffers=True, zbuffer_default=1.0)
    call.shader = _hip.SHADER_VISIBILITY
    res = run_oracle(call, prim_id=False, bary=False)
    z = transpose_for_display(res["zbuffer"][0])
    assert (z == 1.0).sum() > W * H // 2
    assert z[679, 701] < z[779, 1388]


# ---- shadows ----------------------------------------------------------------

def _plane_box_scene(with_box: bool):
    W = H = 48
    spec = np.full((4, 4), 2.0, f32)
    tex = np.full((4, 4, 3), 0.8, f32)
    plane = jr.Model(verts=np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], f32),
                     norms=np.tile(np.array([0, 0, 1], f32), (4, 1)),
                     uvs=np.array([[0, 0], [1, 0], [1, 1], [0, 1]], f32),
                     faces=np.array([[0, 1, 2], [0, 2, 3]], np.int32),
                     faces_norm=np.array([[0, 1, 2], [0, 2, 3]], np.int32),
                     faces_uv=np.array([[0, 1, 2], [0, 2, 3]], np.int32),
                     diffuse_map=tex, specular_map=spec)
    objs = [jr.ModelObject(model=plane)]
    if with_box:
        box = jr.create_cube(np.full(3, 0.2, f32), np.ones(2, f32), tex, spec)
        t = np.identity(4, dtype=f32)
        t[:3, 3] = (0.0, 0.0, 0.5)
        objs.append(jr.ModelObject(model=box, transform=t))
    cam = jr.Renderer.create_camera_from_parameters(jr.CameraParameters(
        viewWidth=W, viewHeight=H, position=(0.3, -2.5, 2.5), target=(0.0, 0.0, 0.0),
        hfov=50.0, vfov=50.0))
    light = LightParameters(direction=np.array((0.1, 0.2, 1.0), f32))
    return jr.merge_objects(objs), light, cam, W, H


def test_shadow_without_occluder_is_fully_lit():
    model, light, cam, W, H = _plane_box_scene(False)
    lit = run_oracle_plan(jr.Renderer.plan(model, light, cam, W, H, jr.ShadowParameters()))
    plain = run_oracle_plan(jr.Renderer.plan(model, light, cam, W, H, None))
    assert (lit["main"]["prim_id"] >= 0).sum() > 500
    assert np.allclose(lit["main"]["canvas"], plain["main"]["canvas"], atol=1e-6)


def test_shadow_footprint_matches_ray_test():
    """A plane pixel is darkened iff the ray towards the light hits the box
    (float64 slab test), away from the footprint's edges."""
    model, light, cam, W, H = _plane_box_scene(True)
    calls = jr.Renderer.plan(model, light, cam, W, H, jr.ShadowParameters())
    res = run_oracle_plan(calls)["main"]
    pid, bary = res["prim_id"][0], res["bary"][0]
    verts = model.verts.astype(np.float64)
    faces = model.faces
    L = np.asarray(light.direction, np.float64)
    L /= np.linalg.norm(L)
    plain = run_oracle_plan(jr.Renderer.plan(model, light, cam, W, H, None))["main"]["canvas"][0]
    agree = total = 0
    for i in range(W):
        for j in range(H):
            f = pid[i, j]
            if f < 0 or f >= 2:  # plane pixels only
                continue
            P = (bary[i, j][:, None] * verts[faces[f]]).sum(0)
            with np.errstate(divide="ignore", invalid="ignore"):
                t1 = (np.array([-0.2, -0.2, 0.3]) - P) / L
                t2 = (np.array([0.2, 0.2, 0.7]) - P) / L
            tmin = np.nanmax(np.minimum(t1, t2))
            tmax = np.nanmin(np.maximum(t1, t2))
            hit = tmax >= max(tmin, 0.0)
            near_edge = abs(tmax - max(tmin, 0.0)) < 0.05
            if near_edge:
                continue
            shadowed = res["canvas"][0][i, j, 0] < plain[i, j, 0] - 1e-4
            total += 1
            agree += int(shadowed == hit)
    assert total > 300
    assert agree / total > 0.97


def test_display_output_matches_numpy_restatement():
    """The oracle's fused display store equals numpy's clip / cast /
    swapaxes / flip of its own canvas, including values outside [0, 1]."""
    from oracle.oracle import run_oracle
    from tests import scenes
    from tests.parity import display_of

    call = scenes.soup_call(GouraudShader, 4, 40, 30, 40)[0]
    call.colour_default = (1.5, -0.25, 0.5)
    out = run_oracle(call, display=True, prim_id=False, bary=False)
    assert out["display"].shape == (1, 30, 40, 3)
    np.testing.assert_array_equal(out["display"], display_of(out["canvas"]))
    flat = out["canvas"].reshape(-1, 3)
    assert (flat > 1).any() and (flat < 0).any()
    # pixel (i, j) lands at row H-1-j, column i, value truncated toward zero
    for i, j in ((5, 3), (0, 0), (39, 29), (17, 11)):
        for q in range(3):
            v = min(max(float(out["canvas"][0, i, j, q]), 0.0), 1.0)
            assert out["display"][0, 30 - 1 - j, i, q] == int(np.float32(v) * np.float32(255))
    u8 = (np.clip(out["canvas"][0], 0, 1) * np.float32(255)).astype(np.uint8)
    np.testing.assert_array_equal(out["display"][0], transpose_for_display(u8))
