"""Pins the oracle's phong_darboux branch (renderer/shaders/phong_darboux.py:
116-293) without the reference:

* a normal map of (0, 0, 1) everywhere turns B @ nm into the interpolated
  normal itself, so darboux must give PhongTextureShader's image (same
  primitives, colour within 1e-6);
* a flat triangle in the z = 0 plane under an identity camera (clip = world,
  eye-space normal (0, 0, 1)): the tangent i is the gradient of u over the
  triangle and j that of v, found here by Cramer's rule on the 2x2 system
  (p1 - p0).i = du1, (p2 - p0).i = du2; with a normal map of (1, 0, 0),
  (0, 1, 0) or a blend the colour is texture x light colour x dot(the map's
  direction in the frame (i/|i|, j/|j|, n), light) in closed form.  The
  triangle's edges are not orthogonal, so a transposed inverse, swapped
  tangents or a transposed frame matrix all change the colour;
* a float64 torch restatement (tests/torchref.py) on random soups, with
  library inverses, within 1e-4 at every pixel away from decision boundaries.
"""
import numpy as np
import pytest
import torch

from jaxrenderer_amd.geometry import Camera
from jaxrenderer_amd.pipeline import build_call
from jaxrenderer_amd.shaders import (
    PhongTextureDarbouxExtraInput,
    PhongTextureDarbouxShader,
    PhongTextureExtraInput,
    PhongTextureShader,
)
from jaxrenderer_amd.types import LightSource
from oracle.oracle import run_oracle
from tests import scenes, torchref

f32 = np.float32


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_flat_normal_map_reduces_to_phong(seed):
    W, H = 64, 48
    cam, faces, ex, _ = scenes.soup_inputs(PhongTextureDarbouxShader, seed, W, H, n_tris=60)
    nm = np.zeros_like(np.asarray(ex.normal_map))
    nm[..., 2] = 1.0
    ex = ex._replace(normal_map=nm)
    dar = run_oracle(build_call(cam, PhongTextureDarbouxShader, faces, ex, W, H,
                                init_buffers=True))
    ph_ex = PhongTextureExtraInput(position=ex.position, normal=ex.normal, uv=ex.uv,
                                   light=ex.light, texture=ex.texture)
    pho = run_oracle(build_call(cam, PhongTextureShader, faces, ph_ex, W, H, init_buffers=True))
    assert np.array_equal(dar["prim_id"], pho["prim_id"])
    assert np.array_equal(dar["zbuffer"], pho["zbuffer"])
    cov = dar["prim_id"][0] >= 0
    assert cov.sum() > 200
    assert np.isfinite(dar["canvas"]).all()
    np.testing.assert_allclose(dar["canvas"], pho["canvas"], atol=1e-6, rtol=0)
    # the darboux image is not trivially the background
    assert (np.abs(dar["canvas"][0][cov] - 0.1).max()) > 0.05


def _flat_scene(W, H, nm_vec, uv0, duv1, duv2, light_dir):
    # NDC triangle (clip = world): edges e1 = (0.9, 0.1), e2 = (0.35, 1.1)
    p0 = np.array((-0.55, -0.5, 0.0), f32)
    p1 = (p0 + np.array((0.9, 0.1, 0.0), f32)).astype(f32)
    p2 = (p0 + np.array((0.35, 1.1, 0.0), f32)).astype(f32)
    pos = np.stack([p0, p1, p2]).astype(f32)
    uv = np.stack([uv0, uv0 + duv1, uv0 + duv2]).astype(f32)
    normal = np.tile(np.array((0.0, 0.0, 1.0), f32), (3, 1))
    tex = np.array([[[0.3, 0.5, 0.7]]], f32)
    nmap = np.asarray(nm_vec, f32).reshape(1, 1, 3)
    cam = Camera.create(view=np.identity(4, dtype=f32), projection=np.identity(4, dtype=f32),
                        viewport=Camera.viewport_matrix(np.zeros(2, np.int32),
                                                        np.array((W, H)), 1.0))
    light = LightSource(direction=np.asarray(light_dir, f32),
                        colour=np.array((1.0, 0.8, 0.6), f32))
    faces = np.array([[0, 1, 2]], np.int32)
    ex = PhongTextureDarbouxExtraInput(position=pos, normal=normal, uv=uv, light=light,
                                       texture=tex, normal_map=nmap,
                                       id_to_face=np.zeros(3, np.int32), faces_indices=faces)
    return cam, faces, ex, (p1 - p0).astype(np.float64), (p2 - p0).astype(np.float64)


def _cramer(e1, e2, d1, d2):
    """The vector t in the z = 0 plane with e1.t = d1 and e2.t = d2."""
    det = e1[0] * e2[1] - e2[0] * e1[1]
    return np.array(((d1 * e2[1] - d2 * e1[1]) / det, (e1[0] * d2 - e2[0] * d1) / det, 0.0))


@pytest.mark.parametrize("nm_vec,light_dir", [
    ((1.0, 0.0, 0.0), (0.5, 0.3, 0.8)),
    ((0.0, 1.0, 0.0), (-0.2, 0.6, 0.7)),
    ((0.6, -0.3, 0.74), (0.4, -0.5, 0.75)),
    ((1.0, 0.0, 0.0), (-0.9, 0.1, 0.2)),  # facing away from the light: black
])
def test_tangent_frame_closed_form(nm_vec, light_dir):
    W, H = 64, 64
    duv1 = np.array((0.35, -0.05), f32)
    duv2 = np.array((0.1, 0.45), f32)
    cam, faces, ex, e1, e2 = _flat_scene(W, H, nm_vec, np.array((0.1, 0.2), f32), duv1, duv2,
                                         light_dir)
    out = run_oracle(build_call(cam, PhongTextureDarbouxShader, faces, ex, W, H,
                                init_buffers=True, colour_default=(0.0, 0.0, 0.0)))
    cov = out["prim_id"][0] >= 0
    assert cov.sum() > 300
    ti = _cramer(e1, e2, float(duv1[0]), float(duv2[0]))
    tj = _cramer(e1, e2, float(duv1[1]), float(duv2[1]))
    n = np.array((0.0, 0.0, 1.0))
    m = np.asarray(nm_vec, np.float64)
    nn = m[0] * ti / np.linalg.norm(ti) + m[1] * tj / np.linalg.norm(tj) + m[2] * n
    nn /= np.linalg.norm(nn)
    L = np.asarray(light_dir, np.float64)
    d = nn @ (L / np.linalg.norm(L))
    lc = np.array((1.0, 0.8, 0.6)) * d
    want = np.array((0.3, 0.5, 0.7)) * lc if (lc >= 0).all() else np.zeros(3)
    got = out["canvas"][0][cov]
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), atol=2e-6, rtol=0)
    # the plain phong colour differs (the normal map matters here)
    ph = np.array((0.3, 0.5, 0.7)) * np.array((1.0, 0.8, 0.6)) * (L[2] / np.linalg.norm(L))
    if m[2] < 0.9 and (lc >= 0).all():
        assert np.abs(want - ph).max() > 1e-2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_darboux_matches_float64_restatement(seed):
    from tests.test_gpu_backward import _leaves, _ref_params

    W, H = 48, 40
    cam, faces, ex, _ = scenes.soup_inputs(PhongTextureDarbouxShader, seed, W, H)
    ref = run_oracle(build_call(cam, PhongTextureDarbouxShader, faces, ex, W, H,
                                init_buffers=True))
    pid = ref["prim_id"][0]
    L, _, _ = _leaves(cam, ex, torch.float64, torch.device("cpu"))
    P = _ref_params(L, ex, None)
    with torch.no_grad():
        canvas, z, amb = torchref.forward(PhongTextureDarbouxShader._jr_shader_id, W, H, pid,
                                          faces, P)
    ok = (pid >= 0) & ~amb.numpy()
    assert ok.sum() > 50
    assert np.allclose(z.numpy()[ok], ref["zbuffer"][0][ok], atol=2e-5, rtol=2e-5)
    np.testing.assert_allclose(canvas.numpy()[ok], ref["canvas"][0][ok], atol=1e-4, rtol=0)
