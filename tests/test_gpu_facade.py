"""The drop-in entry points of the facade on the GPU against the CPU oracle
(renderer/renderer.py:245-476, renderer/shadow.py:40-125):

* Renderer.get_camera_image / get_camera_buffers from ModelObjects and
  CameraParameters (or a Camera), non-default colour and depth defaults, with
  and without ShadowParameters (shadow map the size of the canvas, through
  the camera's own viewport);
* Renderer.render merging into caller buffers that are not the defaults:
  pixels a primitive keeps take the new values, the others keep the old ones
  (no depth test against the old buffer), and the caller's buffers are left
  as they were;
* Shadow.render_shadow_map with the reference's default offset (0.001) and
  with its own, into a fresh map and merged into a map that holds values.

The expected buffers come from the batch planner's calls run by the oracle,
an independent route from the same inputs to the same jr_forward semantics.
"""
import numpy as np
import pytest
import torch

import jaxrenderer_amd as jr
from jaxrenderer_amd import batch, workloads
from jaxrenderer_amd.renderer import (CameraParameters, LightParameters, Renderer,
                                      ShadowParameters)
from jaxrenderer_amd.shadow import Shadow
from jaxrenderer_amd.types import Buffers
from oracle.oracle import run_oracle, run_oracle_plan
from tests.parity import as_np

pytestmark = pytest.mark.gpu
f32 = np.float32
W, H = 96, 72


def _objects(seed=3):
    rng = np.random.Generator(np.random.PCG64(seed))
    return workloads.brax_like_objects(rng)


def _camera_params():
    return CameraParameters(viewWidth=W, viewHeight=H, position=(6.0, -5.0, 4.0),
                            target=(0.0, 0.0, 0.0), up=(0.0, 0.0, 1.0), hfov=70.0, vfov=55.0)


def _expected(objs, light, cam, shadow: ShadowParameters = None, colour=(1.0, 1.0, 1.0),
              zdef=1.0):
    """Oracle render of the batch planner's calls for the same frame."""
    scene = batch.BatchScene(model=jr.merge_objects(objs), camera=cam, light=light)
    spec = None
    if shadow is not None:
        spec = batch.ShadowMapSpec(width=W, height=H,
                                   centre=tuple(float(v) for v in np.asarray(shadow.centre)),
                                   up=tuple(float(v) for v in np.asarray(shadow.up)),
                                   strength=tuple(float(v) for v in np.asarray(shadow.strength)),
                                   offset=float(shadow.offset))
    calls = batch.plan_batch([scene], W, H, spec, zbuffer_default=zdef, colour_default=colour)
    return run_oracle_plan(calls), calls


def _close(got, want, label):
    g, w = as_np(got), as_np(want)
    assert g.shape == w.shape, label
    assert np.array_equal(np.isnan(g), np.isnan(w)), label
    ok = ~np.isnan(w)
    assert np.abs(g[ok] - w[ok]).max() <= 1e-5, f"{label}: {np.abs(g[ok] - w[ok]).max()}"


@pytest.mark.parametrize("with_shadow", [False, True])
@pytest.mark.parametrize("camera_kind", ["parameters", "camera"])
def test_get_camera_image(with_shadow, camera_kind):
    objs = _objects()
    light = workloads.NOTEBOOK_LIGHT
    params = _camera_params()
    cam = Renderer.create_camera_from_parameters(params)
    shadow = ShadowParameters() if with_shadow else None
    colour, zdef = (0.2, 0.3, 0.4), 0.9
    print(f"facade: get_camera_image shadow={with_shadow} camera={camera_kind}", flush=True)
    image = Renderer.get_camera_image(objs, light, params if camera_kind == "parameters" else cam,
                                      W, H, colour_default=np.array(colour, f32),
                                      zbuffer_default=np.float32(zdef), shadow_param=shadow)
    bufs = Renderer.get_camera_buffers(objs, light, params, W, H, colour, zdef, shadow)
    torch.cuda.synchronize()
    ref, _ = _expected(objs, light, cam, shadow, colour, zdef)
    assert tuple(image.shape) == (W, H, 3)
    _close(image, ref["main"]["canvas"][0], "canvas")
    _close(bufs.targets[0], ref["main"]["canvas"][0], "buffers canvas")
    assert np.array_equal(as_np(bufs.zbuffer).view(np.int32),
                          ref["main"]["zbuffer"][0].view(np.int32))
    pid = ref["main"]["prim_id"][0]
    assert (pid >= 0).sum() > 500 and (pid < 0).sum() > 500
    # untouched pixels hold the defaults
    assert np.all(as_np(image)[pid < 0] == np.array(colour, f32))
    if with_shadow:
        plain, _ = _expected(objs, light, cam, None, colour, zdef)
        # the shadow darkens some pixels the shadowless render lights
        assert (ref["main"]["canvas"][0] < plain["main"]["canvas"][0] - 1e-3).any()


@pytest.mark.parametrize("with_shadow", [False, True])
def test_render_merges_into_caller_buffers(with_shadow):
    objs = _objects(seed=5)
    light = LightParameters()
    cam = Renderer.create_camera_from_parameters(_camera_params())
    shadow = ShadowParameters(strength=np.array((0.5, 0.4, 0.3), f32), offset=0.02) \
        if with_shadow else None
    rng = np.random.default_rng(9)
    # old depths partly in front of the scene: no depth test against them
    zb0 = rng.uniform(-1.0, 1.0, size=(W, H)).astype(f32)
    cv0 = rng.uniform(0.0, 1.0, size=(W, H, 3)).astype(f32)
    zb_dev = torch.as_tensor(zb0).cuda()
    cv_dev = torch.as_tensor(cv0).cuda()
    print(f"facade: render merge shadow={with_shadow}", flush=True)
    out = Renderer.render(jr.merge_objects(objs), light, cam, Buffers(zb_dev, (cv_dev,)), shadow)
    torch.cuda.synchronize()
    assert np.array_equal(as_np(zb_dev), zb0) and np.array_equal(as_np(cv_dev), cv0)
    fresh, _ = _expected(objs, light, cam, shadow)
    pid = fresh["main"]["prim_id"][0]
    kept = pid >= 0
    assert kept.sum() > 500 and (~kept).sum() > 500
    want_z = np.where(kept, fresh["main"]["zbuffer"][0], zb0)
    want_c = np.where(kept[..., None], fresh["main"]["canvas"][0], cv0)
    assert np.array_equal(as_np(out.zbuffer).view(np.int32), want_z.view(np.int32))
    _close(out.targets[0], want_c, "merged canvas")


@pytest.mark.parametrize("offset", [None, 0.03])
@pytest.mark.parametrize("merge", [False, True])
def test_shadow_render_shadow_map(offset, merge):
    objs = _objects(seed=7)
    model = jr.merge_objects(objs)
    light = workloads.NOTEBOOK_LIGHT
    cam = Renderer.create_camera_from_parameters(_camera_params())
    centre, up = np.zeros(3, f32), np.array((0.0, 0.0, 1.0), f32)
    strength = np.array((0.6, 0.6, 0.6), f32)
    rng = np.random.default_rng(11)
    init = (rng.uniform(0.0, 1.0, size=(W, H)).astype(f32) if merge
            else np.full((W, H), np.finfo(f32).max, f32))
    kw = {} if offset is None else {"offset": offset}
    print(f"facade: render_shadow_map offset={offset} merge={merge}", flush=True)
    sh = Shadow.render_shadow_map(shadow_map=init, verts=model.verts, faces=model.faces,
                                  light_direction=light.direction, viewport_matrix=cam.viewport,
                                  centre=centre, up=up, strength=strength, **kw)
    torch.cuda.synchronize()
    used = 0.001 if offset is None else offset  # the reference's default offset
    _, calls = _expected(objs, light, cam, ShadowParameters(centre=centre, up=up,
                                                            strength=strength, offset=used))
    call = calls[0]
    assert call.tag == "shadow"
    if merge:
        call.init_buffers = False
        ref = run_oracle(call, zbuffer=init[None])
    else:
        ref = run_oracle(call)
    assert np.array_equal(as_np(sh.shadow_map).view(np.int32), ref["zbuffer"][0].view(np.int32))
    # the light camera is the one the batch planner derives for the pass
    assert np.array_equal(np.asarray(sh.camera.world_to_clip, f32),
                          np.ctypeslib.as_array(call.scenes[0].world_to_clip).reshape(4, 4))
    covered = ref["prim_id"][0] >= 0
    assert covered.sum() > 200
    # every texel carries the offset, covered or not
    if not merge:
        assert np.all(as_np(sh.shadow_map)[~covered] == np.float32(np.finfo(f32).max) + f32(used))
