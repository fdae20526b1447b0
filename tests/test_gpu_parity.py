"""GPU (libjrhip.so) against the CPU oracle on seeded inputs.

Primitive ids, depths and barycentrics must match bit for bit; colours within
1e-5 absolute."""
import numpy as np
import pytest
import torch

from jaxrenderer_amd import _plan
from jaxrenderer_amd.shaders import BUILTIN, DepthShader, PhongReflectionShadowTextureShader
from oracle.oracle import run_oracle_plan
from tests import scenes
from tests.parity import assert_parity

pytestmark = pytest.mark.gpu

TILES = [(8, 8), (16, 16), (8, 32), (4, 16)]


def _check(make_calls, label, **kw):
    ref = run_oracle_plan(make_calls(), display=True)
    got = _plan.run_plan(make_calls(), prim_id=True, bary=True, display=True, **kw)
    torch.cuda.synchronize()
    for tag in ref:
        assert_parity(got[tag], ref[tag], label=f"{label}/{tag}")
    return got, ref


def test_library_loads_on_gpu():
    from jaxrenderer_amd import _hip
    assert _hip.lib().jr_version() == _hip.ABI_VERSION == 2
    assert torch.cuda.is_available()


def test_config1_cube_phong_reflection():
    got, ref = _check(scenes.c1_plan, "c1")
    pid = got["main"]["prim_id"].cpu().numpy()
    assert (pid >= 0).sum() > 10000


@pytest.mark.parametrize("shader", BUILTIN, ids=lambda s: s.__name__)
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_soup_parity(shader, seed):
    _check(lambda: scenes.soup_call(shader, seed), f"{shader.__name__}/{seed}")


@pytest.mark.parametrize("shader", ["PhongReflectionTextureShader",
                                    "PhongReflectionShadowTextureShader"])
def test_specular_exponents_fast_and_exact_power(shader):
    """The specular power max(cos, 0) ^ s takes the hardware log2 / exp2 path
    where the result is at most 1.42 and powf elsewhere: exponents 0 (powf:
    exactly 1), 1e-3, 0.5, 1, 4, 64 and 200 on one soup keep the colours
    within 1e-5 of the oracle's powf."""
    from jaxrenderer_amd import shaders

    cls = getattr(shaders, shader)
    vals = np.array([0.0, 1e-3, 0.5, 1.0, 4.0, 64.0, 200.0], np.float32)

    def calls():
        out = scenes.soup_call(cls, 5, 64, 48, 150)
        for c in out:
            sm = c.inputs.get("specular_maps")
            if sm is not None:
                c.inputs["specular_maps"] = vals[np.arange(sm.size) % vals.size].reshape(sm.shape)
        return out

    got, _ = _check(calls, f"spec/{shader}")
    assert (got["main"]["prim_id"].cpu().numpy() >= 0).sum() > 500


@pytest.mark.parametrize("tile", [(8, 8), (16, 16)], ids=str)
@pytest.mark.parametrize("seed", [0, 1])
def test_unbounded_and_edge_on_triangles(tile, seed, monkeypatch):
    """Triangles whose pixel region has no bounded range: a tenth of the soup
    reaches behind the eye, and a fifth lies in planes through the eye
    (edge-on, projecting to a line).  Their range is the whole viewport only
    where the exact rectangle test cannot reject it; ids and depth stay
    bit-exact against the oracle, colours within 1e-5."""
    soup = scenes._soup
    eye = np.array((0.3, 0.2, 3.0), np.float32)

    def wild(rng, n_tris, behind=0.1, degenerate=0.05):
        pos = soup(rng, n_tris, behind=behind, degenerate=degenerate).reshape(n_tris, 3, 3)
        # in front of the eye (the camera looks down -z from z = 3), in a
        # plane through it
        k = rng.random(n_tris) < 0.2
        d = rng.normal(size=(2, n_tris, 3)).astype(np.float32)
        d[..., 2] = -np.abs(d[..., 2]) - 0.5
        ab = rng.uniform(0.5, 3.0, size=(n_tris, 3, 2)).astype(np.float32)
        plane = eye + ab[..., :1] * d[0][:, None, :] + ab[..., 1:] * d[1][:, None, :]
        pos[k] = plane[k]
        return pos.reshape(-1, 3).astype(np.float32)

    monkeypatch.setattr(scenes, "_soup", wild)
    _check(lambda: scenes.soup_call(PhongReflectionShadowTextureShader, 40 + seed, 72, 56, 300),
           f"unbounded/{seed}/{tile}", tile=tile)


@pytest.mark.parametrize("tile", TILES, ids=str)
@pytest.mark.parametrize("persistent", [False, True])
def test_tiles_and_persistent(tile, persistent):
    _check(lambda: scenes.soup_call(PhongReflectionShadowTextureShader, 7, 70, 50, 120),
           f"tile{tile}/p{persistent}", tile=tile, persistent=persistent)


@pytest.mark.parametrize("tile", [(16, 16), (8, 32)], ids=str)
@pytest.mark.parametrize("form", [0, 1, 3])
@pytest.mark.parametrize("shader", [DepthShader, PhongReflectionShadowTextureShader],
                         ids=lambda s: s.__name__)
def test_four_pixels_per_lane(tile, form, shader):
    """One wavefront per 256-pixel tile, each lane owning 4 pixels, in the
    persistent (0) and 1 / 3 tiles-per-group forms, on a frame whose size is
    not a multiple of the tile."""
    _check(lambda: scenes.soup_call(shader, 8, 70, 50, 120), f"ppl4/{tile}/{form}", tile=tile,
           persistent=form == 0, tiles_per_group=max(form, 1), pixels_per_lane=4)


@pytest.mark.parametrize("capacity", [0, 1, 17])
def test_bin_overflow_falls_back(capacity):
    _check(lambda: scenes.soup_call(DepthShader, 3, 50, 40, 80), f"cap{capacity}",
           capacity=capacity)


def test_merge_into_existing_buffers():
    calls = lambda: scenes.soup_call(DepthShader, 5, init=False)  # noqa: E731
    _check(calls, "merge")


def test_reference_smoke_scene_small():
    _check(lambda: [scenes.smoke_gouraud_call(192, 108)], "smoke192")


def test_dense_sphere_small_triangles():
    """Config-2 sphere (9.2k triangles, many smaller than a pixel) at 240x135:
    exercises the per-pixel atomic path of small triangles."""
    from jaxrenderer_amd import batch, workloads

    sc = workloads.config2(seed=1, width=240, height=135)
    _check(lambda: batch.plan_batch([sc], 240, 135), "sphere240")


def test_config3_batch_against_oracle():
    """Three uneven Brax-like scenes in one batched call, 256x256 shadow maps."""
    from jaxrenderer_amd import batch, workloads

    scs, sm = workloads.config3(n_scenes=3, seed=11)
    _check(lambda: batch.plan_batch(scs, 84, 84, sm), "c3x3")


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
@pytest.mark.parametrize("tile", [(8, 8), (16, 16)], ids=str)
def test_tiny_and_pixel_aligned_triangles(seed, tile):
    """Sub-pixel triangles, slivers and vertices on exact sample positions:
    the pixel-range and tile culling must keep every pixel the brute-force
    oracle covers."""
    from jaxrenderer_amd.shaders import GouraudShader

    shader = DepthShader if seed % 2 == 0 else GouraudShader
    _check(lambda: scenes.tiny_soup_call(seed, shader=shader), f"tiny{seed}", tile=tile)


@pytest.mark.parametrize("tiles_per_group,ppl", [(1, 1), (3, 1), (4, 1), (2, 4), (4, 4), (6, 4)],
                         ids=str)
def test_many_scenes_xcd_mapping_and_tiles_per_group(tiles_per_group, ppl):
    """70 scenes (>= 64: the raster keeps each scene on one XCD) at 40x30 with
    16x16 tiles (partial edge tiles) and several tiles per workgroup (with four
    pixels per lane and at most 4 tiles, one wavefront per tile side by side):
    every (scene, tile) is rendered exactly once and matches the oracle."""
    from jaxrenderer_amd import batch, workloads

    scs, sm = workloads.config3(n_scenes=70, seed=12, width=40, height=30, shadow_size=24)
    _check(lambda: batch.plan_batch(scs, 40, 30, sm), f"xcd{tiles_per_group}x{ppl}",
           tile=(16, 16), tiles_per_group=tiles_per_group, pixels_per_lane=ppl)


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("tile,ppl", [((8, 8), 1), ((16, 16), 4)], ids=str)
def test_depth_only_offset_pass(seed, tile, ppl):
    """A depth-shader pass with an offset and no id outputs (a shadow map):
    small triangles keep 32-bit depth keys, tagged per binning, in the
    workspace.  The depth must still match the oracle bit for bit, on tiny
    and pixel-aligned triangles and on config-3 shadow maps."""
    from jaxrenderer_amd import batch, workloads

    def calls():
        c = scenes.tiny_soup_call(seed)[0]
        c.apply_offset, c.depth_offset, c.zbuffer_default = True, 0.05, 3.0e38
        return [c]

    for make, label in ((calls, f"tinyshadow{seed}"),):
        ref = run_oracle_plan(make(), prim_id=False, bary=False)
        got = _plan.run_plan(make(), tile=tile, pixels_per_lane=ppl)
        torch.cuda.synchronize()
        for tag in ref:
            assert_parity(got[tag], ref[tag], label=f"{label}/{tag}")
    scs, sm = workloads.config3(n_scenes=3, seed=20 + seed)
    shadow = [c for c in batch.plan_batch(scs, 84, 84, sm) if c.tag == "shadow"]
    ref = run_oracle_plan(shadow, prim_id=False, bary=False)
    got = _plan.run_plan(shadow, tile=tile, pixels_per_lane=ppl)
    torch.cuda.synchronize()
    assert_parity(got["shadow"], ref["shadow"], label=f"c3shadow{seed}")


def test_depth_pass_keys_carry_over_calls():
    """The narrow form never clears its small-triangle keys: they stay in the
    workspace, each binning's tagged apart from the older ones.  On one
    workspace, shadow passes of scenes A, then B (same map size, other
    triangles), a whole config-3 frame of A (its main pass bins in the wide
    form on the same workspace), A, B and A again each match the oracle bit
    for bit."""
    from jaxrenderer_amd import batch, workloads

    def calls(seed, full):
        scs, sm = workloads.config3(n_scenes=3, seed=seed)
        out = batch.plan_batch(scs, 84, 84, sm)
        return out if full else [c for c in out if c.tag == "shadow"]

    refs = {}
    for step, (seed, full) in enumerate(((41, False), (42, False), (41, True), (41, False),
                                         (42, False), (41, False))):
        label = f"step{step}/seed{seed}{'/frame' if full else ''}"
        print(label, flush=True)
        if (seed, full) not in refs:
            refs[(seed, full)] = run_oracle_plan(calls(seed, full), prim_id=False, bary=False)
        ref = refs[(seed, full)]
        got = _plan.run_plan(calls(seed, full), tile=(16, 16), pixels_per_lane=4)
        torch.cuda.synchronize()
        for tag in ref:
            assert_parity(got[tag], ref[tag], label=f"{label}/{tag}")


def test_depth_pass_on_a_garbage_workspace():
    """A workspace whose bytes are random (memory that held something else)
    does not describe its tagged keys: the depth pass refills them, so two
    calls on it, the second after the keys were overwritten with other random
    bytes again, both match the oracle bit for bit."""
    from jaxrenderer_amd import _raster, batch, workloads

    scs, sm = workloads.config3(n_scenes=2, seed=51)
    call = [c for c in batch.plan_batch(scs, 84, 84, sm) if c.tag == "shadow"][0]
    ref = run_oracle_plan([call], prim_id=False, bary=False)[call.tag]
    dev = torch.device("cuda", 0)
    B, W, H = call.n_scenes, call.width, call.height
    need = _raster.workspace_size(width=W, height=H, n_scenes=B, n_tris=call.n_tris,
                                  capacity=_raster.default_capacity(call.n_tris,
                                                                    _raster.n_tiles(W, H, B)))
    gen = torch.Generator(device=dev).manual_seed(5)
    ws = torch.randint(0, 256, (need,), dtype=torch.uint8, device=dev, generator=gen)
    inputs = _plan.device_inputs(call, dev)
    scenes_dev = _raster.scenes_to_device(call.scenes, dev)
    for k in range(2):
        if k == 1:
            ws.copy_(torch.randint(0, 256, (need,), dtype=torch.uint8, device=dev,
                                   generator=gen))
        zbuffer = torch.empty((B, W, H), dtype=torch.float32, device=dev)
        _raster.forward(shader=call.shader, width=W, height=H, scenes=scenes_dev, n_scenes=B,
                        inputs=inputs, n_slots=call.n_slots, n_tris=call.n_tris,
                        zbuffer=zbuffer, canvas=None, prim_id=None, bary=None, display=None,
                        shadow_shape=call.shadow_shape, init_buffers=call.init_buffers,
                        zbuffer_default=call.zbuffer_default,
                        colour_default=call.colour_default, apply_offset=call.apply_offset,
                        depth_offset=call.depth_offset, workspace=ws)
        torch.cuda.synchronize()
        assert_parity({"zbuffer": zbuffer}, ref, label=f"garbage workspace, call {k}")


@pytest.mark.parametrize("shader", [DepthShader, PhongReflectionShadowTextureShader],
                         ids=lambda s: s.__name__)
def test_ragged_and_empty_scenes(shader):
    """One call whose scenes own 0, 1, 3, 63, 64, 65, ... triangles, empty
    scenes among them (first and last included): each triangle belongs to
    the last scene whose range begins at or before it, wavefronts of 64
    triangles span several scenes, and an empty scene draws nothing."""
    sizes = [0, 3, 64, 0, 0, 130, 1, 63, 65, 0, 200, 171, 0]

    def calls():
        out = scenes.soup_call(shader, 9, 40, 30, sum(sizes))
        for c in out:
            base = c.scenes[0]
            c.scenes = []
            begin = 0
            for n in sizes:
                sc = type(base).from_buffer_copy(base)
                sc.tri_begin, sc.tri_end = begin, begin + n
                begin += n
                c.scenes.append(sc)
        return out

    got, _ = _check(calls, f"ragged/{shader.__name__}", tile=(8, 8))
    pid = got["main"]["prim_id"].cpu().numpy()
    for s, n in enumerate(sizes):
        if n == 0:
            assert (pid[s] == -1).all()
