"""Shadow depths on demand (jr_shadow_source): the shadow pass runs only its
binning stages and the phong_reflection_shadow raster computes each texel it
reads from that binning.  The main pass's depth and colour must equal the
render with a full shadow map bit for bit (and the oracle's), whatever the
tile shapes, launch forms or bin capacities of either pass; every texel the
main pass computed must equal the full map's; HIP-graph replays must equal
eager renders; the backward, reading the computed texels, must give the
gradients of the full-map render."""
import numpy as np
import pytest
import torch

from jaxrenderer_amd import batch, workloads
from oracle.oracle import run_oracle_plan
from tests.parity import as_np, assert_parity

pytestmark = pytest.mark.gpu


def _calls(n=24, seed=31, W=84, H=84, shadow=256):
    scs, sm = workloads.config3(n_scenes=n, seed=seed, width=W, height=H, shadow_size=shadow)
    return batch.plan_batch(scs, W, H, sm)


def _renderer(calls, on_demand, tile=(16, 16), shadow_tile=(16, 16), ppl=4, form=1,
              keep=True):
    br = batch.BatchRenderer(calls, tile=tile, shadow_tile=shadow_tile, persistent=form == 0,
                             shadow_on_demand=on_demand, keep_shadow_texels=keep)
    for p in br.passes.values():
        p.pixels_per_lane = ppl if p.tile[0] * p.tile[1] == 256 else 1
        p.tiles_per_group = max(1, form)
    br.calibrate()
    return br


def _bits(t):
    return t.contiguous().view(torch.int32)


@pytest.mark.parametrize("tile,shadow_tile,ppl,form", [
    ((16, 16), (16, 16), 4, 1), ((8, 8), (8, 8), 1, 1), ((8, 32), (16, 16), 4, 2),
    ((16, 16), (8, 32), 4, 0), ((4, 16), (4, 16), 1, 4)], ids=str)
def test_on_demand_equals_full_shadow_map(tile, shadow_tile, ppl, form):
    calls = _calls()
    print(f"on-demand: tiles {tile}/{shadow_tile} ppl {ppl} form {form}", flush=True)
    full = _renderer(calls, False, tile, shadow_tile, ppl, form)
    want = {k: v.clone() for k, v in full.render().items()}
    od = _renderer(calls, True, tile, shadow_tile, ppl, form)
    got = od.render()
    torch.cuda.synchronize()
    assert torch.equal(_bits(got["zbuffer"]), _bits(want["zbuffer"]))
    assert torch.equal(_bits(got["canvas"]), _bits(want["canvas"]))
    tex = got["shadow_texels"]
    written = tex != 0
    assert int(written.sum()) > 1000
    assert torch.equal(_bits(tex[written]), _bits(want["shadow_map"][written]))


@pytest.mark.parametrize("capacity", [0, 1, 97])
def test_on_demand_with_overflowing_shadow_bins(capacity):
    """Shadow tile lists that overflow the bin capacity: the lookup scans the
    scene's triangles that list the tile, like the shadow raster."""
    calls = _calls(n=6, seed=32)
    full = _renderer(calls, False)
    want = {k: v.clone() for k, v in full.render().items()}
    od = _renderer(calls, True)
    od.passes["shadow"].set_capacity(capacity)
    got = od.render()
    torch.cuda.synchronize()
    assert torch.equal(_bits(got["zbuffer"]), _bits(want["zbuffer"]))
    assert torch.equal(_bits(got["canvas"]), _bits(want["canvas"]))


def test_on_demand_against_oracle():
    calls = _calls(n=3, seed=33)
    ref = run_oracle_plan(calls)
    od = _renderer(calls, True)
    od.passes["main"].prim_id = torch.empty((3, 84, 84), dtype=torch.int32, device="cuda")
    od.passes["main"].bary = torch.empty((3, 84, 84, 3), dtype=torch.float32, device="cuda")
    got = od.render()
    torch.cuda.synchronize()
    assert_parity({"zbuffer": got["zbuffer"], "canvas": got["canvas"],
                   "prim_id": got["prim_id"], "bary": got["bary"]}, ref["main"], label="od")
    tex = as_np(got["shadow_texels"])
    written = tex != 0
    assert np.array_equal(tex[written].view(np.int32),
                          ref["shadow"]["zbuffer"][written].view(np.int32))


def test_on_demand_graph_replays_equal_eager():
    calls = _calls(n=20, seed=34)
    od = _renderer(calls, True)
    want = {k: v.clone() for k, v in od.render().items()}
    od.capture()
    for _ in range(3):
        od.passes["main"].zbuffer.fill_(5.0)
        od.passes["main"].canvas.fill_(-1.0)
        od.shadow_texels.fill_(0.0)
        torch.cuda.synchronize()
        out = od.replay()
        torch.cuda.synchronize()
        assert torch.equal(_bits(out["zbuffer"]), _bits(want["zbuffer"]))
        assert torch.equal(_bits(out["canvas"]), _bits(want["canvas"]))


def test_on_demand_backward_equals_full_map_backward():
    calls = _calls(n=10, seed=35)
    full = _renderer(calls, False)
    full.enable_backward()
    target = torch.rand_like(full.render()["canvas"])
    want = {k: v.clone() for k, v in full.train_step(target).items()}
    od = _renderer(calls, True)
    od.enable_backward()
    got = od.train_step(target)
    torch.cuda.synchronize()
    for k in want:
        a, b = got[k].double(), want[k].double()
        assert torch.isfinite(a).all(), k
        scale = max(float(b.norm()), 1e-6)
        assert float((a - b).norm()) <= 1e-5 * scale, k


def test_on_demand_without_kept_texels():
    """A forward-only renderer (texels not kept) renders the same image and
    writes no texel."""
    calls = _calls(n=8, seed=36)
    kept = _renderer(calls, True)
    want = {k: v.clone() for k, v in kept.render().items()}
    od = _renderer(calls, True, keep=False)
    got = od.render()
    torch.cuda.synchronize()
    assert "shadow_texels" not in got
    assert torch.equal(_bits(got["zbuffer"]), _bits(want["zbuffer"]))
    assert torch.equal(_bits(got["canvas"]), _bits(want["canvas"]))
    assert int((od.shadow_texels != 0).sum()) == 0
    od.enable_backward()  # the backward needs them: kept from here on
    assert "shadow_texels" in od.render()
