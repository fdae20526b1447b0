"""Batched rendering through HIP-graph replays (the timed path): the graph
runs the main pass's binning on a second stream beside the shadow pass and
the main raster after the shadow map is complete.  Every replay must equal
the eager pass-after-pass render bit for bit (depth, colour, shadow map),
and a pass split into its binning and raster parts must equal the whole
pass."""
import pytest
import torch

from jaxrenderer_amd import _hip, batch, workloads

pytestmark = pytest.mark.gpu


def _renderer(n=40, seed=4):
    scs, sm = workloads.config3(n_scenes=n, seed=seed)
    calls = batch.plan_batch(scs, 84, 84, sm)
    br = batch.BatchRenderer(calls, tile=(16, 16), shadow_tile=(16, 16), persistent=False)
    br.passes["shadow"].pixels_per_lane = 4
    br.calibrate()
    return br


def _snapshot(br):
    return {tag: (p.zbuffer.clone(), None if p.canvas is None else p.canvas.clone())
            for tag, p in br.passes.items()}


def _same(a, b):
    for tag in a:
        assert torch.equal(a[tag][0].view(torch.int32), b[tag][0].view(torch.int32)), tag
        if a[tag][1] is not None:
            assert torch.equal(a[tag][1].view(torch.int32), b[tag][1].view(torch.int32)), tag


def test_graph_replays_equal_eager_render():
    br = _renderer()
    br.render()
    torch.cuda.synchronize()
    want = _snapshot(br)
    br.capture()
    for k in range(4):
        # eager work between replays overwrites the outputs
        for p in br.passes.values():
            p.zbuffer.fill_(7.0)
            if p.canvas is not None:
                p.canvas.fill_(-1.0)
        torch.cuda.synchronize()
        br.replay()
        torch.cuda.synchronize()
        _same(want, _snapshot(br))


def test_bin_and_raster_parts_equal_whole_pass():
    br = _renderer(n=12, seed=6)
    br.render()
    torch.cuda.synchronize()
    want = _snapshot(br)
    for p in br.passes.values():
        p.zbuffer.fill_(3.0)
    for tag in ("shadow", "main"):
        br.passes[tag].launch(_hip.PART_BIN)
        br.passes[tag].launch(_hip.PART_RASTER)
    torch.cuda.synchronize()
    _same(want, _snapshot(br))
