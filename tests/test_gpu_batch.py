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


def test_batch_backward_without_face_table_matches_explicit_table():
    """The batch's main pass passes no face table (identity slots); its
    gradients must equal those of the same step with the explicit table (up
    to the order of the atomic sums)."""
    br = _renderer(n=16, seed=8)
    br.enable_backward()
    out = br.render()
    target = torch.rand_like(out["canvas"])
    torch.cuda.synchronize()
    main = br.passes["main"]
    assert main.inputs["faces"] is None
    got = {k: v.clone() for k, v in br.train_step(target).items()}
    torch.cuda.synchronize()
    n = main.call.n_tris
    main.inputs["faces"] = torch.arange(3 * n, dtype=torch.int32, device=target.device).view(n, 3)
    want = br.train_step(target)
    torch.cuda.synchronize()
    for k in want:
        a, b = got[k].double(), want[k].double()
        assert torch.isfinite(b).all(), k
        scale = max(float(b.norm()), 1e-6)
        assert float((a - b).norm()) <= 1e-5 * scale, k


def test_fused_l2_backward_matches_explicit_cotangent():
    """jr_backward_l2 (the canvas cotangent 2 (canvas - target) formed inside
    the kernel) against jr_backward given that cotangent as a tensor; equal up
    to the order of the atomic sums."""
    br = _renderer(n=16, seed=9)
    br.enable_backward()
    out = br.render()
    target = torch.rand_like(out["canvas"])
    torch.cuda.synchronize()
    got = {k: v.clone() for k, v in br.backward_l2(target).items()}
    torch.cuda.synchronize()
    want = br.backward(2.0 * (out["canvas"] - target))
    torch.cuda.synchronize()
    for k in want:
        a, b = got[k].double(), want[k].double()
        assert torch.isfinite(b).all(), k
        scale = max(float(b.norm()), 1e-6)
        assert float((a - b).norm()) <= 1e-5 * scale, k
    assert float(want["d_scene"].norm()) > 0
