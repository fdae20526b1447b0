"""plan_batch_device (jr_pack on the GPU plus batched records) against the
per-scene host planner: every input array and scene record equal bit for bit,
renders identical, and a face index outside its scene's vertices refused."""
import numpy as np
import pytest
import torch

from jaxrenderer_amd import batch, packing, workloads
from jaxrenderer_amd.model import MergedModel

pytestmark = pytest.mark.gpu


def _host(x):
    return x.cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


@pytest.mark.parametrize("workload", ["config3", "ant", "config1"])
def test_device_plan_equals_host_plan(workload):
    if workload == "config3":
        scs, sm = workloads.config3(n_scenes=50, seed=21)
        W = H = 84
    elif workload == "ant":
        scs, sm = workloads.ant_like(n_scenes=8, seed=5)
        W = H = 84
    else:
        scs, sm, W, H = [workloads.config1()], None, 320, 240
    print(f"packing: {workload}", flush=True)
    host = batch.plan_batch(scs, W, H, sm)
    dev = packing.plan_batch_device(scs, W, H, sm)
    torch.cuda.synchronize()
    assert [c.tag for c in host] == [c.tag for c in dev]
    for h, d in zip(host, dev):
        assert (h.shader, h.width, h.height, h.n_slots, h.n_tris) == \
            (d.shader, d.width, d.height, d.n_slots, d.n_tris)
        assert b"".join(bytes(s) for s in h.scenes) == b"".join(bytes(s) for s in d.scenes)
        assert set(h.inputs) == set(d.inputs), (h.tag, set(h.inputs) ^ set(d.inputs))
        for k in h.inputs:
            a, b = _host(h.inputs[k]), _host(d.inputs[k])
            assert a.shape == b.shape and a.dtype == b.dtype, (h.tag, k)
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (h.tag, k)
        assert h.extras == d.extras
    a = batch.BatchRenderer(host, tile=(16, 16), persistent=False).render()
    b = batch.BatchRenderer(dev, tile=(16, 16), persistent=False).render()
    torch.cuda.synchronize()
    assert torch.equal(a["zbuffer"].view(torch.int32), b["zbuffer"].view(torch.int32))
    assert torch.equal(a["canvas"].view(torch.int32), b["canvas"].view(torch.int32))


def test_device_plan_refuses_out_of_range_faces():
    scs, sm = workloads.config3(n_scenes=4, seed=22)
    m = scs[2].model
    faces = np.array(m.faces, copy=True)
    faces[5, 1] = len(m.verts)  # one past the scene's vertices
    scs[2] = scs[2]._replace(model=MergedModel(*([m.verts, m.norms, m.uvs, faces] + list(m[4:]))))
    with pytest.raises(ValueError):
        packing.plan_batch_device(scs, 84, 84, sm)
    with pytest.raises(ValueError):
        batch.plan_batch(scs, 84, 84, sm)
