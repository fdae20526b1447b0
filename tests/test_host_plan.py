"""Host-side planning (CPU): the per-corner flattened calls mark their face
table as the identity, which the device inputs then leave out (the library
takes slots 3t, 3t+1, 3t+2 for row t without reading a table)."""
import numpy as np
import torch

from jaxrenderer_amd import _plan, batch, workloads


def test_flattened_main_call_has_identity_faces():
    scs, sm = workloads.config3(n_scenes=3, seed=2)
    calls = batch.plan_batch(scs, 84, 84, sm)
    main = [c for c in calls if c.tag == "main"][0]
    shadow = [c for c in calls if c.tag == "shadow"][0]
    assert main.extras.get("identity_faces")
    f = np.asarray(main.inputs["faces"]).reshape(-1)
    assert np.array_equal(f, np.arange(3 * main.n_tris, dtype=np.int32))
    assert main.n_slots >= 3 * main.n_tris
    assert not shadow.extras.get("identity_faces")
    dev = _plan.device_inputs(main, torch.device("cpu"))
    assert dev["faces"] is None and dev["positions"] is not None
    assert _plan.device_inputs(shadow, torch.device("cpu"))["faces"] is not None


def test_batched_scene_records_equal_the_per_scene_planner():
    """packing.scene_records (all scenes at once in numpy) gives the bytes
    plan_batch's per-scene code puts in its calls, for the main and the
    shadow pass, with batch-wide and per-scene shadow centres."""
    from jaxrenderer_amd import batch, packing, workloads

    for scs, sm, W, H in ((*workloads.config3(n_scenes=40, seed=3), 84, 84),
                          (*workloads.ant_like(n_scenes=6, seed=2), 84, 84),
                          ([workloads.config1()], None, 320, 240)):
        host = batch.plan_batch(scs, W, H, sm)
        rec, srec = packing.scene_records(scs, W, H, sm)
        assert b"".join(bytes(s) for s in host[-1].scenes) == rec.tobytes()
        if sm is not None:
            assert b"".join(bytes(s) for s in host[0].scenes) == srec.tobytes()
        else:
            assert srec is None
