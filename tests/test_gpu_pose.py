"""jr_pose (ScenePoser) against the host merge_objects: world positions and
normals bit for bit, and renders of a moved batch equal to renders of the
batch planned from scratch with the moved objects."""
import numpy as np
import pytest
import torch

from jaxrenderer_amd import batch, workloads
from jaxrenderer_amd.geometry import transform_matrix_from_rotation
from jaxrenderer_amd.model import Model, ModelObject, merge_objects
from jaxrenderer_amd.pose import ScenePoser

pytestmark = pytest.mark.gpu

f32 = np.float32


def bits(a):
    return np.ascontiguousarray(a, f32).view(np.uint32)


def poses(n, rng):
    """Rigid poses, plus every 5th with a non-uniform scale and shear and
    every 7th with a projective last row (w != 1 for to_cartesian)."""
    out = np.zeros((n, 4, 4), f32)
    for k in range(n):
        t = np.identity(4, dtype=f32)
        q = rng.standard_normal(4).astype(f32)
        t[:3, :3] = transform_matrix_from_rotation(q)
        t[:3, 3] = rng.uniform(-3.0, 3.0, 3).astype(f32)
        if k % 5 == 1:
            t[:3, :3] = t[:3, :3] @ np.array([[1.3, 0.2, 0.0], [0.0, 0.7, 0.1], [0.0, 0.0, 1.1]],
                                             f32)
        if k % 7 == 3:
            t[3, :3] = rng.uniform(-0.05, 0.05, 3).astype(f32)
        out[k] = t
    return out


def moved(objects, T):
    k = 0
    out = []
    for scene in objects:
        row = []
        for o in scene:
            row.append(o._replace(transform=T[k]))
            k += 1
        out.append(row)
    return out


def scene_objects(big=True):
    objects = []
    workloads.config3(n_scenes=5, seed=31, objects=objects)
    rng = np.random.default_rng(3)
    tex = rng.random((8, 8, 3)).astype(f32)
    spec = np.full((8, 8), 2.0, f32)
    # a dense sphere (a deep pairwise tree: ~14k normal components), a scaled
    # box, and a one-triangle object with a single normal
    if big:
        dense = workloads.mesh_model(*workloads.uv_sphere(48, 96, 0.8), tex, spec)
        objects[0].append(ModelObject(model=dense))
    objects[1].append(ModelObject(model=objects[1][1].model,
                                  local_scaling=np.array([1.5, 0.5, 2.0], f32)))
    one = Model(verts=np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], f32),
                norms=np.array([[0.0, 0.0, 2.0]], f32), uvs=np.zeros((3, 2), f32),
                faces=np.array([[0, 1, 2]], np.int32), faces_norm=np.zeros((1, 3), np.int32),
                faces_uv=np.array([[0, 1, 2]], np.int32), diffuse_map=tex, specular_map=spec)
    objects[2].append(ModelObject(model=one))
    return objects


@pytest.mark.parametrize("big", [False, True], ids=["staged", "recomputed"])
def test_pose_matches_merge_objects(big):
    """Objects of at most 2730 normals keep their transformed normals in LDS;
    with a larger one in the batch every object recomputes them."""
    objects = scene_objects(big)
    poser = ScenePoser(objects)
    assert (poser.max_object_norms > 4000) == big
    dev = poser.device
    for seed in (0, 1):
        T = poses(poser.n_objects, np.random.default_rng(seed))
        pos = torch.empty((poser.n_verts, 3), dtype=torch.float32, device=dev)
        nrm = torch.empty((poser.n_norms, 3), dtype=torch.float32, device=dev)
        poser.pose(T, pos, nrm)
        torch.cuda.synchronize()
        ref = [merge_objects(s) for s in moved(objects, T)]
        rv = np.concatenate([np.asarray(m.verts, f32) for m in ref])
        rn = np.concatenate([np.asarray(m.norms, f32) for m in ref])
        gv, gn = pos.cpu().numpy(), nrm.cpu().numpy()
        assert np.array_equal(bits(gv), bits(rv)), np.abs(gv - rv).max()
        assert np.array_equal(bits(gn), bits(rn)), np.abs(gn - rn).max()


def test_positions_only_and_argument_checks():
    objects = scene_objects()[:2]
    poser = ScenePoser(objects)
    T = poses(poser.n_objects, np.random.default_rng(7))
    pos = torch.empty((poser.n_verts, 3), dtype=torch.float32, device=poser.device)
    poser.pose(torch.from_numpy(T).to(poser.device), pos)
    torch.cuda.synchronize()
    rv = np.concatenate([np.asarray(merge_objects(s).verts, f32) for s in moved(objects, T)])
    assert np.array_equal(bits(pos.cpu().numpy()), bits(rv))
    with pytest.raises(ValueError):
        poser.pose(T, pos[:-1])
    with pytest.raises(ValueError):
        poser.pose(T, pos.double())


def test_pose_update_renders_like_a_replanned_batch():
    """A BatchRenderer moved by ScenePoser.update renders what a renderer
    planned from merge_objects of the moved objects renders, eagerly and in
    replays of a graph captured before the move."""
    W, H = 40, 30
    objects = []
    scs, sm = workloads.config3(n_scenes=8, seed=32, width=W, height=H, shadow_size=32,
                                objects=objects)
    br = batch.BatchRenderer(batch.plan_batch(scs, W, H, sm), tile=(8, 8), persistent=False)
    br.calibrate()
    br.capture()
    poser = ScenePoser(objects)
    for seed in (11, 12):
        T = poses(poser.n_objects, np.random.default_rng(seed))
        for o in range(poser.n_objects):
            if o in poser.object_begin:  # keep each scene's ground plane
                T[o] = poser.initial_transforms[o]
        poser.update(br, T)
        scs2 = [batch.BatchScene(model=merge_objects(m), camera=s.camera, light=s.light)
                for m, s in zip(moved(objects, T), scs)]
        ref = batch.BatchRenderer(batch.plan_batch(scs2, W, H, sm), tile=(8, 8),
                                  persistent=False)
        ref.calibrate()
        want = {k: v.clone() for k, v in ref.render().items()}
        for run in (br.replay, br.render):
            run()
            torch.cuda.synchronize()
            got = br._outputs()
            for k in ("zbuffer", "canvas", "shadow_map"):
                assert torch.equal(got[k], want[k]), (seed, k)


def test_limits_are_checked_before_any_launch():
    """An object above the normal limit is refused by ScenePoser, and a
    max_object_norms outside [0, JR_POSE_MAX_NORMALS] by jr_pose itself."""
    import ctypes

    from jaxrenderer_amd import _hip

    n = _hip.POSE_MAX_NORMALS + 1
    huge = Model(verts=np.zeros((3, 3), f32), norms=np.ones((n, 3), f32),
                 uvs=np.zeros((3, 2), f32), faces=np.array([[0, 1, 2]], np.int32),
                 faces_norm=np.zeros((1, 3), np.int32), faces_uv=np.array([[0, 1, 2]], np.int32),
                 diffuse_map=np.zeros((2, 2, 3), f32), specular_map=np.zeros((2, 2), f32))
    with pytest.raises(ValueError):
        ScenePoser([[ModelObject(model=huge)]])

    poser = ScenePoser(scene_objects(False)[:1])
    dev = poser.device
    T = torch.from_numpy(poser.initial_transforms.reshape(-1, 16).copy()).to(dev)
    pos = torch.empty((poser.n_verts, 3), dtype=torch.float32, device=dev)
    nrm = torch.empty((poser.n_norms, 3), dtype=torch.float32, device=dev)
    a = _hip.JrPoseArgs()
    a.n_objects, a.n_verts, a.n_norms = poser.n_objects, poser.n_verts, poser.n_norms
    a.max_object_norms = _hip.POSE_MAX_NORMALS + 1
    for name in ("vert_begin", "norm_begin", "vert_object", "local_verts", "unit_norms",
                 "scaling", "normal_matrices"):
        setattr(a, name, getattr(poser, name).data_ptr())
    a.transforms, a.positions, a.normals = T.data_ptr(), pos.data_ptr(), nrm.data_ptr()
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    lib = _hip.lib()
    assert lib.jr_pose(ctypes.byref(a), stream) == -1
    assert b"max_object_norms" in lib.jr_last_error()
    a.max_object_norms = poser.max_object_norms
    assert lib.jr_pose(ctypes.byref(a), stream) == 0
    torch.cuda.synchronize()
