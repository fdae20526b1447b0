"""Object poses of a planned batch, moved on the GPU (libjrhip.so jr_pose).

merge_objects (renderer/model.py:447-555) bakes every object's pose into its
scene's world-space vertices and normals.  A batch whose objects only move
(a physics rollout: new positions and orientations every step) would redo
that merge on the host for every scene and upload the result.  `ScenePoser`
keeps the objects' model-space vertices and unit normals on the device and
recomputes the world-space arrays of the whole batch from a [n_objects, 4, 4]
stack of poses with one jr_pose call, bit for bit as merge_objects computes
them, straight into a BatchRenderer's input buffers.
"""

from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _hip, _raster
from .geometry import normalise
from .model import ModelObject

f32 = np.float32


class ScenePoser:
    """Poses of the objects behind a batch of merged scenes.

    `objects[s]` is the list of ModelObject that merge_objects combined into
    scene s's model, in the same order; the scenes are in the order
    plan_batch received them.  Objects are numbered scene by scene
    (`object_begin[s]` is scene s's first).  The models and local scalings
    are fixed; only the 4x4 transforms change."""

    def __init__(self, objects: Sequence[Sequence[ModelObject]],
                 device: Optional[torch.device] = None):
        dev = device or _raster.require_device()
        verts: List[np.ndarray] = []
        units: List[np.ndarray] = []
        vcount: List[int] = []
        ncount: List[int] = []
        scaling: List[np.ndarray] = []
        transforms: List[np.ndarray] = []
        self.object_begin = [0]
        self.scene_verts: List[int] = []
        self.scene_norms: List[int] = []
        for scene in objects:
            self.scene_verts.append(sum(len(np.asarray(o.model.verts).reshape(-1, 3))
                                        for o in scene))
            self.scene_norms.append(sum(len(np.asarray(o.model.norms).reshape(-1, 3))
                                        for o in scene))
            for o in scene:
                v = np.asarray(o.model.verts, f32).reshape(-1, 3)
                n = np.asarray(o.model.norms, f32).reshape(-1, 3)
                if len(n) > _hip.POSE_MAX_NORMALS:
                    raise ValueError(f"an object has {len(n)} normals; jr_pose takes at most "
                                     f"{_hip.POSE_MAX_NORMALS} per object")
                verts.append(v)
                # the first normalise of Camera.apply_vec does not depend on the pose
                units.append(normalise(n).reshape(-1, 3) if len(n) else n)
                vcount.append(len(v))
                ncount.append(len(n))
                scaling.append(np.asarray(o.local_scaling, f32).reshape(3))
                transforms.append(np.asarray(o.transform, f32).reshape(4, 4))
            self.object_begin.append(len(vcount))
        self.n_objects = len(vcount)
        self.n_verts = int(sum(vcount))
        self.n_norms = int(sum(ncount))
        self.max_object_norms = int(max(ncount, default=0))
        vb = np.zeros(self.n_objects + 1, np.int64)
        vb[1:] = np.cumsum(vcount)
        nb = np.zeros(self.n_objects + 1, np.int64)
        nb[1:] = np.cumsum(ncount)
        if vb[-1] >= 2**31 or 3 * nb[-1] >= 2**31:
            raise ValueError("jr_pose indexes vertices and normal components with int32")
        self.device = dev

        def put(a, dtype):
            return torch.from_numpy(np.ascontiguousarray(a, dtype=dtype)).to(dev)

        self.vert_begin = put(vb, np.int32)
        self.norm_begin = put(nb, np.int32)
        self.vert_object = put(np.repeat(np.arange(self.n_objects), vcount), np.int32)
        self.local_verts = put(np.concatenate(verts) if verts else np.zeros((0, 3)), f32)
        self.unit_norms = put(np.concatenate(units) if units else np.zeros((0, 3)), f32)
        self.scaling = put(np.stack(scaling) if scaling else np.zeros((0, 3)), f32)
        self.initial_transforms = np.stack(transforms) if transforms else np.zeros((0, 4, 4), f32)
        self.normal_matrices = torch.empty((self.n_objects, 16), dtype=torch.float32, device=dev)

    def pose(self, transforms, positions: torch.Tensor,
             normals: Optional[torch.Tensor] = None) -> None:
        """Writes world-space `positions` [n_verts, 3] and (when given)
        `normals` [n_norms, 3] for `transforms` [n_objects, 4, 4] (host array
        or device tensor), on the current stream."""
        T = torch.as_tensor(np.asarray(transforms, f32) if not torch.is_tensor(transforms)
                            else transforms)
        T = T.to(device=self.device, dtype=torch.float32).reshape(self.n_objects, 16).contiguous()
        self._check(positions, self.n_verts, "positions")
        if normals is not None:
            self._check(normals, self.n_norms, "normals")
        a = _hip.JrPoseArgs()
        a.n_objects, a.n_verts, a.n_norms = self.n_objects, self.n_verts, self.n_norms
        a.max_object_norms = self.max_object_norms
        a.vert_begin = self.vert_begin.data_ptr()
        a.norm_begin = self.norm_begin.data_ptr()
        a.vert_object = self.vert_object.data_ptr()
        a.local_verts = self.local_verts.data_ptr()
        a.unit_norms = self.unit_norms.data_ptr()
        a.scaling = self.scaling.data_ptr()
        a.transforms = T.data_ptr()
        a.normal_matrices = self.normal_matrices.data_ptr()
        a.positions = positions.data_ptr()
        a.normals = normals.data_ptr() if normals is not None else None
        stream = torch.cuda.current_stream(self.device)
        _hip.check(_hip.lib().jr_pose(ctypes.byref(a), ctypes.c_void_p(stream.cuda_stream)),
                   "jr_pose")
        self._keep = T  # the launch reads it asynchronously

    def update(self, renderer, transforms) -> None:
        """Moves the objects of a BatchRenderer (built on the plan of the
        same scenes) to `transforms`: its position buffer, shared by the
        shadow pass, and the main pass's normals are rewritten in place, so a
        captured graph replays the new poses."""
        main = renderer.passes["main"]
        ex = main.call.extras
        if (ex.get("scene_verts") != self.scene_verts
                or ex.get("scene_norms") != self.scene_norms):
            raise ValueError("the renderer was planned from scenes whose per-scene vertex or "
                             "normal counts differ from this poser's objects")
        self.pose(transforms, main.inputs["positions"], main.inputs.get("normals"))

    def _check(self, t: torch.Tensor, n: int, name: str) -> None:
        if (not torch.is_tensor(t) or t.dtype != torch.float32 or t.device != self.device
                or tuple(t.shape) != (n, 3) or not t.is_contiguous()):
            raise ValueError(f"{name} must be a contiguous float32 [{n}, 3] tensor on "
                             f"{self.device}")


__all__ = ["ScenePoser"]
