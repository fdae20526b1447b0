"""Torch-side driver of jr_forward: turns device tensors into the C-ABI
structs, owns the per-stream workspace and launches on torch's current stream.

Only device tensors go through here; there is no CPU path.
"""

from __future__ import annotations

import ctypes
import threading
from dataclasses import dataclass, field
from typing import Dict, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _hip

_INPUT_FIELDS = (
    ("positions", torch.float32),
    ("pos_index", torch.int32),
    ("normals", torch.float32),
    ("nrm_index", torch.int32),
    ("uvs", torch.float32),
    ("uv_index", torch.int32),
    ("tex_index", torch.int32),
    ("tex_index_index", torch.int32),
    ("colours", torch.float32),
    ("col_index", torch.int32),
    ("faces", torch.int32),
    ("id_to_face", torch.int32),
    ("face_slots", torch.int32),
    ("textures", torch.float32),
    ("specular_maps", torch.float32),
    ("normal_maps", torch.float32),
    ("texture_shapes", torch.int32),
    ("shadow_maps", torch.float32),
    ("clip", torch.float32),
)

DEFAULT_TILE = (8, 8)


def require_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "jaxrenderer_amd renders on a ROCm GPU through libjrhip.so; no GPU is visible"
        )
    return torch.device("cuda", torch.cuda.current_device())


def _ptr(t: Optional[torch.Tensor], name: str, dtype: torch.dtype, dev: torch.device) -> int:
    if t is None:
        return 0
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if t.device != dev:
        raise ValueError(f"{name} must live on {dev}, got {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t.data_ptr()


def scene_struct(
    *,
    world_to_clip,
    viewport,
    world_to_eye_norm=None,
    shadow_world_to_clip=None,
    shadow_viewport=None,
    light_direction=(0.0, 0.0, -1.0),
    light_colour=(1.0, 1.0, 1.0),
    light_dir_eye=(0.0, 0.0, 1.0),
    ambient=(0.0, 0.0, 0.0),
    diffuse=(0.0, 0.0, 0.0),
    specular=(0.0, 0.0, 0.0),
    shadow_strength=(0.0, 0.0, 0.0),
    tri_begin=0,
    tri_end=0,
    tex_base=0,
    tex_shape=(1, 1, 3),
    spec_base=0,
    spec_shape=(1, 1),
    nmap_base=0,
    nmap_shape=(1, 1),
    texshape_base=0,
    texshape_count=1,
    tex_offset=0,
) -> _hip.JrScene:
    """Fills one jr_scene (include/jrhip.h) from host values."""

    def m16(m):
        if m is None:
            return (ctypes.c_float * 16)(*np.identity(4, dtype=np.float32).reshape(-1))
        a = np.asarray(m, dtype=np.float32).reshape(-1)
        return (ctypes.c_float * 16)(*a)

    def v3(v):
        a = np.asarray(v, dtype=np.float32).reshape(-1)
        if a.size == 1:
            a = np.repeat(a, 3)
        return (ctypes.c_float * 3)(*a[:3])

    s = _hip.JrScene()
    s.tex_base, s.spec_base, s.nmap_base = int(tex_base), int(spec_base), int(nmap_base)
    s.world_to_clip = m16(world_to_clip)
    s.viewport = m16(viewport)
    s.world_to_eye_norm = m16(world_to_eye_norm)
    s.shadow_world_to_clip = m16(shadow_world_to_clip)
    s.shadow_viewport = m16(shadow_viewport)
    s.light_direction = v3(light_direction)
    s.light_colour = v3(light_colour)
    s.light_dir_eye = v3(light_dir_eye)
    s.ambient = v3(ambient)
    s.diffuse = v3(diffuse)
    s.specular = v3(specular)
    s.shadow_strength = v3(shadow_strength)
    s.tri_begin, s.tri_end = int(tri_begin), int(tri_end)
    s.tex_w, s.tex_h, s.tex_c = (int(v) for v in tex_shape)
    s.spec_w, s.spec_h = (int(v) for v in spec_shape)
    s.nmap_w, s.nmap_h = (int(v) for v in nmap_shape)
    s.texshape_base, s.texshape_count = int(texshape_base), int(texshape_count)
    s.tex_offset = int(tex_offset)
    return s


def scenes_to_device(scenes: Sequence[_hip.JrScene], dev: torch.device) -> torch.Tensor:
    arr = (_hip.JrScene * len(scenes))(*scenes)
    host = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8)
    return host.to(dev)


@dataclass
class _Workspace:
    buf: Optional[torch.Tensor] = None
    # observed triangle-tile pair counts per (W, H, tile, n_scenes, n_tris)
    pairs: Dict[Tuple, int] = field(default_factory=dict)


_ws_lock = threading.Lock()
_workspaces: Dict[Tuple[int, int], _Workspace] = {}


def workspace_for(stream: torch.cuda.Stream, nbytes: int, dev: torch.device,
                  kind: str = "forward") -> torch.Tensor:
    """A cached workspace per (device, stream, kind).  The backward pass keeps
    its own: it carries zeroed accumulators from one call to the next."""
    key = (dev.index, stream.cuda_stream, kind)
    with _ws_lock:
        ws = _workspaces.setdefault(key, _Workspace())
        if ws.buf is None or ws.buf.numel() < nbytes:
            # zeroed once: the backward's header and the forward's key state
            # then name the state they hold; recycled memory must not pass
            # for a cleared workspace
            ws.buf = torch.zeros(max(nbytes, 1 << 20), dtype=torch.uint8, device=dev)
        return ws.buf


def default_capacity(n_tris: int, n_tiles: int) -> int:
    """Bin capacity heuristic: a few tiles per triangle plus every tile once.
    Lists that overflow it are still rendered exactly (brute force per tile)."""
    return int(min(0x7FFFFFFF, 4 * n_tris + 2 * n_tiles + 65536))


def make_frame(*, width, height, n_scenes, shader, init_buffers=False, apply_offset=False,
               zbuffer_default=1.0, colour_default=(1.0, 1.0, 1.0), depth_offset=0.0,
               tile=DEFAULT_TILE, persistent=False, capacity=0,
               tiles_per_group: int = 1, pixels_per_lane: int = 1,
               part: int = _hip.PART_ALL) -> _hip.JrFrame:
    fr = _hip.JrFrame()
    fr.bin_capacity = int(capacity)
    fr.width, fr.height, fr.n_scenes, fr.shader = int(width), int(height), int(n_scenes), int(shader)
    fr.init_buffers = int(bool(init_buffers))
    fr.apply_offset = int(bool(apply_offset))
    fr.zbuffer_default = float(zbuffer_default)
    cd = np.asarray(colour_default, dtype=np.float32).reshape(-1)
    if cd.size == 1:
        cd = np.repeat(cd, 3)
    fr.colour_default = (ctypes.c_float * 3)(*cd[:3])
    fr.depth_offset = float(depth_offset)
    fr.tile_w, fr.tile_h = int(tile[0]), int(tile[1])
    fr.persistent = int(bool(persistent))
    fr.tiles_per_group = int(tiles_per_group)
    fr.pixels_per_lane = int(pixels_per_lane)
    fr.part = int(part)
    return fr


def n_tiles(width: int, height: int, n_scenes: int, tile=DEFAULT_TILE) -> int:
    return ((width + tile[0] - 1) // tile[0]) * ((height + tile[1] - 1) // tile[1]) * n_scenes


def forward(
    *,
    shader: int,
    width: int,
    height: int,
    scenes: torch.Tensor,
    n_scenes: int,
    inputs: Dict[str, Optional[torch.Tensor]],
    n_slots: int,
    n_tris: int,
    zbuffer: torch.Tensor,
    canvas: Optional[torch.Tensor] = None,
    prim_id: Optional[torch.Tensor] = None,
    bary: Optional[torch.Tensor] = None,
    display: Optional[torch.Tensor] = None,
    shadow_shape: Tuple[int, int] = (0, 0),
    init_buffers: bool = False,
    zbuffer_default: float = 1.0,
    colour_default=(1.0, 1.0, 1.0),
    apply_offset: bool = False,
    depth_offset: float = 0.0,
    tile=DEFAULT_TILE,
    persistent: bool = False,
    tiles_per_group: int = 1,
    pixels_per_lane: int = 1,
    capacity: Optional[int] = None,
    stream: Optional[torch.cuda.Stream] = None,
    workspace: Optional[torch.Tensor] = None,
    part: int = _hip.PART_ALL,
    noalloc: bool = False,
    shadow_source: Optional["ShadowSource"] = None,
) -> None:
    """One jr_forward call on `stream` (torch's current stream by default).
    `part`: every stage, or only the binning stages (PART_BIN) or only the
    raster (PART_RASTER, on the workspace of a PART_BIN call of the same
    frame).  `noalloc`: call jr_forward_noalloc instead, whose workspace is
    the library's own per-GPU scratch (no allocation at all; frames up to
    SCRATCH_BYTES of workspace, whole passes only).  `shadow_source`
    (phong_reflection_shadow): compute the shadow depths from that shadow
    pass's binning instead of reading inputs["shadow_maps"]."""
    L = _hip.lib()
    dev = zbuffer.device
    if dev.type != "cuda":
        raise ValueError("buffers must be GPU tensors")
    stream = stream or torch.cuda.current_stream(dev)
    T = n_tiles(width, height, n_scenes, tile)
    cap = default_capacity(n_tris, T) if capacity is None else int(capacity)
    fr = make_frame(width=width, height=height, n_scenes=n_scenes, shader=shader,
                    init_buffers=init_buffers, apply_offset=apply_offset,
                    zbuffer_default=zbuffer_default, colour_default=colour_default,
                    depth_offset=depth_offset, tile=tile, persistent=persistent, capacity=cap,
                    tiles_per_group=tiles_per_group, pixels_per_lane=pixels_per_lane,
                    part=part)
    jin = _hip.JrInputs()
    for name, dt in _INPUT_FIELDS:
        setattr(jin, name, _ptr(inputs.get(name), name, dt, dev))
    jin.n_slots, jin.n_tris = int(n_slots), int(n_tris)
    jin.shadow_w, jin.shadow_h = int(shadow_shape[0]), int(shadow_shape[1])
    src = None
    if shadow_source is not None:
        src, _keep = shadow_source.struct(dev)
        jin.shadow_source = ctypes.addressof(src)
    jout = _hip.JrOutputs()
    jout.zbuffer = _ptr(zbuffer, "zbuffer", torch.float32, dev)
    jout.canvas = _ptr(canvas, "canvas", torch.float32, dev)
    jout.prim_id = _ptr(prim_id, "prim_id", torch.int32, dev)
    jout.bary = _ptr(bary, "bary", torch.float32, dev)
    jout.display = _ptr(display, "display", torch.uint8, dev)
    if display is not None and tuple(display.shape) != (n_scenes, height, width, 3):
        raise ValueError(f"display must be [{n_scenes}, {height}, {width}, 3] uint8")
    if scenes.device != dev or scenes.dtype != torch.uint8:
        raise ValueError("scenes must be a uint8 device tensor of jr_scene structs")
    if scenes.numel() != n_scenes * ctypes.sizeof(_hip.JrScene):
        raise ValueError("scenes tensor size does not match n_scenes")
    if noalloc:
        rc = L.jr_forward_noalloc(ctypes.byref(fr), ctypes.byref(jin),
                                  ctypes.c_void_p(scenes.data_ptr()), ctypes.byref(jout),
                                  ctypes.c_void_p(stream.cuda_stream))
        _hip.check(rc, "jr_forward_noalloc")
        return
    need = L.jr_workspace_size(ctypes.byref(fr), int(n_tris))
    if need < 0:
        _hip.check(-1, "jr_workspace_size")
    ws = workspace if workspace is not None else workspace_for(stream, need, dev)
    if ws.numel() < need:
        raise ValueError(f"workspace holds {ws.numel()} bytes, {need} needed")
    rc = L.jr_forward(ctypes.byref(fr), ctypes.byref(jin), ctypes.c_void_p(scenes.data_ptr()),
                      ctypes.byref(jout), ctypes.c_void_p(ws.data_ptr()), ctypes.c_int64(ws.numel()),
                      ctypes.c_void_p(stream.cuda_stream))
    _hip.check(rc, "jr_forward")


@dataclass
class ShadowSource:
    """A shadow pass (depth shader, offset, no id outputs) whose binning stages
    ran on `workspace`, for a phong_reflection_shadow pass that computes the
    texel depths it reads (jr_shadow_source): the pass's small-triangle keys
    and tile lists stay in its workspace.  `texels` [n_scenes, width, height]
    receives every depth so computed (None: they are not kept).  `keys` is
    not read (the shadow pass's zbuffer, which its binning leaves as it was)."""
    frame: _hip.JrFrame
    scenes: torch.Tensor
    workspace: torch.Tensor
    keys: Optional[torch.Tensor]
    texels: Optional[torch.Tensor]
    n_tris: int

    def struct(self, dev: torch.device):
        s = _hip.JrShadowSource()
        s.frame = ctypes.addressof(self.frame)
        s.scenes = _ptr(self.scenes, "shadow scenes", torch.uint8, dev)
        s.workspace = _ptr(self.workspace, "shadow workspace", torch.uint8, dev)
        s.workspace_bytes = int(self.workspace.numel())
        s.keys = _ptr(self.keys, "shadow keys", torch.float32, dev)
        s.texels = _ptr(self.texels, "shadow texels", torch.float32, dev)
        n, w, h = self.frame.n_scenes, self.frame.width, self.frame.height
        if self.texels is not None and tuple(self.texels.shape) != (n, w, h):
            raise ValueError(f"shadow texels must be [{n}, {w}, {h}]")
        s.n_tris = int(self.n_tris)
        return s, (self.frame,)


def binned_pairs(*, width, height, n_scenes, n_tris, tile=DEFAULT_TILE, capacity,
                 workspace: torch.Tensor, stream=None) -> int:
    """Triangle-tile pairs the last jr_forward on `workspace` produced (syncs)."""
    L = _hip.lib()
    stream = stream or torch.cuda.current_stream(workspace.device)
    fr = make_frame(width=width, height=height, n_scenes=n_scenes, shader=0, tile=tile,
                    capacity=capacity)
    out = ctypes.c_int64(0)
    rc = L.jr_binned_pairs(ctypes.c_void_p(workspace.data_ptr()), ctypes.byref(fr), int(n_tris),
                           ctypes.byref(out), ctypes.c_void_p(stream.cuda_stream))
    _hip.check(rc, "jr_binned_pairs")
    return int(out.value)


def workspace_size(*, width, height, n_scenes, n_tris, tile=DEFAULT_TILE, capacity) -> int:
    fr = make_frame(width=width, height=height, n_scenes=n_scenes, shader=0, tile=tile,
                    capacity=capacity)
    n = _hip.lib().jr_workspace_size(ctypes.byref(fr), int(n_tris))
    if n < 0:
        _hip.check(-1, "jr_workspace_size")
    return int(n)


_GRAD_FIELDS = ("d_positions", "d_colours", "d_normals", "d_textures", "d_specular_maps",
                "d_scene", "d_uvs", "d_normal_maps")


def backward(
    *,
    shader: int,
    width: int,
    height: int,
    scenes: torch.Tensor,
    n_scenes: int,
    inputs: Dict[str, Optional[torch.Tensor]],
    n_slots: int,
    n_tris: int,
    prim_id: torch.Tensor,
    d_canvas: Optional[torch.Tensor],
    d_zbuffer: Optional[torch.Tensor],
    grads: Dict[str, Optional[torch.Tensor]],
    shadow_shape: Tuple[int, int] = (0, 0),
    stream: Optional[torch.cuda.Stream] = None,
    workspace: Optional[torch.Tensor] = None,
    l2_canvas: Optional[torch.Tensor] = None,
    l2_target: Optional[torch.Tensor] = None,
    l2_scale: float = 2.0,
) -> None:
    """One jr_backward call: accumulates into the float32 tensors of `grads`
    (keys d_positions, d_colours, d_normals, d_textures, d_specular_maps,
    d_scene, d_uvs, d_normal_maps).  With l2_target (and the forward's canvas
    l2_canvas) in place of d_canvas it is one jr_backward_l2 call: the canvas
    cotangent l2_scale * (canvas - target) is formed inside the kernel."""
    L = _hip.lib()
    dev = prim_id.device
    stream = stream or torch.cuda.current_stream(dev)
    fr = make_frame(width=width, height=height, n_scenes=n_scenes, shader=shader)
    jin = _hip.JrInputs()
    for name, dt in _INPUT_FIELDS:
        setattr(jin, name, _ptr(inputs.get(name), name, dt, dev))
    jin.n_slots, jin.n_tris = int(n_slots), int(n_tris)
    jin.shadow_w, jin.shadow_h = int(shadow_shape[0]), int(shadow_shape[1])
    jg = _hip.JrGrads()
    for name in _GRAD_FIELDS:
        setattr(jg, name, _ptr(grads.get(name), name, torch.float32, dev))
    need = backward_workspace_size(width=width, height=height, n_scenes=n_scenes,
                                   shader=shader, n_tris=n_tris)
    ws = workspace if workspace is not None else workspace_for(stream, need, dev, "backward")
    if ws.numel() < need:
        raise ValueError(f"backward workspace holds {ws.numel()} bytes, {need} needed")
    if l2_target is not None:
        if d_canvas is not None or l2_canvas is None:
            raise ValueError("l2_target needs l2_canvas and no d_canvas")
        shape = (n_scenes, width, height, 3)
        for nm, t in (("l2_canvas", l2_canvas), ("l2_target", l2_target)):
            if tuple(t.shape) != shape:
                raise ValueError(f"{nm} must have shape {shape}, got {tuple(t.shape)}")
        rc = L.jr_backward_l2(
            ctypes.byref(fr), ctypes.byref(jin), ctypes.c_void_p(scenes.data_ptr()),
            ctypes.c_void_p(_ptr(l2_canvas, "l2_canvas", torch.float32, dev)),
            ctypes.c_void_p(_ptr(l2_target, "l2_target", torch.float32, dev)),
            ctypes.c_float(l2_scale),
            ctypes.c_void_p(_ptr(d_zbuffer, "d_zbuffer", torch.float32, dev)),
            ctypes.c_void_p(_ptr(prim_id, "prim_id", torch.int32, dev)), ctypes.byref(jg),
            ctypes.c_void_p(ws.data_ptr()), ctypes.c_int64(ws.numel()),
            ctypes.c_void_p(stream.cuda_stream))
        _hip.check(rc, "jr_backward_l2")
        return
    rc = L.jr_backward(
        ctypes.byref(fr), ctypes.byref(jin), ctypes.c_void_p(scenes.data_ptr()),
        ctypes.c_void_p(_ptr(d_canvas, "d_canvas", torch.float32, dev)),
        ctypes.c_void_p(_ptr(d_zbuffer, "d_zbuffer", torch.float32, dev)),
        ctypes.c_void_p(_ptr(prim_id, "prim_id", torch.int32, dev)), ctypes.byref(jg),
        ctypes.c_void_p(ws.data_ptr()), ctypes.c_int64(ws.numel()),
        ctypes.c_void_p(stream.cuda_stream))
    _hip.check(rc, "jr_backward")


def backward_workspace_size(*, width, height, n_scenes, shader, n_tris) -> int:
    fr = make_frame(width=width, height=height, n_scenes=n_scenes, shader=shader)
    n = _hip.lib().jr_backward_workspace_size(ctypes.byref(fr), int(n_tris))
    if n < 0:
        raise ValueError("jr_backward_workspace_size rejected the frame")
    return int(n)
