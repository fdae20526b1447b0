"""Batched rendering of many scenes of uneven size in one pass.

The reference batches with jax.vmap over equally-shaped scenes
(renderer/model.py:424-442, examples/batch_rendering.py:83-95).  Here scenes
are packed CSR-style -- triangles, vertices and texture maps of every scene
concatenated, with one jr_scene record per scene giving its triangle range,
map offsets, camera and light -- so scenes of 50 and 8000 triangles share one
launch of each kernel stage.

`plan_batch` builds the calls (shadow-map pass through the light camera with a
map of its own size, as renderer/shadow.py:40-125 does through
Shadow.render_shadow_map, then the phong_reflection(_shadow) pass of
renderer/renderer.py:277-385).  A batch of one scene is exactly the plain
per-scene forward.  `BatchRenderer` keeps everything on the device, reuses
its outputs and workspaces, and can capture the whole sequence in a HIP graph.
"""

from __future__ import annotations

from typing import Dict, List, NamedTuple, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _hip, _plan, _raster
from ._plan import RasterCall
from .geometry import Camera, asf32, normalise
from .model import MergedModel
from .renderer import LightParameters
from .shadow import Shadow

f32 = np.float32


class BatchScene(NamedTuple):
    model: MergedModel
    camera: Camera
    light: LightParameters
    # ShadowParameters.centre of this scene (renderer/renderer.py:120-130);
    # None: the batch's ShadowMapSpec.centre
    shadow_centre: Optional[Tuple[float, float, float]] = None


class ShadowMapSpec(NamedTuple):
    """Shadow.render_shadow_map arguments shared by the batch."""

    width: int = 256
    height: int = 256
    centre: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    up: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    strength: Tuple[float, float, float] = (0.6, 0.6, 0.6)
    offset: float = 0.05
    distance: float = 10.0


def plan_batch(scenes: Sequence[BatchScene], width: int, height: int,
               shadow: Optional[ShadowMapSpec] = None, *, zbuffer_default: float = 1.0,
               colour_default=(1.0, 1.0, 1.0)) -> List[RasterCall]:
    pos, nrm, uvs = [], [], []
    pos_index, nrm_index, uv_index, slot_tex, shadow_faces = [], [], [], [], []
    tex, spec, tshape = [], [], []
    main_sc, shadow_sc = [], []
    scene_verts: List[int] = []
    scene_norms: List[int] = []
    voff = noff = uoff = toff = tex_off = spec_off = ts_off = 0
    if shadow is not None:
        svp = Camera.viewport_matrix(np.zeros(2, np.int32),
                                     np.array((shadow.width, shadow.height)), 1.0)
    for s in scenes:
        m = s.model
        faces = np.asarray(m.faces, np.int32)
        F = int(faces.shape[0])
        fuv = np.asarray(m.faces_uv, np.int32).reshape(-1)
        ti = np.asarray(m.texture_index, np.int32)
        # indices become global below: one out of range would read another
        # scene's rows (or past the arrays) on the device
        _plan.check_index(faces, len(m.verts), "faces")
        _plan.check_index(m.faces_norm, len(m.norms), "faces_norm")
        _plan.check_index(fuv, len(m.uvs), "faces_uv")
        if np.asarray(m.faces_norm).shape != faces.shape or fuv.size != faces.size:
            raise ValueError("faces, faces_norm and faces_uv must all be [F, 3]")
        pos.append(asf32(m.verts))
        nrm.append(asf32(m.norms))
        uvs.append(asf32(m.uvs))
        pos_index.append(faces.reshape(-1) + voff)
        nrm_index.append(np.asarray(m.faces_norm, np.int32).reshape(-1) + noff)
        uv_index.append(fuv + uoff)
        slot_tex.append(ti[np.clip(fuv, 0, max(len(ti) - 1, 0))])
        shadow_faces.append(faces + voff)
        dm = asf32(m.diffuse_map)
        sm = asf32(m.specular_map)
        ts = np.asarray(m.texture_shape, np.int32)
        tex.append(dm.reshape(-1))
        spec.append(sm.reshape(-1))
        tshape.append(ts)
        light = s.light
        light_dir = normalise(asf32(light.direction))
        sc = dict(
            world_to_clip=s.camera.world_to_clip,
            viewport=s.camera.viewport,
            world_to_eye_norm=s.camera.world_to_eye_norm,
            light_direction=light_dir,
            light_colour=asf32(light.colour),
            light_dir_eye=Camera.apply_vec(light_dir, s.camera.view),
            ambient=asf32(light.ambient),
            diffuse=asf32(light.diffuse),
            specular=asf32(light.specular),
            tri_begin=toff,
            tri_end=toff + F,
            tex_base=tex_off,
            tex_shape=(dm.shape[0], dm.shape[1], dm.shape[2]),
            spec_base=spec_off,
            spec_shape=sm.shape[:2],
            texshape_base=ts_off,
            texshape_count=int(ts.shape[0]),
            tex_offset=int(m.offset),
        )
        if shadow is not None:
            centre = shadow.centre if s.shadow_centre is None else s.shadow_centre
            scam = Shadow.camera_for(light.direction, svp, centre, shadow.up,
                                     shadow.distance)
            sc.update(shadow_world_to_clip=scam.world_to_clip, shadow_viewport=scam.viewport,
                      shadow_strength=asf32(shadow.strength))
            shadow_sc.append(_raster.scene_struct(world_to_clip=scam.world_to_clip,
                                                  viewport=scam.viewport, tri_begin=toff,
                                                  tri_end=toff + F))
        main_sc.append(_raster.scene_struct(**sc))
        scene_verts.append(len(m.verts))
        scene_norms.append(len(m.norms))
        voff += len(m.verts)
        noff += len(m.norms)
        uoff += len(m.uvs)
        toff += F
        tex_off += dm.size
        spec_off += sm.size
        ts_off += int(ts.shape[0])
    positions = np.concatenate(pos)
    inputs = {
        "positions": positions,
        "pos_index": np.concatenate(pos_index),
        "normals": np.concatenate(nrm),
        "nrm_index": np.concatenate(nrm_index),
        "uvs": np.concatenate(uvs),
        "uv_index": np.concatenate(uv_index),
        "tex_index": np.concatenate(slot_tex),
        "faces": np.arange(3 * toff, dtype=np.int32).reshape(toff, 3),
        "textures": np.concatenate(tex),
        "specular_maps": np.concatenate(spec),
        "texture_shapes": np.concatenate(tshape),
    }
    calls: List[RasterCall] = []
    shader = _hip.SHADER_PHONG_REFLECTION
    shadow_shape = (0, 0)
    if shadow is not None:
        calls.append(RasterCall(
            shader=_hip.SHADER_DEPTH, width=shadow.width, height=shadow.height, scenes=shadow_sc,
            inputs={"positions": positions, "faces": np.concatenate(shadow_faces)},
            n_slots=len(positions), n_tris=toff, init_buffers=True,
            zbuffer_default=float(np.finfo(f32).max), apply_offset=True,
            depth_offset=float(shadow.offset), want_canvas=False, tag="shadow"))
        shader = _hip.SHADER_PHONG_REFLECTION_SHADOW
        shadow_shape = (shadow.width, shadow.height)
    calls.append(RasterCall(
        shader=shader, width=width, height=height, scenes=main_sc, inputs=inputs,
        n_slots=3 * toff, n_tris=toff, shadow_shape=shadow_shape, init_buffers=True,
        zbuffer_default=float(zbuffer_default),
        colour_default=tuple(float(c) for c in np.asarray(colour_default, f32).reshape(-1)[:3]),
        shadow_from="shadow" if shadow is not None else None, tag="main",
        extras={"identity_faces": True, "scene_verts": scene_verts,
                "scene_norms": scene_norms}))
    return calls


class _Pass:
    def __init__(self, call: RasterCall, dev: torch.device, tile, persistent: bool,
                 want_ids: bool, want_display: bool = False):
        self.call = call
        self.tile = tile
        self.persistent = persistent
        self.tiles_per_group = 1
        self.pixels_per_lane = 1
        self.inputs = _plan.device_inputs(call, dev)
        self.scenes = _raster.scenes_to_device(call.scenes, dev)
        B, W, H = call.n_scenes, call.width, call.height
        self.zbuffer = torch.empty((B, W, H), dtype=torch.float32, device=dev)
        self.canvas = (torch.empty((B, W, H, 3), dtype=torch.float32, device=dev)
                       if call.want_canvas else None)
        self.prim_id = torch.empty((B, W, H), dtype=torch.int32, device=dev) if want_ids else None
        self.bary = torch.empty((B, W, H, 3), dtype=torch.float32, device=dev) if want_ids else None
        self.display = (torch.empty((B, H, W, 3), dtype=torch.uint8, device=dev)
                        if want_display and call.want_canvas else None)
        self.set_capacity(_raster.default_capacity(call.n_tris, self.n_tiles))

    @property
    def n_tiles(self) -> int:
        c = self.call
        return _raster.n_tiles(c.width, c.height, c.n_scenes, self.tile)

    def set_capacity(self, cap: int) -> None:
        c = self.call
        self.capacity = int(cap)
        need = _raster.workspace_size(width=c.width, height=c.height, n_scenes=c.n_scenes,
                                      n_tris=c.n_tris, tile=self.tile, capacity=self.capacity)
        # zeroed once (the depth pass's key state then describes what it holds)
        self.workspace = torch.zeros(need, dtype=torch.uint8, device=self.zbuffer.device)

    def frame(self, part: int = _hip.PART_ALL) -> _hip.JrFrame:
        """The jr_frame of this pass's launches."""
        c = self.call
        return _raster.make_frame(
            width=c.width, height=c.height, n_scenes=c.n_scenes, shader=c.shader,
            init_buffers=c.init_buffers, apply_offset=c.apply_offset,
            zbuffer_default=c.zbuffer_default, colour_default=c.colour_default,
            depth_offset=c.depth_offset, tile=self.tile, persistent=self.persistent,
            capacity=self.capacity, tiles_per_group=self.tiles_per_group,
            pixels_per_lane=self.pixels_per_lane, part=part)

    def launch(self, part: int = _hip.PART_ALL, shadow_source=None) -> None:
        c = self.call
        _raster.forward(
            shader=c.shader, width=c.width, height=c.height, scenes=self.scenes,
            n_scenes=c.n_scenes, inputs=self.inputs, n_slots=c.n_slots, n_tris=c.n_tris,
            zbuffer=self.zbuffer, canvas=self.canvas, prim_id=self.prim_id, bary=self.bary,
            display=self.display, shadow_shape=c.shadow_shape, init_buffers=c.init_buffers,
            zbuffer_default=c.zbuffer_default, colour_default=c.colour_default,
            apply_offset=c.apply_offset, depth_offset=c.depth_offset, tile=self.tile,
            persistent=self.persistent, tiles_per_group=self.tiles_per_group,
            pixels_per_lane=self.pixels_per_lane,
            capacity=self.capacity, workspace=self.workspace, part=part,
            shadow_source=shadow_source)

    def pairs(self) -> int:
        c = self.call
        return _raster.binned_pairs(width=c.width, height=c.height, n_scenes=c.n_scenes,
                                    n_tris=c.n_tris, tile=self.tile, capacity=self.capacity,
                                    workspace=self.workspace)


# texels per main-pass pixel from which "auto" computes shadow texels on
# demand: measured on config 3 (256x256 maps, 84x84 frames, 9.3 texels per
# pixel: 1.165 -> 1.016 ms per frame) and config 5 (512x512, 480x270, 2.0:
# forward 2.17 -> 2.16 ms, forward + backward 5.92 -> 5.97 ms, so off)
ON_DEMAND_RATIO = 4


class BatchRenderer:
    """Device-resident batched renderer over a plan from `plan_batch`."""

    def __init__(self, calls: List[RasterCall], *, tile=(8, 8), shadow_tile=None,
                 persistent: bool = True, want_ids: bool = False, display: bool = False,
                 shadow_on_demand=False, keep_shadow_texels: bool = False):
        """`display`: also write the main pass's uint8 display images
        [B, H, W, 3] (clamp to [0, 1], x 255, truncate, transpose_for_display).

        `shadow_on_demand`: the shadow pass runs only its binning stages and
        the main pass computes the depth of each shadow-map texel it reads
        from that binning (jr_shadow_source), with the shadow raster's own
        arithmetic, so the image is the same bit for bit.  The depths it
        computed are kept in `shadow_texels` when `keep_shadow_texels` is set
        or once enable_backward() has been called (the backward reads them
        there); a forward-only renderer skips those stores.  The other texels
        are not computed, so no whole shadow map is output.
        "auto": on when the shadow map has at least ON_DEMAND_RATIO times as
        many texels as the main pass has pixels (then most texels are never
        read, and computing only the read ones is the cheaper way)."""
        dev = _raster.require_device()
        self.passes: Dict[str, _Pass] = {}
        for c in calls:
            t = shadow_tile if (c.tag == "shadow" and shadow_tile is not None) else tile
            main = c.tag == "main"
            self.passes[c.tag] = _Pass(c, dev, t, persistent, want_ids and main,
                                       display and main)
        main = self.passes["main"]
        if shadow_on_demand == "auto" and main.call.shadow_from is not None:
            sc = self.passes[main.call.shadow_from].call
            shadow_on_demand = (sc.width * sc.height >=
                                ON_DEMAND_RATIO * main.call.width * main.call.height)
        self.shadow_on_demand = bool(shadow_on_demand) and main.call.shadow_from is not None
        self.keep_shadow_texels = bool(keep_shadow_texels)
        self.shadow_texels: Optional[torch.Tensor] = None
        if self.shadow_on_demand:
            sc = self.passes[main.call.shadow_from].call
            # zeros: texels no pixel reads are never written
            self.shadow_texels = torch.zeros((sc.n_scenes, sc.width, sc.height),
                                             dtype=torch.float32, device=dev)
            main.inputs["shadow_maps"] = self.shadow_texels
        elif main.call.shadow_from is not None:
            main.inputs["shadow_maps"] = self.passes[main.call.shadow_from].zbuffer
        for p in self.passes.values():
            # one device copy of the world positions both passes read (and a
            # pose update rewrites)
            if p is not main and p.call.inputs.get("positions") is main.call.inputs["positions"]:
                p.inputs["positions"] = main.inputs["positions"]
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self._side: Optional[torch.cuda.Stream] = None

    def _shadow_source(self) -> Optional[_raster.ShadowSource]:
        """The main pass's view of the shadow pass's binning (on demand)."""
        if not self.shadow_on_demand:
            return None
        sp = self.passes[self.passes["main"].call.shadow_from]
        return _raster.ShadowSource(frame=sp.frame(), scenes=sp.scenes,
                                    workspace=sp.workspace, keys=sp.zbuffer,
                                    texels=self.shadow_texels if self.keep_shadow_texels
                                    else None, n_tris=sp.call.n_tris)

    def _launch(self, tag: str, part: int = _hip.PART_ALL) -> None:
        """One pass, as a render runs it: with shadows on demand the shadow
        pass stops after binning and the main raster reads that binning."""
        p = self.passes[tag]
        if self.shadow_on_demand and tag != "main":
            if part != _hip.PART_RASTER:
                p.launch(_hip.PART_BIN)
            return
        src = self._shadow_source() if tag == "main" and part != _hip.PART_BIN else None
        p.launch(part, shadow_source=src)

    def render(self) -> Dict[str, torch.Tensor]:
        for tag in self.passes:
            self._launch(tag)
        return self._outputs()

    def calibrate(self, slack: float = 1.25) -> Dict[str, int]:
        """Runs once and sizes every pass's bin capacity to the binned pairs
        (times `slack`), so no tile falls back to the brute-force path."""
        self.render()
        torch.cuda.synchronize()
        got = {}
        for tag, p in self.passes.items():
            n = p.pairs()
            got[tag] = n
            p.set_capacity(int(n * slack) + 1024)
        self.graph = None
        return got

    def autotune(self, **kw) -> Dict[str, dict]:
        """Chooses tile shape, launch form and bin capacity per pass by timing
        them on this batch (jaxrenderer_amd.autotune.tune)."""
        from . import autotune

        return autotune.tune(self, **kw)

    def _outputs(self) -> Dict[str, torch.Tensor]:
        main = self.passes["main"]
        out = {"zbuffer": main.zbuffer, "canvas": main.canvas}
        if main.prim_id is not None:
            out.update(prim_id=main.prim_id, bary=main.bary)
        if main.display is not None:
            out["display"] = main.display
        if self.shadow_on_demand:
            if self.keep_shadow_texels:
                out["shadow_texels"] = self.shadow_texels
        elif "shadow" in self.passes:
            out["shadow_map"] = self.passes["shadow"].zbuffer
        return out

    def render_overlapped(self) -> Dict[str, torch.Tensor]:
        """render() with the main pass's binning on a second stream, beside
        the shadow pass (it reads no shadow map); only the main raster waits
        for the shadow map.  Same results as render()."""
        if "shadow" not in self.passes or len(self.passes) != 2:
            return self.render()
        cur = torch.cuda.current_stream()
        if self._side is None:
            self._side = torch.cuda.Stream()  # default priority
        side = self._side
        side.wait_stream(cur)
        with torch.cuda.stream(side):
            self._launch("main", _hip.PART_BIN)
        self._launch("shadow")
        side.wait_stream(cur)  # the shadow map (or its binning) is complete
        with torch.cuda.stream(side):
            self._launch("main", _hip.PART_RASTER)
        cur.wait_stream(side)
        return self._outputs()

    def capture(self) -> None:
        """Captures render_overlapped() into a HIP graph
        (torch.cuda.CUDAGraph)."""
        self.render_overlapped()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.render_overlapped()
        self.graph = g

    def replay(self) -> Dict[str, torch.Tensor]:
        if self.graph is None:
            self.capture()
        self.graph.replay()
        main = self.passes["main"]
        out = {"zbuffer": main.zbuffer, "canvas": main.canvas}
        if main.prim_id is not None:
            out.update(prim_id=main.prim_id, bary=main.bary)
        if main.display is not None:
            out["display"] = main.display
        return out

    # ---- reverse mode -----------------------------------------------------
    def enable_backward(self) -> Dict[str, torch.Tensor]:
        """Allocates the main pass's primitive-id buffer, the gradient buffers
        (world-space vertices, normals, merged diffuse and specular maps, and
        the per-scene parameter rows of include/jrhip.h's JR_DSCENE layout)
        and the backward workspace.  The shadow pass needs no backward: the
        shadow map only enters through comparisons."""
        main = self.passes["main"]
        dev = main.zbuffer.device
        c = main.call
        if main.prim_id is None:
            main.prim_id = torch.empty((c.n_scenes, c.width, c.height), dtype=torch.int32,
                                       device=dev)
        # the backward reads the shadow depths the forward computed on demand
        self.keep_shadow_texels = True
        self.grads = {}
        for name, key in (("d_positions", "positions"), ("d_normals", "normals"),
                          ("d_colours", "colours"), ("d_textures", "textures"),
                          ("d_specular_maps", "specular_maps")):
            if main.inputs.get(key) is not None:
                self.grads[name] = torch.zeros_like(main.inputs[key])
        self.grads["d_scene"] = torch.zeros((c.n_scenes, _hip.DSCENE_FLOATS),
                                            dtype=torch.float32, device=dev)
        need = _raster.backward_workspace_size(width=c.width, height=c.height,
                                               n_scenes=c.n_scenes, shader=c.shader,
                                               n_tris=c.n_tris)
        # zeroed once: recycled memory whose header happened to match a cleared
        # workspace would otherwise skip the first full clear
        self.bwd_workspace = torch.zeros(need, dtype=torch.uint8, device=dev)
        self.graph = None
        return self.grads

    def backward(self, d_canvas: torch.Tensor,
                 d_zbuffer: Optional[torch.Tensor] = None) -> Dict[str, torch.Tensor]:
        """Gradients of the last render() for upstream gradients of the canvas
        [B, W, H, 3] (and optionally the zbuffer), written into self.grads."""
        main = self.passes["main"]
        c = main.call
        for g in self.grads.values():
            g.zero_()
        _raster.backward(
            shader=c.shader, width=c.width, height=c.height, scenes=main.scenes,
            n_scenes=c.n_scenes, inputs=main.inputs, n_slots=c.n_slots, n_tris=c.n_tris,
            prim_id=main.prim_id, d_canvas=d_canvas, d_zbuffer=d_zbuffer, grads=self.grads,
            shadow_shape=c.shadow_shape, workspace=self.bwd_workspace)
        return self.grads

    def backward_l2(self, target: torch.Tensor) -> Dict[str, torch.Tensor]:
        """Gradients of the L2 image loss sum((canvas - target)^2) of the last
        render(), written into self.grads: one jr_backward_l2 call, which
        forms the canvas cotangent 2 (canvas - target) per pixel itself."""
        main = self.passes["main"]
        c = main.call
        for g in self.grads.values():
            g.zero_()
        _raster.backward(
            shader=c.shader, width=c.width, height=c.height, scenes=main.scenes,
            n_scenes=c.n_scenes, inputs=main.inputs, n_slots=c.n_slots, n_tris=c.n_tris,
            prim_id=main.prim_id, d_canvas=None, d_zbuffer=None, grads=self.grads,
            shadow_shape=c.shadow_shape, workspace=self.bwd_workspace,
            l2_canvas=main.canvas, l2_target=target, l2_scale=2.0)
        return self.grads

    def train_step(self, target: torch.Tensor) -> Dict[str, torch.Tensor]:
        """Forward, the gradient of the L2 image loss sum((canvas - target)^2),
        and backward."""
        self.render_overlapped()
        return self.backward_l2(target)

    def capture_train(self, target: torch.Tensor) -> torch.cuda.CUDAGraph:
        """Captures train_step(target) in one HIP graph."""
        self.train_step(target)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.train_step(target)
        self.train_graph = g
        return g


__all__ = ["BatchRenderer", "BatchScene", "ShadowMapSpec", "plan_batch"]
