"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the renderer package never imports
anything from here.

`run_oracle(call)` executes a jaxrenderer_amd._plan.RasterCall on the host with
the brute-force restatement in jr_oracle.c and returns numpy buffers shaped like
the GPU's.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, List, Optional

import numpy as np

from jaxrenderer_amd import _hip

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_INT_INPUTS = {"pos_index", "nrm_index", "uv_index", "tex_index", "tex_index_index",
               "col_index", "faces", "id_to_face", "face_slots", "texture_shapes"}

_lib: Optional[ctypes.CDLL] = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        L.jr_forward_cpu.argtypes = [P(_hip.JrFrame), P(_hip.JrInputs), ctypes.c_void_p,
                                     P(_hip.JrOutputs)]
        L.jr_forward_cpu.restype = ctypes.c_int32
        L.jr_forward_cpu_counted.argtypes = [P(_hip.JrFrame), P(_hip.JrInputs), ctypes.c_void_p,
                                             P(_hip.JrOutputs), ctypes.c_void_p]
        L.jr_forward_cpu_counted.restype = ctypes.c_int32
        L.jr_oracle_inv3.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.jr_oracle_det3.argtypes = [ctypes.c_void_p]
        L.jr_oracle_det3.restype = ctypes.c_float
        _lib = L
    return _lib


def _host(x, name):
    if x is None:
        return None
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    dt = np.int32 if name in _INT_INPUTS else np.float32
    return np.ascontiguousarray(x, dtype=dt)


def _p(a: Optional[np.ndarray]) -> int:
    return 0 if a is None else a.ctypes.data


def run_oracle(call, zbuffer: Optional[np.ndarray] = None, canvas: Optional[np.ndarray] = None,
               prim_id: bool = True, bary: bool = True,
               display: bool = False, complexity: bool = False) -> Dict[str, np.ndarray]:
    """`complexity`: also return, per pixel, how many primitives pass
    keep & inside & front-facing (the depth complexity)."""
    L = lib()
    B, W, H = call.n_scenes, call.width, call.height
    if zbuffer is None:
        zbuffer = np.full((B, W, H), call.zbuffer_default, dtype=np.float32)
    else:
        zbuffer = np.ascontiguousarray(zbuffer, dtype=np.float32).copy()
    if call.want_canvas:
        if canvas is None:
            canvas = np.empty((B, W, H, 3), dtype=np.float32)
            canvas[:] = np.asarray(call.colour_default, dtype=np.float32)
        else:
            canvas = np.ascontiguousarray(canvas, dtype=np.float32).copy()
    else:
        canvas = None
    pid = np.full((B, W, H), -1, dtype=np.int32) if prim_id else None
    bar = np.zeros((B, W, H, 3), dtype=np.float32) if bary else None
    disp = np.zeros((B, H, W, 3), dtype=np.uint8) if display else None
    keep_alive: List[np.ndarray] = []
    jin = _hip.JrInputs()
    for name, _ in _hip.JrInputs._fields_:
        if name in ("n_slots", "n_tris", "shadow_w", "shadow_h"):
            continue
        a = _host(call.inputs.get(name), name)
        if a is not None:
            keep_alive.append(a)
        setattr(jin, name, _p(a))
    jin.n_slots, jin.n_tris = call.n_slots, call.n_tris
    jin.shadow_w, jin.shadow_h = call.shadow_shape
    fr = _hip.JrFrame()
    fr.width, fr.height, fr.n_scenes, fr.shader = W, H, B, call.shader
    fr.init_buffers = int(bool(call.init_buffers))
    fr.apply_offset = int(bool(call.apply_offset))
    fr.zbuffer_default = call.zbuffer_default
    fr.colour_default = (ctypes.c_float * 3)(*np.asarray(call.colour_default, np.float32))
    fr.depth_offset = call.depth_offset
    fr.tile_w, fr.tile_h = 8, 8
    out = _hip.JrOutputs()
    out.zbuffer, out.canvas, out.prim_id, out.bary = _p(zbuffer), _p(canvas), _p(pid), _p(bar)
    out.display = _p(disp)
    scenes = (_hip.JrScene * B)(*call.scenes)
    cx = np.zeros((B, W, H), dtype=np.int32) if complexity else None
    rc = L.jr_forward_cpu_counted(ctypes.byref(fr), ctypes.byref(jin),
                                  ctypes.cast(scenes, ctypes.c_void_p), ctypes.byref(out), _p(cx))
    if rc != 0:
        raise RuntimeError(f"jr_forward_cpu failed: {rc}")
    res = {"zbuffer": zbuffer}
    if canvas is not None:
        res["canvas"] = canvas
    if pid is not None:
        res["prim_id"] = pid
    if bar is not None:
        res["bary"] = bar
    if disp is not None:
        res["display"] = disp
    if cx is not None:
        res["complexity"] = cx
    return res


def run_oracle_plan(calls, **kw) -> Dict[str, Dict[str, np.ndarray]]:
    results: Dict[str, Dict[str, np.ndarray]] = {}
    for c in calls:
        inputs = dict(c.inputs)
        if c.shadow_from is not None:
            inputs["shadow_maps"] = results[c.shadow_from]["zbuffer"]
        c2 = type(c)(**{**c.__dict__, "inputs": inputs})
        results[c.tag] = run_oracle(c2, **kw)
    return results


def inv3(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float32)
    x = np.empty((3, 3), dtype=np.float32)
    lib().jr_oracle_inv3(a.ctypes.data, x.ctypes.data)
    return x


def det3(a: np.ndarray) -> np.float32:
    a = np.ascontiguousarray(a, dtype=np.float32)
    return np.float32(lib().jr_oracle_det3(a.ctypes.data))
