"""ctypes binding of libjrhip.so, the HIP rasterisation library.

The structs mirror include/jrhip.h field for field; ``check_layout`` compares
their sizes with the library's own ``jr_sizeof`` so a mismatched build fails
at load time instead of corrupting memory.

There is no fallback: if the library is missing the import of this module
raises, and every renderer entry point that needs it fails loudly.
"""

from __future__ import annotations

import ctypes
import os
from typing import Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjrhip.so")

# include/jrhip.h's JR_ABI_VERSION this binding is written for
ABI_VERSION = 2

# shader ids (include/jrhip.h)
SHADER_DEPTH = 0
SHADER_GOURAUD = 1
SHADER_GOURAUD_TEXTURE = 2
SHADER_PHONG = 3
SHADER_PHONG_DARBOUX = 4
SHADER_PHONG_REFLECTION = 5
SHADER_PHONG_REFLECTION_SHADOW = 6
SHADER_VISIBILITY = 7

# jr_stage_times: JR_STAGES timings per jr_forward call, in this order
STAGE_NAMES = ("clear", "setup", "wide", "scan", "bin", "raster")
# jr_frame.part: every stage, the binning stages 1-5 only, the raster only
PART_ALL, PART_BIN, PART_RASTER = 0, 1, 2
STAGES = len(STAGE_NAMES)

c_f = ctypes.c_float
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_vp = ctypes.c_void_p


class JrScene(ctypes.Structure):
    _fields_ = [
        ("tex_base", c_i64),
        ("spec_base", c_i64),
        ("nmap_base", c_i64),
        ("world_to_clip", c_f * 16),
        ("viewport", c_f * 16),
        ("world_to_eye_norm", c_f * 16),
        ("shadow_world_to_clip", c_f * 16),
        ("shadow_viewport", c_f * 16),
        ("light_direction", c_f * 3),
        ("light_colour", c_f * 3),
        ("light_dir_eye", c_f * 3),
        ("ambient", c_f * 3),
        ("diffuse", c_f * 3),
        ("specular", c_f * 3),
        ("shadow_strength", c_f * 3),
        ("pad_f", c_f),
        ("tri_begin", c_i32),
        ("tri_end", c_i32),
        ("tex_w", c_i32),
        ("tex_h", c_i32),
        ("tex_c", c_i32),
        ("nmap_w", c_i32),
        ("nmap_h", c_i32),
        ("texshape_base", c_i32),
        ("texshape_count", c_i32),
        ("tex_offset", c_i32),
        ("spec_w", c_i32),
        ("spec_h", c_i32),
    ]


class JrInputs(ctypes.Structure):
    _fields_ = [
        ("positions", c_vp),
        ("pos_index", c_vp),
        ("normals", c_vp),
        ("nrm_index", c_vp),
        ("uvs", c_vp),
        ("uv_index", c_vp),
        ("tex_index", c_vp),
        ("tex_index_index", c_vp),
        ("colours", c_vp),
        ("col_index", c_vp),
        ("faces", c_vp),
        ("id_to_face", c_vp),
        ("face_slots", c_vp),
        ("textures", c_vp),
        ("specular_maps", c_vp),
        ("normal_maps", c_vp),
        ("texture_shapes", c_vp),
        ("shadow_maps", c_vp),
        ("clip", c_vp),
        ("n_slots", c_i32),
        ("n_tris", c_i32),
        ("shadow_w", c_i32),
        ("shadow_h", c_i32),
        ("shadow_source", c_vp),  # host pointer to a JrShadowSource, or NULL
    ]


class JrOutputs(ctypes.Structure):
    _fields_ = [
        ("zbuffer", c_vp),
        ("canvas", c_vp),
        ("prim_id", c_vp),
        ("bary", c_vp),
        ("display", c_vp),
    ]


class JrFrame(ctypes.Structure):
    _fields_ = [
        ("bin_capacity", c_i64),
        ("width", c_i32),
        ("height", c_i32),
        ("n_scenes", c_i32),
        ("shader", c_i32),
        ("init_buffers", c_i32),
        ("apply_offset", c_i32),
        ("zbuffer_default", c_f),
        ("colour_default", c_f * 3),
        ("depth_offset", c_f),
        ("tile_w", c_i32),
        ("tile_h", c_i32),
        ("persistent", c_i32),
        ("tiles_per_group", c_i32),
        ("pixels_per_lane", c_i32),
        ("part", c_i32),
    ]


class JrShadowSource(ctypes.Structure):
    """jr_shadow_source: a shadow pass's binning, read by a
    phong_reflection_shadow pass that computes the shadow depths it needs."""
    _fields_ = [
        ("frame", c_vp),  # host pointer to the shadow pass's JrFrame
        ("scenes", c_vp),
        ("workspace", c_vp),
        ("workspace_bytes", c_i64),
        ("keys", c_vp),
        ("texels", c_vp),
        ("n_tris", c_i32),
        ("pad", c_i32),
    ]


class JrPackArgs(ctypes.Structure):
    """jr_pack_args: the per-corner flattening of a batch on the device."""
    _fields_ = [
        ("n_scenes", c_i32),
        ("pad", c_i32),
        ("n_faces", c_i64),
        ("tri_off", c_vp),
        ("vert_off", c_vp),
        ("norm_off", c_vp),
        ("uv_off", c_vp),
        ("texidx_off", c_vp),
        ("faces", c_vp),
        ("faces_norm", c_vp),
        ("faces_uv", c_vp),
        ("texture_index", c_vp),
        ("pos_index", c_vp),
        ("nrm_index", c_vp),
        ("uv_index", c_vp),
        ("tex_index", c_vp),
        ("error", c_vp),
    ]


class JrGrads(ctypes.Structure):
    _fields_ = [
        ("d_positions", c_vp),
        ("d_colours", c_vp),
        ("d_normals", c_vp),
        ("d_textures", c_vp),
        ("d_specular_maps", c_vp),
        ("d_scene", c_vp),
        ("d_uvs", c_vp),
        ("d_normal_maps", c_vp),
    ]


class JrPoseArgs(ctypes.Structure):
    _fields_ = [
        ("n_objects", c_i32),
        ("n_verts", c_i32),
        ("n_norms", c_i32),
        ("max_object_norms", c_i32),
        ("vert_begin", c_vp),
        ("norm_begin", c_vp),
        ("vert_object", c_vp),
        ("local_verts", c_vp),
        ("unit_norms", c_vp),
        ("scaling", c_vp),
        ("transforms", c_vp),
        ("normal_matrices", c_vp),
        ("positions", c_vp),
        ("normals", c_vp),
    ]


# bytes of jr_forward_noalloc's per-GPU scratch (include/jrhip.h)
SCRATCH_BYTES = 64 << 20

# largest normal count of one object jr_pose accepts (include/jrhip.h)
POSE_MAX_NORMALS = 43690

DSCENE_FLOATS = 69
DSCENE = {
    "light_direction": 0,
    "light_colour": 3,
    "light_dir_eye": 6,
    "ambient": 9,
    "diffuse": 12,
    "specular": 15,
    "world_to_clip": 18,
    "viewport": 34,
    "world_to_eye_norm": 50,
    "shadow_strength": 66,
}

STRUCTS = (JrScene, JrInputs, JrOutputs, JrFrame, JrGrads, JrPoseArgs, JrShadowSource,
           JrPackArgs)


class HipError(RuntimeError):
    """A libjrhip.so entry point returned an error code."""


def _declare(lib: ctypes.CDLL) -> None:
    P = ctypes.POINTER
    lib.jr_version.restype = c_i32
    lib.jr_last_error.restype = ctypes.c_char_p
    lib.jr_sizeof.argtypes = [c_i32]
    lib.jr_sizeof.restype = c_i64
    lib.jr_workspace_size.argtypes = [P(JrFrame), c_i32]
    lib.jr_workspace_size.restype = c_i64
    lib.jr_forward.argtypes = [P(JrFrame), P(JrInputs), c_vp, P(JrOutputs), c_vp, c_i64, c_vp]
    lib.jr_forward.restype = c_i32
    lib.jr_forward_noalloc.argtypes = [P(JrFrame), P(JrInputs), c_vp, P(JrOutputs), c_vp]
    lib.jr_forward_noalloc.restype = c_i32
    lib.jr_binned_pairs.argtypes = [c_vp, P(JrFrame), c_i32, P(c_i64), c_vp]
    lib.jr_binned_pairs.restype = c_i32
    lib.jr_stage_timing.argtypes = [c_i32]
    lib.jr_stage_timing.restype = c_i32
    lib.jr_stage_times.argtypes = [P(c_f), c_i32]
    lib.jr_stage_times.restype = c_i32
    lib.jr_backward.argtypes = [
        P(JrFrame), P(JrInputs), c_vp, c_vp, c_vp, c_vp, P(JrGrads), c_vp, c_i64, c_vp
    ]
    lib.jr_backward.restype = c_i32
    lib.jr_backward_l2.argtypes = [
        P(JrFrame), P(JrInputs), c_vp, c_vp, c_vp, c_f, c_vp, c_vp, P(JrGrads), c_vp, c_i64, c_vp
    ]
    lib.jr_backward_l2.restype = c_i32
    lib.jr_backward_workspace_size.argtypes = [P(JrFrame), c_i32]
    lib.jr_backward_workspace_size.restype = c_i64
    lib.jr_backward_last_error.restype = ctypes.c_char_p
    lib.jr_pose.argtypes = [P(JrPoseArgs), c_vp]
    lib.jr_pose.restype = c_i32
    lib.jr_pack.argtypes = [P(JrPackArgs), c_vp]
    lib.jr_pack.restype = c_i32
    lib.jr_pack_last_error.restype = ctypes.c_char_p


def check_layout(lib: ctypes.CDLL) -> None:
    for which, st in enumerate(STRUCTS):
        n = lib.jr_sizeof(which)
        if n != ctypes.sizeof(st):
            raise ImportError(
                f"{LIB_PATH}: sizeof({st.__name__}) is {n} in the library but "
                f"{ctypes.sizeof(st)} in the binding; rebuild libjrhip.so"
            )


_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises ImportError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: the HIP rasteriser is not built "
                "(run __graft_entry__.build() or `make -C jaxrenderer_amd/csrc`)"
            )
        handle = ctypes.CDLL(LIB_PATH)
        _declare(handle)
        check_layout(handle)
        if handle.jr_version() != ABI_VERSION:
            raise ImportError(f"{LIB_PATH}: ABI version {handle.jr_version()} != {ABI_VERSION}")
        _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        getter = {"jr_backward": lib().jr_backward_last_error,
                  "jr_backward_l2": lib().jr_backward_last_error,
                  "jr_pack": lib().jr_pack_last_error}.get(what, lib().jr_last_error)
        msg = getter().decode(errors="replace")
        raise HipError(f"{what} failed with code {rc}: {msg}")
