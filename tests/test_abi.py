"""The C ABI of libjrhip.so without a GPU: the library loads, exports every
function include/jrhip.h declares, and its struct sizes match the ctypes
binding."""
import ctypes
import os
import re

from jaxrenderer_amd import _hip

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "jrhip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(jr_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("jr_version", "jr_forward", "jr_backward", "jr_pose", "jr_workspace_size"):
        assert n in names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(_hip.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_binding():
    lib = _hip.lib()  # raises on a size mismatch
    assert lib.jr_version() == _hip.ABI_VERSION
    with open(HEADER) as f:
        assert f"#define JR_ABI_VERSION {_hip.ABI_VERSION}\n" in f.read()
    for which, st in enumerate(_hip.STRUCTS):
        assert lib.jr_sizeof(which) == ctypes.sizeof(st), st.__name__
    assert lib.jr_sizeof(len(_hip.STRUCTS)) == -1


def test_frame_checks_without_a_gpu():
    """jr_workspace_size validates the frame on the host: a frame of 2^22 or
    more tiles per scene (the tile-index quotients assume fewer) or an
    unsupported tile shape is refused with -1; a valid one gets a size that
    grows with the triangle count."""
    from jaxrenderer_amd import _raster

    lib = _hip.lib()
    ok = _raster.make_frame(width=84, height=84, n_scenes=2, shader=0, tile=(8, 8))
    n0 = lib.jr_workspace_size(ctypes.byref(ok), 10)
    n1 = lib.jr_workspace_size(ctypes.byref(ok), 100000)
    assert 0 < n0 < n1
    big = _raster.make_frame(width=16384, height=16384, n_scenes=1, shader=0, tile=(8, 8))
    assert lib.jr_workspace_size(ctypes.byref(big), 10) == -1  # 4,194,304 tiles
    edge = _raster.make_frame(width=16384, height=16376, n_scenes=1, shader=0, tile=(8, 8))
    assert lib.jr_workspace_size(ctypes.byref(edge), 10) > 0  # 4,193,792 tiles
    bad = _raster.make_frame(width=84, height=84, n_scenes=1, shader=0, tile=(8, 8))
    bad.tile_w = 12
    assert lib.jr_workspace_size(ctypes.byref(bad), 10) == -1
