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
