"""Regenerates mesh_tables.npz from the reference's mesh constants.

The cube and capsule meshes of renderer/shapes/cube.py:16-140 and
renderer/shapes/capsule.py:19-1956 are data tables (vertex positions, normals,
uvs, face indices).  A drop-in must reproduce them exactly (face order decides
argmin ties), so this script reads the numbers out of those files and stores
them as arrays.  It needs the reference checkout and is run by hand; the
product only loads the resulting jaxrenderer_amd/shapes/mesh_tables.npz.

    python tests/golden/extract_tables.py /root/reference
"""

import os
import re
import sys

import numpy as np

_NUM = r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?"


def _block(text: str, name: str) -> str:
    start = text.index(f"{name}:")
    start = text.index("(", text.index("jnp.array(", start) + len("jnp.array("))
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return text[start : i + 1]
    raise ValueError(name)


def _rows(block: str, width: int):
    block = re.sub(r"#[^\n]*", "", block)
    pat = r"\(\s*" + r"\s*,\s*".join([f"({_NUM})"] * width) + r"\s*,?\s*\)"
    return [tuple(float(v) for v in m.groups()) for m in re.finditer(pat, block)]


def extract(ref_root: str) -> dict:
    out = {}
    for mesh in ("cube", "capsule"):
        text = open(os.path.join(ref_root, "renderer", "shapes", f"{mesh}.py")).read()
        out[f"{mesh}_verts"] = np.array(_rows(_block(text, "_verts"), 3), dtype=np.float32)
        out[f"{mesh}_normals"] = np.array(_rows(_block(text, "_normals"), 3), dtype=np.float32)
        out[f"{mesh}_uvs"] = np.array(_rows(_block(text, "_uvs"), 2), dtype=np.float32)
        out[f"{mesh}_faces"] = np.array(_rows(_block(text, "_faces"), 3), dtype=np.int32)
    return out


if __name__ == "__main__":
    tables = extract(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    for k, v in tables.items():
        print(k, v.shape)
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    np.savez(os.path.join(root, "jaxrenderer_amd", "shapes", "mesh_tables.npz"), **tables)
