"""OBJ / TGA loaders (test_resources/utils.py:47-114).  The african-head mesh
is a data file of the reference's test resources, kept under tests/golden;
the diffuse TGA is read from the reference checkout when it is present."""
import os

import numpy as np
import pytest

from jaxrenderer_amd.resources import load_obj, load_tga, make_model

HERE = os.path.dirname(os.path.abspath(__file__))
HEAD = os.path.join(HERE, "golden", "african_head.obj")
REF_TGA = "/root/reference/test_resources/tga/african_head_diffuse.tga"


def test_obj_head_counts_and_ranges():
    m = load_obj(HEAD)
    assert (m.nverts, m.nfaces) == (1258, 2492)
    assert m.norms.shape == (1258, 3) and m.uv.shape == (1339, 2)
    for f, n in ((m.faces, m.nverts), (m.faces_norm, len(m.norms)), (m.faces_uv, len(m.uv))):
        assert f.dtype == np.int32 and f.min() >= 0 and f.max() < n
    # the file's first vertex, first uv and first face, read by hand
    with open(HEAD) as fh:
        lines = fh.readlines()
    first_v = next(ln for ln in lines if ln.startswith("v "))
    assert np.array_equal(m.verts[0], np.array([float(t) for t in first_v.split()[1:4]],
                                               np.float32))
    first_f = next(ln for ln in lines if ln.startswith("f "))
    v, vt, vn = zip(*[tuple(int(t) for t in c.split("/")) for c in first_f.split()[1:4]])
    assert m.faces[0].tolist() == [i - 1 for i in v]
    assert m.faces_uv[0].tolist() == [i - 1 for i in vt]
    assert m.faces_norm[0].tolist() == [i - 1 for i in vn]


def test_obj_parser_rules():
    m = make_model(["# c", "v 1 -2.5 3e-2", "vt 0.5 0.25 0", "vn 0 0 1",
                    "v 0 1 0", "v 1 1 1", "f 1/1/1 2/1/1 3/1/1"])
    assert np.allclose(m.verts[0], [1.0, -2.5, 0.03])
    assert m.faces.tolist() == [[0, 1, 2]] and m.faces_uv.tolist() == [[0, 0, 0]]
    with pytest.raises(ValueError):
        make_model(["f 1/1/1 2/1/1 3/1/1 4/1/1"])


def test_tga_layout_matches_getpixel(tmp_path):
    from PIL import Image

    rng = np.random.default_rng(0)
    pix = rng.integers(0, 256, size=(8, 8, 3), dtype=np.uint8)
    img = Image.fromarray(pix, "RGB")
    path = str(tmp_path / "t.tga")
    img.save(path)
    tex = load_tga(path)
    assert tex.shape == (8, 8, 3) and tex.dtype == np.float32
    back = Image.open(path)
    for x, y in ((0, 0), (7, 0), (3, 5), (7, 7)):
        # test_resources/utils.py:108: buffer[y, x] = image.getpixel((x, y))
        assert tex[y, x].tolist() == list(back.getpixel((x, y)))


@pytest.mark.skipif(not os.path.exists(REF_TGA), reason="reference checkout not present")
def test_reference_head_diffuse_map():
    from PIL import Image

    tex = load_tga(REF_TGA)
    assert tex.shape == (1024, 1024, 3)
    img = Image.open(REF_TGA)
    for x, y in ((0, 0), (512, 300), (1023, 1023), (100, 900)):
        assert tex[y, x].tolist() == list(img.getpixel((x, y)))
    assert 20.0 < float(tex.mean()) < 235.0
