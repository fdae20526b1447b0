"""Host-side input checks (CPU): index arrays that would make a kernel read
past an array, or another scene's rows, are rejected before any launch, and
the host Shadow.get follows the kernel's float -> int conversion."""
import numpy as np
import pytest
import torch

from jaxrenderer_amd import batch, pipeline, workloads
from jaxrenderer_amd._custom import render_custom
from jaxrenderer_amd.shaders import GouraudExtraInput, GouraudShader
from jaxrenderer_amd.shadow import Shadow
from jaxrenderer_amd.types import Buffers, LightSource
from tests import scenes

f32 = np.float32


def _smoke_extra(n_vertices=6):
    light = LightSource()
    return GouraudExtraInput(
        position=scenes.SMOKE_POSITIONS[:n_vertices],
        colour=scenes.SMOKE_COLOURS[:n_vertices],
        normal=np.tile(np.asarray(light.direction, f32), (n_vertices, 1)),
        light=light,
    )


@pytest.mark.parametrize("bad", [6, -1, 2**20])
def test_build_call_rejects_face_out_of_range(bad):
    faces = scenes.SMOKE_FACES.copy()
    faces[3, 1] = bad
    with pytest.raises(ValueError, match="face_indices out of bound"):
        pipeline.build_call(scenes.smoke_camera(32, 32), GouraudShader, faces, _smoke_extra(),
                            32, 32)


def test_build_call_rejects_short_per_vertex_array():
    ex = _smoke_extra()._replace(colour=scenes.SMOKE_COLOURS[:4])
    with pytest.raises(ValueError, match="extra.colour has 4 rows"):
        pipeline.build_call(scenes.smoke_camera(32, 32), GouraudShader, scenes.SMOKE_FACES, ex,
                            32, 32)


def test_build_call_accepts_valid_faces():
    call = pipeline.build_call(scenes.smoke_camera(32, 32), GouraudShader, scenes.SMOKE_FACES,
                               _smoke_extra(), 32, 32)
    assert call.n_tris == 5 and call.n_slots == 6


@pytest.mark.parametrize("field,delta", [("faces", 1), ("faces_norm", 1), ("faces_uv", 1),
                                         ("faces", -10**6)])
def test_plan_batch_rejects_index_out_of_range(field, delta):
    scs, sm = workloads.config3(n_scenes=2, seed=5)
    m = scs[1].model
    arr = np.asarray(getattr(m, field)).copy()
    n = {"faces": len(m.verts), "faces_norm": len(m.norms), "faces_uv": len(m.uvs)}[field]
    arr[0, 0] = n - 1 + delta if delta > 0 else delta
    scs[1] = scs[1]._replace(model=m._replace(**{field: arr}))
    with pytest.raises(ValueError, match=f"{field} out of bound"):
        batch.plan_batch(scs, 84, 84, sm)


def test_render_custom_rejects_face_out_of_range():
    faces = scenes.SMOKE_FACES.copy()
    faces[0, 2] = 7
    buffers = Buffers(zbuffer=torch.ones((16, 16)), targets=(torch.zeros((16, 16, 3)),))

    def never(call):  # the check must fire before the visibility pass runs
        raise AssertionError("visibility pass reached")

    class Hook(GouraudShader):
        @staticmethod
        def vertex(gl_VertexID, gl_InstanceID, camera, extra):
            return GouraudShader.vertex(gl_VertexID, gl_InstanceID, camera, extra)

    with pytest.raises(ValueError, match="face_indices out of bound"):
        render_custom(scenes.smoke_camera(16, 16), Hook, buffers, faces, _smoke_extra(),
                      visibility=never)


def test_shadow_get_nan_and_huge_positions():
    """XLA converts NaN to 0 and saturates at the int32 range
    (renderer/shadow.py:127-153): NaN reads texel 0, huge values are out of
    range (+inf), negative indices count from the end."""
    sm = torch.arange(12, dtype=torch.float32).reshape(4, 3)
    sh = Shadow(shadow_map=sm, strength=np.zeros(3, f32), camera=None)
    pos = torch.tensor([[float("nan"), float("nan"), 0.0],
                        [1e30, 0.0, 0.0],
                        [-1e30, 1.0, 0.0],
                        [-1.0, -1.0, 0.0],
                        [1.5, 0.49, 0.0],
                        [float("inf"), 0.0, 0.0]])
    got = sh.get(pos)
    assert float(got[0]) == 0.0
    assert float(got[1]) == float("inf")
    assert float(got[2]) == float("inf")
    assert float(got[3]) == float(sm[3, 2])
    assert float(got[4]) == float(sm[2, 0])
    assert float(got[5]) == float("inf")


def test_plan_batch_rejects_what_value_checks_rejects():
    scs, _ = workloads.config3(n_scenes=1, seed=1)
    m = scs[0].model
    faces = np.asarray(m.faces).copy()
    faces[0, 0] = len(m.verts)
    with pytest.raises(ValueError):
        batch.plan_batch([scs[0]._replace(model=m._replace(faces=faces))], 8, 8)
