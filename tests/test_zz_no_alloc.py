"""jr_forward_noalloc, whose workspace is the library's per-GPU scratch array,
against jr_forward (a caller workspace) on the same calls, through the C ABI.

Two calls in a row, the second with a smaller input, so that the second runs
on scratch the first left full: every output buffer must equal jr_forward's
bit for bit (same kernels, same inputs), and the primitive ids and depths must
also equal the CPU oracle's.  A frame whose workspace exceeds the scratch is
refused before any launch."""

import numpy as np
import pytest
import torch

from jaxrenderer_amd import _hip, _plan, _raster
from jaxrenderer_amd.shaders import PhongReflectionShadowTextureShader
from oracle.oracle import run_oracle
from tests import scenes
from tests.parity import as_np, assert_parity

pytestmark = pytest.mark.gpu


def _scratch_need(call) -> int:
    return _raster.workspace_size(width=call.width, height=call.height,
                                  n_scenes=call.n_scenes, n_tris=call.n_tris,
                                  capacity=_raster.default_capacity(
                                      call.n_tris, _raster.n_tiles(call.width, call.height,
                                                                   call.n_scenes)))


def _bits_equal(a: dict, b: dict, label: str) -> None:
    assert a.keys() == b.keys(), label
    for k in a:
        x, y = as_np(a[k]), as_np(b[k])
        assert x.shape == y.shape, f"{label}: {k} shape"
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f"{label}: {k} differs"


def _check_pair(call, label: str) -> None:
    need = _scratch_need(call)
    assert need <= _hip.SCRATCH_BYTES, f"{label}: test call needs {need} bytes"
    print(f"no-alloc: {label} ({call.width}x{call.height}, {call.n_scenes} scene(s), "
          f"{call.n_tris} triangles, {need} workspace bytes)", flush=True)
    want = _plan.run(call, prim_id=True, bary=True)
    got = _plan.run(call, prim_id=True, bary=True, noalloc=True)
    torch.cuda.synchronize()
    _bits_equal(got, want, label)
    assert_parity(got, run_oracle(call), label=label)


def test_noalloc_matches_forward_two_calls_second_smaller():
    torch.cuda.init()
    big = scenes.tiny_soup_call(3, width=320, height=240, n_tris=6000)[0]
    small = scenes.soup_call(PhongReflectionShadowTextureShader, 5, width=40, height=30,
                             n_tris=25)
    assert big.n_tris > small[-1].n_tris
    _check_pair(big, "call 1: 6000-triangle depth soup")
    # the shadow pass feeds the main pass: both through the scratch, in turn
    ref = _plan.run_plan(small, prim_id=True, bary=True)
    shadow = _plan.run(small[0], prim_id=True, bary=True, noalloc=True)
    small[1].inputs["shadow_maps"] = shadow["zbuffer"]
    main = _plan.run(small[1], prim_id=True, bary=True, noalloc=True)
    torch.cuda.synchronize()
    print("no-alloc: call 2: 25-triangle shadowed soup (shadow pass, main pass)", flush=True)
    _bits_equal(shadow, ref["shadow"], "call 2 shadow pass")
    _bits_equal(main, ref["main"], "call 2 main pass")


def test_noalloc_refuses_a_frame_beyond_the_scratch():
    call = scenes.soup_call(scenes.DepthShader, 7, width=4096, height=2304, n_tris=10)[0]
    assert _scratch_need(call) > _hip.SCRATCH_BYTES
    with pytest.raises(_hip.HipError, match="JR_SCRATCH_BYTES"):
        _plan.run(call, noalloc=True)
    # a frame that fits still renders afterwards
    _check_pair(scenes.soup_call(scenes.DepthShader, 8, width=64, height=48, n_tris=30)[0],
                "after a refusal")
