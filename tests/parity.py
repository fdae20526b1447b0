"""Comparison helpers: GPU buffers against the oracle's."""
import numpy as np


def as_np(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def display_of(canvas: np.ndarray) -> np.ndarray:
    """transpose_for_display((clamp(0, canvas, 1) * 255).astype(uint8)) of a
    [B, W, H, 3] float32 canvas (renderer/utils.py:72-96 with the output step
    of notebooks/84x84/A100.ipynb cell 7), written with numpy's own clip,
    cast, swapaxes and flip; NaN -> 0 as XLA converts it."""
    c = np.clip(canvas.astype(np.float32), np.float32(0), np.float32(1))
    v = c * np.float32(255)
    v = np.where(np.isnan(v), np.float32(0), v)
    return np.swapaxes(v.astype(np.uint8), 1, 2)[:, ::-1]


def assert_parity(gpu: dict, ref: dict, atol: float = 1e-5, label: str = ""):
    """prim_id and zbuffer bit-exact; canvas and barycentrics within atol."""
    gz, rz = as_np(gpu["zbuffer"]), as_np(ref["zbuffer"])
    if "prim_id" in gpu and "prim_id" in ref:
        gp, rp = as_np(gpu["prim_id"]), as_np(ref["prim_id"])
        bad = np.argwhere(gp != rp)
        assert bad.size == 0, f"{label}: {len(bad)} prim_id mismatches, first {bad[:5].tolist()} gpu={gp[tuple(bad[0])]} ref={rp[tuple(bad[0])]}"
    gb, rb = gz.view(np.int32), rz.view(np.int32)
    bad = np.argwhere(gb != rb)
    assert bad.size == 0, f"{label}: {len(bad)} zbuffer bit mismatches, first {bad[:5].tolist()} gpu={gz[tuple(bad[0])]!r} ref={rz[tuple(bad[0])]!r}"
    if "canvas" in ref:
        gc, rc = as_np(gpu["canvas"]), as_np(ref["canvas"])
        assert np.array_equal(np.isnan(gc), np.isnan(rc)), f"{label}: NaN pattern differs"
        ok = ~np.isnan(rc)
        err = np.abs(gc[ok] - rc[ok]).max() if ok.any() else 0.0
        assert err <= atol, f"{label}: canvas max abs err {err}"
    if "bary" in gpu and "bary" in ref:
        gb_, rb_ = as_np(gpu["bary"]), as_np(ref["bary"])
        ok = ~np.isnan(rb_)
        assert np.array_equal(np.isnan(gb_), np.isnan(rb_)), f"{label}: bary NaN pattern differs"
        err = np.abs(gb_[ok] - rb_[ok]).max() if ok.any() else 0.0
        assert err == 0.0, f"{label}: bary differs by {err}"
    if "display" in gpu and "canvas" in gpu:
        gd = as_np(gpu["display"])
        want = display_of(as_np(gpu["canvas"]))
        assert np.array_equal(gd, want), f"{label}: display differs from its canvas"
        if "display" in ref and "canvas" in ref:
            rd = as_np(ref["display"])
            assert np.array_equal(rd, display_of(as_np(ref["canvas"]))), f"{label}: oracle display"
            d = np.abs(gd.astype(np.int16) - rd.astype(np.int16)).max()
            assert d <= 1, f"{label}: display differs from the oracle's by {d}"
