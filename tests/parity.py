"""Comparison helpers: GPU buffers against the oracle's."""
import numpy as np


def as_np(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


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
