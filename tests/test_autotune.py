"""The launch-shape autotuner: cache handling on the CPU, tuning on the GPU
(every candidate must reproduce the same image; the choice must be one of
the candidates and be reused from the cache)."""

from __future__ import annotations

import json

import pytest
import torch

from jaxrenderer_amd import autotune, batch, workloads


def test_cache_roundtrip(tmp_path):
    path = str(tmp_path / "sub" / "tune.json")
    assert autotune._load(path) == {}
    autotune._store(path, {"k": {"tile": [8, 8], "form": 0, "ms": 1.0}})
    assert autotune._load(path)["k"]["tile"] == [8, 8]
    with open(path, "w") as f:
        f.write("not json")
    assert autotune._load(path) == {}


def test_candidates_are_supported_tiles():
    assert (8, 8) in autotune.TILES
    for (tw, th), form, ppl in autotune.DEFAULT_CANDIDATES:
        assert (tw * th // ppl) % 64 == 0  # whole wavefronts per workgroup
        assert form in autotune.FORMS and 0 <= form <= 64
        assert ppl == 1 or tw * th == 256


def test_cache_path_env(monkeypatch, tmp_path):
    monkeypatch.setenv("JR_AUTOTUNE_CACHE", str(tmp_path / "x.json"))
    assert autotune.cache_path() == str(tmp_path / "x.json")


@pytest.mark.gpu
def test_autotune_small_batch(tmp_path):
    scenes, smspec = workloads.config3(n_scenes=24, seed=3, width=84, height=84)
    calls = batch.plan_batch(scenes, 84, 84, smspec)
    ref = batch.BatchRenderer(calls, tile=(8, 8), shadow_tile=(8, 8), persistent=False)
    ref.calibrate()
    want = {k: v.clone() for k, v in ref.render().items()}
    torch.cuda.synchronize()

    br = batch.BatchRenderer(calls)
    path = str(tmp_path / "tune.json")
    got = br.autotune(path=path, reps=2)
    assert set(got) == {"shadow", "main"}
    for tag, c in got.items():
        assert not c["cached"]
        assert len(c["timings"]) == len(autotune.DEFAULT_CANDIDATES)
        assert (tuple(c["tile"]), c["form"], c["ppl"]) in autotune.DEFAULT_CANDIDATES
        assert br.passes[tag].persistent == (c["form"] == 0)
        assert br.passes[tag].tile == tuple(c["tile"])
    out = br.replay()
    torch.cuda.synchronize()
    assert torch.equal(out["zbuffer"].view(torch.int32), want["zbuffer"].view(torch.int32))
    assert float((out["canvas"] - want["canvas"]).abs().max()) <= 1e-5
    with open(path) as f:
        assert len(json.load(f)) == 2

    again = batch.BatchRenderer(calls).autotune(path=path, reps=2)
    for tag in got:
        assert again[tag]["cached"]
        assert again[tag]["tile"] == got[tag]["tile"]
