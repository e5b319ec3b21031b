"""Host side of the batched driver: the system mix and the case builder."""

import math

import numpy as np


def test_batch_specs_mix_and_sizes():
    from kcfd_amd import workloads as W
    specs = W.batch_specs(512)
    n = np.array([s[1] for s in specs])
    assert len(specs) == 512 and n.min() >= 2000 and n.max() <= 250000
    # log-uniform: the mean of log n is the midpoint of the log range
    mid = 0.5 * (math.log(2000) + math.log(250000))
    assert abs(np.log(n).mean() - mid) < 0.15
    frac = {k: sum(s[0] == k for s in specs) / 512 for k in ("quad", "hex", "tet")}
    assert abs(frac["quad"] - 0.5) < 0.06 and abs(frac["hex"] - 0.3) < 0.06 and abs(frac["tet"] - 0.2) < 0.06
    assert [s[2] for s in specs] == list(range(4000, 4512))
    assert W.batch_specs(512) == specs


def test_parallel_build_matches_serial():
    from kcfd_amd import _lib as L, batch as B
    if L._default_ctx is not None:
        import pytest
        pytest.skip("builder processes are forked only before the GPU context exists")
    specs = [("quad", 400, 4000), ("hex", 343, 4001), ("tet", 600, 4002), ("quad", 900, 4003)]
    par = B.build_cases(specs, workers=2)
    ser = B.build_cases(specs, workers=1)
    for a, b in zip(par, ser):
        assert a.name == b.name and a.mesh.n_cells == b.mesh.n_cells
        assert np.array_equal(a.mesh.owner, b.mesh.owner)
        assert np.array_equal(a.rhs, b.rhs) if a.rhs is not None else b.rhs is None
    assert [c.name.split("_")[0] for c in par] == ["quad", "hex", "tet", "quad"]


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    from kcfd_amd import tune
    monkeypatch.setenv("KCFD_TUNE_CACHE", str(tmp_path / "t.json"))
    assert tune.cache_path() == str(tmp_path / "t.json")
    db = {"k": {"best": {"grid_per_cu": 4, "trsv_per_cu": 2, "check_every": 16}}}
    tune._store(db)
    assert tune._load() == db
    assert tune.describe(db["k"]["best"]) == "check_every=16, grid_per_cu=4, trsv_per_cu=2"
