"""The autotuner (hgx_autotune): launch shapes per size class measured on the device, kept in
the context and written to a JSON cache keyed by arch and CU count, loaded back by a new context
(explicitly and through HGX_TUNE_CACHE). Tuning never changes a result: the same batch before and
after tuning gives identical counts and state hashes."""
import json
import os

import numpy as np
import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_autotune_cache_roundtrip_and_results_unchanged(tmp_path):
    ctx = hgx.Context(0)
    reqs = [W.batched_run(i, cap=1024) for i in range(192)]
    reqs += [W.served_run(i, cap=512) for i in (3, 10, 37, 90)]          # 300-2,000 edge roots
    before = ctx.evolve_batch(reqs)
    cache = str(tmp_path / "tune.json")
    rep = ctx.autotune(reqs, reps=1, cache_path=cache)
    chosen = rep["chosen"]
    assert all(x in (16, 32, 64) for x in chosen["expand_lanes"])
    assert all(x in (4, 8, 16) for x in chosen["expand_grid"])
    assert all(x in (0, 1, 4) for x in chosen["join_team"])
    assert all(x in (8, 16, 32) for x in chosen["join_grid"])
    assert rep["measured"]["expand"] and rep["measured"]["join"]
    assert "autotuned" in ctx.tuning()["source"]
    after = ctx.evolve_batch(reqs)
    for i in range(len(reqs)):
        assert after[i].counts == before[i].counts
        assert np.array_equal(after[i].states()[0], before[i].states()[0])
    # the cache: one entry under this device's key, read back by a new context
    text = json.load(open(cache))
    assert len(text) == 1 and next(iter(text)).startswith("gfx950:")
    ctx2 = hgx.Context(0)
    ctx2.tune_load(cache)
    got = ctx2.tuning()
    for k in ("expand_lanes", "expand_grid", "join_team", "join_grid"):
        assert got[k] == chosen[k]
    os.environ["HGX_TUNE_CACHE"] = cache
    try:
        ctx3 = hgx.Context(0)
        assert cache in ctx3.tuning()["source"]
        ctx3.close()
    finally:
        del os.environ["HGX_TUNE_CACHE"]
    with pytest.raises(hgx.HGXError):
        ctx2.tune_load(str(tmp_path / "missing.json"))
    ctx2.close()
    ctx.close()
