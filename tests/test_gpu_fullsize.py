"""Full-size parity of the benchmark workloads against committed fixtures.

The fixtures (tests/golden/make_config_fixtures.py) hold the oracle's counts and array digests,
cross-checked in the build container against the reference engine's counts at the same
generation count:

  * the single large evolution (rule {{x,y},{x,z}}->{{x,z},{x,w},{y,w},{z,w}} from {{1,1},{1,1}},
    6 steps): every count and the WL fingerprint, transitive reduction off and on;
  * all 4,096 runs of the batched workload at the 4,096-state cap, from a prepared batch: every
    count, and per-state hashes/steps/creators, events, causal pairs with reduction tags and
    branchial pairs through their digest;
  * 32 draws of the served workload at the 4,096-state cap, 12 of them with 1,500-2,048 initial
    edges and 6-8 steps, submitted together through the served path (continuous batching), and
    the largest ones again alone.
"""
import json
import os

import numpy as np
import pytest

from hypergraphrewritingengine_amd import workloads as W
from tests.parity import gpu_digest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return json.load(open(os.path.join(GOLD, name)))


def _check_row(name, counts, row, columns, digest=None):
    want = dict(zip(columns, row))
    for k in columns[:-1]:
        assert counts[k] == want[k], f"{name}: {k} = {counts[k]}, fixture {want[k]}"
    if digest is not None:
        assert digest == want["digest"], f"{name}: array digest differs"


def test_config2_full_counts_fixture(ctx):
    fx = _load("config2_counts.json")["counts"]
    for key, flags in (("plain", 0x3), ("tr", 0x103)):
        res = ctx.evolve(W.config2(steps=6, flags=flags))
        for k, v in fx[key].items():
            assert res.counts[k] == v, f"config2 ({key}): {k} = {res.counts[k]}, fixture {v}"


def test_batched_4096_runs_exact(ctx):
    fx = _load("batched4096.json")
    cols, rows = fx["columns"], fx["runs"]
    assert len(rows) == 4096
    reqs = [W.batched_run(i, cap=fx["cap"], flags=fx["flags"]) for i in range(4096)]
    pb = ctx.prepare(reqs)
    res = ctx.evolve_batch(pb)
    bad = []
    for i in range(4096):
        r = res[i]
        try:
            _check_row(f"run {i}", r.counts, rows[i], cols, gpu_digest(r))
        except AssertionError as e:
            bad.append(str(e))
        r.close()
    assert not bad, f"{len(bad)} runs differ, first: {bad[0]}"
    pb.close()


def test_served_mix_at_cap_with_largest_draws(ctx):
    fx = _load("served_mix.json")
    cols, ids, rows = fx["columns"], fx["ids"], fx["runs"]
    reqs = {i: W.served_run(i, cap=fx["cap"], flags=fx["flags"]) for i in ids}
    assert sum(len(reqs[i].init) >= 1500 for i in ids) >= 12
    rid = {}
    # arrivals in three waves: the later ones join while the first are on the device
    waves = [ids[0::3], ids[1::3], ids[2::3]]
    done = {}
    for w in waves:
        for i in w:
            rid[ctx.submit(reqs[i])] = i
        ctx.serve_step()
        for k, r in ctx.drain().items():
            done[rid[k]] = r
    while len(done) < len(ids):
        ctx.serve_step()
        for k, r in ctx.drain().items():
            done[rid[k]] = r
    for i, row in zip(ids, rows):
        _check_row(f"served draw {i}", done[i].counts, row, cols, gpu_digest(done[i]))
    # the largest draws alone give what they gave in the mix
    for i in fx["large"][:3]:
        alone = ctx.evolve(reqs[i])
        assert alone.counts == done[i].counts
        assert np.array_equal(alone.states()[0], done[i].states()[0])
