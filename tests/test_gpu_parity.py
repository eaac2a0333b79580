"""GPU parity: the HIP path through the C ABI against the oracle, on the same seeded inputs.

Bit-exact everywhere (integer arithmetic): per-state WL hashes, event numbering, causal pairs
with their transitive-reduction tags, branchial pairs, and every count. At full workload sizes
the checks are size-independent properties (per-step event counts from brute-force
enumeration, the reference-published totals, mix independence).
"""
import json
import os
import random

import numpy as np
import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd import workloads as W
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from oracle import pyoracle as po
from tests.parity import assert_exact, normalized_gpu, normalized_oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden_matrix_none.json")))
REL = 0x1 | 0x2 | 0x4          # causal + branchial + materialised relations
TR = 0x100


def _case(c, flags):
    rules = [rule_from_lists(r["lhs"], r["rhs"], i) for i, r in enumerate(c["rules"])]
    return Request(rules, c["init"], c["steps"], flags=flags)


def _random_states(seed, n):
    rnd = random.Random(seed)
    out = []
    for _ in range(n):
        kind = rnd.randrange(4)
        ne = rnd.randrange(0, 40)
        nv = rnd.randrange(1, 30)
        if kind == 0:
            st = [[rnd.randrange(nv), rnd.randrange(nv)] for _ in range(ne)]
        elif kind == 1:
            st = [[rnd.randrange(nv) for _ in range(rnd.choice([1, 2, 3]))] for _ in range(ne)]
        elif kind == 2:
            st = [[rnd.randrange(4) for _ in range(rnd.choice([3, 4, 5]))] for _ in range(ne)]
        else:
            st = [[rnd.randrange(nv)] * rnd.choice([1, 2, 3]) for _ in range(ne)]
        out.append(st)
    return out


def test_wl_hash_batch_matches_oracle(ctx):
    states = _random_states(7, 3000)
    got = ctx.wl_hash(states)
    want = np.array([po.wl_hash(s) for s in states], dtype=np.uint64)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"{len(bad)} mismatches, first state {states[bad[0]]}"


def test_wl_hash_large_states(ctx):
    rnd = random.Random(11)
    states = [[[rnd.randrange(n), rnd.randrange(n)] for _ in range(m)]
              for (m, n) in [(200, 150), (500, 300), (1000, 900), (64, 8), (300, 2)]]
    got = ctx.wl_hash(states)
    want = np.array([po.wl_hash(s) for s in states], dtype=np.uint64)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("tr", [False, True])
@pytest.mark.parametrize("case", GOLD["cases"], ids=[c["name"] for c in GOLD["cases"]])
def test_golden_corpus_exact(ctx, case, tr):
    flags = REL | (TR if tr else 0)
    res = ctx.evolve(_case(case, flags))
    o = po.evolve(_case(case, flags))
    assert_exact(res, o)
    if not tr:
        for k, v in case["expect"].items():
            assert res.counts[k] == v, k


@pytest.mark.parametrize("tr", [False, True])
def test_config1_exact(ctx, tr):
    req = W.config1(flags=REL | (TR if tr else 0))
    res = ctx.evolve(req)
    o = po.evolve(req)
    assert_exact(res, o)
    assert normalized_gpu(res, tr) == normalized_oracle(o, tr)


def test_config2_prefix_exact(ctx):
    req = W.config2(steps=4, flags=REL | TR)
    assert_exact(ctx.evolve(req), po.evolve(req))


def test_config2_full_counts(ctx):
    # 6 steps: 10,309,587 raw states (2, 24, 408, 9504, 280080, 10019568 events per step)
    res = ctx.evolve(W.config2(steps=6, flags=0x3))
    assert res.counts["raw_states"] == 10_309_587
    assert res.counts["events"] == 10_309_586
    o = po.evolve(W.config2(steps=5, flags=0x3))
    # the first five generations are a prefix: canonical count cannot shrink
    assert res.counts["canonical_states"] >= o.counts["canonical_states"]


def test_multi_root_exact(ctx):
    req = Request([rule_from_lists([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])],
                  [[[0, 1], [1, 2]], [[5, 5], [5, 6]], [[1, 2], [2, 3]], []], 3,
                  flags=REL | TR, multi_root=True)
    assert_exact(ctx.evolve(req), po.evolve(req))


def test_batched_runs_exact_and_mix_independent(ctx):
    reqs = [W.batched_run(i, cap=2048, flags=REL | TR) for i in range(48)]
    batch = ctx.evolve_batch(reqs)
    for i, (r, req) in enumerate(zip(batch, reqs)):
        assert_exact(r, po.evolve(req))
    # the same request alone, and in a different mix, gives the same result
    for i in (0, 5, 17):
        alone = ctx.evolve(reqs[i])
        assert alone.counts == batch[i].counts
        assert np.array_equal(alone.states()[0], batch[i].states()[0])
    other = ctx.evolve_batch([reqs[17], reqs[5], W.config1(flags=REL), reqs[0]])
    assert other[0].counts == batch[17].counts and other[3].counts == batch[0].counts
    assert np.array_equal(other[1].events(), batch[5].events())


def test_state_cap_contract(ctx):
    req = W.config1(steps=5, flags=0x3)
    req.state_cap = 1000
    res = ctx.evolve(req)
    assert res.counts["truncated"] == 1 and res.counts["steps_done"] == 3
    assert res.counts["raw_states"] == 136


def test_counts_only_mode_matches_relations_mode(ctx):
    a = ctx.evolve(W.config1(flags=0x3)).counts
    b = ctx.evolve(W.config1(flags=REL)).counts
    assert a == b


def test_edge_cases(ctx):
    # zero steps, empty initial state, a rule that never matches, an RHS with no edges
    r = rule_from_lists([[0, 1]], [])
    res = ctx.evolve(Request([r], [[1, 2], [2, 3]], 2, flags=REL))
    o = po.evolve(Request([r], [[1, 2], [2, 3]], 2, flags=REL))
    assert_exact(res, o)
    res = ctx.evolve(Request([r], [[1, 2]], 0, flags=REL))
    assert res.counts["raw_states"] == 1 and res.counts["events"] == 0
    res = ctx.evolve(Request([rule_from_lists([[0, 0, 0]], [[0]])], [[1, 2]], 3, flags=REL))
    assert res.counts["events"] == 0
    with pytest.raises(hgx.HGXError):
        ctx.evolve(Request([r], [[1, 2]], 99))


def test_served_results_equal_alone_and_in_the_mix(ctx):
    """Continuous batching: requests arriving while others are on the device are admitted into
    the next generation; every result equals the same request evolved alone (and the oracle)."""
    reqs = [W.served_run(i, cap=1024, flags=REL | TR) for i in range(24)]
    reqs[3] = W.config1(steps=4, flags=REL | TR)
    ids = {}
    got = {}
    # staggered arrivals: a few requests per generation step
    pending = list(range(len(reqs)))
    while pending or ids:
        for _ in range(3):
            if pending:
                i = pending.pop(0)
                ids[ctx.submit(reqs[i])] = i
        ctx.serve_step()
        for rid, res in ctx.drain().items():
            got[ids.pop(rid)] = res
    assert sorted(got) == list(range(len(reqs)))
    for i, req in enumerate(reqs):
        assert_exact(got[i], po.evolve(req))
        alone = ctx.evolve(req)
        assert alone.counts == got[i].counts
    # wait() drives the loop by itself
    rid = ctx.submit(reqs[5])
    assert ctx.wait(rid).counts == got[5].counts


def test_prepared_batch_equals_host_batch(ctx):
    """hgx_batch_prepare stages rules and initial states in device memory once; evolving the
    prepared batch (twice) gives what hgx_evolve_batch gives on the same requests."""
    reqs = [W.batched_run(i, cap=1024, flags=REL | TR) for i in range(40)]
    reqs += [W.config1(steps=4, flags=REL),
             Request([rule_from_lists([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])],
                     [[[0, 1], [1, 2]], [[5, 5], [5, 6]], []], 3, flags=REL | TR, multi_root=True),
             Request([rule_from_lists([[0, 1, 2]], [[0, 1], [1, 2, 3]])], [[1, 2, 3], [3, 4, 5]], 3,
                     flags=REL)]
    host = ctx.evolve_batch(reqs)
    prep = ctx.prepare(reqs)
    for _ in range(2):
        dev = ctx.evolve_batch(prep)
        for a, b in zip(dev, host):
            assert a.counts == b.counts
            for x, y in zip(a.states(), b.states()):
                assert np.array_equal(x, y)
            assert np.array_equal(a.events(), b.events())
    for i in (0, 40, 41, 42):
        assert_exact(dev[i], po.evolve(reqs[i]))
    prep.close()


@pytest.mark.parametrize("seed,n,nv,steps,cap", [(0, 70, 40, 2, 0), (1, 110, 40, 2, 4000), (2, 150, 40, 1, 0),
                                                (3, 58, 80, 3, 0), (4, 125, 40, 2, 0)])
def test_branchial_large_parents_exact(ctx, seed, n, nv, steps, cap):
    """Branchial pairs of parents past one 64-bit consumed set: LDS bitsets with two-word sets up
    to 128 edges, the per-edge match-list path above, and runs whose states cross 64 and 128
    edges between generations. Exact pairs against the oracle, with hyperedges, repeated edges
    and a two-edge rule whose matches share both edges in reverse order."""
    rnd = random.Random(seed)
    init = [[rnd.randrange(nv), rnd.randrange(nv)] for _ in range(n)]
    init += [[3, 4, 5], [5, 4, 3], [7, 7]]
    rules = [rule_from_lists([[0, 1], [0, 2]], [[0, 2], [0, 3], [1, 3], [2, 3]], 0),
             rule_from_lists([[0, 1, 2]], [[0, 1], [1, 2]], 1)]
    req = Request(rules, init, steps, state_cap=cap, flags=REL | TR)
    res = ctx.evolve(req)
    assert res.counts["branchial_pairs"] > 0
    assert_exact(res, po.evolve(req))


def test_result_list_views(ctx):
    """A batch's ResultList: length, negative and slice indexing, one wrapper per request, and
    results taken from it stay readable after the list itself is gone."""
    import gc
    reqs = [W.batched_run(i, cap=512, flags=REL) for i in range(6)]
    res = ctx.evolve_batch(reqs)
    assert len(res) == 6
    assert res[-1] is res[5] and res[0] is res[0]
    assert [r.counts for r in res[1:4]] == [res[k].counts for k in (1, 2, 3)]
    with pytest.raises(IndexError):
        res[6]
    assert [int(x) for x in res.counts()["raw_states"]] == [r.counts["raw_states"] for r in res]
    keep = res[2]
    want = ctx.evolve(reqs[2])
    del res
    gc.collect()
    assert keep.counts == want.counts
    assert np.array_equal(keep.states()[0], want.states()[0])
