"""Failure handling of the engine under a small HBM budget (host-side paths only; every kernel
runs normally): a generation that does not fit finishes its request with a partial result and
the capacity status, the request is never dropped or left spinning, and the other requests of a
served mix finish exactly as they would alone (the reference's capacity warning and partial
result, gpu/src/evolve.cu:573-700, gpu/include/hg_gpu/errors.hpp:31-69)."""
import numpy as np
import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd import abi
from hypergraphrewritingengine_amd import workloads as W
from oracle import pyoracle as po
from tests.parity import arrays_digest, gpu_digest

pytestmark = pytest.mark.gpu

BUDGET = 2 << 30            # holds five generations of the large evolution, not the sixth
FLAGS = 0x1 | 0x2 | 0x4 | 0x100


@pytest.fixture(scope="module")
def small():
    c = hgx.Context(0, budget_bytes=BUDGET)
    yield c
    c.close()


def _oracle_digest(req):
    o = po.evolve(req)
    return o.counts, arrays_digest(o.state_hash, o.state_step, o.state_creator, o.events, o.causal,
                                   o.causal_tr, o.branchial)


def _same_counts(got, want):
    for k in ("raw_states", "events", "canonical_states", "causal_triples", "causal_pairs",
              "causal_pairs_tr", "causal_triples_tr", "branchial_pairs", "fingerprint"):
        assert got[k] == want[k], (k, got[k], want[k])


def test_generation_beyond_budget_gives_partial_result(small):
    res = small.evolve(W.config2(steps=6, flags=0x3 | 0x100))
    c = res.counts
    code, why = res.status()
    assert code == abi.HGX_ERR_CAPACITY and "HBM budget" in why, (code, why)
    assert c["truncated"] == 2 and 1 <= c["steps_done"] < 6, (c, why)
    # the partial result is exactly the evolution to the generations it completed
    want = po.evolve(W.config2(steps=c["steps_done"], flags=0x3 | 0x100)).counts
    _same_counts(c, want)


def test_served_mix_survives_an_oversize_request(small):
    smalls = [W.batched_run(i, cap=4096, flags=FLAGS) for i in range(12)]
    big = W.config2(steps=6, flags=FLAGS)
    order = smalls[:6] + [big] + smalls[6:]
    ids = [small.submit(r) for r in order]
    got = {}
    guard = 0
    while len(got) < len(ids):
        small.serve_step()
        got.update(small.drain())
        guard += 1
        assert guard < 200, "the served engine stopped making progress"
    for rid, req in zip(ids, order):
        r = got[rid]
        if req is big:
            code, why = r.status()
            assert code == abi.HGX_ERR_CAPACITY and r.counts["truncated"] == 2, (code, why)
            continue
        assert r.status()[0] == abi.HGX_OK, r.status()
        oc, od = _oracle_digest(req)
        _same_counts(r.counts, oc)
        assert r.counts["steps_done"] == oc["steps_done"]
        assert gpu_digest(r) == od


def test_wait_on_oversize_request_returns(small):
    rid = small.submit(W.config2(steps=6, flags=0x3))
    r = small.wait(rid)
    assert r.counts["truncated"] == 2 and r.status()[0] == abi.HGX_ERR_CAPACITY
    # a request too large to start finishes with no states instead of being retried forever
    huge = hgx.Context(0, budget_bytes=200 << 20)
    rid = huge.submit(W.config2(steps=6, flags=0x3))
    r = huge.wait(rid)
    code, why = r.status()
    assert code == abi.HGX_ERR_CAPACITY, (code, why)
    assert r.counts["steps_done"] < 6 and r.counts["truncated"] == 2
    huge.close()


def test_batch_path_partial_and_others_complete(small):
    reqs = [W.batched_run(i, cap=4096, flags=FLAGS) for i in range(20, 28)] + \
           [W.config2(steps=6, flags=FLAGS)]
    res = small.evolve_batch(reqs)
    for req, r in zip(reqs, res):
        if req.steps == 6 and len(req.init) == 2:
            assert r.counts["truncated"] == 2 and r.status()[0] == abi.HGX_ERR_CAPACITY
        else:
            oc, od = _oracle_digest(req)
            _same_counts(r.counts, oc)
            assert gpu_digest(r) == od
