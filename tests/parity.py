"""Comparison helpers shared by the parity tests: the reference differential suite's
normalisation (gpu/tests/test_gpu_vs_cpu_differential.cpp:64-160) and exact array checks."""
from __future__ import annotations

import hashlib

import numpy as np

from oracle import pyoracle as po


def normalized_gpu(res, tr=False):
    h, s, c = res.states()
    ev = res.events()
    sh = {i: int(x) for i, x in enumerate(h)}
    evl = [(e, *map(int, ev[e])) for e in range(len(ev))]
    cp, ct = res.causal() if res.counts["causal_pairs"] else (np.zeros((0, 2), np.uint32), np.zeros(0, np.uint8))
    if tr:
        cp = cp[ct.astype(bool)]
    br = res.branchial() if res.counts["branchial_pairs"] else np.zeros((0, 2), np.uint32)
    return po.normalize(sh, evl, [tuple(map(int, p)) for p in cp], [tuple(map(int, p)) for p in br])


def normalized_oracle(o, tr=False):
    sh = {i: int(x) for i, x in enumerate(o.state_hash)}
    evl = [(e, *map(int, o.events[e])) for e in range(len(o.events))]
    cp = o.causal[o.causal_tr.astype(bool)] if tr else o.causal
    return po.normalize(sh, evl, [tuple(map(int, p)) for p in cp], [tuple(map(int, p)) for p in o.branchial])


def sorted_pairs(p):
    p = np.asarray(p, dtype=np.uint64).reshape(-1, 2)
    if len(p) == 0:
        return p
    key = (p[:, 0] << np.uint64(32)) | p[:, 1]
    return p[np.argsort(key, kind="stable")]


def assert_exact(res, o, relations=True, tr=False):
    """GPU result identical to the oracle's: counts, per-state hashes/steps/creators, events
    (same numbering), causal pairs with TR tags, branchial pairs."""
    gc = dict(res.counts)
    oc = dict(o.counts)
    assert gc == oc, f"counts differ:\n gpu {gc}\n cpu {oc}"
    h, s, c = res.states()
    assert np.array_equal(h, o.state_hash), "state hashes differ"
    assert np.array_equal(s, o.state_step), "state steps differ"
    assert np.array_equal(c, o.state_creator), "state creators differ"
    assert np.array_equal(res.events(), o.events), "events differ"
    if relations:
        if gc["causal_pairs"]:
            cp, ct = res.causal()
            ga = np.concatenate([cp.astype(np.uint64), ct.astype(np.uint64)[:, None]], axis=1)
            oa = np.concatenate([o.causal.astype(np.uint64), o.causal_tr.astype(np.uint64)[:, None]], axis=1)
            gk = sorted(map(tuple, ga.tolist()))
            ok = sorted(map(tuple, oa.tolist()))
            assert gk == ok, "causal pairs differ"
        if gc["branchial_pairs"]:
            gb = sorted_pairs(res.branchial())
            ob = sorted_pairs(o.branchial)
            assert np.array_equal(gb, ob), "branchial pairs differ"


def _rows(a, cols):
    a = np.asarray(a, dtype=np.uint64).reshape(-1, cols)
    return a[np.lexsort(a.T[::-1])] if len(a) else a


def arrays_digest(hashes, steps, creators, events, causal, causal_tr, branchial) -> str:
    """sha256 over one result's arrays (16 hex digits): per-state hashes, steps and creators
    and the event records in the deterministic numbering; the causal (producer, consumer,
    reduction tag) and branchial (smaller, larger) pairs as sorted sets of rows. The fixtures of
    tests/golden/make_config_fixtures.py hold the oracle's digests."""
    h = hashlib.sha256()
    for a, dt in ((hashes, np.uint64), (steps, np.uint32), (creators, np.uint32),
                  (events, np.uint32)):
        h.update(np.ascontiguousarray(np.asarray(a, dtype=dt)).tobytes())
    cz = np.concatenate([np.asarray(causal, np.uint64).reshape(-1, 2),
                         np.asarray(causal_tr, np.uint64).reshape(-1, 1)], axis=1)
    h.update(_rows(cz, 3).tobytes())
    h.update(_rows(branchial, 2).tobytes())
    return h.hexdigest()[:16]


def gpu_digest(res) -> str:
    """arrays_digest of a device result (needs causal, branchial and relations recorded)."""
    h, s, c = res.states()
    cnt = res.counts
    if cnt["causal_pairs"]:
        cp, ct = res.causal()
    else:
        cp, ct = np.zeros((0, 2), np.uint32), np.zeros(0, np.uint8)
    br = res.branchial() if cnt["branchial_pairs"] else np.zeros((0, 2), np.uint32)
    return arrays_digest(h, s, c, res.events(), cp, ct, br)
