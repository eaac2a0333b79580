"""Comparison helpers shared by the parity tests: the reference differential suite's
normalisation (gpu/tests/test_gpu_vs_cpu_differential.cpp:64-160) and exact array checks."""
from __future__ import annotations

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
