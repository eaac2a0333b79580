"""CPU tests that pin the oracle (oracle/hgx_oracle.c) before anything is compared with it.

Pins, in order of strength:
  1. reference/golden_matrix.txt 'None None' rows (17 corpus cases): raw states, events, causal
     triples, causal pairs, branchial pairs and the WL fingerprint (sum of mix64 over the
     distinct WL hashes -- a value-level pin of the WL core on every reachable state).
  2. Reference-engine golden vectors (tests/golden/ref_vectors.json, written by
     tests/golden/make_ref_vectors.py from the reference engine built by oracle/Makefile):
     digests of the normalised state-hash set and event / causal / branchial multisets.
  3. Closed forms and brute force: per-step event counts of the two evolution workloads
     (brute-force enumeration), match counts against itertools enumeration, transitive
     reduction against brute-force reachability, branchial against a direct pair scan.
  4. Invariants: WL is invariant under vertex relabelling and edge order when no unary edge
     truncates the initial colour; an empty state hashes to EMPTY_STATE_CANONICAL_HASH.
"""
import itertools
import json
import os
import random

import numpy as np
import pytest

from hypergraphrewritingengine_amd import workloads as W
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden_matrix_none.json")))


def _case_request(c, flags=0x3):
    rules = [rule_from_lists(r["lhs"], r["rhs"], i) for i, r in enumerate(c["rules"])]
    return Request(rules, c["init"], c["steps"], flags=flags)


@pytest.mark.parametrize("case", GOLD["cases"], ids=[c["name"] for c in GOLD["cases"]])
def test_golden_matrix_none_rows(case):
    res = po.evolve(_case_request(case))
    for k, v in case["expect"].items():
        assert res.counts[k] == v, (k, res.counts[k], v)


def test_config1_events_per_step():
    # brute-force enumeration of config 1: 3, 15, 117, 1329, 20739 events per step
    res = po.evolve(W.config1())
    per_step = np.bincount(res.events[:, 3], minlength=6)[1:]
    assert per_step.tolist() == [3, 15, 117, 1329, 20739]
    assert res.counts["raw_states"] == 22204


def test_config2_events_per_step_prefix():
    # brute-force enumeration of config 2: 2, 24, 408, 9504, 280080 (, 10019568) events per step
    res = po.evolve(W.config2(steps=5))
    per_step = np.bincount(res.events[:, 3], minlength=6)[1:]
    assert per_step.tolist() == [2, 24, 408, 9504, 280080]


def _brute_matches(lhs, edges):
    """Every ordered edge-injective assignment with a consistent (non-injective) binding."""
    out = []
    for tup in itertools.permutations(range(len(edges)), len(lhs)):
        bind = {}
        ok = True
        for pe, ei in zip(lhs, tup):
            e = edges[ei]
            if len(e) != len(pe):
                ok = False
                break
            for var, v in zip(pe, e):
                if bind.setdefault(var, v) != v:
                    ok = False
                    break
            if not ok:
                break
        if ok:
            out.append(tup)
    return out


@pytest.mark.parametrize("seed", range(12))
def test_match_against_brute_force(seed):
    rnd = random.Random(seed)
    lhs_pool = [
        [[0, 1], [1, 2]], [[0, 1], [0, 2]], [[0, 0], [0, 1]], [[0, 1], [1, 2], [2, 0]],
        [[0, 1, 2], [2, 3, 4]], [[0], [0, 1]], [[0, 1], [2, 3]], [[0, 1], [1, 0]],
    ]
    lhs = lhs_pool[seed % len(lhs_pool)]
    arities = sorted({len(e) for e in lhs})
    edges = [[rnd.randrange(5) for _ in range(rnd.choice(arities))] for _ in range(rnd.randrange(3, 9))]
    rule = rule_from_lists(lhs, [[0]])
    n, got = po.match([rule], 0, edges, want_edges=True, cap=100000)
    want = _brute_matches(lhs, edges)
    assert n == len(want)
    assert sorted(map(tuple, got.tolist())) == sorted(want)


def test_match_big_state_uses_index_and_agrees():
    # > 64 edges switches the oracle to its incidence index; the count must not change
    rule, edges = W.config3a(n_edges=400, n_vertices=60)
    n_idx = po.match([rule], 0, edges)
    want = sum(1 for i, a in enumerate(edges) for j, b in enumerate(edges) if i != j and b[0] == a[2])
    assert n_idx == want


def test_match_order_follows_reference_rule():
    # pattern.cpp:43-75: seed = max self-constraint score, then most shared variables
    r = rule_from_lists([[0, 1], [1, 1], [2, 0]], [[0]])
    assert po.match_order([r], 0) == [1, 0, 2]
    r = rule_from_lists([[0, 1], [2, 3], [1, 2]], [[0]])
    assert po.match_order([r], 0) == [2, 0, 1]


def _relabel(edges, rnd):
    verts = sorted({v for e in edges for v in e})
    perm = verts[:]
    rnd.shuffle(perm)
    m = dict(zip(verts, [p * 7 + 3 for p in perm]))
    out = [[m[v] for v in e] for e in edges]
    rnd.shuffle(out)
    return out


@pytest.mark.parametrize("seed", range(20))
def test_wl_invariant_under_relabel_and_edge_order(seed):
    rnd = random.Random(100 + seed)
    edges = [[rnd.randrange(6) for _ in range(rnd.choice([2, 3]))] for _ in range(rnd.randrange(1, 10))]
    h = po.wl_hash(edges)
    for _ in range(4):
        assert po.wl_hash(_relabel(edges, rnd)) == h


def test_wl_special_cases():
    assert po.wl_hash([]) == 0x9E3779B97F4A7C15
    # different global structure, same local structure: 6-cycle vs two triangles
    c6 = [[i, (i + 1) % 6] for i in range(6)]
    t2 = [[0, 1], [1, 2], [2, 0], [3, 4], [4, 5], [5, 3]]
    assert po.wl_hash(c6) != po.wl_hash(t2)
    # reversing every edge of a directed 2-path gives an isomorphic state
    assert po.wl_hash([[0, 1], [1, 2]]) == po.wl_hash([[1, 0], [2, 1]])
    # orientation is structure: a 2-path is not a 2-star
    assert po.wl_hash([[0, 1], [1, 2]]) != po.wl_hash([[0, 1], [0, 2]])


def _brute_tr(pairs):
    adj = {}
    for p, c in pairs:
        adj.setdefault(p, set()).add(c)

    def reach(a, b, skip):
        stack = [x for x in adj.get(a, ()) if (a, x) != skip]
        seen = set()
        while stack:
            x = stack.pop()
            if x == b:
                return True
            if x in seen:
                continue
            seen.add(x)
            stack.extend(adj.get(x, ()))
        return False

    return {(p, c) for (p, c) in pairs if not reach(p, c, (p, c))}


@pytest.mark.parametrize("name", ["cycle4-automorphic", "disconnected-lhs", "reductive-2to1"])
def test_transitive_reduction_brute_force(name):
    case = next(c for c in GOLD["cases"] if c["name"] == name)
    res = po.evolve(_case_request(case, flags=0x103))
    pairs = [tuple(map(int, p)) for p in res.causal]
    kept = {tuple(map(int, p)) for p, k in zip(res.causal, res.causal_tr) if k}
    assert kept == _brute_tr(pairs)


def test_branchial_direct_scan():
    """Branchial pairs rebuilt from each parent's match list (positions overlap) agree with
    the oracle's relation: same input state, intersecting consumed sets, one pair each."""
    req = W.config1(steps=3)
    res = po.evolve(req, with_edges=True)
    so, eo, vv = res.state_edges
    ev = res.events
    want = set()
    first_event = {}
    for e in range(len(ev)):
        first_event.setdefault(int(ev[e][0]), e)
    for parent, e0 in first_event.items():
        edges = [vv[eo[k]:eo[k + 1]].tolist() for k in range(so[parent], so[parent + 1])]
        n, m = po.match(req.rules, 0, edges, want_edges=True, cap=10**6)
        for i in range(n):
            for j in range(i + 1, n):
                if set(m[i].tolist()) & set(m[j].tolist()):
                    want.add((e0 + i, e0 + j))
    got = {tuple(map(int, p)) for p in res.branchial}
    assert got == want and res.counts["branchial_pairs"] == len(want)


def test_state_cap_is_a_step_prefix():
    req = W.config1(steps=5)
    req.state_cap = 1000
    res = po.evolve(req)
    assert res.counts["truncated"] == 1
    assert res.counts["steps_done"] == 3 and res.counts["raw_states"] == 1 + 3 + 15 + 117


REF_VECTORS = os.path.join(HERE, "golden", "ref_vectors.json")


@pytest.mark.skipif(not os.path.exists(REF_VECTORS), reason="reference vectors not generated")
def test_reference_vectors():
    from tests.golden.make_ref_vectors import digest_oracle, load_cases

    doc = json.load(open(REF_VECTORS))
    cases = dict(load_cases())
    for entry in doc["cases"]:
        req = cases[entry["name"]]
        tr = bool(entry["tr"])
        if tr:
            req.flags |= 0x100
        got = digest_oracle(po.evolve(req), tr)
        assert got == entry["digest"], entry["name"]


def test_oracle_reproduces_fullsize_fixtures_sample():
    """The committed full-size fixtures (whose counts the reference engine confirmed) are the
    oracle's: a sample of batched runs and the small served draws recomputed here."""
    import json
    import os

    from hypergraphrewritingengine_amd import workloads as W
    from tests.parity import arrays_digest

    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for name, make, ids in (("batched4096.json", W.batched_run, range(0, 4096, 97)),
                            ("served_mix.json", W.served_run, None)):
        fx = json.load(open(os.path.join(gold, name)))
        pairs = list(zip(fx["ids"], fx["runs"])) if ids is None else [(i, fx["runs"][i]) for i in ids]
        for i, row in pairs:
            req = make(i, cap=fx["cap"], flags=fx["flags"])
            if ids is None and len(req.init) > 400:
                continue                      # the large served draws take seconds each
            r = po.evolve(req)
            want = dict(zip(fx["columns"], row))
            for k in fx["columns"][:-1]:
                assert r.counts[k] == want[k], (name, i, k)
            d = arrays_digest(r.state_hash, r.state_step, r.state_creator, r.events, r.causal,
                              r.causal_tr, r.branchial)
            assert d == want["digest"], (name, i)
