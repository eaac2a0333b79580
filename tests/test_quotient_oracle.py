"""CPU: the oracle's restatement of quotient exploration and of the raw unfolding beneath it.

Pinned two ways. The golden matrix's quotient rows (reference/golden_matrix.txt, '<case> Full
<mode> 1'): the number of exact classes, the event count under each identity mode, causal triples,
causal pairs and branchial pairs, for every corpus case. And what the mathematics fixes: the
replay of the class expansions is the full multiway system, so its counts (events, triples,
pairs, pairs and triples after transitive reduction, branchial pairs) equal the C oracle's full
evolution on every case -- rule sets with automorphic states, several rules, several roots.
"""
import json
import os
import random

import pytest

from hypergraphrewritingengine_amd import workloads as W
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from oracle import pyoracle as po
from oracle import quotient as Q

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "golden_matrix_none.json")))["cases"]
QROWS = json.load(open(os.path.join(HERE, "golden", "golden_matrix_quotient.json")))["rows"]
AUTO = Q.K_IN | Q.K_OUT | Q.K_STEP | Q.K_CONS | Q.K_PROD
KEYS = {"None": 0, "Full": Q.K_IN | Q.K_OUT, "Automatic": AUTO}


def _rules(c):
    return [rule_from_lists(r["lhs"], r["rhs"], k) for k, r in enumerate(c["rules"])]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_quotient_rows_of_the_golden_matrix(case):
    rules = _rules(case)
    x = Q.explore(rules, [case["init"]], case["steps"])
    u = Q.unfold(x, case["steps"])
    for mode, keys in KEYS.items():
        want = QROWS[f"{case['name']} {mode}"]
        events = u["raw_events"] if not keys else Q.identity_events(x, keys)
        got = [Q.classes(x), events, u["causal_triples"], u["causal_pairs"], u["branchial_pairs"]]
        assert got == want, (case["name"], mode)


def _full(rules, roots, steps):
    o = po.evolve(Request(rules, roots, steps, flags=0x1 | 0x2 | 0x100, multi_root=True))
    c = o.counts
    return {k: c[k] if k != "raw_events" else c["events"] for k in
            ("raw_events", "causal_triples", "causal_pairs", "causal_pairs_tr", "causal_triples_tr",
             "branchial_pairs")}


RULESETS = [
    [([[0, 1]], [[0, 2], [2, 1]])],
    [([[0, 1], [0, 2]], [[0, 1], [0, 3], [1, 3], [2, 3]])],
    [([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]]), ([[0, 1]], [[1, 0]])],
    [([[0, 1], [2, 3]], [[0, 2], [1, 3]])],                         # disconnected LHS
    [([[0, 1, 2]], [[0, 1], [1, 2], [2, 0]]), ([[0, 1], [1, 2]], [[0, 1, 2]])],
    [([[0, 1], [1, 2], [2, 0]], [[0, 1], [1, 2], [2, 3], [3, 0]])],
]
ROOTS = [[[0, 1], [1, 2], [2, 0]], [[0, 0], [0, 0]], [[0, 1], [0, 2], [0, 3]], [[0, 1, 2], [2, 3, 4]]]


@pytest.mark.parametrize("k", range(len(RULESETS)))
@pytest.mark.parametrize("r", range(len(ROOTS)))
def test_unfolding_is_the_full_multiway_system(k, r):
    rules = [rule_from_lists(l, h, i) for i, (l, h) in enumerate(RULESETS[k])]
    steps = 3
    x = Q.explore(rules, [ROOTS[r]], steps)
    u = Q.unfold(x, steps)
    full = _full(rules, [ROOTS[r]], steps)
    assert {k2: u[k2] for k2 in full} == full


@pytest.mark.parametrize("i", range(8))
def test_unfolding_of_random_runs(i):
    req = W.batched_run(i)
    rnd = random.Random(i)
    init = [list(e) for e in req.init[: rnd.randint(3, 6)]]
    steps = 3
    if po.evolve(Request(req.rules, [init], steps, flags=0x0, multi_root=True)).counts["raw_states"] > 6000:
        pytest.skip("full evolution too large for the brute-force replay")
    x = Q.explore(req.rules, [init], steps)
    assert {k: Q.unfold(x, steps)[k] for k in _full(req.rules, [init], steps)} == _full(req.rules, [init], steps)


@pytest.mark.parametrize("merge", [False, True])
def test_isomorphic_roots(merge):
    """Two isomorphic roots: both are expanded unless they merge; the unfolding replays both
    either way (every root seeds an instance, hypergraph.cpp:977-1003), so it equals the full
    evolution of both roots."""
    rules = [rule_from_lists([[0, 1], [0, 2]], [[0, 1], [0, 3], [1, 3], [2, 3]], 0)]
    roots = [[[0, 1], [0, 2]], [[5, 6], [5, 7]]]
    x = Q.explore(rules, roots, 3, merge_roots=merge)
    expanded_roots = sum(x.expanded[i] for i in range(2))
    assert expanded_roots == (1 if merge else 2)
    assert Q.unfold(x, 3)["raw_events"] == _full(rules, roots, 3)["raw_events"]
    assert {k: Q.unfold(x, 3)[k] for k in _full(rules, roots, 3)} == _full(rules, roots, 3)


def test_explored_states_are_one_per_class_and_minimal_depth():
    """Every class of the full evolution is met, at its first step, and expanded once."""
    rules = [rule_from_lists([[0, 1], [0, 2]], [[0, 1], [0, 3], [1, 3], [2, 3]], 0)]
    import hypergraphrewritingengine_amd as hgx
    steps = 4
    x = Q.explore(rules, [[[0, 1], [0, 2]]], steps)
    o = po.evolve(Request(rules, [[[0, 1], [0, 2]]], steps, flags=0x8, multi_root=True), with_edges=True)
    so, eo, vv = o.state_edges
    full = [[list(map(int, vv[eo[e]:eo[e + 1]])) for e in range(int(so[s]), int(so[s + 1]))]
            for s in range(len(so) - 1)]
    first_step = {}
    for h, s in zip((int(h) for h in hgx.ir_hash(full)), o.state_step):
        first_step.setdefault(h, int(s))
    mine = {}
    for st in x.states:
        mine.setdefault(st[2], st[1])
    assert mine == first_step
    expanded = [x.states[i][2] for i in range(len(x.states)) if x.expanded[i]]
    assert len(expanded) == len(set(expanded))
    assert set(expanded) == {h for h, s in first_step.items() if s < steps}


def test_config2_exact_classes_and_the_wl_collisions():
    """Config 2 (6 steps) has 26,673 exact classes where the WL hash sees 26,661: it merges 12
    pairs of states. Each merged pair is provably non-isomorphic -- the multisets of edge
    multiplicities differ -- so exploration by exact class must expand both."""
    import collections
    req = W.config2(6)
    x = Q.explore(req.rules, [req.init], 6)
    assert Q.classes(x) == 26673 and len(x.events) == 66476
    by_wl = collections.defaultdict(dict)
    for st in x.states:
        by_wl[po.wl_hash(st[0])].setdefault(st[2], st[0])
    assert len(by_wl) == 26661
    merged = [v for v in by_wl.values() if len(v) > 1]
    assert len(merged) == 12

    def multiplicities(state):
        return sorted(collections.Counter(tuple(e) for e in state).values())
    for v in merged:
        a, b = list(v.values())
        assert multiplicities(a) != multiplicities(b)
