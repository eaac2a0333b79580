"""The large-state join (hgx_match_count): device CSR index + seeded join over one big state,
against the oracle's indexed DFS -- exact match lists, in the same order."""
import random

import numpy as np
import pytest

from hypergraphrewritingengine_amd import workloads as W
from hypergraphrewritingengine_amd.rules import rule_from_lists
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def test_ternary_match_join_full_size(ctx):
    rule, edges = W.config3a()                       # 50,000 ternary edges, 10,000 vertices
    n, m, ms = ctx.match([rule], 0, edges, want_edges=True, cap=10**7)
    on, om = po.match([rule], 0, edges, want_edges=True, cap=10**7)
    print(f"[bigjoin] ternary: {n} matches, {ms:.3f} ms")
    assert n == on
    assert np.array_equal(m, om)
    # closed form: sum over e1 of |{e2 != e1 : e2[0] == e1[2]}|
    first = np.bincount(np.array([e[0] for e in edges]), minlength=10_000)
    want = sum(int(first[e[2]]) - (1 if e[0] == e[2] else 0) for e in edges)
    assert n == want


def test_triangle_match_join_full_size(ctx):
    rule, edges = W.config3b()                       # 200,000 binary edges, 20,000 vertices
    n, m, ms = ctx.match([rule], 0, edges, want_edges=True, cap=10**7)
    on, om = po.match([rule], 0, edges, want_edges=True, cap=10**7)
    print(f"[bigjoin] triangle: {n} matches, {ms:.3f} ms")
    assert n == on
    assert np.array_equal(m, om)
    # every match is a directed 3-cycle x->y->z->x
    e = np.array(edges, dtype=np.int64)
    for a, b, c in m[:200]:
        assert e[a][1] == e[b][0] and e[b][1] == e[c][0] and e[c][1] == e[a][0]


@pytest.mark.parametrize("seed", range(8))
def test_random_rules_small_states(ctx, seed):
    rnd = random.Random(seed)
    pool = [[[0, 1], [1, 2]], [[0, 1], [0, 2]], [[0, 0], [0, 1]], [[0, 1], [1, 2], [2, 0]],
            [[0, 1, 2], [2, 3, 4]], [[0], [0, 1]], [[0, 1], [2, 3]], [[0, 1], [1, 0]]]
    lhs = pool[seed % len(pool)]
    arities = sorted({len(x) for x in lhs})
    edges = [[rnd.randrange(30) for _ in range(rnd.choice(arities))] for _ in range(rnd.randrange(50, 400))]
    rule = rule_from_lists(lhs, [[0]])
    n, m, _ = ctx.match([rule], 0, edges, want_edges=True, cap=10**6)
    on, om = po.match([rule], 0, edges, want_edges=True, cap=10**6)
    assert n == on and np.array_equal(m, om)


def test_single_edge_rule_empty_state_and_disconnected_lhs(ctx):
    rnd = random.Random(11)
    edges = [[rnd.randrange(50), rnd.randrange(50)] for _ in range(3000)]
    for lhs in ([[0, 1]], [[0, 0]], [[0, 1], [2, 3]], [[0, 1], [1, 1]]):
        rule = rule_from_lists(lhs, [[0]])
        n, m, _ = ctx.match([rule], 0, edges, want_edges=True, cap=10**7)
        on, om = po.match([rule], 0, edges, want_edges=True, cap=10**7)
        assert n == on and np.array_equal(m, om), lhs
    rule = rule_from_lists([[0, 1], [1, 2]], [[0]])
    n, m, _ = ctx.match([rule], 0, [], want_edges=True, cap=10)
    assert n == 0


def test_deep_rule_and_cap_prefix(ctx):
    # a four-edge path over a random digraph: three breadth-first depths after the seeds
    rnd = random.Random(5)
    edges = [[rnd.randrange(400), rnd.randrange(400)] for _ in range(4000)]
    rule = rule_from_lists([[0, 1], [1, 2], [2, 3], [3, 4]], [[0, 4]])
    n, m, ms = ctx.match([rule], 0, edges, want_edges=True, cap=10**7)
    on, om = po.match([rule], 0, edges, want_edges=True, cap=10**7)
    assert n == on and np.array_equal(m, om)
    k = n // 3
    n2, m2, _ = ctx.match([rule], 0, edges, want_edges=True, cap=k)
    assert n2 == n and np.array_equal(m2, om[:k])
    # mixed arities with a repeated variable and a ternary pivot
    edges3 = [[rnd.randrange(60) for _ in range(rnd.choice([2, 3]))] for _ in range(5000)]
    rule3 = rule_from_lists([[0, 1, 1], [1, 2], [2, 0, 3]], [[0]])
    n3, m3, _ = ctx.match([rule3], 0, edges3, want_edges=True, cap=10**7)
    on3, om3 = po.match([rule3], 0, edges3, want_edges=True, cap=10**7)
    assert n3 == on3 and np.array_equal(m3, om3)
