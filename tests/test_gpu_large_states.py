"""GPU: evolution of states too large for the LDS-resident join (above about 6,000 binary edges):
the join runs over per-workgroup global-memory workspaces (k_match's global variant, eight-wave
teams) and the children, as large, are rewritten and hashed in global workspaces. Every result
equals the oracle's element by element (per-state WL hashes, events, causal pairs with reduction
tags, branchial pairs), alone and in a batch with small requests. States above 65,535 edges
(16-bit edge positions in the match list) are refused, never evolved differently.
"""
import random

import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from hypergraphrewritingengine_amd import workloads as W
from oracle import pyoracle as po
from tests.parity import assert_exact

pytestmark = pytest.mark.gpu

REL = 0x1 | 0x2 | 0x4
TR = 0x100


def _state(m, n, a, seed):
    rnd = random.Random(seed)
    return [[rnd.randrange(n) for _ in range(a)] for _ in range(m)]


REVERSE = rule_from_lists([[0, 1], [1, 2], [2, 0]], [[0, 2], [2, 1], [1, 0]], 0)      # reverse a 3-cycle
SHRINK = rule_from_lists([[0, 1], [1, 2], [2, 0]], [[0, 1], [1, 2]], 0)               # drop its last edge
TERNARY = rule_from_lists([[0, 1, 2], [2, 3, 4], [4, 5, 0]], [[0, 1, 2], [2, 3, 4]], 0)

CASES = {
    # 7,000 binary edges, 2 steps: 343 states, the second generation's matches consume the
    # first's produced edges (54 causal pairs)
    "binary-7k-2steps": lambda: Request([REVERSE], _state(7000, 2500, 2, 1), 2, flags=REL | TR),
    # 30,000 ternary edges, a ternary 3-cycle LHS: 120 children of 29,999 edges
    "ternary-30k": lambda: Request([TERNARY], _state(30000, 6000, 3, 2), 1, flags=REL | TR),
    # 60,000 binary edges: six children just below the 16-bit position limit
    "binary-60k": lambda: Request([SHRINK], _state(60000, 40000, 2, 4), 1, flags=REL | TR),
}


@pytest.mark.parametrize("name", list(CASES))
def test_large_state_evolution_equals_oracle(ctx, name):
    req = CASES[name]()
    res = ctx.evolve(req)
    o = po.evolve(req)
    assert res.counts["events"] > 0
    assert_exact(res, o, tr=True)


def test_large_state_in_a_batch_with_small_requests(ctx):
    """One large request among small ones: its join launch takes global workspaces while the
    small requests' classes stay in LDS; every result equals the oracle and the request alone."""
    reqs = [W.batched_run(i, cap=1024, flags=REL | TR) for i in range(6)]
    reqs.insert(3, CASES["binary-7k-2steps"]())
    res = ctx.evolve_batch(reqs)
    for i, req in enumerate(reqs):
        assert_exact(res[i], po.evolve(req), tr=True)


def test_state_above_16_bit_positions_is_refused(ctx):
    req = Request([SHRINK], _state(70000, 50000, 2, 9), 1, flags=REL)
    with pytest.raises(hgx.HGXError, match="65535"):
        ctx.evolve(req)
