"""GPU parity of the team sizes: one child rewritten and hashed (HGX_OPTION_EXPAND_TEAM), or
one parent state joined (HGX_OPTION_JOIN_TEAM), by one wave or by the waves of a workgroup
sharing one workspace. Every event, state hash and relation must be the same whichever team
size runs it, and the same as the oracle's. The cases cover the WL hash's uniform-arity front end (binary and ternary states
above 256 edges included), its general front end (mixed and unary arities), states whose
workspace exceeds the LDS (global workspace) and the automatic choice.
"""
import random

import pytest

from hypergraphrewritingengine_amd import workloads as W
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from oracle import pyoracle as po
from tests.parity import assert_exact
from tests.test_gpu_forward import REL, TR, _random_request, _same

pytestmark = pytest.mark.gpu

TEAMS = (1, 2, 4, 8)


def _by_team(ctx, reqs, teams=TEAMS):
    out = {}
    try:
        for t in (0,) + tuple(teams):
            ctx.set_expand_team(t)
            out[t] = ctx.evolve_batch(reqs)
    finally:
        ctx.set_expand_team(0)
    return out


def _by_join_team(ctx, reqs):
    """Join team sizes 1, 4, 8 and automatic, each with forwarding on and off."""
    out = {}
    try:
        for fwd in (1, 0):
            ctx.set_forwarding(fwd)
            for t in (0, 1, 4, 8):
                ctx.set_join_team(t)
                out[(fwd, t)] = ctx.evolve_batch(reqs)
    finally:
        ctx.set_join_team(0)
        ctx.set_forwarding(-1)
    return out


def _ternary_state(seed, m):
    rnd = random.Random(seed)
    nv = 1
    out = []
    for _ in range(m):
        e = [rnd.randrange(nv)]
        for _ in range(2):
            if rnd.random() < 0.5:
                e.append(nv)
                nv += 1
            else:
                e.append(rnd.randrange(nv))
        rnd.shuffle(e)
        out.append(e)
    return out


def _mixed_state(seed, m):
    rnd = random.Random(seed)
    nv = 2
    out = []
    for _ in range(m):
        a = rnd.choice([1, 2, 2, 3])
        e = [rnd.randrange(nv) for _ in range(a)]
        if rnd.random() < 0.4:
            e[rnd.randrange(a)] = nv
            nv += 1
        out.append(e)
    return out


def test_team_random_rules_exact(ctx):
    reqs = [_random_request(1000 + s) for s in range(48)]
    r = _by_team(ctx, reqs)
    for i, req in enumerate(reqs):
        for t in TEAMS:
            _same(r[t][i], r[1][i])
        _same(r[0][i], r[1][i])
        assert_exact(r[8][i], po.evolve(req))


def test_team_large_binary_states(ctx):
    """Served-like runs on 400-1500 edge binary states: the uniform-arity front end above 256
    edges, and the workgroup teams the automatic choice takes for them (the first: a second
    generation over the cap, so the run stops truncated after the first)."""
    rb = rule_from_lists([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])
    reqs = [Request([rb], W._connected_init(W.Stream(5), 400), 2, state_cap=3000, flags=REL | TR)]
    for k, m, steps in ((1, 700, 2), (2, 1500, 1)):
        st = W.Stream(W.splitmix64(0xAB00 + k))
        rule = W._random_rule(st)
        reqs.append(Request([rule], W._connected_init(st, m), steps, state_cap=20000, flags=REL | TR))
    r = _by_team(ctx, reqs)
    assert r[1][0].counts["truncated"] and r[1][1].counts["causal_pairs"]
    for i, req in enumerate(reqs):
        for t in (0,) + TEAMS:
            _same(r[t][i], r[1][i])
        assert_exact(r[0][i], po.evolve(req))


def test_team_large_ternary_and_mixed_states(ctx):
    rt = rule_from_lists([[0, 1, 2], [2, 3, 4]], [[0, 3, 5], [1, 4, 5], [2, 5, 5]])
    rb = rule_from_lists([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])
    ru = rule_from_lists([[0], [0, 1]], [[1], [0, 1], [1, 2]])
    reqs = [Request([rt], _ternary_state(7, 400), 2, state_cap=3000, flags=REL | TR),
            Request([rb, ru], _mixed_state(8, 450), 2, state_cap=3000, flags=REL | TR),
            Request([ru], _mixed_state(9, 300), 3, state_cap=3000, flags=REL | TR)]
    r = _by_team(ctx, reqs)
    for i, req in enumerate(reqs):
        assert r[1][i].counts["events"] > 100
        for t in (0,) + TEAMS:
            _same(r[t][i], r[1][i])
        assert_exact(r[4][i], po.evolve(req))


def test_team_global_workspace(ctx):
    """A 3000-edge state: its child's workspace is larger than a CU's LDS, so every team size
    runs on the global workspace."""
    st = W.Stream(W.splitmix64(0xC0FFEE))
    rule = W._random_rule(st)
    req = Request([rule], W._connected_init(st, 3000), 1, state_cap=200, flags=REL | TR)
    r = _by_team(ctx, [req], teams=(1, 8))
    _same(r[8][0], r[1][0])
    _same(r[0][0], r[1][0])
    assert_exact(r[1][0], po.evolve(req))


def test_join_team_random_rules_exact(ctx):
    """Join teams on random rule sets (forwarding's delta joins and merge included)."""
    reqs = [_random_request(2000 + s) for s in range(40)]
    r = _by_join_team(ctx, reqs)
    for i, req in enumerate(reqs):
        for key in r:
            _same(r[key][i], r[(0, 1)][i])
        assert_exact(r[(1, 8)][i], po.evolve(req))


def test_join_team_large_states(ctx):
    """Join teams on 400-1500 edge binary, ternary and mixed-arity states."""
    rb = rule_from_lists([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])
    rt = rule_from_lists([[0, 1, 2], [2, 3, 4]], [[0, 3, 5], [1, 4, 5], [2, 5, 5]])
    ru = rule_from_lists([[0], [0, 1]], [[1], [0, 1], [1, 2]])
    reqs = [Request([rb], W._connected_init(W.Stream(5), 400), 2, state_cap=3000, flags=REL | TR),
            Request([rt], _ternary_state(7, 400), 2, state_cap=3000, flags=REL | TR),
            Request([rb, ru], _mixed_state(8, 450), 2, state_cap=3000, flags=REL | TR)]
    st = W.Stream(W.splitmix64(0xAB01))
    reqs.append(Request([W._random_rule(st)], W._connected_init(st, 700), 2, state_cap=20000, flags=REL | TR))
    r = _by_join_team(ctx, reqs)
    for i, req in enumerate(reqs):
        for key in r:
            _same(r[key][i], r[(0, 1)][i])
        assert_exact(r[(1, 0)][i], po.evolve(req))


def test_team_option_rejects_bad_values(ctx):
    import hypergraphrewritingengine_amd as hgx
    for bad in (3, 16, -1):
        with pytest.raises(hgx.HGXError):
            ctx.set_expand_team(bad)
    for bad in (2, 3, 16, -1):
        with pytest.raises(hgx.HGXError):
            ctx.set_join_team(bad)
    ctx.set_expand_team(0)
    ctx.set_join_team(0)


def _by_lanes(ctx, reqs, lanes=(64, 32, 16)):
    out = {}
    try:
        for t in (0,) + tuple(lanes):
            ctx.set_expand_lanes(t)
            out[t] = ctx.evolve_batch(reqs)
    finally:
        ctx.set_expand_lanes(0)
    return out


def test_subwave_teams_random_rules_exact(ctx):
    """Sub-wave expansion teams: 16 and 32 lanes per child (four or two children per wave, each
    with its own workspace), forced for every size class whose workspaces fit, against one wave
    per child and the oracle -- random rule sets (unary to ternary arities, repeated variables,
    general and uniform-arity hash front ends) and batched runs of 3-64 edges."""
    reqs = [_random_request(3000 + s) for s in range(40)]
    reqs += [W.batched_run(i, cap=2048, flags=REL | TR) for i in range(0, 400, 25)]
    r = _by_lanes(ctx, reqs)
    for i, req in enumerate(reqs):
        for t in (0, 32, 16):
            _same(r[t][i], r[64][i])
        assert_exact(r[16][i], po.evolve(req))


def test_subwave_teams_mixed_class_batch(ctx):
    """One batch whose children fall in every expansion size class (up to 32, 64, 128 and 512
    edges), so sub-wave launches and wave launches run side by side in one generation; the
    children of 65-300 edges take 16 or 32 lanes only where their workspace still fits."""
    rb = rule_from_lists([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])
    reqs = []
    for k, m in enumerate((5, 20, 40, 60, 90, 200, 300)):
        reqs.append(Request([rb], W._connected_init(W.Stream(70 + k), m), 3, state_cap=1500, flags=REL | TR))
    reqs.append(Request([rule_from_lists([[0], [0, 1]], [[1], [0, 1], [1, 2]])], _mixed_state(3, 40), 3,
                        state_cap=1500, flags=REL | TR))
    r = _by_lanes(ctx, reqs)
    for i, req in enumerate(reqs):
        for t in (0, 32, 16):
            _same(r[t][i], r[64][i])
        assert_exact(r[0][i], po.evolve(req))


def test_expand_lanes_option_rejects_bad_values(ctx):
    import hypergraphrewritingengine_amd as hgx
    for bad in (8, 24, 128, -1):
        with pytest.raises(hgx.HGXError):
            ctx.set_expand_lanes(bad)
    ctx.set_expand_lanes(0)


def _by_join_lanes(ctx, reqs):
    """Join lanes 64, 32, 16 and automatic (one-wave teams), each with forwarding on and off."""
    out = {}
    try:
        ctx.set_join_team(1)
        for fwd in (1, 0):
            ctx.set_forwarding(fwd)
            for lanes in (0, 64, 32, 16):
                ctx.set_join_lanes(lanes)
                out[(fwd, lanes)] = ctx.evolve_batch(reqs)
    finally:
        ctx.set_join_lanes(0)
        ctx.set_join_team(0)
        ctx.set_forwarding(-1)
    return out


def test_subwave_join_random_rules_exact(ctx):
    """Sub-wave join teams: 16 and 32 lanes per parent state (four or two states per wave, each
    with its own workspace), forced for every size class whose workspaces fit, with forwarding
    (the delta joins and the merge) on and off, against one wave per state and the oracle."""
    reqs = [_random_request(4000 + s) for s in range(40)]
    reqs += [W.batched_run(i, cap=2048, flags=REL | TR) for i in range(0, 400, 25)]
    r = _by_join_lanes(ctx, reqs)
    for i, req in enumerate(reqs):
        for key in r:
            _same(r[key][i], r[(0, 64)][i])
        assert_exact(r[(1, 16)][i], po.evolve(req))


def test_subwave_join_mixed_class_batch(ctx):
    """Parents in every join size class (up to 32, 64, 128, 512 edges) in one batch, so sub-wave
    launches and wave launches run side by side in one generation."""
    rb = rule_from_lists([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])
    reqs = []
    for k, m in enumerate((5, 20, 40, 60, 90, 200, 300)):
        reqs.append(Request([rb], W._connected_init(W.Stream(90 + k), m), 3, state_cap=1500, flags=REL | TR))
    reqs.append(Request([rule_from_lists([[0], [0, 1]], [[1], [0, 1], [1, 2]])], _mixed_state(4, 40), 3,
                        state_cap=1500, flags=REL | TR))
    r = _by_join_lanes(ctx, reqs)
    for i, req in enumerate(reqs):
        for key in r:
            _same(r[key][i], r[(0, 64)][i])
        assert_exact(r[(0, 0)][i], po.evolve(req))


def test_join_lanes_option_rejects_bad_values(ctx):
    import hypergraphrewritingengine_amd as hgx
    for bad in (8, 24, 128, -1):
        with pytest.raises(hgx.HGXError):
            ctx.set_join_lanes(bad)
    ctx.set_join_lanes(0)


def test_counter_flush_changes_nothing(ctx):
    """HGX_OPTION_COUNTER_FLUSH (hardware-counter runs) streams a buffer through the L2 around
    every timed kernel; every result stays the oracle's, and bad sizes are refused."""
    import hypergraphrewritingengine_amd as hgx
    reqs = [W.batched_run(i, cap=1024, flags=REL | TR) for i in range(0, 96, 12)]
    for bad in (1, 15, 2048, -1):
        with pytest.raises(hgx.HGXError):
            ctx.set_counter_flush(bad)
    try:
        ctx.set_counter_flush(64)
        r = ctx.evolve_batch(reqs)
    finally:
        ctx.set_counter_flush(0)
    for i, req in enumerate(reqs):
        assert_exact(r[i], po.evolve(req))
