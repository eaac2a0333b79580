"""CPU: the oracle's restatement of the sampling decisions, pinned by what the mathematics fixes.

The draw (sampling_core.hpp:42-57) is a threshold on a uniform 53-bit value: certain at rate 1,
impossible at rate 0, monotone in the rate, and survives at the rate's frequency; the cap's rank
(:66-71) is a different mix. The cut-out evolution (`sampled`) agrees with an independent
brute-force evolution that draws as it goes (`brute`), and the caps give closed-form counts.
"""
import random

import pytest

from oracle import sampling as SM

SPLIT = [([[0, 1]], [[0, 2], [2, 1]])]
SPLIT_FLIP = SPLIT + [([[0, 1]], [[1, 0]])]
TRI = [([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])]


def test_draw_bounds_monotonicity_and_frequency():
    rnd = random.Random(1)
    keys = [rnd.getrandbits(64) for _ in range(20000)]
    for seed in (0, 1, 0xABCDEF):
        assert all(SM.transition_survives(k, seed, 1.0) for k in keys[:100])
        assert not any(SM.transition_survives(k, seed, 0.0) for k in keys[:100])
        for rate in (0.1, 0.5, 0.9):
            n = sum(SM.transition_survives(k, seed, rate) for k in keys)
            sd = (len(keys) * rate * (1 - rate)) ** 0.5
            assert abs(n - rate * len(keys)) < 5 * sd, (seed, rate, n)
        # a transition that survives at a rate survives at every higher one
        for k in keys[:2000]:
            if SM.transition_survives(k, seed, 0.3):
                assert SM.transition_survives(k, seed, 0.31)
    # the seed changes the draw, seed 0 included
    a = [SM.transition_survives(k, 0, 0.5) for k in keys[:1000]]
    b = [SM.transition_survives(k, 1, 0.5) for k in keys[:1000]]
    assert a != b and 300 < sum(a) < 700


def test_rank_is_a_different_mix_from_the_draw():
    rnd = random.Random(2)
    keys = [rnd.getrandbits(64) for _ in range(20000)]
    surv = [k for k in keys if SM.transition_survives(k, 5, 0.5)]
    low = sum(SM.transition_rank(k, 5) < (1 << 63) for k in surv)
    assert abs(low - len(surv) / 2) < 5 * (len(surv) / 4) ** 0.5
    assert len({SM.transition_rank(k, 5) for k in keys}) == len(keys)


def test_rate_for_rule_and_threshold():
    assert SM.rate_for_rule(0.5, None, 0) == 0.5
    assert SM.rate_for_rule(0.5, [2.0], 0) == 1.0
    assert SM.rate_for_rule(0.5, [2.0], 1) == 0.5      # past the list: weight 1
    assert SM.rate_for_rule(1.0, [-1.0], 0) == 0.0
    assert SM.explore_threshold(1.0) == 0xFFFFFFFF and SM.explore_threshold(0.0) == 0
    assert SM.explore_threshold(0.5) == 1 << 31
    assert SM.explore_threshold(7.0) == 0xFFFFFFFF and SM.explore_threshold(-1.0) == 0


@pytest.mark.parametrize("rules,init,steps", [(SPLIT, [[0, 1]], 5), (SPLIT_FLIP, [[0, 1]], 4),
                                              (TRI, [[0, 1], [1, 2], [2, 0]], 3)])
@pytest.mark.parametrize("kw", [dict(rate=0.75, seed=1), dict(rate=0.5, seed=2), dict(per_state_rule=1, seed=0x5EED),
                                dict(per_state_rule=2, seed=3), dict(rate=0.6, per_state_rule=2, seed=3),
                                dict(weights=[1.0, 0.0], seed=7), dict(weights=[0.5], rate=0.9, seed=11)])
def test_cut_out_evolution_equals_brute_force(rules, init, steps, kw):
    import collections
    import hypergraphrewritingengine_amd as hgx
    states, events, _, _ = SM.sampled(rules, init, steps, **kw)
    hs, rk = hgx.ir_rank(states)
    got_s = collections.Counter(int(h) for h in hs)
    got_e = collections.Counter((int(hs[i]), int(hs[o]), r, tuple(rk[i][q] for q in c)) for i, o, r, c in events)
    assert (got_s, got_e) == SM.brute(rules, init, steps, **kw)


def test_caps_closed_forms():
    """A state at step s of {{x,y}} -> {{x,z},{z,y}} from one edge has s + 1 edges and as many
    children (see tests/test_gpu_sampling.py::test_caps_closed_forms)."""
    assert len(SM.sampled(SPLIT, [[0, 1]], 5)[0]) == 1 + 1 + 2 + 6 + 24 + 120
    assert len(SM.sampled(SPLIT, [[0, 1]], 5, per_step=2)[0]) == 1 + 1 + 2 + 6 + 8 + 10
    assert len(SM.sampled(SPLIT, [[0, 1]], 5, per_parent=1)[0]) == 1 + 1 + 2 + 3 + 4 + 5
    assert len(SM.sampled(SPLIT, [[0, 1]], 5, per_state_rule=1, seed=9)[0]) == 6
    assert len(SM.sampled(SPLIT, [[0, 1]], 5, explore_p=0.0)[0]) == 2
    assert len(SM.sampled(SPLIT, [[0, 1]], 5, rate=0.0)[0]) == 1
