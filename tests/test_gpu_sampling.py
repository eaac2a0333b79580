"""GPU: the sampling and capping options, through the job envelope (hgx_run_wxf) and through the
C ABI (hgx_run_desc.sampling) in batches and the served path; the engine takes the decisions
generation by generation and expands only what survives.

TransitionRate, RuleWeights and MatchesPerStateRule thin transitions by a draw keyed on the
transition's isomorphism class (sampling_core.hpp, event_core.hpp:55); ExplorationProbability,
MaxStatesPerStep and MaxSuccessorStatesPerParent decide which new states are expanded
(exploration.hpp:66-122). Each sampled result is compared, state by state and event by event,
with the oracle's restatement of the same decisions over the C oracle's evolution
(oracle/sampling.py `sampled`), whose order-free part is itself checked against an independent
brute-force evolution (`brute`); the caps are also checked against closed forms, and the draw
against invariance under relabelling.
"""
import collections

import pytest

import hypergraphrewritingengine_amd as hgx
from oracle import sampling as SM
from tests.wxf_py import Rule, Symbol, dumps, loads

pytestmark = pytest.mark.gpu

SPLIT = [([[0, 1]], [[0, 2], [2, 1]])]                                 # binary edge splitting
SPLIT_FLIP = SPLIT + [([[0, 1]], [[1, 0]])]
TRI = [([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])]
CASES = {"split": (SPLIT, [[0, 1]], 5), "split_flip": (SPLIT_FLIP, [[0, 1]], 4),
         "path": (TRI, [[0, 1], [1, 2], [2, 0]], 3)}


def _run(ctx, rules, init, steps, **opts):
    o = {"CanonicalizeStates": Symbol("None"), "IncludeCanonicalHashes": Symbol("True"),
         "CausalTransitiveReduction": Symbol("False")}
    o.update(opts)
    job = {"InitialStates": [init], "Rules": {f"R{k}": Rule(l, r) for k, (l, r) in enumerate(rules)},
           "Steps": steps, "Options": o}
    return loads(ctx.run_wxf(dumps(job)))


def _decoded(out):
    states = [[e[1:] for e in out["States"][k]["Edges"]] for k in sorted(out["States"])]
    events = [(e["InputState"], e["OutputState"], e["RuleIndex"], tuple(e["ConsumedEdges"]))
              for _, e in sorted(out["Events"].items())]
    causal = sorted((x["From"], x["To"]) for x in out["CausalEdges"])
    branchial = sorted((x["From"], x["To"]) for x in out["BranchialEdges"])
    return states, events, causal, branchial


def _assert_matches_oracle(out, rules, init, steps, **kw):
    states, events, causal, branchial = _decoded(out)
    w_states, w_events, w_causal, w_branchial = SM.sampled(rules, init, steps, **kw)
    assert states == w_states
    assert events == w_events
    assert causal == sorted((a, b) for a, b, _ in w_causal)
    assert branchial == sorted(w_branchial)
    assert out["NumStates"] == len(states) and out["NumEvents"] == len(events)
    assert out["NumCausalEdges"] == len(causal) and out["NumBranchialEdges"] == len(branchial)


OPTS = {"rate": "TransitionRate", "weights": "RuleWeights", "seed": "RandomSeed",
        "per_state_rule": "MatchesPerStateRule", "explore_p": "ExplorationProbability",
        "per_step": "MaxStatesPerStep", "per_parent": "MaxSuccessorStatesPerParent"}


def _opts(kw):
    return {OPTS[k]: v for k, v in kw.items()}


SAMPLED = [("split", dict(rate=0.75, seed=1)), ("split", dict(rate=0.75, seed=0xABCDEF)),
           ("split", dict(rate=0.5, seed=2)), ("split_flip", dict(rate=0.75, seed=1)),
           ("split_flip", dict(rate=0.5, seed=2)), ("split_flip", dict(weights=[1.0, 0.0], seed=7)),
           ("split_flip", dict(weights=[0.5, 1.0], seed=7)), ("path", dict(rate=0.5, seed=2)),
           ("path", dict(rate=0.75, seed=11)), ("split", dict(per_state_rule=1, seed=0x5EED)),
           ("split_flip", dict(per_state_rule=2, seed=0x5EED)), ("path", dict(per_state_rule=1, seed=3)),
           ("path", dict(rate=0.75, per_state_rule=2, seed=3)), ("split", dict(explore_p=0.5, seed=2)),
           ("split_flip", dict(explore_p=0.5, seed=1)), ("path", dict(explore_p=0.5, seed=7)),
           ("split", dict(per_step=2)), ("split_flip", dict(per_parent=1)),
           ("split", dict(explore_p=0.7, per_step=3, per_parent=2, seed=1)),
           ("path", dict(rate=0.75, explore_p=0.7, per_step=3, per_parent=2, seed=11))]


@pytest.mark.parametrize("case,kw", SAMPLED, ids=[f"{c}-{'-'.join(f'{k}{v}' for k, v in kw.items())}" for c, kw in SAMPLED])
def test_sampled_run_equals_oracle(ctx, case, kw):
    rules, init, steps = CASES[case]
    out = _run(ctx, rules, init, steps, **_opts(kw))
    _assert_matches_oracle(out, rules, init, steps, **kw)


@pytest.mark.parametrize("case,kw", [("split", dict(rate=0.75, seed=1)), ("split_flip", dict(weights=[1.0, 0.0], seed=7)),
                                     ("path", dict(rate=0.5, per_state_rule=2, seed=3))])
def test_order_free_decisions_equal_the_brute_force_evolution(ctx, case, kw):
    """The kept states and transitions, as multisets of exact hashes and (input, output, rule,
    consumed ranks), against an evolution that draws as it goes (as the reference's engines do,
    compared across engines as sets: test_gpu_vs_cpu_differential.cpp:1380-1434)."""
    rules, init, steps = CASES[case]
    out = _run(ctx, rules, init, steps, **_opts(kw))
    states, events, _, _ = _decoded(out)
    hs, rk = hgx.ir_rank(states)
    got_s = collections.Counter(int(h) for h in hs)
    got_e = collections.Counter((int(hs[i]), int(hs[o]), r, tuple(rk[i][q] for q in c)) for i, o, r, c in events)
    want_s, want_e = SM.brute(rules, init, steps, **kw)
    assert got_s == want_s and got_e == want_e


def test_caps_closed_forms(ctx):
    """{{x,y}} -> {{x,z},{z,y}} from one edge, 5 steps: a state at step s has s + 1 edges and as
    many children. Unsampled: 1 + 1 + 2 + 6 + 24 + 120 states. MaxStatesPerStep 2 expands the
    first two states of each step: 1 + 1 + 2 + 6 + 8 + 10. MaxSuccessorStatesPerParent 1
    expands one child of each expanded state: 1 + 1 + 2 + 3 + 4 + 5. MatchesPerStateRule 1
    keeps one transition per state: 6. ExplorationProbability 0 expands nothing past the root:
    2. TransitionRate 0 keeps the root alone."""
    rules, init, steps = CASES["split"]
    assert _run(ctx, rules, init, steps)["NumStates"] == 154
    assert _run(ctx, rules, init, steps, MaxStatesPerStep=2)["NumStates"] == 28
    assert _run(ctx, rules, init, steps, MaxSuccessorStatesPerParent=1)["NumStates"] == 16
    assert _run(ctx, rules, init, steps, MatchesPerStateRule=1, RandomSeed=9)["NumStates"] == 6
    assert _run(ctx, rules, init, steps, ExplorationProbability=0.0)["NumStates"] == 2
    assert _run(ctx, rules, init, steps, TransitionRate=0.0)["NumStates"] == 1
    # MatchesPerStep is MaxStatesPerStep when UniformRandom selects it
    assert _run(ctx, rules, init, steps, MatchesPerStep=2, UniformRandom=Symbol("True"))["NumStates"] == 28
    assert _run(ctx, rules, init, steps, MatchesPerStep=2)["NumStates"] == 154


def test_neutral_options_change_nothing(ctx):
    """Rate 1 with every weight 1, or a cap above every count, is the unsampled evolution."""
    rules, init, steps = CASES["split_flip"]
    plain = _decoded(_run(ctx, rules, init, steps))
    for opts in ({"TransitionRate": 1.0, "RuleWeights": [1.0, 1.0], "RandomSeed": 5},
                 {"RuleWeights": [1.0]}, {"MatchesPerStateRule": 100}, {"MaxStatesPerStep": 10**6},
                 {"ExplorationProbability": 1.0}):
        assert _decoded(_run(ctx, rules, init, steps, **opts)) == plain, opts


def test_draw_is_invariant_under_relabelling(ctx):
    """The draw is keyed on exact hashes and canonical ranks, so the same initial state written
    with other labels and another edge order keeps the same transitions: equal multisets of
    state hashes and event counts."""
    rules = TRI
    a = _run(ctx, rules, [[0, 1], [1, 2], [2, 0]], 3, TransitionRate=0.6, RandomSeed=21)
    b = _run(ctx, rules, [[5, 9], [7, 5], [9, 7]], 3, TransitionRate=0.6, RandomSeed=21)
    ha = collections.Counter(s["CanonicalHash"] for s in a["States"].values())
    hb = collections.Counter(s["CanonicalHash"] for s in b["States"].values())
    assert ha == hb and a["NumEvents"] == b["NumEvents"] and a["NumStates"] > 1


def test_sampling_with_state_canonicalization(ctx):
    """Full and Automatic group the sampled states: the class counts are those of the sampled
    raw states."""
    rules, init, steps = CASES["split"]
    kw = dict(rate=0.75, seed=1)
    raw = _run(ctx, rules, init, steps, **_opts(kw))
    full = _run(ctx, rules, init, steps, CanonicalizeStates=Symbol("Full"), **_opts(kw))
    states, _, _, _ = _decoded(raw)
    assert full["NumStates"] == len(set(int(h) for h in hgx.ir_hash(states)))
    assert full["NumEvents"] == raw["NumEvents"]


# ---- the engine's sampling through the C ABI (hgx_run_desc.sampling) ----

def _request(case, kw):
    from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
    rules, init, steps = CASES[case]
    return Request([rule_from_lists(l, r, k) for k, (l, r) in enumerate(rules)], [init], steps,
                   flags=0x1 | 0x2 | 0x4 | 0x8, multi_root=True, sampling=dict(kw) if kw else None)


def _arrays(res):
    so, eo, vv = res.state_edges()
    states = [[[int(x) for x in vv[eo[e]:eo[e + 1]]] for e in range(int(so[s]), int(so[s + 1]))]
              for s in range(len(so) - 1)]
    off, pos = res.consumed()
    events = [(int(a), int(b), int(r), tuple(int(x) for x in pos[off[e]:off[e + 1]]))
              for e, (a, b, r, _) in enumerate(res.events())]
    cp, _ = res.causal()
    return states, events, sorted((int(a), int(b)) for a, b in cp), sorted((int(a), int(b)) for a, b in res.branchial())


def _want(case, kw):
    rules, init, steps = CASES[case]
    st, ev, ca, br = SM.sampled(rules, init, steps, **kw)
    return st, ev, sorted((a, b) for a, b, _ in ca), sorted(br)


MIX = [("split", dict(rate=0.75, seed=1)), ("split", {}), ("split_flip", dict(per_state_rule=2, seed=0x5EED)),
       ("path", dict(rate=0.75, explore_p=0.7, per_step=3, per_parent=2, seed=11)), ("path", {}),
       ("split_flip", dict(weights=[1.0, 0.0], seed=7)), ("split", dict(explore_p=0.5, seed=2)),
       ("split", dict(per_step=2)), ("split_flip", dict(rate=0.5, seed=2))]


def test_engine_sampling_in_a_mixed_batch(ctx):
    """Sampled and unsampled requests in one batch: the thinned match list is compacted across
    requests, and every result equals its own restatement (the unsampled ones the full
    evolution)."""
    res = ctx.evolve_batch([_request(c, kw) for c, kw in MIX])
    for i, (c, kw) in enumerate(MIX):
        got = _arrays(res[i])
        assert got == _want(c, kw), (c, kw)
        assert res[i].counts["raw_states"] == len(got[0]) and res[i].counts["events"] == len(got[1])


def test_engine_sampling_served(ctx):
    """The same requests through the served path, arriving a few per generation."""
    ids, got = {}, {}
    pending = list(range(len(MIX)))
    while pending or ids:
        for _ in range(2):
            if pending:
                i = pending.pop(0)
                ids[ctx.submit(_request(*MIX[i]))] = i
        ctx.serve_step()
        for rid, r in ctx.drain().items():
            got[ids.pop(rid)] = r
    for i, (c, kw) in enumerate(MIX):
        assert _arrays(got[i]) == _want(c, kw), (c, kw)


def test_sampled_batch_is_not_captured(ctx):
    batch = ctx.prepare([_request("split", dict(rate=0.75, seed=1)), _request("split", {})])
    with pytest.raises(hgx.HGXError, match="not recorded into graphs"):
        ctx.capture(batch)
    batch.close()
