"""GPU: quotient exploration (HGX_OPT_QUOTIENT) and the raw unfolding beneath it
(hgx_result_unfold), against the oracle (oracle/quotient.py), the C oracle's full evolution and
the golden matrix's quotient rows.

* The explored evolution equals the oracle's element by element: every state's edges and step,
  every event (input, output, rule, consumed positions), and the relations among the explored
  events (causal pairs with their reduction tags, branchial pairs).
* The unfolding replayed on the device is the full multiway system: its counts equal the C
  oracle's full evolution (events, causal triples and pairs with and without transitive
  reduction, branchial pairs) -- at full size on config 2 (10,309,586 raw events from 26,661
  classes), and on a full-capture result it gives that result's own counts.
* Through the job envelope, ExploreFromCanonicalStatesOnly with exact states serves the golden
  quotient rows (states, and events under None / Full / Automatic identity, causal pairs,
  branchial pairs), and Automatic identity under full capture serves the golden 'Full Automatic'
  rows.
"""
import json
import os

import numpy as np
import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd import workloads as W
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from oracle import pyoracle as po
from oracle import quotient as Q
from tests.wxf_py import Rule, Symbol, dumps, loads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "golden_matrix_none.json")))["cases"]
QROWS = json.load(open(os.path.join(HERE, "golden", "golden_matrix_quotient.json")))["rows"]
C2 = json.load(open(os.path.join(HERE, "golden", "config2_counts.json")))["counts"]
REL = 0x1 | 0x2 | 0x4 | 0x8
TR = 0x100
QUO = hgx.abi.HGX_OPT_QUOTIENT
QROOTS = hgx.abi.HGX_OPT_QUOTIENT_ROOTS
UNFOLD_KEYS = ("raw_events", "causal_triples", "causal_pairs", "causal_pairs_tr", "causal_triples_tr",
               "branchial_pairs")

RULESETS = [
    [([[0, 1]], [[0, 2], [2, 1]])],
    [([[0, 1], [0, 2]], [[0, 1], [0, 3], [1, 3], [2, 3]])],
    [([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]]), ([[0, 1]], [[1, 0]])],
    [([[0, 1], [2, 3]], [[0, 2], [1, 3]])],
    [([[0, 1, 2]], [[0, 1], [1, 2], [2, 0]]), ([[0, 1], [1, 2]], [[0, 1, 2]])],
]
ROOTS = [[[0, 1], [1, 2], [2, 0]], [[0, 0], [0, 0]], [[0, 1], [0, 2], [0, 3]]]


def _rules(spec):
    return [rule_from_lists(l, h, i) for i, (l, h) in enumerate(spec)]


def _corpus():
    out = [(c["name"], [(r["lhs"], r["rhs"]) for r in c["rules"]], [c["init"]], c["steps"]) for c in CASES]
    for k, rs in enumerate(RULESETS):
        for r, root in enumerate(ROOTS):
            out.append((f"set{k}-root{r}", rs, [root], 3))
    return out


def _edges(res):
    so, eo, vv = res.state_edges()
    return [[list(map(int, vv[eo[e]:eo[e + 1]])) for e in range(int(so[s]), int(so[s + 1]))]
            for s in range(len(so) - 1)]


def _full_counts(rules, roots, steps):
    c = po.evolve(Request(rules, roots, steps, flags=0x1 | 0x2 | 0x100, multi_root=True)).counts
    return {k: c["events"] if k == "raw_events" else c[k] for k in UNFOLD_KEYS}


def _assert_explored_equal(res, x, relations=True):
    c = res.counts
    assert c["raw_states"] == len(x.states) and c["events"] == len(x.events)
    _, steps, _ = res.states()
    assert _edges(res) == [s[0] for s in x.states]
    assert [int(v) for v in steps] == [s[1] for s in x.states]
    ev = res.events()
    off, pos = res.consumed()
    for e, (i, o, ri, p) in enumerate(x.events):
        assert (int(ev[e][0]), int(ev[e][1]), int(ev[e][2])) == (i, o, ri), e
        assert [int(v) for v in pos[off[e]:off[e + 1]]] == p, e
    if not relations:
        return
    causal, branchial = Q.relations(x)
    cp, ct = res.causal()
    assert {(int(a), int(b)): bool(t) for (a, b), t in zip(cp, ct)} == causal
    assert {(int(a), int(b)) for a, b in res.branchial()} == branchial


@pytest.mark.parametrize("name,rules,roots,steps", _corpus(), ids=lambda v: v if isinstance(v, str) else "")
def test_quotient_exploration_equals_oracle(ctx, name, rules, roots, steps):
    R = _rules(rules)
    res = ctx.evolve(Request(R, roots, steps, flags=REL | TR | QUO, multi_root=True))
    x = Q.explore(R, roots, steps)
    _assert_explored_equal(res, x)
    # the raw unfolding beneath it: the full multiway system
    u = res.unfold()
    full = _full_counts(R, roots, steps)
    assert {k: u[k] for k in UNFOLD_KEYS} == full
    ou = Q.unfold(x, steps)
    assert {k: u[k] for k in ou} == ou
    assert u["class_matches"] == sum(len(v) for v in Q.record(x).values())


@pytest.mark.parametrize("merge", [False, True])
def test_isomorphic_roots(ctx, merge):
    R = _rules([([[0, 1], [0, 2]], [[0, 1], [0, 3], [1, 3], [2, 3]])])
    roots = [[[0, 1], [0, 2]], [[5, 6], [5, 7]]]
    res = ctx.evolve(Request(R, roots, 3, flags=REL | TR | QUO | (QROOTS if merge else 0), multi_root=True))
    x = Q.explore(R, roots, 3, merge_roots=merge)
    _assert_explored_equal(res, x)
    u = res.unfold()
    assert {k: u[k] for k in UNFOLD_KEYS} == _full_counts(R, roots, 3)


def test_unfold_of_a_full_capture_is_its_own_counts(ctx):
    for i in range(6):
        b = W.batched_run(i)
        req = Request(b.rules, [list(e) for e in b.init[:6]], 3, flags=REL | TR)
        res = ctx.evolve(req)
        c = res.counts
        u = res.unfold()
        assert u["raw_events"] == c["events"]
        assert {k: u[k] for k in UNFOLD_KEYS[1:]} == {k: c[k] for k in UNFOLD_KEYS[1:]}, i


def test_config2_quotient_unfolds_to_the_full_system(ctx):
    """Config 2 at its full 6 steps, explored once per exact class: 26,673 classes (the WL hash
    merges 12 pairs of them -- states that differ in edge multiplicities -- into its 26,661), the
    explored evolution equal to the oracle's state by state and event by event, and the replay of
    the classes' expansions reproduces the 10,309,586-event multiway system's relations exactly."""
    req = W.config2(6, flags=0x1 | 0x2 | 0x8 | TR | QUO)
    res = ctx.evolve(req)
    x = Q.explore(req.rules, [req.init], 6)
    assert Q.classes(x) == 26673
    _assert_explored_equal(res, x, relations=False)
    hs = hgx.ir_hash(_edges(res))
    assert len({int(h) for h in hs}) == 26673
    wl, _, _ = res.states()
    assert len({int(h) for h in wl}) == C2["tr"]["canonical_states"]
    u = res.unfold()
    want = C2["tr"]
    assert u["raw_events"] == want["events"]
    assert {k: u[k] for k in UNFOLD_KEYS[1:]} == {k: want[k] for k in UNFOLD_KEYS[1:]}
    assert res.counts["events"] < want["events"] // 100


def test_unfold_needs_the_state_edges(ctx):
    R = _rules(RULESETS[0])
    res = ctx.evolve(Request(R, [ROOTS[0]], 2, flags=0x3 | QUO, multi_root=True))
    with pytest.raises(hgx.HGXError, match="HGX_REC_STATES"):
        res.unfold()


def _job(case, mode, quotient, tr=False, roots=None, **extra):
    opts = {"CanonicalizeStates": Symbol("Full"), "CanonicalizeEvents": Symbol(mode),
            "CausalTransitiveReduction": Symbol("True" if tr else "False")}
    if quotient:
        opts["ExploreFromCanonicalStatesOnly"] = Symbol("True")
    opts.update(extra)
    return dumps({"InitialStates": roots or [case["init"]],
                  "Rules": {f"R{k}": Rule(r["lhs"], r["rhs"]) for k, r in enumerate(case["rules"])},
                  "Steps": case["steps"], "Options": opts})


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_wxf_quotient_rows(ctx, case):
    for mode in ("None", "Full", "Automatic"):
        out = loads(ctx.run_wxf(_job(case, mode, True)))
        want = QROWS[f"{case['name']} {mode}"]
        got = [out["NumStates"], out["NumEvents"], None, out["NumCausalEdges"], out["NumBranchialEdges"]]
        assert got == [want[0], want[1], None, want[3], want[4]], (case["name"], mode)
        # full capture under the same identity: the golden full-capture row
        out = loads(ctx.run_wxf(_job(case, mode, False)))
        want = QROWS[f"{case['name']} {mode} full"]
        got = [out["NumStates"], out["NumEvents"], None, out["NumCausalEdges"], out["NumBranchialEdges"]]
        assert got == [want[0], want[1], None, want[3], want[4]], (case["name"], mode, "full")


def test_wxf_quotient_with_reduction_and_lists(ctx):
    case = next(c for c in CASES if c["name"] == "cycle4-automorphic")
    rules = _rules([(r["lhs"], r["rhs"]) for r in case["rules"]])
    full = _full_counts(rules, [case["init"]], case["steps"])
    out = loads(ctx.run_wxf(_job(case, "None", True, tr=True)))
    assert out["NumCausalEdges"] == full["causal_pairs_tr"]
    # explored events only in the lists; each event's CanonicalId is a recorded expansion
    x = Q.explore(rules, [case["init"]], case["steps"])
    assert len(out["Events"]) == len(x.events)
    assert len(out["States"]) == Q.classes(x)
    out = loads(ctx.run_wxf(_job(case, "Automatic", True)))
    assert len({e["CanonicalId"] for e in out["Events"].values()}) == out["NumEvents"]


def test_wxf_isomorphic_roots(ctx):
    case = {"rules": [{"lhs": [[0, 1], [0, 2]], "rhs": [[0, 1], [0, 3], [1, 3], [2, 3]]}], "steps": 3}
    roots = [[[0, 1], [0, 2]], [[5, 6], [5, 7]]]
    R = _rules([([[0, 1], [0, 2]], [[0, 1], [0, 3], [1, 3], [2, 3]])])
    full = _full_counts(R, roots, 3)
    for merge in (False, True):
        out = loads(ctx.run_wxf(_job(case, "None", True, roots=roots,
                                     QuotientInitialStates=Symbol("True" if merge else "False"))))
        assert out["NumEvents"] == full["raw_events"]
        assert len(out["Events"]) == len(Q.explore(R, roots, 3, merge_roots=merge).events)


def test_wxf_quotient_refusals_and_warning(ctx):
    case = CASES[0]
    with pytest.raises(hgx.HGXError, match="content identity|CanonicalizeStates -> Automatic"):
        ctx.run_wxf(_job(case, "None", True, CanonicalizeStates=Symbol("Automatic")))
    out = loads(ctx.run_wxf(_job(case, "None", True, CanonicalizeStates=Symbol("None"))))
    assert any(w["Kind"] == "Engine" for w in out["Warnings"])
    # None states: every state its own class, so the full evolution is explored
    assert out["NumEvents"] == QROWS[f"{case['name']} None full"][1]
    with pytest.raises(hgx.HGXError, match="ConsumedEdges needs InputState"):
        ctx.run_wxf(_job(case, "None", True, CanonicalizeEvents=[Symbol("ConsumedEdges")]))


@pytest.mark.parametrize("steps,states,events", [(5, 401, 400), (6, 3867, 3866), (7, 45317, 45316)])
def test_published_wpp_quotient_sizes(ctx, steps, states, events):
    """The reference's published quotient workload (paper/tables/t7_gpu.tex:8-10: the WPP rule
    from {{0,1},{0,2}}, exact states): the explored evolution has the table's raw states and
    events at every depth; at depth 5 it equals the oracle's state by state, and its unfolding is
    the full multiway system's."""
    R = _rules([([[0, 1], [0, 2]], [[0, 1], [0, 3], [1, 3], [2, 3]])])
    res = ctx.evolve(Request(R, [[0, 1], [0, 2]], steps, flags=0x1 | 0x2 | 0x8 | TR | QUO))
    assert (res.counts["raw_states"], res.counts["events"]) == (states, events)
    if steps == 5:
        _assert_explored_equal(res, Q.explore(R, [[[0, 1], [0, 2]]], steps), relations=False)
        u = res.unfold()
        assert {k: u[k] for k in UNFOLD_KEYS} == _full_counts(R, [[[0, 1], [0, 2]]], steps)


def test_quotient_requests_served_and_batched_in_a_mix(ctx):
    """Quotient requests beside full-capture and sampled ones, in one batch and on the served
    path: each result equals the same request evolved alone (and the quotient ones the oracle)."""
    qreqs = []
    for k, rs in enumerate(RULESETS):
        qreqs.append(Request(_rules(rs), [ROOTS[k % len(ROOTS)]], 3, flags=REL | TR | QUO, multi_root=True))
    others = [W.batched_run(i, cap=1024, flags=REL | TR) for i in range(4)]
    sampled = W.config2(4, flags=REL | TR)
    sampled.sampling = {"rate": 0.5, "seed": 3}
    mix = [qreqs[0], others[0], qreqs[1], sampled, qreqs[2], others[1], qreqs[3], others[2], qreqs[4], others[3]]
    alone = [ctx.evolve(r) for r in mix]
    batch = ctx.evolve_batch(mix)
    ids = [ctx.submit(r) for r in mix]
    served = [ctx.wait(i) for i in ids]
    for k, r in enumerate(mix):
        for got in (batch[k], served[k]):
            assert got.counts == alone[k].counts, k
            assert _edges(got) == _edges(alone[k]), k
            assert (got.events() == alone[k].events()).all(), k
        if r.flags & QUO:
            x = Q.explore(r.rules, r.init, r.steps)
            _assert_explored_equal(batch[k], x)
            assert batch[k].unfold() == served[k].unfold() == alone[k].unfold()


def test_unfold_of_a_thinned_full_capture_is_its_own_counts(ctx):
    """A transition draw keyed on the transition's exact class keeps the same matches on every
    state of a class, so the replay of a thinned full evolution is that evolution: its counts."""
    req = W.config2(5, flags=REL | TR)
    req.sampling = {"rate": 0.5, "seed": 11}
    res = ctx.evolve(req)
    c = res.counts
    u = res.unfold()
    assert u["raw_events"] == c["events"] > 0
    assert {k: u[k] for k in UNFOLD_KEYS[1:]} == {k: c[k] for k in UNFOLD_KEYS[1:]}
