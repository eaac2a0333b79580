"""GPU: the paclet's job envelope (hgx_run_wxf) and the hg_evolve-style worker.

A WXF job is evolved through the envelope and its decoded result is checked against the same
request evolved through the plain API and the oracle: counts, every state's edges, events,
consumed/produced edges (applying an event's consumed positions to its input state and appending
its produced edges gives its output state), causal pairs (with and without transitive reduction)
and branchial pairs. With CanonicalizeStates -> Full the classes are the exact (IR) ones: the golden matrix's Full
rows fix their number, each representative is emitted in its canonical form, and every event
names its states' representatives. The worker binary answers one-shot and framed (--serve)
jobs with the same bytes, and a failing job with a zero-length frame.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from oracle import pyoracle as po
from tests.wxf_py import Rule, Symbol, dumps, loads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
WORKER = os.path.join(os.path.dirname(HERE), "hypergraphrewritingengine_amd", "hgx_evolve")

RULES = [([[0, 1], [0, 2]], [[0, 2], [0, 3], [1, 3], [2, 3]]), ([[0, 1, 2]], [[0, 1], [1, 2]])]
STATES = [[[1, 1], [1, 1]], [[5, 6, 7], [7, 8], [8, 5]]]


def _job(tr=True, steps=3, requested=None):
    opts = {"CausalTransitiveReduction": Symbol("True" if tr else "False"), "ShowProgress": Symbol("False")}
    if requested is not None:
        opts["RequestedData"] = requested
    return dumps({"InitialStates": STATES,
                  "Rules": {f"Rule{k + 1}": Rule(l, r) for k, (l, r) in enumerate(RULES)},
                  "Steps": steps, "Options": opts})


def _request(tr, steps=3):
    rules = [rule_from_lists(l, r, k) for k, (l, r) in enumerate(RULES)]
    return Request(rules, STATES, steps, flags=0x1 | 0x2 | 0x4 | 0x8 | (0x100 if tr else 0), multi_root=True)


@pytest.mark.parametrize("tr", [True, False])
def test_wxf_envelope_matches_direct_evolution(ctx, tr):
    out = loads(ctx.run_wxf(_job(tr)))
    req = _request(tr)
    res = ctx.evolve(req)
    c = res.counts
    o = po.evolve(req)
    # CanonicalizeStates -> None: every state its own class (hg_gpu_backend.cpp:278-305)
    assert out["NumStates"] == c["raw_states"] == o.counts["raw_states"]
    assert out["NumEvents"] == c["events"]
    assert out["NumBranchialEdges"] == c["branchial_pairs"]
    assert out["NumCausalEdges"] == (c["causal_pairs_tr"] if tr else c["causal_pairs"])
    assert out["Warnings"][0]["Context"] == "ShowProgress"
    # states: edges as {index, vertices...}
    so, eo, vv = res.state_edges()
    hashes, steps, _ = res.states()
    S = out["States"]
    assert len(S) == c["raw_states"]
    edges_of = {}
    for sid, st in S.items():
        assert st["Id"] == sid and st["Step"] == int(steps[sid])
        assert st["IsInitial"] == ("True" if sid < len(STATES) else "False")
        want = [list(map(int, vv[eo[e]:eo[e + 1]])) for e in range(int(so[sid]), int(so[sid + 1]))]
        got = [ed[1:] for ed in st["Edges"]]
        assert [ed[0] for ed in st["Edges"]] == list(range(len(want)))
        assert got == want
        assert st["CanonicalId"] == st["ContentStateId"] == sid
        edges_of[sid] = got
    # events, with consumed/produced edges reproducing each output state
    ev = res.events()
    E = out["Events"]
    assert len(E) == c["events"]
    for eid, e in E.items():
        assert (e["InputState"], e["OutputState"], e["RuleIndex"]) == tuple(int(x) for x in ev[eid][:3])
        inp, outp = edges_of[e["InputState"]], edges_of[e["OutputState"]]
        kept = [inp[k] for k in range(len(inp)) if k not in set(e["ConsumedEdges"])]
        assert len(e["ConsumedEdges"]) == len(RULES[e["RuleIndex"]][0])
        assert len(e["ProducedEdges"]) == len(RULES[e["RuleIndex"]][1])
        assert outp[:len(kept)] == kept
        assert e["ProducedEdges"] == list(range(len(kept), len(outp)))
    # relations
    cp, ct = res.causal()
    want_c = {(int(a), int(b)) for (a, b), t in zip(cp, ct) if t or not tr}
    assert {(x["From"], x["To"]) for x in out["CausalEdges"]} == want_c
    assert len(out["CausalEdges"]) == len(want_c)
    want_b = sorted((int(a), int(b)) for a, b in res.branchial())
    assert sorted((x["From"], x["To"]) for x in out["BranchialEdges"]) == want_b


GOLD_NONE = __import__("json").load(open(os.path.join(HERE, "golden", "golden_matrix_none.json")))["cases"]
GOLD_FULL = __import__("json").load(open(os.path.join(HERE, "golden", "golden_matrix_full.json")))["cases"]


@pytest.mark.parametrize("case", [c for c in GOLD_NONE if c["name"] in GOLD_FULL], ids=lambda c: c["name"])
def test_wxf_full_canonicalization(ctx, case):
    job = dumps({"InitialStates": [case["init"]],
                 "Rules": {f"R{k}": Rule(r["lhs"], r["rhs"]) for k, r in enumerate(case["rules"])},
                 "Steps": case["steps"],
                 "Options": {"CanonicalizeStates": Symbol("Full"), "IncludeCanonicalHashes": Symbol("True")}})
    out = loads(ctx.run_wxf(job))
    want = GOLD_FULL[case["name"]]
    assert out["NumStates"] == want["ir_states"]
    assert out["NumEvents"] == want["events"]
    # the same evolution through the plain API: the raw states and their exact hashes
    rules = [rule_from_lists(r["lhs"], r["rhs"], k) for k, r in enumerate(case["rules"])]
    res = ctx.evolve(Request(rules, [case["init"]], case["steps"], flags=0x1 | 0x2 | 0x4 | 0x8 | 0x100,
                             multi_root=True))
    so, eo, vv = res.state_edges()
    edges = [[list(map(int, vv[eo[e]:eo[e + 1]])) for e in range(int(so[s]), int(so[s + 1]))]
             for s in range(len(so) - 1)]
    ir = [int(h) for h in hgx.ir_hash(edges)]
    reps = {}
    for s, h in enumerate(ir):
        reps.setdefault(h, s)
    assert sorted(out["States"]) == sorted(reps.values())
    for sid, st in out["States"].items():
        h, form = hgx.ir_canonical_form(edges[sid])
        # written as a signed 64-bit integer, as the reference writes it
        assert st["CanonicalHash"] % (1 << 64) == h and st["CanonicalId"] == sid
        assert [e[1:] for e in st["Edges"]] == form
    for e in out["Events"].values():
        assert e["CanonicalInputState"] == reps[ir[e["InputState"]]]
        assert e["CanonicalOutputState"] == reps[ir[e["OutputState"]]]


@pytest.mark.parametrize("mode", ["Full", "Automatic"])
@pytest.mark.parametrize("case", [c for c in GOLD_NONE if c["name"] in GOLD_FULL], ids=lambda c: c["name"])
def test_wxf_event_canonicalization(ctx, case, mode):
    """CanonicalizeEvents: events identified by their signature over exact state hashes (and,
    for Automatic, the step and the canonical ranks of the consumed and produced edges); the
    golden matrix's 'None Full' / 'None Automatic' rows fix the event counts, and the relations
    are the raw events' (the same counts as without event canonicalization)."""
    base = {"InitialStates": [case["init"]],
            "Rules": {f"R{k}": Rule(r["lhs"], r["rhs"]) for k, r in enumerate(case["rules"])},
            "Steps": case["steps"]}
    plain = loads(ctx.run_wxf(dumps({**base, "Options": {}})))
    out = loads(ctx.run_wxf(dumps({**base, "Options": {"CanonicalizeEvents": Symbol(mode)}})))
    assert out["NumEvents"] == GOLD_FULL[case["name"]]["events_" + mode.lower()]
    assert plain["NumEvents"] == GOLD_FULL[case["name"]]["events"] == len(out["Events"])
    for k in ("NumCausalEdges", "NumBranchialEdges", "NumStates"):
        assert out[k] == plain[k]
    ids = {e["CanonicalId"] for e in out["Events"].values()}
    assert len(ids) == out["NumEvents"]
    for eid, e in out["Events"].items():
        assert e["CanonicalId"] <= eid and out["Events"][e["CanonicalId"]]["CanonicalId"] == e["CanonicalId"]


@pytest.mark.parametrize("case", [c for c in GOLD_NONE if c["name"] in GOLD_FULL], ids=lambda c: c["name"])
def test_wxf_automatic_state_identity(ctx, case):
    """CanonicalizeStates -> Automatic: states of equal content (edges in order, vertex names
    compared as such) are one class. The golden matrix's 'Automatic None' row fixes the class
    count; its other columns, and the 'Automatic Automatic|Full' rows, equal the None rows'."""
    base = {"InitialStates": [case["init"]],
            "Rules": {f"R{k}": Rule(r["lhs"], r["rhs"]) for k, r in enumerate(case["rules"])},
            "Steps": case["steps"]}
    want = GOLD_FULL[case["name"]]
    plain = loads(ctx.run_wxf(dumps({**base, "Options": {}})))
    out = loads(ctx.run_wxf(dumps({**base, "Options": {"CanonicalizeStates": Symbol("Automatic")}})))
    assert out["NumStates"] == want["content_states"] == len(out["States"])
    for k in ("NumEvents", "NumCausalEdges", "NumBranchialEdges"):
        assert out[k] == plain[k]
    # a class's members hold its representative's edges (one initial state: equal names are
    # equal device labels), and a representative is its class's first state
    for e in out["Events"].values():
        rep = e["CanonicalOutputState"]
        assert rep <= e["OutputState"] and rep in out["States"]
        assert plain["States"][e["OutputState"]]["Edges"] == plain["States"][rep]["Edges"]
    for mode in ("Automatic", "Full"):
        both = loads(ctx.run_wxf(dumps({**base, "Options": {"CanonicalizeStates": Symbol("Automatic"),
                                                            "CanonicalizeEvents": Symbol(mode)}})))
        assert both["NumStates"] == want["content_states"]
        assert both["NumEvents"] == want["events_" + mode.lower()]


@pytest.mark.parametrize("k", [2, 3, 5])
def test_wxf_automatic_fresh_vertices_are_distinct_names(ctx, k):
    """{{x}} -> {{x, y}} over k unary edges on one vertex, two steps. Every event creates a vertex
    no other event creates (the reference draws fresh vertices from one counter), so no two
    states have equal content: 1 + k + k(k - 1) classes. Up to isomorphism there are 3."""
    base = {"InitialStates": [[[0]] * k], "Rules": {"R": Rule([[0]], [[0, 1]])}, "Steps": 2}
    out = loads(ctx.run_wxf(dumps({**base, "Options": {"CanonicalizeStates": Symbol("Automatic")}})))
    assert out["NumStates"] == 1 + k + k * (k - 1)
    full = loads(ctx.run_wxf(dumps({**base, "Options": {"CanonicalizeStates": Symbol("Full")}})))
    assert full["NumStates"] == 3


def test_wxf_event_signature_keys(ctx):
    """A key list: {"InputState", "OutputState"} is Full's signature."""
    c = next(c for c in GOLD_NONE if c["name"] == "cycle4-automorphic")
    base = {"InitialStates": [c["init"]],
            "Rules": {f"R{k}": Rule(r["lhs"], r["rhs"]) for k, r in enumerate(c["rules"])}, "Steps": c["steps"]}
    out = loads(ctx.run_wxf(dumps({**base, "Options": {"CanonicalizeEvents": ["InputState", "OutputState"]}})))
    assert out["NumEvents"] == GOLD_FULL[c["name"]]["events_full"]


def test_wxf_requested_data_and_errors(ctx):
    out = loads(ctx.run_wxf(_job(requested=["NumStates", "EventsMinimal"])))
    assert set(out) == {"Events", "NumStates", "Warnings"}
    e0 = out["Events"][0]
    assert set(e0) == {"Id", "InputState", "OutputState"}
    with pytest.raises(hgx.HGXError, match="ShowGenesisEvents"):
        ctx.run_wxf(dumps({"InitialEdges": [[1, 2]], "Rules": {"R": Rule([[0, 1]], [[0, 1], [1, 2]])},
                           "Steps": 2, "Options": {"ShowGenesisEvents": Symbol("True")}}))


def _frame(b):
    return struct.pack("<Q", len(b)) + b


def test_worker_one_shot_and_frames(ctx):
    job = _job(steps=2)
    want = ctx.run_wxf(job)
    p = subprocess.run([WORKER], input=job, capture_output=True, timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    assert p.stdout == want
    bad = dumps({"InitialEdges": [[1, 2]], "Steps": 1})
    p = subprocess.run([WORKER, "--serve"], input=_frame(job) + _frame(bad) + _frame(job),
                       capture_output=True, timeout=120)
    assert p.returncode == 0, p.stderr.decode()
    frames, b = [], p.stdout
    while b:
        n = struct.unpack("<Q", b[:8])[0]
        frames.append(b[8:8 + n])
        b = b[8 + n:]
    assert frames == [want, b"", want]
    assert b"no Rules" in p.stderr


def test_worker_serve_socket(ctx, tmp_path):
    """--serve-socket: the port is published through a file, frames go over one loopback
    connection (the paclet's persistent worker, HypergraphRewriting.wl:197-230)."""
    import socket
    import time
    job = _job(steps=2)
    want = ctx.run_wxf(job)
    portfile = str(tmp_path / "port")
    proc = subprocess.Popen([WORKER, "--serve-socket", portfile], stdout=subprocess.PIPE,
                            stderr=subprocess.PIPE)
    try:
        t0 = time.time()
        while not os.path.exists(portfile):
            assert proc.poll() is None, proc.stderr.read().decode()
            assert time.time() - t0 < 60, "no port file"
            time.sleep(0.05)
        port = int(open(portfile).read())
        with socket.create_connection(("127.0.0.1", port), timeout=60) as s:
            def recv_exact(n):
                b = b""
                while len(b) < n:
                    c = s.recv(n - len(b))
                    assert c, "connection closed"
                    b += c
                return b
            got = []
            for frame in (job, b"not wxf", job):
                s.sendall(_frame(frame))
                n = struct.unpack("<Q", recv_exact(8))[0]
                got.append(recv_exact(n) if n else b"")
        assert got == [want, b"", want]
        assert proc.wait(timeout=60) == 0
    finally:
        if proc.poll() is None:
            proc.kill()
            proc.wait()
