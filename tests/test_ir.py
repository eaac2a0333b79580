"""The host exact canonicalizer (hgx_ir_hash_batch / hgx_ir_canonical_form), pinned three ways:

* against the reference's own IRCanonicalizer on 600 states (tests/golden/ir_vectors.json,
  written by tests/golden/make_ir_vectors.py from oracle/_ref/irref): hash, canonical form and
  per-edge canonical ranks;
* against the reference's golden matrix 'Full' rows: every raw state of each corpus case's
  multiway evolution (from the oracle) hashed here, the number of distinct hashes and the sum of
  mix64 over them equal the reference's exact-canonical state count and fingerprint;
* against what the mathematics fixes: the hash and the form do not move under any relabelling
  of the vertices or reordering of the edges, the form is a relabelling of the state, and
  states that differ in an isomorphism invariant get different forms.
Host code only: no device is used.
"""
import json
import os
import random

import numpy as np
import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from oracle import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
VEC = json.load(open(os.path.join(HERE, "golden", "ir_vectors.json")))["vectors"]
FULL = json.load(open(os.path.join(HERE, "golden", "golden_matrix_full.json")))["cases"]
NONE = json.load(open(os.path.join(HERE, "golden", "golden_matrix_none.json")))["cases"]
M64 = (1 << 64) - 1


def mix64(x):
    x ^= x >> 33; x = (x * 0xff51afd7ed558ccd) & M64
    x ^= x >> 33; x = (x * 0xc4ceb9fe1a85ec53) & M64
    x ^= x >> 33
    return x


def test_reference_vectors_hash():
    got = hgx.ir_hash([v["edges"] for v in VEC])
    bad = [i for i, v in enumerate(VEC) if int(got[i]) != v["hash"]]
    assert not bad, f"{len(bad)} hashes differ, first {bad[:5]}"


def test_reference_vectors_rank():
    _, ranks = hgx.ir_rank([v["edges"] for v in VEC])
    bad = [i for i, v in enumerate(VEC) if ranks[i] != v["rank"]]
    assert not bad, f"{len(bad)} rank vectors differ, first {bad[:5]}"


def test_reference_vectors_form():
    for v in VEC:
        h, form = hgx.ir_canonical_form(v["edges"])
        assert h == v["hash"]
        assert form == v["form"]
        assert len({x for e in form for x in e}) == v["vertex_count"]


@pytest.mark.parametrize("case", [c for c in NONE if c["name"] in FULL], ids=lambda c: c["name"])
def test_golden_matrix_full_rows(case):
    rules = [rule_from_lists(r["lhs"], r["rhs"], i) for i, r in enumerate(case["rules"])]
    req = Request(rules, case["init"], case["steps"], flags=0x3)
    o = po.evolve(req, with_edges=True)
    so, eo, vv = o.state_edges
    states = [[vv[eo[e]:eo[e + 1]].tolist() for e in range(int(so[s]), int(so[s + 1]))]
              for s in range(len(so) - 1)]
    want = FULL[case["name"]]
    assert o.counts["events"] == want["events"]
    distinct = set(int(h) for h in hgx.ir_hash(states))
    assert len(distinct) == want["ir_states"]
    assert sum(mix64(h) for h in distinct) & M64 == want["ir_fingerprint"]


def _relabel(edges, rnd):
    verts = sorted({x for e in edges for x in e})
    perm = rnd.sample(range(10 * len(verts) + 5), len(verts))
    m = dict(zip(verts, perm))
    out = [[m[x] for x in e] for e in edges]
    rnd.shuffle(out)
    return out


def test_invariant_under_relabelling_and_edge_order():
    rnd = random.Random(5)
    for v in VEC[::3]:
        for _ in range(3):
            h, form = hgx.ir_canonical_form(_relabel(v["edges"], rnd))
            assert (h, form) == (v["hash"], v["form"])


def test_form_is_a_relabelling():
    for v in VEC[::7]:
        _, form = hgx.ir_canonical_form(v["edges"])
        verts = sorted({x for e in v["edges"] for x in e})
        assert sorted(len(e) for e in form) == sorted(len(e) for e in v["edges"])
        # some bijection maps the state onto the form: degree sequences and multiplicities agree
        deg = lambda es: sorted(sum(e.count(x) for e in es) for x in {y for e in es for y in e})
        assert deg(form) == deg(v["edges"])
        assert sorted({x for e in form for x in e}) == list(range(len(verts)))


def test_non_isomorphic_states_differ():
    pairs = [
        ([[0, 1], [1, 2], [2, 0]], [[0, 1], [1, 2], [0, 2]]),          # directed 3-cycle vs not
        ([[0, 1], [1, 0]], [[0, 1], [0, 1]]),                          # 2-cycle vs double edge
        ([[0, 0], [0, 1]], [[0, 1], [1, 1]]),                          # loop at the tail / head
        ([[0, 1], [2, 3]], [[0, 1], [1, 2]]),                          # disjoint vs path
        ([[0, 1, 2], [2, 3, 4]], [[0, 1, 2], [1, 3, 4]]),             # shared vertex position
    ]
    for a, b in pairs:
        assert hgx.ir_canonical_form(a) != hgx.ir_canonical_form(b)
    # isomorphic (relabelled) pair
    assert hgx.ir_hash([[[0, 1, 2]], [[5, 7, 9]]])[0] == hgx.ir_hash([[[5, 7, 9]]])[0]


def test_empty_and_invalid():
    assert int(hgx.ir_hash([[]])[0]) == 0
    assert hgx.ir_canonical_form([]) == (0, [])
    with pytest.raises((hgx.HGXError, ValueError)):
        hgx.ir_hash([[[0] * 17]])
