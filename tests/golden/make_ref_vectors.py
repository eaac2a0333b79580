"""Writes tests/golden/ref_vectors.json from the REFERENCE engine (oracle/_ref/hgref, built
from the reference sources by oracle/Makefile). Run in the build container:

    make -C oracle ref && python tests/golden/make_ref_vectors.py

Each vector is an input (named below) and digests of the reference's normalised output, the
comparison of gpu/tests/test_gpu_vs_cpu_differential.cpp:64-160: the set of WL state hashes,
the multiset of event keys fnv(in, out, rule, step), the causal pair multiset (pairs of event
keys, deduplicated by event pair; TR view when tr=1) and the branchial pair multiset.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from hypergraphrewritingengine_amd import workloads as W  # noqa: E402
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def _digest_multiset(values) -> dict:
    vals = sorted(int(v) for v in values)
    h = hashlib.sha256()
    for v in vals:
        h.update(v.to_bytes(8, "little"))
    return {"n": len(vals), "sha256": h.hexdigest()[:32]}


def digest_normalized(norm) -> dict:
    hs, evk, ck, bk = norm
    ex = lambda c: [k for k, n in c.items() for _ in range(n)]  # noqa: E731
    return {"states": _digest_multiset(hs), "events": _digest_multiset(ex(evk)),
            "causal": _digest_multiset(ex(ck)), "branchial": _digest_multiset(ex(bk))}


def digest_oracle(res, tr: bool) -> dict:
    sh = {i: int(x) for i, x in enumerate(res.state_hash)}
    evl = [(e, *map(int, res.events[e])) for e in range(len(res.events))]
    cp = res.causal[res.causal_tr.astype(bool)] if tr else res.causal
    return digest_normalized(po.normalize(sh, evl, [tuple(map(int, p)) for p in cp],
                                          [tuple(map(int, p)) for p in res.branchial]))


def digest_reference(d: dict) -> dict:
    sh = {sid: h for sid, (h, st) in d["states"].items()}
    evl = [tuple(e) for e in d["events"]]
    pairs = sorted(set(d["causal"]))   # the reference lists causal TRIPLES; pairs dedup them
    return digest_normalized(po.normalize(sh, evl, pairs, d["branchial"]))


def load_cases():
    gold = json.load(open(os.path.join(HERE, "golden_matrix_none.json")))
    for c in gold["cases"]:
        rules = [rule_from_lists(r["lhs"], r["rhs"], i) for i, r in enumerate(c["rules"])]
        yield c["name"], Request(rules, c["init"], c["steps"], flags=0x3)
    yield "config1", W.config1()
    yield "config2-s4", W.config2(steps=4)
    yield "multi-root", Request([rule_from_lists([[0, 1], [1, 2]], [[0, 2], [2, 3], [3, 1]])],
                                [[[0, 1], [1, 2]], [[5, 5], [5, 6]], [[1, 2], [2, 3]]], 3,
                                flags=0x3, multi_root=True)
    for i in range(40):
        req = W.batched_run(i, cap=0)
        yield f"batched-{i}", req


def main():
    if not po.ref_available():
        sys.exit("build the reference first: make -C oracle ref")
    out = []
    for name, req in load_cases():
        if name.startswith("batched-"):
            # keep runs whose uncapped unfolding is small enough for a fixture
            probe = Request(req.rules, req.init, req.steps, state_cap=20000, flags=0x3)
            if po.evolve(probe).counts["truncated"]:
                continue
        for tr in (False, True):
            rc, d = po.ref_evolve(req, threads=4, tr=tr, dump=True)
            out.append({"name": name, "tr": int(tr), "counts": rc, "digest": digest_reference(d)})
            print(name, tr, rc["raw_states"], flush=True)
    json.dump({"source": "reference engine (oracle/_ref/hgref) built from the reference sources",
               "cases": out}, open(os.path.join(HERE, "ref_vectors.json"), "w"), indent=0)


if __name__ == "__main__":
    main()
