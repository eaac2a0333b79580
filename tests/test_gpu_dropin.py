"""The C++ surface as a drop-in: the reference engine's own driver (oracle/ref_driver.cpp, the
code that drives hypergraph::ParallelEvolutionEngine for the golden vectors), compiled against
include/hgx_engine.hpp with only its include and namespace changed (tests/cpp/ref_driver_hgx),
runs on the device and must print the reference engine's counts and dump the reference engine's
normalised outputs: the digests in tests/golden/ref_vectors.json were written from the
reference engine itself (state-hash set, event keys, causal pairs from get_causal_edges(),
branchial pairs from get_branchial_edges(); gpu/tests/test_gpu_vs_cpu_differential.cpp:64-160)."""
import json
import os
import subprocess
import tempfile

import pytest

from oracle import pyoracle as po
from tests.golden.make_ref_vectors import digest_reference, load_cases

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "ref_driver_hgx")
VECTORS = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))["cases"]


def _parse_dump(path):
    S, E, Cz, B = {}, [], [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if t[0] == "S":
                S[int(t[1])] = (int(t[2]), int(t[3]))
            elif t[0] == "E":
                E.append(tuple(int(x) for x in t[1:]))
            elif t[0] == "C":
                Cz.append((int(t[1]), int(t[2])))
            elif t[0] == "B":
                B.append((int(t[1]), int(t[2])))
    return {"states": S, "events": E, "causal": Cz, "branchial": B}


def test_reference_driver_over_the_device_engine():
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/ref_driver_hgx is not built; run __graft_entry__.build()")
    cases = dict(load_cases())
    # every golden case, and 8 of the batched draws, with reduction on and off
    picked = [v for v in VECTORS if not v["name"].startswith("batched-") or
              int(v["name"].split("-")[1]) < 8]
    assert len(picked) >= 40
    with tempfile.TemporaryDirectory() as td:
        for v in picked:
            req = cases[v["name"]]
            job = os.path.join(td, "job.txt")
            dump = os.path.join(td, "dump.txt")
            with open(job, "w") as f:
                f.write(req.job_text(tr=bool(v["tr"])))
            p = subprocess.run([DRIVER, job, "--dump", dump], capture_output=True, text=True, timeout=120)
            assert p.returncode == 0, (v["name"], p.stderr)
            got = json.loads(p.stdout.strip().splitlines()[-1])
            for k in ("raw_states", "events", "canonical_states", "causal_triples", "causal_pairs",
                      "branchial_pairs", "fingerprint"):
                assert got[k] == v["counts"][k], (v["name"], v["tr"], k, got[k], v["counts"][k])
            d = _parse_dump(dump)
            assert len(d["causal"]) == got["causal_triples"]          # one entry per triple
            assert digest_reference(d) == v["digest"], (v["name"], v["tr"])
