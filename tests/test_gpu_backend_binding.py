"""run_gpu_evolution (include/hgx_gpu_backend.hpp, the counterpart of
paclet_source/hg_gpu_backend.hpp:119) called from a compiled C++ caller (tests/cpp/
gpu_backend_test): each GpuJob's WXF result is byte-for-byte the result of the same job sent as a
WXF envelope to hgx_run_wxf (run_rewriting_core), whose content tests/test_gpu_wxf.py checks
against the oracle (tests/test_gpu_quotient.py for quotient exploration); a job asking for genesis
events is refused with its reason."""
import os
import subprocess
import tempfile

import pytest

from tests.wxf_py import Rule, Symbol, dumps, loads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CALLER = os.path.join(HERE, "cpp", "gpu_backend_test")


def _jobs():
    r1 = {"Rule1": Rule([[0, 1], [1, 2]], [[0, 1], [1, 2], [2, 0]])}
    i1 = [[[1, 2], [2, 3], [3, 1]]]
    r2 = {"a": Rule([[0, 1], [0, 2]], [[0, 2], [0, 3], [1, 3], [2, 3]]), "b": Rule([[0, 0]], [[0, 1], [1, 0]])}
    i2 = [[[1, 1], [1, 1]], [[5, 6], [6, 5]]]
    yield 0, {"InitialStates": i1, "Rules": r1, "Steps": 3,
              "Options": {"CanonicalizeStates": Symbol("Full")}}
    yield 1, {"InitialStates": i1, "Rules": r1, "Steps": 3,
              "Options": {"CanonicalizeStates": Symbol("None"), "CanonicalizeEvents": Symbol("Automatic"),
                          "IncludeCanonicalHashes": Symbol("True"),
                          "CausalTransitiveReduction": Symbol("False")}}
    yield 2, {"InitialStates": i2, "Rules": r2, "Steps": 3,
              "Options": {"CanonicalizeStates": Symbol("Full")}}
    yield 3, {"InitialStates": i2, "Rules": r2, "Steps": 3,
              "Options": {"CanonicalizeStates": Symbol("Full"), "ExplorationProbability": 0.5,
                          "TransitionRate": 0.75, "RandomSeed": 12345}}
    yield 4, {"InitialStates": i2, "Rules": r2, "Steps": 3,
              "Options": {"CanonicalizeStates": Symbol("Automatic"), "CanonicalizeEvents": Symbol("Full")}}
    yield 5, {"InitialStates": i2, "Rules": r2, "Steps": 3,
              "Options": {"CanonicalizeStates": Symbol("Full"), "CanonicalizeEvents": Symbol("Automatic"),
                          "ExploreFromCanonicalStatesOnly": Symbol("True"),
                          "QuotientInitialStates": Symbol("True")}}


def test_run_gpu_evolution_equals_the_envelope(ctx):
    if not os.path.exists(CALLER):
        pytest.fail("tests/cpp/gpu_backend_test is not built; run __graft_entry__.build()")
    with tempfile.TemporaryDirectory() as td:
        p = subprocess.run([CALLER, td], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        assert "refused 6:" in p.stdout and "ShowGenesisEvents" in p.stdout, p.stdout
        assert "progress 6" in p.stdout, p.stdout
        for i, job in _jobs():
            got = open(os.path.join(td, f"case{i}.wxf"), "rb").read()
            want = ctx.run_wxf(dumps(job))
            assert got == want, f"case {i}: run_gpu_evolution and the envelope differ"
            out = loads(got)
            assert out["NumEvents"] > 0 and len(out["States"]) == out["NumStates"]
