"""The WXF job envelope without a device: the library's WXF reader (through hgx_wxf_validate)
on byte strings produced by the Wolfram kernel, the test encoder's round trip, and the job
checks of the FFI (hypergraph_ffi.cpp:136-376, 520-636)."""
import pytest

import hypergraphrewritingengine_amd as hgx
from tests.wxf_py import Rule, Symbol, dumps, loads


def _job(**kw):
    job = {"InitialEdges": [[1, 2], [2, 3], [3, 1]],
           "Rules": {"Rule1": Rule([[0, 1], [1, 2]], [[0, 1], [1, 2], [2, 0]])},
           "Steps": 3}
    job.update(kw)
    return job


def test_kernel_generated_association_parses():
    # BinarySerialize[<|int8 -> -42, int64 -> 123456789, positive -> 100, zero -> 0|>] with
    # symbol keys, as the reference's WXF tests hold it (wxf/tests/test_wxf.cpp:746-766). It is a
    # well-formed WXF value but not a job: the reader accepts it and the job check refuses it.
    hexs = ("383a41042d730b476c6f62616c60696e743843d62d730c476c6f62616c60696e7436346915cd5b072d73"
            "0f476c6f62616c60706f73697469766543642d730b476c6f62616c607a65726f4300")
    b = bytes.fromhex(hexs)
    d = loads(b)
    assert d == {"Global`int8": -42, "Global`int64": 123456789, "Global`positive": 100, "Global`zero": 0}
    ok, msg = hgx.wxf_validate(b)
    assert not ok and "no Rules" in msg
    ok, msg = hgx.wxf_validate(b[:-3])
    assert not ok and ("end of WXF data" in msg or "exceeds the input" in msg), msg
    ok, msg = hgx.wxf_validate(b[:-1])
    assert not ok and "end of WXF data" in msg, msg


def test_encoder_roundtrip_and_valid_job():
    job = _job(Options={"CausalTransitiveReduction": Symbol("False"),
                        "RequestedData": ["States", "NumStates"]})
    b = dumps(job)
    back = loads(b)
    assert back["Steps"] == 3 and back["InitialEdges"] == [[1, 2], [2, 3], [3, 1]]
    assert back["Options"]["RequestedData"] == ["States", "NumStates"]
    assert hgx.wxf_validate(b) == (True, "")
    # packed arrays are read as nested lists (BinarySerialize packs numeric lists)
    packed = bytes.fromhex("383a") + b"A\x03" + b"-" + b"S\x0cInitialEdges" + \
        bytes([0xC1, 0x00, 0x02, 0x02, 0x02, 1, 2, 2, 3]) + b"-S\x05Steps" + b"C\x02" + \
        b"-S\x05Rules" + dumps({"R": Rule([[0, 1]], [[0, 1], [1, 2]])})[2:]
    assert hgx.wxf_validate(packed) == (True, "")


@pytest.mark.parametrize("job,needle", [
    (_job(Steps=-1), "non-negative"),
    (_job(Rules={"R": Rule([[0, 40]], [[0]])}), "maximum is 31"),
    (_job(Rules={"R": Rule([[0, -1]], [[0]])}), "negative pattern variable"),
    (_job(Rules={"R": Rule([[]], [[0]])}), "empty left-hand side"),
    (_job(Rules={"R": [[0, 1], [1, 2]]}), "Rule[lhs, rhs]"),
    (_job(Options={"CanonicalizeStates": Symbol("Sometimes")}), "unknown state canonicalization"),
    (_job(Options={"MaxStatesPerStep": -1}), "must be non-negative"),
    (_job(Options={"ExplorationProbability": Symbol("High")}), "must be a number"),
    (_job(Options={"RuleWeights": 0.5}), "RuleWeights"),
    (_job(Options={"ShowGenesisEvents": Symbol("True")}), "True is not served"),
    (_job(Options={"ExploreFromCanonicalStatesOnly": Symbol("True"), "CanonicalizeStates": Symbol("Automatic")}),
     "content identity"),
    (_job(Op="Open"), "sessions"),
    ({"InitialEdges": [[1, 2]], "Steps": 2}, "no Rules"),
    (_job(InitialEdges=[[-1, -2]]), "no non-empty initial state"),
])
def test_job_checks(job, needle):
    ok, msg = hgx.wxf_validate(dumps(job))
    assert not ok and needle in msg, msg


def test_compressed_and_garbage_are_refused():
    assert not hgx.wxf_validate(b"8C:x\x9c")[0]
    assert not hgx.wxf_validate(b"\xff\xff\x43\x00")[0]
    assert not hgx.wxf_validate(dumps(_job()) + b"C\x00")[0]          # trailing bytes
    # a list length the input cannot hold is refused before anything is allocated
    assert not hgx.wxf_validate(b"8:f\xff\xff\xff\xff\x0fs\x04List")[0]


def test_array_dims_with_no_data_are_refused_before_allocation():
    # PackedArray int8 of dims {2^28, 0}: no element bytes, 2^28 empty lists. The reader bounds
    # the list nodes by the input left, so it refuses this before reserving anything.
    ok, msg = hgx.wxf_validate(bytes.fromhex("383ac10002808080800100"))
    assert not ok and "larger than the input" in msg, msg
    # a small array of empty rows ({{}, {}}: dims {2, 0}) still parses; it is no job
    ok, msg = hgx.wxf_validate(bytes.fromhex("383ac100020200"))
    assert not ok and "must be an Association" in msg, msg


def test_int64_vertex_labels_are_renumbered():
    # labels above 2^32-1 are renumbered by first appearance, as the FFI does
    # (hypergraph_ffi.cpp:598-636), so the job is accepted
    big = 1 << 40
    ok, msg = hgx.wxf_validate(dumps(_job(InitialEdges=[[big, big + 1], [big + 1, 7], [7, big]])))
    assert ok, msg
