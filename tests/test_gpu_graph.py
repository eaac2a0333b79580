"""A prepared batch's pass recorded into one HIP graph (hgx_batch_capture / hgx_replay_launch):
every replay gives exactly the eager pass's results -- counts, per-state hashes, steps and
creators, events, causal pairs with reduction tags and branchial pairs -- replayed several times
with the program's own eager work and synchronisations between the replays."""
import numpy as np
import pytest

from hypergraphrewritingengine_amd import workloads as W
from tests.parity import gpu_digest

pytestmark = pytest.mark.gpu

FLAGS = 0x1 | 0x2 | 0x4 | 0x100


def test_graph_replays_equal_the_eager_pass(ctx):
    reqs = [W.batched_run(i, cap=1024, flags=FLAGS) for i in range(384)]
    reqs += [W.served_run(i, cap=512, flags=FLAGS) for i in (3, 10, 37, 90)]     # 300-2,000 edge roots
    pb = ctx.prepare(reqs)
    eager = ctx.evolve_batch(pb)
    want = [(eager[i].counts, gpu_digest(eager[i])) for i in range(len(reqs))]
    assert sum(c["raw_states"] for c, _ in want) > 50_000
    rp = ctx.capture(pb)
    try:
        for k in range(3):
            res = rp.launch()
            for i in range(len(reqs)):
                assert res[i].counts == want[i][0], (k, i)
                assert gpu_digest(res[i]) == want[i][1], (k, i)
            del res
            # the program's own eager work between replays
            other = ctx.evolve_batch(reqs[:24])
            for i in range(24):
                assert other[i].counts == want[i][0]
    finally:
        rp.close()
    pb.close()


def test_graph_of_the_batched_workload_counts(ctx):
    reqs = [W.batched_run(i) for i in range(4096)]
    pb = ctx.prepare(reqs)
    eager = ctx.evolve_batch(pb).counts()
    rp = ctx.capture(pb)
    try:
        for _ in range(2):
            got = rp.launch().counts()
            assert np.array_equal(got, eager)
    finally:
        rp.close()
    pb.close()
