"""hgx_match_count_noalloc against hgx_match_count: the same matches, in the same order, from a
state and an output buffer in the caller's device memory and no device allocation by the library
(its working memory is the fixed scratch array). Two calls in a row on one device, the second
with a smaller input, so the second runs over what the first left in the scratch; and a state
whose index cannot fit the scratch is refused with HGX_ERR_CAPACITY before anything runs."""
import random

import numpy as np
import pytest

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd import abi
from hypergraphrewritingengine_amd import workloads as W
from hypergraphrewritingengine_amd.rules import rule_from_lists

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers through the HIP runtime the library itself uses (no second runtime in the
    process): hipMalloc / hipMemcpy / hipFree by ctypes."""

    def __init__(self):
        import ctypes as C
        import os

        self.C = C
        self.hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, arr: np.ndarray) -> int:
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), max(arr.nbytes, 1)) == 0
        if arr.nbytes:
            assert self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0      # host to device
        self.live.append(p)
        return p.value

    def fill(self, n_bytes: int, byte: int) -> int:
        p = self.C.c_void_p()
        assert self.hip.hipMalloc(self.C.byref(p), max(n_bytes, 1)) == 0
        assert self.hip.hipMemset(p, byte, max(n_bytes, 1)) == 0
        self.live.append(p)
        return p.value

    def get(self, ptr: int, n: int, dtype) -> np.ndarray:
        out = np.zeros(n, dtype)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0     # device to host
        return out

    def free(self):
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


def _on_device(dev, edges, stride):
    ev = np.zeros((max(len(edges), 1), stride), np.uint32)
    ar = np.zeros(max(len(edges), 1), np.uint8)
    for i, e in enumerate(edges):
        ev[i, :len(e)] = e
        ar[i] = len(e)
    nv = int(ev.max()) + 1 if len(edges) else 1
    return dev.put(ev), dev.put(ar), nv


def test_noalloc_join_equals_the_allocating_join():
    ctx = hgx.Context(0)
    dev = _Dev()
    try:
        rnd = random.Random(9)
        small = [[rnd.randrange(300), rnd.randrange(300)] for _ in range(3000)]
        cases = [W.config3b(),                                # 200,000 binary edges: triangle LHS
                 (rule_from_lists([[0, 1], [1, 2], [2, 3], [3, 4]], [[0]]), small)]
        for call, (rule, edges) in enumerate(cases):
            k = len(rule.lhs)
            n_ref, m_ref, _ = ctx.match([rule], 0, edges, want_edges=True, cap=10**7)
            d_ev, d_ar, nv = _on_device(dev, edges, 2)
            cap = n_ref + 5
            out = dev.fill(cap * k * 4, 0xFF)
            n = ctx.match_noalloc([rule], 0, d_ev, d_ar, len(edges), 2, nv, out, cap)
            got = dev.get(out, cap * k, np.uint32)
            assert n == n_ref, (call, n, n_ref)
            assert np.array_equal(got[:n * k].reshape(-1, k), m_ref), call
            assert (got[n * k:] == 0xFFFFFFFF).all()             # nothing past the matches
            # count only, and a cap below the count: the prefix
            assert ctx.match_noalloc([rule], 0, d_ev, d_ar, len(edges), 2, nv) == n_ref
            half = max(1, n_ref // 2)
            out2 = dev.fill(half * k * 4, 0)
            assert ctx.match_noalloc([rule], 0, d_ev, d_ar, len(edges), 2, nv, out2, half) == n_ref
            assert np.array_equal(dev.get(out2, half * k, np.uint32).reshape(-1, k), m_ref[:half])
        # an index beyond the fixed scratch is refused with its reason
        m = 4_000_000
        big = np.random.default_rng(1).integers(0, 1000, size=(m, 2), dtype=np.uint32)
        d_ev = dev.put(big)
        d_ar = dev.put(np.full(m, 2, np.uint8))
        with pytest.raises(hgx.HGXError) as ei:
            ctx.match_noalloc([cases[0][0]], 0, d_ev, d_ar, m, 2, 1000)
        assert ei.value.code == abi.HGX_ERR_CAPACITY and "scratch" in str(ei.value), ei.value
    finally:
        dev.free()
        ctx.close()
