"""hypergraphrewritingengine_amd -- MI355X-native multiway hypergraph evolution.

Python face of the C ABI in include/hgx.h (libhgx.so, HIP kernels for gfx950). The C++ face
is include/hgx_engine.hpp. Nothing here falls back to a CPU path: without the built library
or without a gfx950 device every entry point raises.
"""
from __future__ import annotations

import collections.abc
import ctypes as C
import os

import numpy as np

from . import abi
from .rules import (Request, Rule, RuleBuilder, encode_roots, encode_rules, encode_state,
                    make_rule, rule_from_lists)

__all__ = ["Context", "Result", "ResultList", "Request", "Rule", "RuleBuilder", "make_rule", "rule_from_lists",
           "encode_rules", "encode_state", "encode_roots", "library_path", "load_library",
           "HGXError", "EXPORTED_SYMBOLS", "wxf_validate", "ir_hash", "ir_rank", "ir_canonical_form"]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# Every symbol include/hgx.h declares.
EXPORTED_SYMBOLS = [
    "hgx_ctx_create", "hgx_ctx_destroy", "hgx_last_error", "hgx_device_info",
    "hgx_evolve_batch", "hgx_request_submit", "hgx_serve_step", "hgx_request_poll",
    "hgx_request_wait", "hgx_request_footprint", "hgx_result_counts", "hgx_result_states",
    "hgx_result_events", "hgx_result_state_edges", "hgx_result_causal", "hgx_result_branchial",
    "hgx_result_timing", "hgx_result_release", "hgx_result_release_many", "hgx_wl_hash_batch",
    "hgx_match_count",
    "hgx_profile_enable", "hgx_profile_read", "hgx_result_counts_many", "hgx_serve_drain",
    "hgx_ctx_set_option", "hgx_batch_prepare", "hgx_batch_evolve", "hgx_batch_release",
    "hgx_result_consumed", "hgx_run_wxf", "hgx_buffer_free", "hgx_wxf_validate",
    "hgx_ir_hash_batch", "hgx_ir_rank_batch", "hgx_ir_canonical_form", "hgx_result_status",
    "hgx_result_causal_triples", "hgx_match_count_noalloc", "hgx_autotune", "hgx_tune_load",
    "hgx_tune_get", "hgx_batch_capture", "hgx_replay_launch", "hgx_replay_release",
    "hgx_calibration_read", "hgx_result_unfold",
]

OPTION_FORWARD = 1     # HGX_OPTION_FORWARD
OPTION_DELTA_CAP = 2   # HGX_OPTION_DELTA_CAP
OPTION_EXPAND_TEAM = 3  # HGX_OPTION_EXPAND_TEAM
OPTION_JOIN_TEAM = 4    # HGX_OPTION_JOIN_TEAM
OPTION_EXPAND_LANES = 5 # HGX_OPTION_EXPAND_LANES
OPTION_JOIN_LANES = 6   # HGX_OPTION_JOIN_LANES
OPTION_COUNTER_FLUSH = 7  # HGX_OPTION_COUNTER_FLUSH


class HGXError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hgx error {code}: {msg}")
        self.code = code


def library_path() -> str:
    return os.path.join(_HERE, "libhgx.so")


def load_library():
    """Load libhgx.so (built in-tree by __graft_entry__.build()). Raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise HGXError(abi.HGX_ERR_NO_DEVICE,
                       f"{path} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.hgx_ctx_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
    L.hgx_ctx_destroy.argtypes = [vp]
    L.hgx_ctx_destroy.restype = None
    L.hgx_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.hgx_last_error.restype = C.c_size_t
    L.hgx_device_info.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.hgx_device_info.restype = C.c_size_t
    L.hgx_evolve_batch.argtypes = [vp, C.POINTER(abi.RunDesc), C.c_uint32, C.POINTER(vp)]
    L.hgx_request_submit.argtypes = [vp, C.POINTER(abi.RunDesc), C.POINTER(C.c_uint64)]
    L.hgx_serve_step.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.hgx_request_poll.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.hgx_request_wait.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.hgx_request_footprint.argtypes = [C.POINTER(abi.RunDesc)]
    L.hgx_request_footprint.restype = C.c_uint64
    L.hgx_result_counts.argtypes = [vp, C.POINTER(abi.Counts)]
    L.hgx_result_states.argtypes = [vp, vp, vp, vp, C.c_uint64]
    L.hgx_result_events.argtypes = [vp, vp, C.c_uint64]
    L.hgx_result_state_edges.argtypes = [vp, vp, vp, vp, vp, vp]
    L.hgx_result_causal.argtypes = [vp, vp, vp, C.c_uint64]
    L.hgx_result_branchial.argtypes = [vp, vp, C.c_uint64]
    L.hgx_result_timing.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.hgx_result_status.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.hgx_result_causal_triples.argtypes = [vp, vp, vp, C.c_uint64]
    L.hgx_autotune.argtypes = [vp, vp, C.c_uint32, C.c_char_p, C.c_char_p, C.c_size_t]
    L.hgx_batch_capture.argtypes = [vp, vp, C.POINTER(vp)]
    L.hgx_calibration_read.argtypes = [vp, C.c_uint64]
    L.hgx_replay_launch.argtypes = [vp, vp, C.POINTER(vp)]
    L.hgx_replay_release.argtypes = [vp]
    L.hgx_replay_release.restype = None
    L.hgx_tune_load.argtypes = [vp, C.c_char_p]
    L.hgx_tune_get.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.hgx_tune_get.restype = C.c_size_t
    L.hgx_match_count_noalloc.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, C.c_uint32,
                                          C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), vp, C.c_uint64]
    L.hgx_result_release.argtypes = [vp]
    L.hgx_result_release.restype = None
    L.hgx_result_release_many.argtypes = [vp, C.c_uint32]
    L.hgx_result_release_many.restype = None
    L.hgx_wl_hash_batch.argtypes = [vp, vp, C.c_uint32, vp, vp, vp]
    L.hgx_ir_hash_batch.argtypes = [vp, vp, C.c_uint32, vp, vp, vp]
    L.hgx_ir_rank_batch.argtypes = [vp, vp, C.c_uint32, vp, vp, vp, vp]
    L.hgx_ir_canonical_form.argtypes = [vp, vp, C.c_uint32, vp, vp, C.POINTER(C.c_uint64)]
    L.hgx_match_count.argtypes = [vp, vp, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp, C.c_uint32,
                                  C.POINTER(C.c_uint64), vp, C.c_uint64, C.POINTER(C.c_float)]
    L.hgx_result_counts_many.argtypes = [vp, C.c_uint32, C.POINTER(abi.Counts)]
    L.hgx_serve_drain.argtypes = [vp, C.POINTER(C.c_uint64), vp, C.c_uint32, C.POINTER(C.c_uint32)]
    L.hgx_profile_enable.argtypes = [vp, C.c_int]
    L.hgx_ctx_set_option.argtypes = [vp, C.c_uint32, C.c_int64]
    L.hgx_batch_prepare.argtypes = [vp, C.POINTER(abi.RunDesc), C.c_uint32, C.POINTER(vp)]
    L.hgx_batch_evolve.argtypes = [vp, vp, C.POINTER(vp)]
    L.hgx_batch_release.argtypes = [vp]
    L.hgx_batch_release.restype = None
    L.hgx_result_consumed.argtypes = [vp, vp, vp, C.c_uint64]
    L.hgx_result_unfold.argtypes = [vp, vp, C.POINTER(abi.Unfolding)]
    L.hgx_run_wxf.argtypes = [vp, C.c_char_p, C.c_size_t, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t)]
    L.hgx_buffer_free.argtypes = [C.POINTER(C.c_uint8)]
    L.hgx_buffer_free.restype = None
    L.hgx_wxf_validate.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]
    L.hgx_profile_read.argtypes = [vp, C.POINTER(abi.KernelStat), C.c_uint32, C.POINTER(C.c_uint32)]
    _LIB = L
    return L


_COUNTS_DTYPE = np.dtype([(k, np.uint64 if t is C.c_uint64 else np.uint32) for k, t in abi.Counts._fields_])


class _BatchHandles:
    """The hgx_result* handles of one batch, released together when the last holder goes."""

    def __init__(self, lib, outs, n, ctx):
        self._lib, self.outs, self.n, self._ctx = lib, outs, n, ctx

    def __del__(self):
        try:
            if self.outs is not None:
                self._lib.hgx_result_release_many(self.outs, self.n)
                self.outs = None
        except Exception:
            pass


class Result:
    """One request's result (hgx_result*). Arrays are read back lazily."""

    def __init__(self, lib, handle, ctx, counts=None, row=None, owner=None):
        self._lib = lib
        self._h = handle
        self._ctx = ctx           # keep the context alive while results exist
        self._row = row           # (structured array, index) of a batch's counts, read lazily
        self._counts = counts
        self._owner = owner       # a batch's handles: released with the batch, not one by one
        if counts is None and row is None:
            c = abi.Counts()
            self._check(lib.hgx_result_counts(handle, C.byref(c)))
            self._counts = c.as_dict()

    @property
    def counts(self) -> dict:
        if self._counts is None:
            arr, i = self._row
            self._counts = {k: int(arr[k][i]) for k in arr.dtype.names}
        return self._counts

    def _check(self, rc):
        if rc != abi.HGX_OK:
            raise HGXError(rc, "result readback failed")

    def states(self):
        n = self.counts["raw_states"]
        h = np.zeros(n, np.uint64)
        s = np.zeros(n, np.uint32)
        c = np.zeros(n, np.uint32)
        self._check(self._lib.hgx_result_states(self._h, h.ctypes.data, s.ctypes.data, c.ctypes.data, n))
        return h, s, c

    def events(self):
        n = self.counts["events"]
        ev = np.zeros((n, 4), np.uint32)
        if n:
            self._check(self._lib.hgx_result_events(self._h, ev.ctypes.data, n))
        return ev

    def causal(self):
        n = self.counts["causal_pairs"]
        p = np.zeros((n, 2), np.uint32)
        t = np.zeros(n, np.uint8)
        self._check(self._lib.hgx_result_causal(self._h, p.ctypes.data if n else None,
                                                t.ctypes.data if n else None, n))
        return p, t

    def branchial(self):
        n = self.counts["branchial_pairs"]
        p = np.zeros((n, 2), np.uint32)
        if n:
            self._check(self._lib.hgx_result_branchial(self._h, p.ctypes.data, n))
        return p

    def state_edges(self):
        te, tv = C.c_uint64(), C.c_uint64()
        self._check(self._lib.hgx_result_state_edges(self._h, None, None, None, C.byref(te), C.byref(tv)))
        n = self.counts["raw_states"]
        so = np.zeros(n + 1, np.uint64)
        eo = np.zeros(te.value + 1, np.uint32)
        vv = np.zeros(max(tv.value, 1), np.uint32)
        self._check(self._lib.hgx_result_state_edges(self._h, so.ctypes.data, eo.ctypes.data,
                                                     vv.ctypes.data, None, None))
        return so, eo, vv

    def consumed(self):
        """(offsets, positions): event e consumed the edges at positions[offsets[e]:offsets[e+1]]
        of its input state's edge list, in pattern order (needs HGX_REC_STATES)."""
        n = self.counts["events"]
        off = np.zeros(n + 1, np.uint32)
        self._check(self._lib.hgx_result_consumed(self._h, off.ctypes.data, None, 0))
        pos = np.zeros(max(1, int(off[-1])), np.uint16)
        self._check(self._lib.hgx_result_consumed(self._h, off.ctypes.data, pos.ctypes.data, int(off[-1])))
        return off, pos[:int(off[-1])]

    def unfold(self) -> dict:
        """The raw unfolding beneath this result (hgx_result_unfold): every class expansion
        replayed on every raw instance, on the device. Needs HGX_REC_STATES."""
        u = abi.Unfolding()
        rc = self._lib.hgx_result_unfold(self._ctx._h, self._h, C.byref(u))
        if rc != abi.HGX_OK:
            raise self._ctx._err(rc)
        return u.as_dict()

    def timing(self):
        d, l = C.c_uint64(), C.c_uint64()
        self._check(self._lib.hgx_result_timing(self._h, C.byref(d), C.byref(l)))
        return d.value, l.value

    def status(self):
        """(status, reason): HGX_OK for a complete result, HGX_ERR_CAPACITY for a partial one (a
        generation did not fit the context's HBM budget; the result holds the ones before it)."""
        buf = C.create_string_buffer(1024)
        rc = self._lib.hgx_result_status(self._h, buf, 1024)
        return rc, buf.value.decode()

    def close(self):
        if self._h:
            if self._owner is None:
                self._lib.hgx_result_release(self._h)
            self._h = None
            self._owner = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ResultList(collections.abc.Sequence):
    """The results of one batch, in request order. Each Result is wrapped on first access; the
    batch's handles are released together once the list and every Result taken from it are gone."""

    def __init__(self, handles: _BatchHandles, counts):
        self._hd = handles
        self._arr = counts
        self._items = [None] * handles.n

    def __len__(self):
        return self._hd.n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(self._hd.n))]
        n = self._hd.n
        if i < 0:
            i += n
        if not 0 <= i < n:
            raise IndexError("result index out of range")
        r = self._items[i]
        if r is None:
            hd = self._hd
            r = Result(hd._lib, C.c_void_p(hd.outs[i]), hd._ctx, row=(self._arr, i), owner=hd)
            self._items[i] = r
        return r

    def counts(self):
        """Every result's counts as one numpy structured array (one row per request)."""
        return self._arr


def wxf_validate(job: bytes):
    """Parse and check a WXF job without a device: (True, "") or (False, reason)."""
    L = load_library()
    buf = C.create_string_buffer(1024)
    rc = L.hgx_wxf_validate(job, len(job), buf, 1024)
    return rc == abi.HGX_OK, buf.value.decode()


def _flatten_states(states):
    offs = [0]
    edges: list = []
    for st in states:
        edges.extend(st)
        offs.append(len(edges))
    eo, vv = encode_state(edges) if edges else (np.zeros(1, np.uint32), np.zeros(1, np.uint32))
    if vv.size == 0:
        vv = np.zeros(1, np.uint32)
    return np.asarray(offs, np.uint64), eo, vv


def ir_hash(states) -> np.ndarray:
    """Exact (individualization-refinement) canonical hashes of a list of states (each a list
    of edges), computed on the host -- the identity of CanonicalizeStates -> Full. No device
    needed."""
    L = load_library()
    so, eo, vv = _flatten_states(states)
    out = np.zeros(max(1, len(states)), np.uint64)
    rc = L.hgx_ir_hash_batch(None, so.ctypes.data, len(states), eo.ctypes.data, vv.ctypes.data,
                             out.ctypes.data)
    if rc != abi.HGX_OK:
        raise HGXError(rc, "invalid state")
    return out[:len(states)]


def ir_rank(states):
    """(hashes, ranks): the exact canonical hash of each state and, per state, each edge's
    canonical rank (its place in the state's canonical edge order)."""
    L = load_library()
    so, eo, vv = _flatten_states(states)
    out = np.zeros(max(1, len(states)), np.uint64)
    rk = np.zeros(max(1, int(so[-1]) if len(so) else 1), np.uint32)
    rc = L.hgx_ir_rank_batch(None, so.ctypes.data, len(states), eo.ctypes.data, vv.ctypes.data,
                             out.ctypes.data, rk.ctypes.data)
    if rc != abi.HGX_OK:
        raise HGXError(rc, "invalid state")
    return out[:len(states)], [rk[int(so[i]):int(so[i + 1])].tolist() for i in range(len(states))]


def ir_canonical_form(edges):
    """(hash, canonical edges): one state relabelled 0..n-1, its edges in canonical order."""
    L = load_library()
    eo, vv = encode_state(edges) if len(edges) else (np.zeros(1, np.uint32), np.zeros(1, np.uint32))
    if vv.size == 0:
        vv = np.zeros(1, np.uint32)
    oe = np.zeros(len(edges) + 1, np.uint32)
    ov = np.zeros(max(1, int(eo[-1])), np.uint32)
    h = C.c_uint64()
    rc = L.hgx_ir_canonical_form(eo.ctypes.data, vv.ctypes.data, len(edges), oe.ctypes.data,
                                 ov.ctypes.data, C.byref(h))
    if rc != abi.HGX_OK:
        raise HGXError(rc, "invalid state" if rc == abi.HGX_ERR_INVALID else "search depth exceeded")
    return int(h.value), [ov[oe[k]:oe[k + 1]].tolist() for k in range(len(edges))]


class PreparedBatch:
    """hgx_run_desc records of a list of requests, with the buffers they point into, and (when
    made by Context.prepare) the requests parsed and staged in device memory (hgx_batch)."""

    def __init__(self, requests, ctx=None):
        self._keep: list = []
        self.n = len(requests)
        self.descs = (abi.RunDesc * max(self.n, 1))()
        for i, r in enumerate(requests):
            self.descs[i] = abi.make_run_desc(r, self._keep)
        self._ctx = ctx
        self._h = None
        if ctx is not None:
            h = C.c_void_p()
            rc = ctx._lib.hgx_batch_prepare(ctx._h, self.descs, self.n, C.byref(h))
            if rc != abi.HGX_OK:
                raise ctx._err(rc)
            self._h = h

    def close(self):
        if self._h:
            self._ctx._lib.hgx_batch_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Replay:
    """A prepared batch's pass recorded into one HIP graph (hgx_batch_capture). launch() runs it
    and returns the results; they are valid until the next launch or close()."""

    def __init__(self, ctx, batch: "PreparedBatch"):
        self._ctx, self._batch = ctx, batch
        h = C.c_void_p()
        rc = ctx._lib.hgx_batch_capture(ctx._h, batch._h, C.byref(h))
        if rc != abi.HGX_OK:
            raise ctx._err(rc)
        self._h = h
        self.n = batch.n

    def launch(self):
        ctx = self._ctx
        outs = (C.c_void_p * max(self.n, 1))()
        rc = ctx._lib.hgx_replay_launch(ctx._h, self._h, outs)
        if rc != abi.HGX_OK:
            raise ctx._err(rc)
        handles = _BatchHandles(ctx._lib, outs, self.n, ctx)
        cnts = (abi.Counts * max(self.n, 1))()
        rc = ctx._lib.hgx_result_counts_many(outs, self.n, cnts)
        if rc != abi.HGX_OK:
            raise ctx._err(rc)
        return ResultList(handles, np.frombuffer(cnts, dtype=_COUNTS_DTYPE, count=self.n))

    def close(self):
        if self._h:
            self._ctx._lib.hgx_replay_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """An hgx_ctx on one device: the HBM budget, the stream, the served-request queue."""

    def __init__(self, device: int = 0, budget_bytes: int = 0):
        self._lib = load_library()
        h = C.c_void_p()
        rc = self._lib.hgx_ctx_create(device, budget_bytes, C.byref(h))
        if rc != abi.HGX_OK:
            buf = C.create_string_buffer(512)
            self._lib.hgx_last_error(None, buf, 512)
            raise HGXError(rc, buf.value.decode())
        self._h = h

    def _err(self, rc):
        buf = C.create_string_buffer(1024)
        self._lib.hgx_last_error(self._h, buf, 1024)
        return HGXError(rc, buf.value.decode())

    @property
    def info(self) -> str:
        buf = C.create_string_buffer(512)
        self._lib.hgx_device_info(self._h, buf, 512)
        return buf.value.decode()

    def set_forwarding(self, mode: int):
        """Incremental match forwarding: 1 on, 0 off (full join for every state), -1 automatic.
        Results never depend on it."""
        rc = self._lib.hgx_ctx_set_option(self._h, OPTION_FORWARD, int(mode))
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def set_delta_cap(self, n: int):
        """Delta matches one forwarded state may stage on chip (1..4096); beyond it the state
        takes the full join. Results never depend on it."""
        rc = self._lib.hgx_ctx_set_option(self._h, OPTION_DELTA_CAP, int(n))
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def set_expand_team(self, waves: int):
        """Waves that rewrite and hash one child: 1, 2, 4 or 8 for every child, 0 automatic.
        Results never depend on it."""
        rc = self._lib.hgx_ctx_set_option(self._h, OPTION_EXPAND_TEAM, int(waves))
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def set_expand_lanes(self, lanes: int):
        """Lanes that rewrite and hash one child when a wave is the team: 64, 32 or 16 (sub-wave
        teams), 0 automatic. Results never depend on it."""
        rc = self._lib.hgx_ctx_set_option(self._h, OPTION_EXPAND_LANES, int(lanes))
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def set_join_lanes(self, lanes: int):
        """Lanes that join one parent state when a wave is the team: 64, 32 or 16 (sub-wave
        teams), 0 automatic. Results never depend on it."""
        rc = self._lib.hgx_ctx_set_option(self._h, OPTION_JOIN_LANES, int(lanes))
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def set_counter_flush(self, mib: int):
        """Hardware-counter runs only: stream a buffer of `mib` MiB through the L2 before and
        after every timed kernel, so each kernel's HBM write-backs are attributable (0 off)."""
        rc = self._lib.hgx_ctx_set_option(self._h, OPTION_COUNTER_FLUSH, int(mib))
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def set_join_team(self, waves: int):
        """Waves that join one parent state: 1, 4 or 8 for every state, 0 automatic. Results
        never depend on it."""
        rc = self._lib.hgx_ctx_set_option(self._h, OPTION_JOIN_TEAM, int(waves))
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def prepare(self, requests) -> "PreparedBatch":
        """Parse requests and stage their rules and initial states in device memory once
        (hgx_batch_prepare), for repeated evolve_batch calls from device-resident inputs."""
        return PreparedBatch(requests, self)

    def capture(self, batch: "PreparedBatch") -> "Replay":
        """Record a prepared batch's pass into one HIP graph (hgx_batch_capture)."""
        return Replay(self, batch)

    def evolve_batch(self, requests):
        """Evolve requests (a list or a PreparedBatch) together on the device: a ResultList,
        one Result per request in request order."""
        import time as _t
        t0 = _t.perf_counter()
        pb = requests if isinstance(requests, PreparedBatch) else PreparedBatch(requests)
        n, descs = pb.n, pb.descs
        outs = (C.c_void_p * max(n, 1))()
        t1 = _t.perf_counter()
        if pb._h is not None:
            rc = self._lib.hgx_batch_evolve(self._h, pb._h, outs)
        else:
            rc = self._lib.hgx_evolve_batch(self._h, descs, n, outs)
        t2 = _t.perf_counter()
        if rc != abi.HGX_OK:
            raise self._err(rc)
        handles = _BatchHandles(self._lib, outs, n, self)
        cnts = (abi.Counts * max(n, 1))()
        rc = self._lib.hgx_result_counts_many(outs, n, cnts)
        if rc != abi.HGX_OK:
            raise self._err(rc)
        out = ResultList(handles, np.frombuffer(cnts, dtype=_COUNTS_DTYPE, count=n))
        self.last_timing = {"prepare_s": t1 - t0, "library_s": t2 - t1,
                            "results_s": _t.perf_counter() - t2}
        return out

    def evolve(self, request):
        return self.evolve_batch([request])[0]

    def wl_hash(self, states):
        """WL hashes of a list of states (each a list of edges)."""
        offs = [0]
        edges: list = []
        for st in states:
            edges.extend(st)
            offs.append(len(edges))
        eo, vv = encode_state(edges) if edges else (np.zeros(1, np.uint32), np.zeros(1, np.uint32))
        so = np.asarray(offs, np.uint64)
        if vv.size == 0:
            vv = np.zeros(1, np.uint32)
        out = np.zeros(len(states), np.uint64)
        rc = self._lib.hgx_wl_hash_batch(self._h, so.ctypes.data, len(states), eo.ctypes.data,
                                         vv.ctypes.data, out.ctypes.data)
        if rc != abi.HGX_OK:
            raise self._err(rc)
        return out

    def match(self, rules, rule_index: int, edges, want_edges: bool = False, cap: int = 0):
        """All matches of one rule in one (large) state on the device: count, and up to `cap`
        matches as edge indices in pattern order. Returns (count, matches or None, device ms)."""
        prog = np.frombuffer(encode_rules(rules), np.uint8).copy()
        offs, verts = encode_state(edges)
        if verts.size == 0:
            verts = np.zeros(1, np.uint32)
        k = len(rules[rule_index].lhs)
        out = np.zeros(max(cap, 1) * k, np.uint32) if want_edges else None
        cnt = C.c_uint64()
        ms = C.c_float()
        rc = self._lib.hgx_match_count(self._h, prog.ctypes.data, len(prog), len(rules), rule_index,
                                       offs.ctypes.data, verts.ctypes.data, len(edges), C.byref(cnt),
                                       out.ctypes.data if want_edges else None,
                                       cap if want_edges else 0, C.byref(ms))
        if rc != abi.HGX_OK:
            raise self._err(rc)
        n = cnt.value
        mat = out[: min(n, cap) * k].reshape(-1, k) if want_edges else None
        return n, mat, ms.value

    def match_noalloc(self, rules, rule_index: int, vertices_ptr: int, arity_ptr: int, n_edges: int,
                      stride: int, n_vertices: int, out_ptr: int = 0, cap: int = 0) -> int:
        """hgx_match_count_noalloc: the large-state join over a state already in device memory
        (vertices [n_edges * stride] uint32, arity [n_edges] uint8, ids below n_vertices), writing
        up to `cap` matches to the device buffer at out_ptr, with no device allocation (its working
        memory is the library's fixed scratch). Returns the match count."""
        prog = np.frombuffer(encode_rules(rules), np.uint8).copy()
        cnt = C.c_uint64()
        rc = self._lib.hgx_match_count_noalloc(self._h, prog.ctypes.data, len(prog), len(rules), rule_index,
                                               C.c_void_p(vertices_ptr), C.c_void_p(arity_ptr), n_edges,
                                               stride, n_vertices, C.byref(cnt),
                                               C.c_void_p(out_ptr) if out_ptr else None, cap)
        if rc != abi.HGX_OK:
            raise self._err(rc)
        return cnt.value

    def run_wxf(self, job: bytes) -> bytes:
        """The paclet's job envelope: a WXF job in, the WXF result out (hgx_run_wxf)."""
        out = C.POINTER(C.c_uint8)()
        n = C.c_size_t()
        rc = self._lib.hgx_run_wxf(self._h, job, len(job), C.byref(out), C.byref(n))
        if rc != abi.HGX_OK:
            raise self._err(rc)
        try:
            return C.string_at(out, n.value)
        finally:
            self._lib.hgx_buffer_free(out)

    # ---- served path: continuous batching ----

    def submit(self, request) -> int:
        """Queue a request; it joins the next generation launch that has room for it."""
        keep: list = []
        d = abi.make_run_desc(request, keep)
        rid = C.c_uint64()
        rc = self._lib.hgx_request_submit(self._h, C.byref(d), C.byref(rid))
        if rc != abi.HGX_OK:
            raise self._err(rc)
        return rid.value

    def serve_step(self) -> int:
        """Advance every admitted request by one generation; returns live + queued requests."""
        live = C.c_uint32()
        rc = self._lib.hgx_serve_step(self._h, C.byref(live))
        if rc != abi.HGX_OK:
            raise self._err(rc)
        return live.value

    def drain(self, cap: int = 4096):
        """{request id: Result} of requests that finished since the last drain."""
        ids = (C.c_uint64 * cap)()
        outs = (C.c_void_p * cap)()
        n = C.c_uint32()
        rc = self._lib.hgx_serve_drain(self._h, ids, outs, cap, C.byref(n))
        if rc != abi.HGX_OK:
            raise self._err(rc)
        return {int(ids[i]): Result(self._lib, C.c_void_p(outs[i]), self) for i in range(n.value)}

    def poll(self, request_id: int):
        out = C.c_void_p()
        rc = self._lib.hgx_request_poll(self._h, request_id, C.byref(out))
        if rc == abi.HGX_ERR_NOT_READY:
            return None
        if rc != abi.HGX_OK:
            raise self._err(rc)
        return Result(self._lib, out, self)

    def wait(self, request_id: int):
        out = C.c_void_p()
        rc = self._lib.hgx_request_wait(self._h, request_id, C.byref(out))
        if rc != abi.HGX_OK:
            raise self._err(rc)
        return Result(self._lib, out, self)

    def autotune(self, requests, reps: int = 2, cache_path: str | None = None) -> dict:
        """Measure launch shapes per size class over `requests` (a list or a PreparedBatch made by
        this context) and keep the fastest (hgx_autotune); with cache_path, write the JSON cache
        hgx_ctx_create reads. Returns {"chosen": ..., "measured": ...}."""
        import json as _json
        pb = requests if isinstance(requests, PreparedBatch) and requests._h is not None else self.prepare(requests)
        buf = C.create_string_buffer(1 << 16)
        rc = self._lib.hgx_autotune(self._h, pb._h, int(reps),
                                    cache_path.encode() if cache_path else None, buf, len(buf))
        if rc != abi.HGX_OK:
            raise self._err(rc)
        return _json.loads(buf.value.decode())

    def tune_load(self, path: str):
        rc = self._lib.hgx_tune_load(self._h, path.encode())
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def tuning(self) -> dict:
        """The launch shapes in effect and where they came from."""
        import json as _json
        buf = C.create_string_buffer(4096)
        self._lib.hgx_tune_get(self._h, buf, len(buf))
        return _json.loads(buf.value.decode())

    def calibration_read(self, nbytes: int):
        """Launches of known traffic for counter calibration (hgx_calibration_read): a fill kernel
        writing nbytes, then a read-only kernel reading them."""
        rc = self._lib.hgx_calibration_read(self._h, int(nbytes))
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def profile(self, on: bool = True):
        """Event-time every generation kernel launch (see hgx_profile_enable)."""
        rc = self._lib.hgx_profile_enable(self._h, 1 if on else 0)
        if rc != abi.HGX_OK:
            raise self._err(rc)

    def profile_read(self) -> dict:
        """{kernel: {launches, total_ms, alg_bytes, items}}; clears the accumulators."""
        n = C.c_uint32()
        buf = (abi.KernelStat * 32)()
        rc = self._lib.hgx_profile_read(self._h, buf, 32, C.byref(n))
        if rc != abi.HGX_OK:
            raise self._err(rc)
        out = {}
        for i in range(min(n.value, 32)):
            s = buf[i]
            out[s.name.decode()] = {"launches": int(s.launches), "total_ms": float(s.total_ms),
                                    "alg_bytes": int(s.alg_bytes), "items": int(s.items)}
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.hgx_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
