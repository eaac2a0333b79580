"""ctypes mirror of include/hgx.h (structs, constants, signatures)."""
from __future__ import annotations

import ctypes as C

HGX_OK = 0
HGX_ERR_INVALID = 1
HGX_ERR_NO_DEVICE = 2
HGX_ERR_DEVICE = 3
HGX_ERR_CAPACITY = 4
HGX_ERR_NOT_READY = 5
HGX_ERR_UNSUPPORTED = 6

HGX_REC_CAUSAL = 0x1
HGX_REC_BRANCHIAL = 0x2
HGX_REC_RELATIONS = 0x4
HGX_REC_STATES = 0x8
HGX_OPT_TR = 0x100
HGX_OPT_QUOTIENT = 0x200
HGX_OPT_QUOTIENT_ROOTS = 0x400
HGX_REC_DEFAULT = HGX_REC_CAUSAL | HGX_REC_BRANCHIAL
HGX_MAX_STEPS = 63


class RunDesc(C.Structure):
    _fields_ = [
        ("rules", C.POINTER(C.c_uint8)),
        ("rules_len", C.c_uint32),
        ("n_rules", C.c_uint32),
        ("edge_offsets", C.POINTER(C.c_uint32)),
        ("vertices", C.POINTER(C.c_uint32)),
        ("n_edges", C.c_uint32),
        ("root_offsets", C.POINTER(C.c_uint32)),
        ("n_roots", C.c_uint32),
        ("steps", C.c_uint32),
        ("state_cap", C.c_uint32),
        ("flags", C.c_uint32),
        ("sampling", C.c_void_p),
    ]


class Sampling(C.Structure):
    """hgx_sampling."""
    _fields_ = [
        ("transition_rate", C.c_double),
        ("rule_weights", C.POINTER(C.c_double)),
        ("n_rule_weights", C.c_uint32),
        ("reserved", C.c_uint32),
        ("seed", C.c_uint64),
        ("exploration_probability", C.c_double),
        ("max_states_per_step", C.c_uint64),
        ("max_successor_states_per_parent", C.c_uint64),
        ("matches_per_state_rule", C.c_uint64),
    ]


def make_sampling(spec: dict, keep: list):
    """An hgx_sampling from a dict with any of: rate, weights, seed, explore_p, per_step,
    per_parent, per_state_rule."""
    import numpy as np
    s = Sampling()
    s.transition_rate = float(spec.get("rate", 1.0))
    w = spec.get("weights")
    if w is not None:
        arr = np.asarray(w, dtype=np.float64)
        keep.append(arr)
        s.rule_weights = arr.ctypes.data_as(C.POINTER(C.c_double))
        s.n_rule_weights = len(arr)
    s.seed = int(spec.get("seed", 0))
    s.exploration_probability = float(spec.get("explore_p", 1.0))
    s.max_states_per_step = int(spec.get("per_step", 0))
    s.max_successor_states_per_parent = int(spec.get("per_parent", 0))
    s.matches_per_state_rule = int(spec.get("per_state_rule", 0))
    keep.append(s)
    return s


class Counts(C.Structure):
    _fields_ = [
        ("raw_states", C.c_uint64),
        ("events", C.c_uint64),
        ("canonical_states", C.c_uint64),
        ("causal_triples", C.c_uint64),
        ("causal_pairs", C.c_uint64),
        ("causal_pairs_tr", C.c_uint64),
        ("causal_triples_tr", C.c_uint64),
        ("branchial_pairs", C.c_uint64),
        ("fingerprint", C.c_uint64),
        ("steps_done", C.c_uint32),
        ("truncated", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Unfolding(C.Structure):
    _fields_ = [(k, C.c_uint64) for k in ("raw_events", "causal_triples", "causal_pairs", "causal_pairs_tr",
                                          "causal_triples_tr", "branchial_pairs", "instances", "classes",
                                          "class_matches")]

    def as_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Event(C.Structure):
    _fields_ = [("input_state", C.c_uint32), ("output_state", C.c_uint32),
                ("rule", C.c_uint32), ("step", C.c_uint32)]


class Pair(C.Structure):
    _fields_ = [("a", C.c_uint32), ("b", C.c_uint32)]


def _ptr(arr, ctype):
    """Pointer into a contiguous numpy array (kept alive by the caller)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))


def make_run_desc(req, keep: list) -> RunDesc:
    """Fill an hgx_run_desc from a rules.Request; `keep` holds the backing buffers."""
    import numpy as np

    enc = req.encoded()
    if "desc" in enc:                   # built before: the arrays it points into live in enc
        keep.append(enc["desc_arrays"])
        d = enc["desc"]
        d.steps, d.state_cap, d.flags = int(req.steps), int(req.state_cap), int(req.flags)
        d.sampling = _sampling_ptr(req, keep)
        return d
    prog = np.frombuffer(enc["prog"], dtype=np.uint8).copy()
    offs, verts, roots = enc["offs"], enc["verts"], enc["roots"]
    if verts.size == 0:
        verts = np.zeros(1, dtype=np.uint32)
    arrays = [prog, offs, verts, roots]
    keep.append(arrays)
    d = RunDesc()
    d.rules = _ptr(prog, C.c_uint8)
    d.rules_len = len(enc["prog"])
    d.n_rules = len(req.rules)
    d.edge_offsets = _ptr(offs, C.c_uint32)
    d.vertices = _ptr(verts, C.c_uint32)
    d.n_edges = len(offs) - 1
    d.root_offsets = _ptr(roots, C.c_uint32) if roots is not None else None
    d.n_roots = (len(roots) - 1) if roots is not None else 1
    d.steps = int(req.steps)
    d.state_cap = int(req.state_cap)
    d.flags = int(req.flags)
    d.sampling = _sampling_ptr(req, keep)
    enc["desc"], enc["desc_arrays"] = d, arrays
    return d


def _sampling_ptr(req, keep):
    spec = getattr(req, "sampling", None)
    if not spec:
        return None
    s = make_sampling(spec, keep)
    return C.cast(C.pointer(s), C.c_void_p)


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double),
                ("alg_bytes", C.c_uint64), ("items", C.c_uint64)]
