"""Rule model and request encoding for the hgx C ABI (include/hgx.h).

Mirrors the reference's rule builder -- ``make_rule(i).lhs({0,1}).rhs({0,2}).build()``
(hypergraph/include/hypergraph/pattern.hpp:180-213) -- and encodes rules into the byte
"rule program" the C ABI takes, and initial states into (edge_offsets, vertices) arrays.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Optional, Iterable, Sequence

import numpy as np

MAX_ARITY = 16
MAX_PATTERN_EDGES = 16
MAX_VARS = 32


@dataclass(frozen=True)
class Rule:
    """A rewrite rule: LHS and RHS as tuples of pattern edges of variable indices.

    Same meaning as ``hypergraph::RewriteRule``: variables are small integers, a variable
    that occurs only on the RHS is fresh (one new vertex per application, ascending by
    variable index, common/include/hgcommon/rewrite_core.hpp:53-62).
    """

    lhs: tuple
    rhs: tuple
    index: int = 0

    def __post_init__(self):
        if not (1 <= len(self.lhs) <= MAX_PATTERN_EDGES) or len(self.rhs) > MAX_PATTERN_EDGES:
            raise ValueError("rule: 1..16 LHS edges and 0..16 RHS edges")
        for e in tuple(self.lhs) + tuple(self.rhs):
            if not (1 <= len(e) <= MAX_ARITY):
                raise ValueError("rule: pattern edge arity must be 1..16")
            if any((not isinstance(v, (int, np.integer))) or v < 0 or v >= MAX_VARS for v in e):
                raise ValueError("rule: pattern variables must be integers in [0, 32)")

    @property
    def lhs_var_mask(self) -> int:
        m = 0
        for e in self.lhs:
            for v in e:
                m |= 1 << v
        return m

    @property
    def rhs_var_mask(self) -> int:
        m = 0
        for e in self.rhs:
            for v in e:
                m |= 1 << v
        return m

    @property
    def new_var_mask(self) -> int:
        return self.rhs_var_mask & ~self.lhs_var_mask


class RuleBuilder:
    """``make_rule(index).lhs(...).rhs(...).build()``, as in pattern.hpp:180-213."""

    def __init__(self, index: int = 0):
        self._index = index
        self._lhs: list = []
        self._rhs: list = []

    def lhs(self, vars_: Iterable[int]) -> "RuleBuilder":
        self._lhs.append(tuple(int(v) for v in vars_))
        return self

    def rhs(self, vars_: Iterable[int]) -> "RuleBuilder":
        self._rhs.append(tuple(int(v) for v in vars_))
        return self

    def build(self) -> Rule:
        return Rule(tuple(self._lhs), tuple(self._rhs), self._index)


def make_rule(index: int = 0) -> RuleBuilder:
    return RuleBuilder(index)


def rule_from_lists(lhs: Sequence[Sequence[int]], rhs: Sequence[Sequence[int]], index: int = 0) -> Rule:
    return Rule(tuple(tuple(int(v) for v in e) for e in lhs),
                tuple(tuple(int(v) for v in e) for e in rhs), index)


def encode_rules(rules: Sequence[Rule]) -> bytes:
    """Rule program: per rule n_lhs, n_rhs, then (arity, vars...) per LHS then RHS edge."""
    out = bytearray()
    for r in rules:
        out += bytes([len(r.lhs), len(r.rhs)])
        for e in tuple(r.lhs) + tuple(r.rhs):
            out += bytes([len(e)]) + bytes(int(v) for v in e)
    return bytes(out)


def encode_state(edges: Sequence[Sequence[int]]):
    """(edge_offsets[n+1] uint32, vertices uint32) of one edge list."""
    n = len(edges)
    lens = np.fromiter(map(len, edges), dtype=np.int64, count=n)
    if n and (lens.min() < 1 or lens.max() > MAX_ARITY):
        raise ValueError("state: edge arity must be 1..16")
    offs = np.zeros(n + 1, dtype=np.uint32)
    offs[1:] = np.cumsum(lens)
    flat = np.fromiter(itertools.chain.from_iterable(edges), dtype=np.int64, count=int(offs[-1]))
    if flat.size and (flat.min() < 0 or flat.max() > 0xFFFFFFFF):
        raise ValueError("state: vertex labels must be 0..2^32-1")
    return offs, flat.astype(np.uint32)


def encode_roots(roots: Sequence[Sequence[Sequence[int]]]):
    """(edge_offsets, vertices, root_offsets) for several initial states."""
    all_edges: list = []
    root_offs = [0]
    for st in roots:
        all_edges.extend(st)
        root_offs.append(len(all_edges))
    offs, verts = encode_state(all_edges)
    return offs, verts, np.asarray(root_offs, dtype=np.uint32)


@dataclass
class Request:
    """One evolution request (the fields of hgx_run_desc)."""

    rules: Sequence[Rule]
    init: Sequence  # one state (list of edges) or, with multi_root, a list of states
    steps: int
    state_cap: int = 0
    flags: int = 0x3
    multi_root: bool = False
    # hgx_sampling as a dict: rate, weights, seed, explore_p, per_step, per_parent, per_state_rule
    sampling: Optional[dict] = None
    _enc: dict = field(default_factory=dict, repr=False)

    def encoded(self):
        if not self._enc:
            prog = encode_rules(self.rules)
            if self.multi_root:
                offs, verts, roots = encode_roots(self.init)
            else:
                offs, verts = encode_state(self.init)
                roots = None
            self._enc.update(prog=prog, offs=offs, verts=verts, roots=roots)
        return self._enc

    def job_text(self, tr: bool = False) -> str:
        """The reference driver's job format (oracle/ref_driver.cpp)."""
        toks = [str(len(self.rules))]
        for r in self.rules:
            toks += [str(len(r.lhs)), str(len(r.rhs))]
            for e in tuple(r.lhs) + tuple(r.rhs):
                toks += [str(len(e))] + [str(v) for v in e]
        roots = self.init if self.multi_root else [self.init]
        toks.append(str(len(roots)))
        for st in roots:
            toks.append(str(len(st)))
            for e in st:
                toks += [str(len(e))] + [str(int(v)) for v in e]
        toks += [str(self.steps), "1" if tr else "0"]
        return " ".join(toks) + "\n"
