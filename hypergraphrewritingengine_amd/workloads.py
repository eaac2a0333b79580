"""Seeded synthetic workloads of the benchmark configurations.

All randomness is one SplitMix64 stream per seed, so every workload is reproducible from
its seed alone, on any host.

  config1()        rule {{x,y},{y,z}}->{{x,y},{y,z},{z,x}} from the triangle, 5 steps
  config2()        rule {{x,y},{x,z}}->{{x,z},{x,w},{y,w},{z,w}} from {{1,1},{1,1}}, 6 steps
  config3a/3b()    one big seeded random state and a 2- / 3-edge LHS (one rewriting step)
  batched_run(i)   run i of the batched workload: a random connected 2-3 edge binary LHS over
                   <= 4 variables, 2-4 RHS edges with <= 2 fresh variables, a connected random
                   initial state of 3-64 edges, 3-6 steps, cap 4,096 states
  served_run(i)    the same draw with a log-uniform 3-2,048 edge initial state and 2-8 steps
"""
from __future__ import annotations

import math

from .rules import Request, rule_from_lists

M64 = (1 << 64) - 1


def splitmix64(z: int) -> int:
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


class Stream:
    """SplitMix64 generator: state += golden gamma, output = mix(state)."""

    def __init__(self, seed: int):
        self.s = seed & M64

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & M64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        """Uniform in [0, n): the top 32 bits scaled, (x >> 32) * n >> 32."""
        return ((self.next() >> 32) * n) >> 32

    def uniform(self, lo: int, hi: int) -> int:
        return lo + self.below(hi - lo + 1)


def config1(steps: int = 5, flags: int = 0x3) -> Request:
    r = rule_from_lists([[0, 1], [1, 2]], [[0, 1], [1, 2], [2, 0]])
    return Request([r], [[1, 2], [2, 3], [3, 1]], steps, flags=flags)


def config2(steps: int = 6, flags: int = 0x3) -> Request:
    r = rule_from_lists([[0, 1], [0, 2]], [[0, 2], [0, 3], [1, 3], [2, 3]])
    return Request([r], [[1, 1], [1, 1]], steps, flags=flags)


def random_state(st: Stream, n_edges: int, n_vertices: int, arity: int):
    return [[st.below(n_vertices) for _ in range(arity)] for _ in range(n_edges)]


def config3a(seed: int = 3, n_edges: int = 50_000, n_vertices: int = 10_000):
    st = Stream(splitmix64(seed))
    rule = rule_from_lists([[0, 1, 2], [2, 3, 4]], [[0, 3, 5], [1, 4, 5], [2, 5, 5]])
    return rule, random_state(st, n_edges, n_vertices, 3)


def config3b(seed: int = 4, n_edges: int = 200_000, n_vertices: int = 20_000):
    st = Stream(splitmix64(seed))
    rule = rule_from_lists([[0, 1], [1, 2], [2, 0]], [[0, 1], [1, 2], [2, 0], [0, 3]])
    return rule, random_state(st, n_edges, n_vertices, 2)


def _random_rule(st: Stream):
    """Connected binary LHS of 2-3 edges over <= 4 variables; 2-4 RHS edges, <= 2 fresh."""
    n_lhs = st.uniform(2, 3)
    lhs = []
    nv = 0
    for i in range(n_lhs):
        if i == 0:
            a = 0
            b = st.below(2)          # a self-loop or a new variable
            nv = 1 + b
        else:
            # one end on an existing variable keeps the LHS connected
            a = st.below(nv)
            b = st.below(min(nv + 1, 4))
            nv = max(nv, b + 1)
        if st.below(2):
            a, b = b, a
        lhs.append([a, b])
    n_fresh = st.below(3)
    pool = nv + n_fresh
    n_rhs = st.uniform(2, 4)
    rhs = [[st.below(pool), st.below(pool)] for _ in range(n_rhs)]
    # every fresh variable occurs on the right
    for k in range(n_fresh):
        v = nv + k
        if not any(v in e for e in rhs):
            rhs[st.below(n_rhs)][st.below(2)] = v
    return rule_from_lists(lhs, rhs)


def _connected_init(st: Stream, m: int):
    edges = []
    nverts = 1
    for _ in range(m):
        u = st.below(nverts)
        v = nverts if st.below(2) == 0 else st.below(nverts)
        if v == nverts:
            nverts += 1
        edges.append([u, v] if st.below(2) else [v, u])
    return edges


def batched_run(i: int, cap: int = 4096, flags: int = 0x3) -> Request:
    st = Stream(splitmix64(i))
    rule = _random_rule(st)
    init = _connected_init(st, st.uniform(3, 64))
    steps = st.uniform(3, 6)
    return Request([rule], init, steps, state_cap=cap, flags=flags)


def served_run(i: int, cap: int = 4096, flags: int = 0x3) -> Request:
    st = Stream(splitmix64(i ^ 0x5E5E5E5E))
    rule = _random_rule(st)
    # log-uniform in [3, 2048]
    u = st.next() / float(1 << 64)
    m = int(round(math.exp(math.log(3) + u * (math.log(2048) - math.log(3)))))
    init = _connected_init(st, max(3, min(2048, m)))
    steps = st.uniform(2, 8)
    return Request([rule], init, steps, state_cap=cap, flags=flags)
