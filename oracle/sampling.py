"""TEST INFRASTRUCTURE: restatement of the reference's sampling and capping decisions.

The functions of common/include/hgcommon/sampling_core.hpp (rate per rule, the transition draw,
the cap's rank), the transition key (event_signature over EVENT_SIG_TRANSITION,
common/include/hgcommon/event_core.hpp:55-84, as gpu/include/hg_gpu/exploration.hpp:48-64 builds
it), the exploration coin and the two arrival-order bounds (exploration.hpp:66-122, threshold
encoding gpu/src/evolve.cu:271-285), and two evolutions that apply them:

* ``sampled(...)``   -- the decisions taken generation by generation over the C oracle's full
  multiway evolution, in its event order (the order the device numbers events in);
* ``brute(...)``     -- an independent multiway evolution in plain Python (brute-force matching
  of every injective assignment of LHS edges) that applies the order-free decisions (rate,
  weights, per-(state, rule) cap) as it goes, for small cases.

Exact state hashes and edge ranks come from the library's host IR canonicalizer
(hypergraphrewritingengine_amd.ir_rank), which tests/test_ir.py pins against the reference's
IRCanonicalizer vectors. Imported only by tests/.
"""
from __future__ import annotations

import collections
import itertools

import numpy as np

import hypergraphrewritingengine_amd as hgx
from hypergraphrewritingengine_amd.rules import Request, rule_from_lists
from oracle import pyoracle as po

M64 = (1 << 64) - 1
FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
GOLDEN = 0x9E3779B97F4A7C15


def fnv(h, v):
    """core.hpp:74-78."""
    return ((h ^ v) * FNV_PRIME) & M64


def splitmix64(z):
    """core.hpp:82-87 (Steele, Lea and Flood's SplitMix64 finalizer)."""
    z = (z + GOLDEN) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def rate_for_rule(rate, weights, rule):
    """sampling_core.hpp:28-34."""
    w = weights[rule] if weights is not None and rule < len(weights) else 1.0
    r = rate * w
    return 0.0 if r < 0.0 else (1.0 if r > 1.0 else r)


def transition_survives(key, seed, rate):
    """sampling_core.hpp:42-57."""
    if rate >= 1.0:
        return True
    if rate <= 0.0:
        return False
    x = splitmix64(key ^ ((seed * GOLDEN) & M64))
    return (x >> 11) * (1.0 / 9007199254740992.0) < rate


def transition_rank(key, seed):
    """sampling_core.hpp:66-71."""
    x = key ^ ((seed * GOLDEN) & M64) ^ 0xA5A5A5A5A5A5A5A5
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def transition_key(input_hash, rule, consumed_ranks):
    """event_signature(EVENT_SIG_TRANSITION, ...): input state, rule, consumed ranks in pattern
    order (event_core.hpp:62-84)."""
    sig = fnv(FNV_OFFSET, input_hash)
    sig = fnv(sig, rule & 0xFFFF)
    for r in consumed_ranks:
        sig = fnv(sig, r)
    return 1 if sig in (0, FNV_OFFSET) else sig


def explore_threshold(p):
    """evolve.cu:271-285: the probability as a float, scaled to 2^32."""
    f = float(np.float32(p))
    if not f > 0.0:
        f = 0.0
    if f > 1.0:
        f = 1.0
    if f >= 1.0:
        return 0xFFFFFFFF
    if f <= 0.0:
        return 0
    return int(f * 4294967296.0) & 0xFFFFFFFF


def coin(seed, step, sid, thr):
    """exploration.hpp:93-100."""
    if thr == 0xFFFFFFFF:
        return True
    if thr == 0:
        return False
    return (splitmix64(seed ^ (step << 32) ^ sid) & 0xFFFFFFFF) < thr


def _states(so, eo, vv):
    return [[[int(x) for x in vv[eo[e]:eo[e + 1]]] for e in range(int(so[s]), int(so[s + 1]))]
            for s in range(len(so) - 1)]


def sampled(rules, init, steps, *, rate=1.0, weights=None, seed=0, per_state_rule=0,
            explore_p=1.0, per_step=0, per_parent=0):
    """The sampled evolution, cut out of the C oracle's full one (event order). Returns
    (states, events, causal, branchial): states as edge lists; events as (input, output, rule,
    consumed positions); causal as (producer, consumer, in_tr); branchial as (a, b); all
    renumbered in order. `rules` as (lhs, rhs) lists; `init` one state."""
    req = Request([rule_from_lists(l, r, k) for k, (l, r) in enumerate(rules)], [init], steps,
                  flags=0x1 | 0x2 | 0x4 | 0x8, multi_root=True)
    res = po.evolve(req, with_edges=True)
    states = _states(*res.state_edges)
    co, cp = res.consumed
    ev = [(int(a), int(b), int(r)) for a, b, r, _ in res.events]
    cons = [[int(x) for x in cp[co[e]:co[e + 1]]] for e in range(len(ev))]
    hashes, ranks = hgx.ir_rank(states)
    draws = rate < 1.0 or weights is not None
    keys = [transition_key(int(hashes[i]), r, [ranks[i][q] for q in cons[e]]) for e, (i, _, r) in enumerate(ev)]
    capped_in = [True] * len(ev)
    if per_state_rule:
        groups = collections.defaultdict(list)
        for e, (i, _, r) in enumerate(ev):
            groups[(i, r)].append((transition_rank(keys[e], seed), e))
        for g in groups.values():
            for _, e in sorted(g)[per_state_rule:]:
                capped_in[e] = False
    thr = explore_threshold(explore_p)
    ns = len(states)
    exists, expanded = [False] * ns, [False] * ns
    new_sid = {}
    for s in range(ns):
        if int(res.state_creator[s]) == 0xFFFFFFFF:
            exists[s] = expanded[s] = True
            new_sid[s] = len(new_sid)
    alive, new_eid = [False] * len(ev), {}
    step_n, parent_n = collections.Counter(), collections.Counter()
    sstep = [int(x) for x in res.state_step]
    for e, (i, c, r) in enumerate(ev):
        ok = expanded[i] and capped_in[e]
        if ok and draws:
            ok = transition_survives(keys[e], seed, rate_for_rule(rate, weights, r))
        if not ok:
            continue
        alive[e] = True
        new_eid[e] = len(new_eid)
        exists[c] = True
        new_sid[c] = len(new_sid)
        if sstep[c] >= steps:
            continue
        go = coin(seed, sstep[c], new_sid[c], thr)
        if go and per_step:
            go = step_n[sstep[c]] < per_step
            step_n[sstep[c]] += 1
        if go and per_parent:
            go = parent_n[i] < per_parent
            parent_n[i] += 1
        expanded[c] = go
    out_states = [states[s] for s in range(ns) if exists[s]]
    out_events = [(new_sid[i], new_sid[c], r, tuple(cons[e])) for e, (i, c, r) in enumerate(ev) if alive[e]]
    causal = [(new_eid[int(a)], new_eid[int(b)], int(t)) for (a, b), t in zip(res.causal, res.causal_tr)
              if alive[int(a)] and alive[int(b)]]
    branchial = [(new_eid[int(a)], new_eid[int(b)]) for a, b in res.branchial if alive[int(a)] and alive[int(b)]]
    return out_states, out_events, causal, branchial


def _matches(state, lhs):
    """Every injective assignment of state edges to the LHS pattern edges with consistent
    variable bindings, as tuples of positions in pattern order."""
    out = []
    for combo in itertools.permutations(range(len(state)), len(lhs)):
        bind, ok = {}, True
        for pe, si in zip(lhs, combo):
            edge = state[si]
            if len(edge) != len(pe):
                ok = False
                break
            for var, v in zip(pe, edge):
                if bind.setdefault(var, v) != v:
                    ok = False
                    break
            if not ok:
                break
        if ok:
            out.append((combo, bind))
    return out


def brute(rules, init, steps, *, rate=1.0, weights=None, seed=0, per_state_rule=0):
    """Independent evolution with the order-free decisions applied in the loop. Returns the
    multiset of exact state hashes and the multiset of events keyed as (input hash, output hash,
    rule, consumed ranks)."""
    draws = rate < 1.0 or weights is not None
    nxt = [max([v for e in init for v in e], default=-1) + 1]
    frontier = [init]
    all_states, events = [init], []
    for _ in range(steps):
        if not frontier:
            break
        hs, rk = hgx.ir_rank(frontier)
        new = []
        for st, h, ranks in zip(frontier, hs, rk):
            cand = []
            for ri, (lhs, rhs) in enumerate(rules):
                found = []
                for combo, bind in _matches(st, lhs):
                    key = transition_key(int(h), ri, [ranks[q] for q in combo])
                    found.append((transition_rank(key, seed), key, combo, bind))
                if per_state_rule:
                    found = sorted(found, key=lambda x: x[0])[:per_state_rule]
                for _, key, combo, bind in found:
                    if draws and not transition_survives(key, seed, rate_for_rule(rate, weights, ri)):
                        continue
                    cand.append((ri, combo, bind, ranks))
            for ri, combo, bind, ranks in cand:
                lhs, rhs = rules[ri]
                b = dict(bind)
                for pe in rhs:
                    for var in pe:
                        if var not in b:
                            b[var] = nxt[0]
                            nxt[0] += 1
                child = [e for q, e in enumerate(st) if q not in combo] + [[b[v] for v in pe] for pe in rhs if pe]
                new.append(child)
                events.append((st, child, ri, tuple(ranks[q] for q in combo)))
        all_states += new
        frontier = new
    hs_all = hgx.ir_hash(all_states)
    ih = {id(s): int(h) for s, h in zip(all_states, hs_all)}
    ev_keys = collections.Counter((ih[id(a)], ih[id(b)], r, rk) for a, b, r, rk in events)
    return collections.Counter(int(h) for h in hs_all), ev_keys
