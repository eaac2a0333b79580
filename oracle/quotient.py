"""TEST INFRASTRUCTURE: restatement of quotient exploration and of the raw unfolding beneath it.

* ``explore(rules, roots, steps, merge_roots=False)`` -- ExploreFromCanonicalStatesOnly
  (hypergraph/src/parallel_evolution.cpp:161-166, :341-356 try_claim_expanded; the device's
  state_survives_dedup with dedup on, gpu/include/hg_gpu/exploration.hpp:66-92; root handling
  gpu/src/persistent.cu:122-135): generation by generation, every state of the frontier that is
  expanded applies every match of every rule (rule order, then the matcher's order); a new state
  is expanded when its step is below ``steps`` and it is the first state of its exact class met so
  far. Every root is expanded, unless ``merge_roots``, when a root isomorphic to an earlier one is
  not. Rewriting follows Rewriter::apply (rewriter.cpp:47-89): the child keeps the parent's
  unconsumed edges in order, then the RHS edges, fresh vertices ascending by variable from a
  per-lineage counter. Matches come from the C oracle's join (join_core.hpp order).
* ``unfold(x, steps)`` -- the per-instance replay of common/include/hgcommon/quotient_replay_core.hpp
  (qr_apply, :112-226; seeding hypergraph.cpp:977-1003; instances hypergraph.cpp:1017-1045): each
  class's expansion is recorded once, from the first of its states that was expanded, in slots
  (here the canonical ranks of the class's edges, a frame equivalent to slot_core.hpp's orbit
  ranks); every raw instance replays it. Counts raw events, causal triples and distinct pairs,
  the pairs surviving transitive reduction (CausalGraph's reduction: a pair (p, c) is redundant
  when p reaches another producer of c), and branchial pairs (sibling applications of one instance
  with overlapping consumed slots). Written independently of the device replay: a depth-first
  walk over explicit producer vectors and an explicit set of causal pairs per branch.
* ``relations(x)`` -- the causal and branchial relations among the explored events themselves
  (what the evolution materialises): producers along each event's own lineage (Rewriter::apply's
  causal block, rewriter.cpp:138-191), the reduction over that relation, and pairs of events of one
  input state with overlapping consumed edges.
* ``identity_events(x, keys)`` -- the event count under an identity mode on the reconstruction
  route: the distinct signatures (event_core.hpp:60-84) of the recorded (class, match) expansions,
  with class hashes, the output class's first step, and slots as ranks.

Exact hashes and ranks come from the library's host IR canonicalizer (hgx.ir_rank), which
tests/test_ir.py pins against the reference's IRCanonicalizer vectors. Imported only by tests/.
"""
from __future__ import annotations

import hypergraphrewritingengine_amd as hgx
from oracle import pyoracle as po

FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
M64 = (1 << 64) - 1
K_IN, K_OUT, K_STEP, K_RULE, K_CONS, K_PROD = 1, 2, 4, 8, 16, 32


def _renumber(state):
    m = {}
    out = []
    for e in state:
        out.append([m.setdefault(v, len(m)) for v in e])
    return out, len(m)


def _apply(rule, state, nxt, pos):
    """Rewriter::apply: survivors in order, then the RHS with fresh vertices ascending by var."""
    b = {}
    for pe, q in zip(rule.lhs, pos):
        for x, y in zip(pe, state[q]):
            b[x] = y
    fresh = sorted({x for e in rule.rhs for x in e if x not in b})
    for x in fresh:
        b[x] = nxt
        nxt += 1
    taken = set(pos)
    child = [list(state[k]) for k in range(len(state)) if k not in taken]
    child += [[b[x] for x in e] for e in rule.rhs]
    return child, nxt


class Explored:
    """states[i] = (edges, step, exact hash, ranks); events[e] = (input, output, rule, positions);
    expanded[i] = whether state i was expanded."""

    def __init__(self):
        self.states, self.events, self.expanded = [], [], []


def explore(rules, roots, steps, merge_roots=False):
    x = Explored()
    seen = set()
    gen, nxts = [], []
    for r in roots:
        st, nv = _renumber(r)
        gen.append(len(x.states))
        x.states.append(st)
        nxts.append(nv)
    x.states = [[s, 0] for s in x.states]
    nxt_of = dict(zip(gen, nxts))

    def classify(ids, roots_gen):
        hs, rk = hgx.ir_rank([x.states[i][0] for i in ids])
        go = []
        for i, h, r in zip(ids, hs, rk):
            h = int(h)
            x.states[i] += [h, [int(v) for v in r]]
            first = h not in seen
            seen.add(h)
            keep = first or (roots_gen and not merge_roots) or h == 0
            go.append(keep and x.states[i][1] < steps)
        return go

    go = classify(gen, True)
    x.expanded = list(go)
    for step in range(1, steps + 1):
        kids = []
        for i, g in zip(gen, go):
            if not g:
                continue
            st = x.states[i][0]
            for ri, rule in enumerate(rules):
                n = po.match(rules, ri, st)
                if not n:
                    continue
                _, pos = po.match(rules, ri, st, want_edges=True, cap=n)
                for p in pos:
                    p = [int(v) for v in p]
                    child, nx = _apply(rule, st, nxt_of[i], p)
                    c = len(x.states)
                    x.states.append([child, step])
                    x.expanded.append(False)
                    nxt_of[c] = nx
                    x.events.append((i, c, ri, p))
                    kids.append(c)
        if not kids:
            break
        go = classify(kids, False)
        for c, g in zip(kids, go):
            x.expanded[c] = g
        gen = kids
    # states at the last generation are never classified above when the loop ran out of steps
    rest = [i for i in range(len(x.states)) if len(x.states[i]) == 2]
    if rest:
        hs, rk = hgx.ir_rank([x.states[i][0] for i in rest])
        for i, h, r in zip(rest, hs, rk):
            x.states[i] += [int(h), [int(v) for v in r]]
    return x


def record(x):
    """Each class's expansion, from the first of its states that was expanded: per class hash, the
    list of (rule, consumed slots, child class, child slot sources (parent slot or None))."""
    rec, first = {}, {}
    for (i, c, ri, pos) in x.events:
        h = x.states[i][2]
        first.setdefault(h, i)
        if first[h] != i:
            continue
        rk_in, rk_out = x.states[i][3], x.states[c][3]
        nin = len(x.states[i][0])
        surv = [k for k in range(nin) if k not in set(pos)]
        src = [None] * len(x.states[c][0])
        for j, k in enumerate(surv):
            src[rk_out[j]] = rk_in[k]
        rec.setdefault(h, []).append((ri, [rk_in[q] for q in pos], x.states[c][2], src,
                                      [rk_out[j] for j in range(len(surv), len(x.states[c][0]))]))
    return rec


def unfold(x, steps):
    """The raw unfolding's counts by depth-first replay (qr_apply): events, causal triples, pairs,
    reduced pairs, reduced triples, branchial pairs, and the instances below the last step."""
    rec = record(x)
    tot = dict(raw_events=0, causal_triples=0, causal_pairs=0, causal_pairs_tr=0, causal_triples_tr=0,
               branchial_pairs=0, instances=0)
    # a branch: the events applied so far (ids are their depths), and the causal pairs among them
    def walk(h, prod, depth, edges_of_branch):
        tot["instances"] += 1
        ms = rec.get(h, [])
        for a in range(len(ms)):
            for b in range(a + 1, len(ms)):
                if set(ms[a][1]) & set(ms[b][1]):
                    tot["branchial_pairs"] += 1
        for (ri, cons, to, src, produced) in ms:
            tot["raw_events"] += 1
            ev = depth
            producers = [prod[s] for s in cons if prod[s] is not None]
            tot["causal_triples"] += len(producers)
            P = set(producers)
            tot["causal_pairs"] += len(P)
            pairs = edges_of_branch | {(p, ev) for p in P}
            # (p, ev) is kept when no other producer q of ev is reachable from p
            def reach(a, b):
                stack, seen = [a], set()
                while stack:
                    u = stack.pop()
                    if u == b:
                        return True
                    for (s, t) in pairs:
                        if s == u and t not in seen and t != ev:
                            seen.add(t)
                            stack.append(t)
                return False
            kept = {p for p in P if not any(q != p and reach(p, q) for q in P)}
            tot["causal_pairs_tr"] += len(kept)
            tot["causal_triples_tr"] += sum(1 for p in producers if p in kept)
            if depth + 1 < steps:
                child = [ev if f is None else prod[f] for f in src]
                walk(to, child, depth + 1, pairs)
    for i in range(len(x.states)):
        if x.states[i][1] == 0:
            walk(x.states[i][2], [None] * len(x.states[i][0]), 0, frozenset())
    return tot


def relations(x):
    """(causal pairs {(p, c): survives reduction}, branchial pairs {(a, b)}) of the explored events."""
    prod = {}
    for i, st in enumerate(x.states):
        if st[1] == 0:
            prod[i] = [None] * len(st[0])
    pairs = set()
    by_input = {}
    for e, (i, c, ri, pos) in enumerate(x.events):
        pp = prod[i]
        for q in pos:
            if pp[q] is not None:
                pairs.add((pp[q], e))
        taken = set(pos)
        prod[c] = [pp[k] for k in range(len(pp)) if k not in taken] + [e] * (len(x.states[c][0]) - len(pp) + len(pos))
        by_input.setdefault(i, []).append((e, taken))
    succ = {}
    for (a, b) in pairs:
        succ.setdefault(a, []).append(b)

    def reach(a, b, skip):
        stack, seen = [a], {a}
        while stack:
            u = stack.pop()
            for v in succ.get(u, []):
                if (u, v) == skip or v in seen:
                    continue
                if v == b:
                    return True
                seen.add(v)
                stack.append(v)
        return False
    causal = {(a, b): not reach(a, b, (a, b)) for (a, b) in pairs}
    branchial = set()
    for evs in by_input.values():
        for k in range(len(evs)):
            for l in range(k + 1, len(evs)):
                if evs[k][1] & evs[l][1]:
                    branchial.add((evs[k][0], evs[l][0]))
    return causal, branchial


def _fnv(h, v):
    return ((h ^ (v & M64)) * FNV_PRIME) & M64


def identity_events(x, keys):
    """Distinct signatures of the recorded (class, match) expansions under `keys`."""
    rec = record(x)
    first_step = {}
    for s in x.states:
        first_step.setdefault(s[2], s[1])
    sigs = set()
    for h, ms in rec.items():
        for (ri, cons, to, src, produced) in ms:
            sig = FNV_OFFSET
            if keys & K_IN:
                sig = _fnv(sig, h)
            if keys & K_OUT:
                sig = _fnv(sig, to)
            if keys & K_STEP:
                sig = _fnv(sig, first_step[to])
            if keys & K_RULE:
                sig = _fnv(sig, ri & 0xFFFF)
            if keys & K_CONS:
                for s in cons:
                    sig = _fnv(sig, s)
            if keys & K_PROD:
                for s in produced:
                    sig = _fnv(sig, s)
            if sig in (0, FNV_OFFSET):
                sig = 1
            sigs.add(sig)
    return len(sigs)


def classes(x):
    return len({s[2] for s in x.states})
