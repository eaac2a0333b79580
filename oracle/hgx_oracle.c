/*
 * hgx_oracle.c -- TEST INFRASTRUCTURE (see hgx_oracle.h). A serial CPU restatement of the
 * reference multiway evolution, written for clarity rather than speed. Used only as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Semantics restated (reference paths relative to the reference repository root):
 *   rule model + join order      hypergraph/src/pattern.cpp:13-75, :151-173
 *   join (edge-injective DFS)    common/include/hgcommon/join_core.hpp:102-180
 *   variable binding             common/include/hgcommon/match_core.hpp:32-49
 *   rewrite                      hypergraph/src/rewriter.cpp:11-103,
 *                                common/include/hgcommon/rewrite_core.hpp:53-87
 *   WL canonical hash            common/include/hgcommon/wl_core.hpp:67-202 with the caller
 *                                sizing of hypergraph/include/hypergraph/wl_hash.hpp:109-239
 *                                and the empty-state rule of hypergraph/src/hypergraph.cpp:740-755
 *   causal relation              hypergraph/src/rewriter.cpp:121-191,
 *                                hypergraph/src/causal_graph.cpp:110-149, :187-234
 *   transitive reduction         common/include/hgcommon/transitive_reduction.hpp:36-96
 *   branchial relation           hypergraph/src/causal_graph.cpp:324-351
 *   full multiway schedule       hypergraph/src/parallel_evolution.cpp:74-180 (every raw
 *                                state expanded; events at generations 1..steps)
 *
 * Deliberate, documented differences from the reference that no observable depends on:
 *   - fresh vertex ids come from a per-lineage counter (the reference uses a global
 *     atomic counter). WL hashes, event keys and all relations are label-invariant.
 *   - events are numbered generation by generation (parent, rule, match order), where
 *     the reference's numbering depends on thread scheduling.
 *   - state_cap applies the deterministic step-prefix contract (the reference's
 *     max_states is schedule-dependent, parallel_evolution.cpp:1616-1619).
 */
#include "hgx_oracle.h"

#include <stdlib.h>
#include <string.h>

#define INV 0xFFFFFFFFu
#define FNV_OFFSET 0xcbf29ce484222325ULL
#define FNV_PRIME 0x100000001b3ULL
#define EMPTY_STATE_HASH 0x9E3779B97F4A7C15ULL
#define WL_MAX_ITERS 100u

/* ---------------- hash primitives: common/include/hgcommon/core.hpp:54-87 ---------------- */

static inline uint64_t fnv(uint64_t h, uint64_t v) { h ^= v; h *= FNV_PRIME; return h; }

static inline uint64_t mix64(uint64_t x) {                       /* core.hpp:64-71 */
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33; return x;
}

static inline uint64_t splitmix(uint64_t z) {                    /* core.hpp:82-87 */
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

static void isort_u64(uint64_t* a, uint32_t n) {                 /* core.hpp:54-61 */
    for (uint32_t i = 1; i < n; ++i) {
        uint64_t key = a[i];
        uint32_t j = i;
        while (j > 0 && a[j - 1] > key) { a[j] = a[j - 1]; --j; }
        a[j] = key;
    }
}

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

static int cmp_pair(const void* a, const void* b) {
    const hgx_pair* x = (const hgx_pair*)a; const hgx_pair* y = (const hgx_pair*)b;
    if (x->a != y->a) return (x->a > y->a) - (x->a < y->a);
    return (x->b > y->b) - (x->b < y->b);
}

static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

static int popc32(uint32_t x) { return __builtin_popcount(x); }

/* ---------------- growable arrays ---------------- */

#define VEC(T) struct { T* p; uint64_t n, cap; }
#define VPUSH(v, x) do { if ((v).n == (v).cap) { (v).cap = (v).cap ? (v).cap * 2 : 64; \
    (v).p = realloc((v).p, (v).cap * sizeof(*(v).p)); } (v).p[(v).n++] = (x); } while (0)
#define VRESERVE(v, m) do { if ((v).cap < (m)) { (v).cap = (m); \
    (v).p = realloc((v).p, (v).cap * sizeof(*(v).p)); } } while (0)

/* ---------------- rules: hypergraph/src/pattern.cpp ---------------- */

typedef struct {
    uint8_t n_lhs, n_rhs;
    uint8_t la[16], lv[16][16];
    uint8_t ra[16], rv[16][16];
    uint8_t order[16];
    uint32_t lhs_mask, rhs_mask, new_mask;
} orule;

static uint32_t edge_var_mask(const uint8_t* vars, uint8_t arity) {   /* pattern.cpp:146-150 */
    uint32_t m = 0;
    for (uint8_t i = 0; i < arity; ++i) m |= 1u << vars[i];
    return m;
}

/* RewriteRule::edge_constraint_score, pattern.cpp:31-41 */
static int constraint_score(const orule* r, uint8_t e) {
    int s = 0;
    for (uint8_t i = 0; i < r->la[e]; ++i)
        for (uint8_t j = (uint8_t)(i + 1); j < r->la[e]; ++j)
            if (r->lv[e][i] == r->lv[e][j]) s += 100;
    for (uint8_t o = 0; o < r->n_lhs; ++o)
        if (o != e && (edge_var_mask(r->lv[e], r->la[e]) & edge_var_mask(r->lv[o], r->la[o])))
            s += 1;
    return s;
}

/* RewriteRule::compute_match_order, pattern.cpp:43-75 */
static void compute_match_order(orule* r) {
    for (uint8_t i = 0; i < 16; ++i) r->order[i] = i;
    if (r->n_lhs <= 1) return;
    int used[16] = {0};
    uint8_t first = 0; int best = -1;
    for (uint8_t e = 0; e < r->n_lhs; ++e) {
        int s = constraint_score(r, e);
        if (s > best) { best = s; first = e; }
    }
    r->order[0] = first; used[first] = 1;
    uint32_t bound = edge_var_mask(r->lv[first], r->la[first]);
    for (uint8_t pos = 1; pos < r->n_lhs; ++pos) {
        uint8_t pick = 0; int pick_shared = -1, pick_self = -1;
        for (uint8_t e = 0; e < r->n_lhs; ++e) {
            if (used[e]) continue;
            int shared = popc32(edge_var_mask(r->lv[e], r->la[e]) & bound);
            int self = constraint_score(r, e);
            if (shared > pick_shared || (shared == pick_shared && self > pick_self)) {
                pick = e; pick_shared = shared; pick_self = self;
            }
        }
        r->order[pos] = pick; used[pick] = 1;
        bound |= edge_var_mask(r->lv[pick], r->la[pick]);
    }
}

/* Decode the rule program (format in include/hgx.h). Masks: pattern.cpp:13-29, :151-169. */
static int parse_rules(const uint8_t* p, uint32_t len, uint32_t n, orule* out) {
    uint32_t k = 0;
    for (uint32_t r = 0; r < n; ++r) {
        orule* R = &out[r];
        memset(R, 0, sizeof(*R));
        if (k + 2 > len) return -1;
        R->n_lhs = p[k++]; R->n_rhs = p[k++];
        if (R->n_lhs < 1 || R->n_lhs > 16 || R->n_rhs > 16) return -1;
        for (int side = 0; side < 2; ++side) {
            uint8_t ne = side ? R->n_rhs : R->n_lhs;
            for (uint8_t e = 0; e < ne; ++e) {
                if (k >= len) return -1;
                uint8_t a = p[k++];
                if (a < 1 || a > 16 || k + a > len) return -1;
                for (uint8_t i = 0; i < a; ++i) {
                    uint8_t v = p[k++];
                    if (v >= 32) return -1;
                    if (side) R->rv[e][i] = v; else R->lv[e][i] = v;
                }
                if (side) R->ra[e] = a; else R->la[e] = a;
            }
        }
        for (uint8_t e = 0; e < R->n_lhs; ++e) R->lhs_mask |= edge_var_mask(R->lv[e], R->la[e]);
        for (uint8_t e = 0; e < R->n_rhs; ++e) R->rhs_mask |= edge_var_mask(R->rv[e], R->ra[e]);
        R->new_mask = R->rhs_mask & ~R->lhs_mask;
        compute_match_order(R);
    }
    return k == len ? 0 : -1;
}

/* ---------------- WL canonical hash ---------------- */

typedef struct {
    VEC(uint32_t) verts;      /* sorted distinct vertex ids */
    VEC(uint8_t)  ea;
    VEC(uint32_t) eoff, ev;
    VEC(uint32_t) nn, occ_off, occ_edge, comp;
    VEC(uint8_t)  occ_pos;
    VEC(uint64_t) cur, nxt, dscr, nbr;
} wl_scratch;

static void wl_scratch_free(wl_scratch* s) {
    free(s->verts.p); free(s->ea.p); free(s->eoff.p); free(s->ev.p); free(s->nn.p);
    free(s->occ_off.p); free(s->occ_edge.p); free(s->comp.p); free(s->occ_pos.p);
    free(s->cur.p); free(s->nxt.p); free(s->dscr.p); free(s->nbr.p);
}

static uint32_t wl_find(uint32_t* parent, uint32_t x) {         /* wl_core.hpp:55-58 */
    while (parent[x] != x) { parent[x] = parent[parent[x]]; x = parent[x]; }
    return x;
}

static uint32_t ndistinct(uint64_t* scratch, const uint64_t* cur, uint32_t n) {
    /* HG_WL_NDISTINCT, wl_core.hpp:111-116 */
    for (uint32_t v = 0; v < n; ++v) scratch[v] = cur[v];
    isort_u64(scratch, n);
    uint32_t d = 0;
    for (uint32_t v = 0; v < n; ++v) if (v == 0 || scratch[v] != scratch[v - 1]) ++d;
    return d;
}

/* hgcommon::wl_canonical_hash, wl_core.hpp:67-202, statement for statement. */
static uint64_t wl_core(const uint8_t* ea, const uint32_t* eoff, const uint32_t* ev,
                        uint32_t n_edges, uint32_t n_verts,
                        uint64_t* cur, uint64_t* nxt, uint32_t* occ_off, uint32_t* occ_edge,
                        uint8_t* occ_pos, uint64_t* nbr, uint32_t nbr_cap, uint64_t* dscr,
                        uint32_t* comp) {
    if (n_verts == 0 || n_edges == 0) return 0;
    for (uint32_t v = 0; v <= n_verts; ++v) occ_off[v] = 0;
    for (uint32_t e = 0; e < n_edges; ++e)
        for (uint8_t p = 0; p < ea[e]; ++p) occ_off[ev[eoff[e] + p] + 1u]++;
    for (uint32_t v = 0; v < n_verts; ++v) occ_off[v + 1] += occ_off[v];
    for (uint32_t v = 0; v < n_verts; ++v) dscr[v] = occ_off[v];
    for (uint32_t e = 0; e < n_edges; ++e)
        for (uint8_t p = 0; p < ea[e]; ++p) {
            uint32_t w = (uint32_t)dscr[ev[eoff[e] + p]]++;
            occ_edge[w] = e; occ_pos[w] = p;
        }
    /* initial colours, wl_core.hpp:93-107 */
    for (uint32_t v = 0; v < n_verts; ++v) {
        uint64_t h = FNV_OFFSET;
        uint32_t deg = occ_off[v + 1] - occ_off[v];
        h = fnv(h, deg);
        uint32_t cnt = 0;
        for (uint32_t j = occ_off[v]; j < occ_off[v + 1] && cnt < nbr_cap; ++j)
            nbr[cnt++] = ((uint64_t)ea[occ_edge[j]] << 8) | (uint64_t)occ_pos[j];
        isort_u64(nbr, cnt);
        for (uint32_t i = 0; i < cnt; ++i) { h = fnv(h, nbr[i] >> 8); h = fnv(h, nbr[i] & 0xFFu); }
        cur[v] = h;
    }
    uint32_t prev_d = ndistinct(dscr, cur, n_verts);
    /* refinement, wl_core.hpp:119-141 */
    for (uint32_t it = 0; it < WL_MAX_ITERS; ++it) {
        for (uint32_t v = 0; v < n_verts; ++v) {
            uint64_t h = cur[v];
            uint32_t nn = 0;
            for (uint32_t j = occ_off[v]; j < occ_off[v + 1]; ++j) {
                uint32_t e = occ_edge[j];
                uint8_t mypos = occ_pos[j];
                for (uint8_t k = 0; k < ea[e]; ++k) {
                    if (k == mypos) continue;
                    if (nn < nbr_cap) nbr[nn++] = fnv(cur[ev[eoff[e] + k]], k);
                }
            }
            isort_u64(nbr, nn);
            for (uint32_t i = 0; i < nn; ++i) h = fnv(h, nbr[i]);
            nxt[v] = h;
        }
        for (uint32_t v = 0; v < n_verts; ++v) cur[v] = nxt[v];
        uint32_t d = ndistinct(dscr, cur, n_verts);
        if (d == prev_d) break;
        prev_d = d;
    }
    /* components, wl_core.hpp:157-168 */
    for (uint32_t v = 0; v < n_verts; ++v) comp[v] = v;
    for (uint32_t e = 0; e < n_edges; ++e) {
        if (ea[e] == 0) continue;
        uint32_t r = wl_find(comp, ev[eoff[e]]);
        for (uint8_t p = 1; p < ea[e]; ++p) {
            uint32_t x = wl_find(comp, ev[eoff[e] + p]);
            if (x != r) comp[x] = r;
        }
    }
    for (uint32_t v = 0; v < n_verts; ++v) comp[v] = wl_find(comp, v);
    /* folds, wl_core.hpp:170-201 */
    for (uint32_t v = 0; v < n_verts; ++v) { nxt[v] = 0; dscr[v] = 0; }
    for (uint32_t v = 0; v < n_verts; ++v) { nxt[comp[v]] += splitmix(cur[v]); dscr[comp[v]] += 1; }
    uint64_t vsum = 0;
    for (uint32_t v = 0; v < n_verts; ++v) vsum += splitmix(cur[v]);
    uint64_t esum = 0;
    for (uint32_t e = 0; e < n_edges; ++e) {
        uint64_t eh = fnv(FNV_OFFSET, ea[e]);
        for (uint8_t p = 0; p < ea[e]; ++p) eh = fnv(eh, cur[ev[eoff[e] + p]]);
        esum += splitmix(eh);
        dscr[comp[ev[eoff[e]]]] += ((uint64_t)1 << 32);
    }
    uint64_t csum = 0;
    for (uint32_t v = 0; v < n_verts; ++v) {
        if (comp[v] != v) continue;
        uint64_t ch = fnv(FNV_OFFSET, dscr[v]);
        ch = fnv(ch, nxt[v]);
        csum += splitmix(ch);
    }
    uint64_t hash = FNV_OFFSET;
    hash = fnv(hash, n_verts);
    hash = fnv(hash, n_edges);
    hash = fnv(hash, vsum);
    hash = fnv(hash, esum);
    hash = fnv(hash, csum);
    return hash;
}

static uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) { uint32_t mid = (lo + hi) / 2; if (a[mid] < x) lo = mid + 1; else hi = mid; }
    return lo;
}

/* Caller side: Hypergraph::compute_wl_hash (hypergraph.cpp:740-755) and
 * WLHash::compute_state_hash_impl (wl_hash.hpp:109-239). Edges arrive in state order
 * (ascending EdgeId). Local index = sorted-vertex position; nbr_cap = max over vertices of
 * the summed (arity - 1) of its occurrences, at least 1. */
static uint64_t wl_state_hash(wl_scratch* s, uint32_t n_edges, const uint8_t* arity,
                              const uint32_t* const* evs) {
    if (n_edges == 0) return EMPTY_STATE_HASH;
    s->verts.n = 0;
    uint64_t total_occ = 0;
    for (uint32_t e = 0; e < n_edges; ++e) {
        for (uint8_t i = 0; i < arity[e]; ++i) VPUSH(s->verts, evs[e][i]);
        total_occ += arity[e];
    }
    qsort(s->verts.p, s->verts.n, sizeof(uint32_t), cmp_u32);
    uint32_t nv = 0;
    for (uint64_t i = 0; i < s->verts.n; ++i)
        if (i == 0 || s->verts.p[i] != s->verts.p[i - 1]) s->verts.p[nv++] = s->verts.p[i];
    if (nv == 0) return 0;
    VRESERVE(s->ea, n_edges); VRESERVE(s->eoff, n_edges); VRESERVE(s->ev, total_occ);
    uint32_t off = 0;
    for (uint32_t e = 0; e < n_edges; ++e) {
        s->eoff.p[e] = off; s->ea.p[e] = arity[e];
        for (uint8_t i = 0; i < arity[e]; ++i)
            s->ev.p[off++] = lower_bound_u32(s->verts.p, nv, evs[e][i]);
    }
    VRESERVE(s->nn, nv);
    for (uint32_t v = 0; v < nv; ++v) s->nn.p[v] = 0;
    for (uint32_t e = 0; e < n_edges; ++e) {
        uint32_t contrib = arity[e] ? (uint32_t)(arity[e] - 1) : 0u;
        for (uint8_t p = 0; p < arity[e]; ++p) s->nn.p[s->ev.p[s->eoff.p[e] + p]] += contrib;
    }
    uint32_t nbr_cap = 1;
    for (uint32_t v = 0; v < nv; ++v) if (s->nn.p[v] > nbr_cap) nbr_cap = s->nn.p[v];
    VRESERVE(s->cur, nv); VRESERVE(s->nxt, nv); VRESERVE(s->dscr, nv);
    VRESERVE(s->occ_off, (uint64_t)nv + 1); VRESERVE(s->occ_edge, total_occ);
    VRESERVE(s->occ_pos, total_occ); VRESERVE(s->nbr, nbr_cap); VRESERVE(s->comp, nv);
    return wl_core(s->ea.p, s->eoff.p, s->ev.p, n_edges, nv, s->cur.p, s->nxt.p, s->occ_off.p,
                   s->occ_edge.p, s->occ_pos.p, s->nbr.p, nbr_cap, s->dscr.p, s->comp.p);
}

/* ---------------- one state, seen by the join ---------------- */

typedef struct {
    uint32_t n;                 /* edges in the state                       */
    const uint8_t* ar;          /* [n] arity                                */
    const uint32_t* const* vs;  /* [n] vertex tuple pointers                */
    /* optional incidence index (vertex -> ascending edge positions), built for big states */
    int indexed;
    VEC(uint32_t) inc_vert;     /* sorted distinct vertices                 */
    VEC(uint32_t) inc_off;      /* [n_vert+1]                               */
    VEC(uint32_t) inc_edge;
} join_view;

static void join_view_index(join_view* jv) {
    /* Restates the role of InvertedVertexIndex (hypergraph/src/index.cpp:45-80): one entry per
     * (vertex, edge), deduplicated within an edge, listed by ascending edge position. */
    jv->indexed = 1;
    VEC(uint64_t) pairs = {0};
    for (uint32_t e = 0; e < jv->n; ++e)
        for (uint8_t i = 0; i < jv->ar[e]; ++i) {
            int dup = 0;
            for (uint8_t j = 0; j < i; ++j) if (jv->vs[e][j] == jv->vs[e][i]) dup = 1;
            if (!dup) VPUSH(pairs, ((uint64_t)jv->vs[e][i] << 32) | e);
        }
    qsort(pairs.p, pairs.n, sizeof(uint64_t), cmp_u64);
    jv->inc_vert.n = jv->inc_off.n = jv->inc_edge.n = 0;
    for (uint64_t i = 0; i < pairs.n; ++i) {
        uint32_t v = (uint32_t)(pairs.p[i] >> 32);
        if (jv->inc_vert.n == 0 || jv->inc_vert.p[jv->inc_vert.n - 1] != v) {
            VPUSH(jv->inc_vert, v); VPUSH(jv->inc_off, (uint32_t)i);
        }
        VPUSH(jv->inc_edge, (uint32_t)(pairs.p[i] & 0xFFFFFFFFu));
    }
    VPUSH(jv->inc_off, (uint32_t)pairs.n);
    free(pairs.p);
}

static void join_view_free(join_view* jv) {
    free(jv->inc_vert.p); free(jv->inc_off.p); free(jv->inc_edge.p);
}

typedef void (*emit_fn)(void* ctx, const uint32_t* matched_by_pattern, const uint32_t* binding);

typedef struct {
    const orule* R;
    const join_view* jv;
    uint32_t matched[16];       /* by depth */
    uint8_t pattern[16];        /* by depth */
    uint32_t binding[32];
    uint32_t bound;
    uint8_t depth;
    emit_fn emit; void* emit_ctx;
} join_st;

/* bind_pattern_edge, match_core.hpp:32-49 */
static int bind_edge(const uint32_t* ev, uint8_t ea, const uint8_t* pv, uint8_t pa,
                     uint32_t* binding, uint32_t* bound) {
    if (ea != pa) return 0;
    for (uint8_t i = 0; i < ea; ++i) {
        uint8_t var = pv[i]; uint32_t bit = 1u << var;
        if (*bound & bit) { if (binding[var] != ev[i]) return 0; }
        else { binding[var] = ev[i]; *bound |= bit; }
    }
    return 1;
}

static void join_try(join_st* st, uint8_t p, uint32_t e);

/* join_dfs, join_core.hpp:143-180. Candidates in ascending edge position: every edge, or
 * the incidence list of the first bound vertex of the pattern edge when the state is indexed
 * (the same candidate set after the bind filter, in the same order). */
static void join_dfs(join_st* st) {
    const orule* R = st->R;
    if (st->depth == R->n_lhs) {
        uint32_t by_pattern[16];
        for (uint8_t d = 0; d < st->depth; ++d) by_pattern[st->pattern[d]] = st->matched[d];
        st->emit(st->emit_ctx, by_pattern, st->binding);
        return;
    }
    uint32_t bpm = 0;
    for (uint8_t d = 0; d < st->depth; ++d) bpm |= 1u << st->pattern[d];
    uint8_t p = 0xFF;                                   /* join_next_position, :102-110 */
    for (uint8_t k = 0; k < R->n_lhs; ++k)
        if (!(bpm & (1u << R->order[k]))) { p = R->order[k]; break; }
    if (p == 0xFF) return;
    const join_view* jv = st->jv;
    if (jv->indexed) {
        for (uint8_t i = 0; i < R->la[p]; ++i) {
            uint8_t var = R->lv[p][i];
            if (st->bound & (1u << var)) {
                uint32_t v = st->binding[var];
                uint32_t k = lower_bound_u32(jv->inc_vert.p, (uint32_t)jv->inc_vert.n, v);
                if (k == jv->inc_vert.n || jv->inc_vert.p[k] != v) return;
                for (uint32_t j = jv->inc_off.p[k]; j < jv->inc_off.p[k + 1]; ++j)
                    join_try(st, p, jv->inc_edge.p[j]);
                return;
            }
        }
    }
    for (uint32_t e = 0; e < jv->n; ++e) join_try(st, p, e);
}

static void join_try(join_st* st, uint8_t p, uint32_t e) {
    for (uint8_t d = 0; d < st->depth; ++d) if (st->matched[d] == e) return;  /* injective */
    uint32_t saved = st->bound;
    if (!bind_edge(st->jv->vs[e], st->jv->ar[e], st->R->lv[p], st->R->la[p],
                   st->binding, &st->bound)) { st->bound = saved; return; }
    st->pattern[st->depth] = p; st->matched[st->depth] = e; ++st->depth;
    join_dfs(st);
    --st->depth; st->bound = saved;
}

static void join_all(const orule* R, const join_view* jv, emit_fn emit, void* ctx) {
    join_st st;
    memset(&st, 0, sizeof(st));
    st.R = R; st.jv = jv; st.emit = emit; st.emit_ctx = ctx;
    for (int v = 0; v < 32; ++v) st.binding[v] = INV;
    join_dfs(&st);
}

/* ---------------- evolution ---------------- */

#define BIG_STATE 64u

struct hgxo_result {
    hgx_counts counts;
    /* edges */
    VEC(uint8_t) e_ar; VEC(uint64_t) e_voff; VEC(uint32_t) e_prod; VEC(uint32_t) vpool;
    /* states */
    VEC(uint64_t) s_eoff; VEC(uint32_t) s_ne; VEC(uint32_t) s_step; VEC(uint32_t) s_next;
    VEC(uint32_t) s_creator; VEC(uint64_t) s_hash; VEC(uint32_t) spool;
    /* events */
    VEC(hgx_event) ev;
    VEC(uint16_t) ev_cq;    /* consumed edges' positions in the input state, pattern order */
    VEC(uint32_t) ev_cqoff; /* per event, the start of its run in ev_cq */
    VEC(uint64_t) ev_anc;   /* lineage-step mask of causal ancestors, for tests of the GPU's */
    /* relations */
    VEC(hgx_pair) causal; VEC(uint8_t) causal_tr;
    VEC(hgx_pair) branchial;
};

typedef struct {
    uint32_t parent;   /* state id */
    uint32_t rule;
    uint32_t edges[16];   /* edge ids, pattern order */
    uint32_t binding[32];
} omatch;

typedef struct {
    VEC(omatch) m;
    uint32_t parent, rule;
    const join_view* jv;
    const uint32_t* ids;   /* state's edge ids by position */
} collect_ctx;

static void collect_emit(void* c, const uint32_t* by_pattern, const uint32_t* binding) {
    collect_ctx* cc = (collect_ctx*)c;
    omatch m;
    m.parent = cc->parent; m.rule = cc->rule;
    memset(m.edges, 0xFF, sizeof(m.edges));
    for (int i = 0; i < 16; ++i) m.edges[i] = by_pattern[i];
    memcpy(m.binding, binding, sizeof(m.binding));
    VPUSH(cc->m, m);
}

static uint32_t add_edge(hgxo_result* R, const uint32_t* vs, uint8_t a, uint32_t producer) {
    uint32_t id = (uint32_t)R->e_ar.n;
    VPUSH(R->e_ar, a); VPUSH(R->e_voff, R->vpool.n); VPUSH(R->e_prod, producer);
    for (uint8_t i = 0; i < a; ++i) VPUSH(R->vpool, vs[i]);
    return id;
}

static uint64_t state_hash(hgxo_result* R, wl_scratch* ws, uint32_t sid) {
    uint32_t ne = R->s_ne.p[sid];
    const uint32_t* ids = R->spool.p + R->s_eoff.p[sid];
    uint8_t* ar = malloc(ne ? ne : 1);
    const uint32_t** vs = malloc(sizeof(uint32_t*) * (ne ? ne : 1));
    for (uint32_t i = 0; i < ne; ++i) { ar[i] = R->e_ar.p[ids[i]]; vs[i] = R->vpool.p + R->e_voff.p[ids[i]]; }
    uint64_t h = wl_state_hash(ws, ne, ar, vs);
    free(ar); free(vs);
    return h;
}

/* distinct-u64 set (open addressing; 0 tracked by a flag) */
typedef struct { uint64_t* k; uint64_t cap, n; int has0; } u64set;
static int u64set_insert(u64set* s, uint64_t x) {
    if (x == 0) { if (s->has0) return 0; s->has0 = 1; return 1; }
    if ((s->n + 1) * 2 > s->cap) {
        uint64_t oc = s->cap; uint64_t* ok = s->k;
        s->cap = oc ? oc * 2 : 1024; s->k = calloc(s->cap, 8); s->n = 0;
        for (uint64_t i = 0; i < oc; ++i) if (ok[i]) {
            uint64_t j = mix64(ok[i]) & (s->cap - 1);
            while (s->k[j]) j = (j + 1) & (s->cap - 1);
            s->k[j] = ok[i]; s->n++;
        }
        free(ok);
    }
    uint64_t j = mix64(x) & (s->cap - 1);
    while (s->k[j]) { if (s->k[j] == x) return 0; j = (j + 1) & (s->cap - 1); }
    s->k[j] = x; s->n++;
    return 1;
}

/* tr_reduce restated (transitive_reduction.hpp:36-96), ids topological. Marks in_tr for the
 * pairs of `pairs` (which must be deduplicated) that survive. */
static void tr_mark(const hgx_pair* pairs_in, uint64_t n, uint8_t* in_tr) {
    if (n == 0) return;
    hgx_pair* all = malloc(n * sizeof(hgx_pair));
    memcpy(all, pairs_in, n * sizeof(hgx_pair));
    qsort(all, n, sizeof(hgx_pair), cmp_pair);
    uint32_t max_id = 0;
    for (uint64_t i = 0; i < n; ++i) { if (all[i].a > max_id) max_id = all[i].a; if (all[i].b > max_id) max_id = all[i].b; }
    uint64_t* succ = calloc((uint64_t)max_id + 2, 8);
    for (uint64_t i = 0; i < n; ++i) succ[all[i].a + 1]++;
    for (uint64_t k = 1; k < (uint64_t)max_id + 2; ++k) succ[k] += succ[k - 1];
    uint32_t* seen = calloc((uint64_t)max_id + 1, 4);
    uint32_t stamp = 0;
    uint32_t* stack = malloc(((uint64_t)max_id + 1) * 4);
    uint8_t* keep = calloc(n, 1);
    for (uint64_t i = 0; i < n;) {
        uint32_t p = all[i].a;
        uint64_t j = succ[p + 1];
        uint32_t cmax = all[j - 1].b;
        ++stamp;
        uint64_t sp = 0;
        for (uint64_t a = succ[p]; a < succ[p + 1]; ++a) {
            uint32_t w = all[a].b;
            for (uint64_t b = succ[w]; b < succ[w + 1]; ++b) {
                uint32_t x = all[b].b;
                if (x > cmax) continue;
                if (seen[x] != stamp) { seen[x] = stamp; stack[sp++] = x; }
            }
        }
        while (sp) {
            uint32_t x = stack[--sp];
            for (uint64_t b = succ[x]; b < succ[x + 1]; ++b) {
                uint32_t y = all[b].b;
                if (y > cmax) continue;
                if (seen[y] != stamp) { seen[y] = stamp; stack[sp++] = y; }
            }
        }
        for (uint64_t k = i; k < j; ++k) keep[k] = (seen[all[k].b] != stamp);
        i = j;
    }
    /* map back: pairs_in order -> sorted position by binary search */
    for (uint64_t i = 0; i < n; ++i) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) { uint64_t mid = (lo + hi) / 2; if (cmp_pair(&all[mid], &pairs_in[i]) < 0) lo = mid + 1; else hi = mid; }
        in_tr[i] = keep[lo];
    }
    free(all); free(succ); free(seen); free(stack); free(keep);
}

static int validate_run(const hgx_run_desc* d, orule* rules) {
    if (!d || !d->rules || d->n_rules == 0 || d->n_rules > 65535) return -1;
    if (parse_rules(d->rules, d->rules_len, d->n_rules, rules) != 0) return -1;
    if (d->steps > HGX_MAX_STEPS) return -1;
    if (d->n_edges && (!d->edge_offsets || !d->vertices)) return -1;
    for (uint32_t e = 0; e < d->n_edges; ++e) {
        uint32_t a = d->edge_offsets[e + 1] - d->edge_offsets[e];
        if (d->edge_offsets[e + 1] < d->edge_offsets[e] || a < 1 || a > 16) return -1;
    }
    if (d->root_offsets) {
        if (d->n_roots == 0) return -1;
        if (d->root_offsets[0] != 0 || d->root_offsets[d->n_roots] != d->n_edges) return -1;
        for (uint32_t r = 0; r < d->n_roots; ++r) if (d->root_offsets[r + 1] < d->root_offsets[r]) return -1;
    }
    return 0;
}

int hgxo_evolve(const hgx_run_desc* d, hgxo_result** out) {
    orule* rules = calloc(d && d->n_rules ? d->n_rules : 1, sizeof(orule));
    if (validate_run(d, rules) != 0) { free(rules); return HGX_ERR_INVALID; }
    hgxo_result* R = calloc(1, sizeof(hgxo_result));
    wl_scratch ws; memset(&ws, 0, sizeof(ws));
    const int rec_causal = (d->flags & HGX_REC_CAUSAL) != 0;
    const int rec_branch = (d->flags & HGX_REC_BRANCHIAL) != 0;

    /* Roots: each renumbered densely by first appearance (hypergraph_ffi.cpp:599-636). */
    uint32_t n_roots = d->root_offsets ? d->n_roots : 1;
    for (uint32_t r = 0; r < n_roots; ++r) {
        uint32_t e0 = d->root_offsets ? d->root_offsets[r] : 0;
        uint32_t e1 = d->root_offsets ? d->root_offsets[r + 1] : d->n_edges;
        VEC(uint32_t) seen = {0};
        uint32_t sid = (uint32_t)R->s_ne.n;
        VPUSH(R->s_eoff, R->spool.n);
        for (uint32_t e = e0; e < e1; ++e) {
            uint32_t vs[16];
            uint8_t a = (uint8_t)(d->edge_offsets[e + 1] - d->edge_offsets[e]);
            for (uint8_t i = 0; i < a; ++i) {
                uint32_t v = d->vertices[d->edge_offsets[e] + i], k;
                for (k = 0; k < seen.n; ++k) if (seen.p[k] == v) break;
                if (k == seen.n) VPUSH(seen, v);
                vs[i] = k;
            }
            VPUSH(R->spool, add_edge(R, vs, a, INV));
        }
        VPUSH(R->s_ne, e1 - e0); VPUSH(R->s_step, 0); VPUSH(R->s_next, (uint32_t)seen.n);
        VPUSH(R->s_creator, INV);
        VPUSH(R->s_hash, state_hash(R, &ws, sid));
        free(seen.p);
    }

    uint64_t lvl_begin = 0, lvl_end = n_roots;
    uint32_t steps_done = 0;
    int truncated = 0;
    for (uint32_t step = 1; step <= d->steps && lvl_begin < lvl_end; ++step) {
        /* all matches of the generation: parent order, rule order, join order */
        collect_ctx cc; memset(&cc, 0, sizeof(cc));
        VEC(uint64_t) par_begin = {0};      /* first match index of each parent */
        for (uint64_t s = lvl_begin; s < lvl_end; ++s) {
            VPUSH(par_begin, cc.m.n);
            join_view jv; memset(&jv, 0, sizeof(jv));
            jv.n = R->s_ne.p[s];
            const uint32_t* ids = R->spool.p + R->s_eoff.p[s];
            uint8_t* ar = malloc(jv.n ? jv.n : 1);
            const uint32_t** vs = malloc(sizeof(uint32_t*) * (jv.n ? jv.n : 1));
            for (uint32_t i = 0; i < jv.n; ++i) { ar[i] = R->e_ar.p[ids[i]]; vs[i] = R->vpool.p + R->e_voff.p[ids[i]]; }
            jv.ar = ar; jv.vs = vs;
            if (jv.n > BIG_STATE) join_view_index(&jv);
            cc.parent = (uint32_t)s; cc.ids = ids; cc.jv = &jv;
            for (uint32_t r = 0; r < d->n_rules; ++r) { cc.rule = r; join_all(&rules[r], &jv, collect_emit, &cc); }
            /* translate positions to edge ids */
            join_view_free(&jv);
            free(ar); free(vs);
        }
        VPUSH(par_begin, cc.m.n);
        /* positions -> edge ids (the join saw positions) */
        for (uint64_t i = 0, p = 0; i < cc.m.n; ++i) {
            while (par_begin.p[p + 1] <= i) ++p;
            const uint32_t* ids = R->spool.p + R->s_eoff.p[lvl_begin + p];
            const orule* rr = &rules[cc.m.p[i].rule];
            for (uint8_t k = 0; k < rr->n_lhs; ++k) cc.m.p[i].edges[k] = ids[cc.m.p[i].edges[k]];
        }
        /* deterministic cap: stop before a generation that would exceed it */
        if (d->state_cap && R->s_ne.n + cc.m.n > d->state_cap) {
            truncated = 1;
            free(cc.m.p); free(par_begin.p);
            break;
        }
        uint64_t next_begin = R->s_ne.n;
        for (uint64_t p = 0; p + 1 < par_begin.n; ++p) {
            uint32_t parent = (uint32_t)(lvl_begin + p);
            uint64_t m0 = par_begin.p[p], m1 = par_begin.p[p + 1];
            uint32_t pne = R->s_ne.p[parent];
            for (uint64_t i = m0; i < m1; ++i) {
                const omatch* m = &cc.m.p[i];
                const orule* rr = &rules[m->rule];
                uint32_t evid = (uint32_t)R->ev.n;
                uint32_t child = (uint32_t)R->s_ne.n;
                /* Rewriter::apply, rewriter.cpp:47-89: child = parent minus consumed, then the
                 * produced edges in RHS order; fresh vertices ascending by variable. */
                uint32_t fresh[32];
                for (int v = 0; v < 32; ++v) fresh[v] = INV;
                uint32_t base = R->s_next.p[parent], nxt = base;
                for (uint32_t mk = rr->new_mask; mk; mk &= mk - 1) fresh[__builtin_ctz(mk)] = nxt++;
                VPUSH(R->s_eoff, R->spool.n);
                /* spool may realloc: copy parent ids first */
                uint32_t* tmp = malloc(sizeof(uint32_t) * (pne ? pne : 1));
                memcpy(tmp, R->spool.p + R->s_eoff.p[parent], sizeof(uint32_t) * pne);
                uint32_t cne = 0;
                for (uint32_t q = 0; q < pne; ++q) {
                    int used = 0;
                    for (uint8_t k = 0; k < rr->n_lhs; ++k) if (m->edges[k] == tmp[q]) used = 1;
                    if (!used) { VPUSH(R->spool, tmp[q]); ++cne; }
                }
                free(tmp);
                for (uint8_t j = 0; j < rr->n_rhs; ++j) {
                    uint32_t vs[16];
                    for (uint8_t i = 0; i < rr->ra[j]; ++i) {       /* resolve_rhs_vertices */
                        uint8_t var = rr->rv[j][i];
                        vs[i] = m->binding[var] != INV && (rr->lhs_mask >> var & 1) ? m->binding[var] : fresh[var];
                    }
                    VPUSH(R->spool, add_edge(R, vs, rr->ra[j], evid));
                    ++cne;
                }
                VPUSH(R->s_ne, cne); VPUSH(R->s_step, step); VPUSH(R->s_next, nxt);
                VPUSH(R->s_creator, evid);
                VPUSH(R->s_hash, state_hash(R, &ws, child));
                /* consumed edges as positions in the parent, pattern order */
                VPUSH(R->ev_cqoff, (uint32_t)R->ev_cq.n);
                for (uint8_t k = 0; k < rr->n_lhs; ++k)
                    for (uint32_t q = 0; q < pne; ++q)
                        if (R->spool.p[R->s_eoff.p[parent] + q] == m->edges[k]) { VPUSH(R->ev_cq, (uint16_t)q); break; }
                hgx_event ev = { parent, child, m->rule, step };
                VPUSH(R->ev, ev);
                /* causal: one triple per consumed edge with a producer; pairs deduplicated
                 * (rewriter.cpp:121-191, causal_graph.cpp:187-234). */
                uint64_t anc = 0;
                if (rec_causal) {
                    uint32_t prods[16]; int np = 0;
                    for (uint8_t k = 0; k < rr->n_lhs; ++k) {
                        uint32_t P = R->e_prod.p[m->edges[k]];
                        if (P == INV) continue;
                        R->counts.causal_triples++;
                        int dup = 0;
                        for (int t = 0; t < np; ++t) if (prods[t] == P) dup = 1;
                        if (!dup) prods[np++] = P;
                    }
                    for (int t = 0; t < np; ++t) {
                        hgx_pair pr = { prods[t], evid };
                        VPUSH(R->causal, pr);
                        anc |= (1ull << R->ev.p[prods[t]].step) | R->ev_anc.p[prods[t]];
                    }
                }
                VPUSH(R->ev_anc, anc);
            }
            /* branchial: sibling events of one input state whose consumed sets intersect,
             * one pair each (causal_graph.cpp:324-351). */
            if (rec_branch) {
                for (uint64_t i = m0; i < m1; ++i)
                    for (uint64_t j = i + 1; j < m1; ++j) {
                        const omatch* a = &cc.m.p[i]; const omatch* b = &cc.m.p[j];
                        int hit = 0;
                        for (uint8_t x = 0; x < rules[a->rule].n_lhs && !hit; ++x)
                            for (uint8_t y = 0; y < rules[b->rule].n_lhs; ++y)
                                if (a->edges[x] == b->edges[y]) { hit = 1; break; }
                        if (hit) {
                            uint32_t base_ev = (uint32_t)(R->ev.n - (m1 - m0));
                            hgx_pair pr = { base_ev + (uint32_t)(i - m0), base_ev + (uint32_t)(j - m0) };
                            VPUSH(R->branchial, pr);
                        }
                    }
            }
        }
        free(cc.m.p); free(par_begin.p);
        if (R->s_ne.n > next_begin) steps_done = step;   /* last generation with an event */
        lvl_begin = next_begin; lvl_end = R->s_ne.n;
    }

    /* counts */
    R->counts.raw_states = R->s_ne.n;
    R->counts.events = R->ev.n;
    u64set hs = {0};
    uint64_t fp = 0;
    for (uint64_t s = 0; s < R->s_hash.n; ++s)
        if (u64set_insert(&hs, R->s_hash.p[s])) fp += mix64(R->s_hash.p[s]);
    R->counts.canonical_states = hs.n + (uint64_t)hs.has0;
    R->counts.fingerprint = fp;
    free(hs.k);
    R->counts.causal_pairs = R->causal.n;
    VRESERVE(R->causal_tr, R->causal.n ? R->causal.n : 1);
    R->causal_tr.n = R->causal.n;
    if (R->causal.n && (d->flags & HGX_OPT_TR)) tr_mark(R->causal.p, R->causal.n, R->causal_tr.p);
    else for (uint64_t i = 0; i < R->causal.n; ++i) R->causal_tr.p[i] = 1;
    for (uint64_t i = 0; i < R->causal.n; ++i) R->counts.causal_pairs_tr += R->causal_tr.p[i];
    /* triples whose pair survives: recount per consumer */
    if (rec_causal) {
        /* each pair (P,C) carries as many triples as C consumed edges produced by P */
        for (uint64_t i = 0; i < R->causal.n; ++i) {
            if (!R->causal_tr.p[i]) continue;
            uint32_t C = R->causal.p[i].b, P = R->causal.p[i].a;
            /* consumed edges of C = parent's edges not in child; count those produced by P */
            uint32_t par = R->ev.p[C].input_state, ch = R->ev.p[C].output_state;
            const uint32_t* pid = R->spool.p + R->s_eoff.p[par];
            const uint32_t* cid = R->spool.p + R->s_eoff.p[ch];
            uint32_t pn = R->s_ne.p[par], cn = R->s_ne.p[ch];
            for (uint32_t q = 0; q < pn; ++q) {
                if (R->e_prod.p[pid[q]] != P) continue;
                int kept = 0;
                for (uint32_t t = 0; t < cn; ++t) if (cid[t] == pid[q]) { kept = 1; break; }
                if (!kept) R->counts.causal_triples_tr++;
            }
        }
    }
    R->counts.branchial_pairs = R->branchial.n;
    R->counts.steps_done = steps_done;
    R->counts.truncated = (uint32_t)truncated;
    if (!(d->flags & HGX_OPT_TR)) R->counts.causal_pairs_tr = R->counts.causal_pairs;
    wl_scratch_free(&ws);
    free(rules);
    *out = R;
    return HGX_OK;
}

void hgxo_free(hgxo_result* R) {
    if (!R) return;
    free(R->e_ar.p); free(R->e_voff.p); free(R->e_prod.p); free(R->vpool.p);
    free(R->s_eoff.p); free(R->s_ne.p); free(R->s_step.p); free(R->s_next.p);
    free(R->s_creator.p); free(R->s_hash.p); free(R->spool.p); free(R->ev.p); free(R->ev_anc.p);
    free(R->ev_cq.p); free(R->ev_cqoff.p);
    free(R->causal.p); free(R->causal_tr.p); free(R->branchial.p);
    free(R);
}

int hgxo_counts(const hgxo_result* R, hgx_counts* out) { *out = R->counts; return HGX_OK; }

int hgxo_states(const hgxo_result* R, uint64_t* hash, uint32_t* step, uint32_t* creator) {
    for (uint64_t s = 0; s < R->s_ne.n; ++s) {
        if (hash) hash[s] = R->s_hash.p[s];
        if (step) step[s] = R->s_step.p[s];
        if (creator) creator[s] = R->s_creator.p[s];
    }
    return HGX_OK;
}

int hgxo_events(const hgxo_result* R, hgx_event* out) {
    if (R->ev.n) memcpy(out, R->ev.p, R->ev.n * sizeof(hgx_event));
    return HGX_OK;
}

uint64_t hgxo_num_causal_pairs(const hgxo_result* R) { return R->causal.n; }
int hgxo_causal(const hgxo_result* R, hgx_pair* pairs, uint8_t* in_tr) {
    for (uint64_t i = 0; i < R->causal.n; ++i) {
        if (pairs) pairs[i] = R->causal.p[i];
        if (in_tr) in_tr[i] = R->causal_tr.p[i];
    }
    return HGX_OK;
}

uint64_t hgxo_num_branchial_pairs(const hgxo_result* R) { return R->branchial.n; }
int hgxo_branchial(const hgxo_result* R, hgx_pair* pairs) {
    if (R->branchial.n) memcpy(pairs, R->branchial.p, R->branchial.n * sizeof(hgx_pair));
    return HGX_OK;
}

int hgxo_state_edges(const hgxo_result* R, uint64_t* soff, uint32_t* eoff, uint32_t* verts,
                     uint64_t* n_edges_total, uint64_t* n_vertices_total) {
    uint64_t ne = 0, nv = 0;
    for (uint64_t s = 0; s < R->s_ne.n; ++s) {
        if (soff) soff[s] = ne;
        const uint32_t* ids = R->spool.p + R->s_eoff.p[s];
        for (uint32_t i = 0; i < R->s_ne.p[s]; ++i) {
            uint8_t a = R->e_ar.p[ids[i]];
            if (eoff) eoff[ne] = (uint32_t)nv;
            if (verts) memcpy(verts + nv, R->vpool.p + R->e_voff.p[ids[i]], a * 4);
            nv += a; ++ne;
        }
    }
    if (soff) soff[R->s_ne.n] = ne;
    if (eoff) eoff[ne] = (uint32_t)nv;
    if (n_edges_total) *n_edges_total = ne;
    if (n_vertices_total) *n_vertices_total = nv;
    return HGX_OK;
}

int hgxo_consumed(const hgxo_result* R, uint32_t* offsets, uint16_t* positions, uint64_t* total) {
    const uint64_t ne = R->ev.n;
    if (offsets) {
        for (uint64_t e = 0; e < ne; ++e) offsets[e] = R->ev_cqoff.p[e];
        offsets[ne] = (uint32_t)R->ev_cq.n;
    }
    if (positions && R->ev_cq.n) memcpy(positions, R->ev_cq.p, R->ev_cq.n * sizeof(uint16_t));
    if (total) *total = R->ev_cq.n;
    return HGX_OK;
}

uint64_t hgxo_wl_hash(const uint32_t* edge_offsets, const uint32_t* vertices, uint32_t n_edges) {
    wl_scratch ws; memset(&ws, 0, sizeof(ws));
    uint8_t* ar = malloc(n_edges ? n_edges : 1);
    const uint32_t** vs = malloc(sizeof(uint32_t*) * (n_edges ? n_edges : 1));
    for (uint32_t e = 0; e < n_edges; ++e) {
        ar[e] = (uint8_t)(edge_offsets[e + 1] - edge_offsets[e]);
        vs[e] = vertices + edge_offsets[e];
    }
    uint64_t h = wl_state_hash(&ws, n_edges, ar, vs);
    free(ar); free(vs); wl_scratch_free(&ws);
    return h;
}

typedef struct { uint64_t n, cap; uint32_t* out; uint8_t k; } count_ctx;
static void count_emit(void* c, const uint32_t* by_pattern, const uint32_t* binding) {
    (void)binding;
    count_ctx* cc = (count_ctx*)c;
    if (cc->out && cc->n < cc->cap)
        for (uint8_t i = 0; i < cc->k; ++i) cc->out[cc->n * cc->k + i] = by_pattern[i];
    cc->n++;
}

int64_t hgxo_match(const uint8_t* rules, uint32_t rules_len, uint32_t n_rules, uint32_t rule,
                   const uint32_t* edge_offsets, const uint32_t* vertices, uint32_t n_edges,
                   uint32_t* match_edges, uint64_t cap) {
    if (rule >= n_rules) return -1;
    orule* rs = calloc(n_rules, sizeof(orule));
    if (parse_rules(rules, rules_len, n_rules, rs) != 0) { free(rs); return -1; }
    join_view jv; memset(&jv, 0, sizeof(jv));
    jv.n = n_edges;
    uint8_t* ar = malloc(n_edges ? n_edges : 1);
    const uint32_t** vs = malloc(sizeof(uint32_t*) * (n_edges ? n_edges : 1));
    for (uint32_t e = 0; e < n_edges; ++e) {
        ar[e] = (uint8_t)(edge_offsets[e + 1] - edge_offsets[e]);
        vs[e] = vertices + edge_offsets[e];
    }
    jv.ar = ar; jv.vs = vs;
    if (n_edges > BIG_STATE) join_view_index(&jv);
    count_ctx cc = { 0, cap, match_edges, rs[rule].n_lhs };
    join_all(&rs[rule], &jv, count_emit, &cc);
    join_view_free(&jv); free(ar); free(vs); free(rs);
    return (int64_t)cc.n;
}

int hgxo_match_order(const uint8_t* rules, uint32_t rules_len, uint32_t n_rules, uint32_t rule,
                     uint8_t* order_out) {
    if (rule >= n_rules) return HGX_ERR_INVALID;
    orule* rs = calloc(n_rules, sizeof(orule));
    if (parse_rules(rules, rules_len, n_rules, rs) != 0) { free(rs); return HGX_ERR_INVALID; }
    memcpy(order_out, rs[rule].order, 16);
    free(rs);
    return HGX_OK;
}
