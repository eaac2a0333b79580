// hgx_kernels.hip -- gfx950 kernels of one multiway generation.
//
//   k_match<false>   SCAN -> EXPAND -> SINK join, counting matches per state
//   k_match<true>    the same join, writing match records in the deterministic order
//   k_expand         rewrite + WL hash + event + causal emission, one team per child state
//   k_dedup          each child's hash into its request's state map
//   k_run_events     a generation's event and causal counters, summed per request
//   k_branchial      sibling-event overlap pairs, one wave per parent state (k_brl_*: the
//                    sorted per-edge lists for parents above 128 edges)
//   k_wl_states      WL hash of given states (roots, and the hgx_wl_hash_batch entry point)
//   small helpers    per-run totals, count masking, map rehash
//
// Every kernel is wave64-native: one team (a wave, or the waves of a workgroup for large
// states) owns one state or one child, stages that state's
// edges in LDS and cooperates through ballots, mbcnt prefix sums and LDS atomics. There is no
// MFMA: nothing here is a dense contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <type_traits>

#include "hgx_device.hpp"
#include "hgx_kernels.hpp"
#include "hgx_wl.hpp"

namespace hgx {

// Logical block index with consecutive indices kept on one XCD (blocks are dealt to the 8
// XCDs round-robin), so neighbouring children -- which share their parent's edges -- hit the
// same L2. gridDim.x is a multiple of 8 for every launch that uses it.
__device__ inline uint32_t xcd_block() {
    const uint32_t g = gridDim.x, b = blockIdx.x;
    return (b % 8u) * (g / 8u) + b / 8u;
}

// ------------------------------------------------------------------------------------------
// The join. A wave owns one parent state, whose edges are staged in LDS. For each rule of the
// state's run:
//   SCAN    lanes test every edge against the first position of the rule's join order
//           (arity + the repetitions the pattern demands, match_core.hpp:32-49) and the seed
//           list is compacted in edge order with a ballot + mbcnt prefix.
//   EXPAND  each lane extends one seed depth-first over the remaining positions in join order,
//           trying candidates in ascending edge order (join_core.hpp:143-180): edge-injective,
//           vertex-non-injective.
//   SINK    per-lane match counts are prefix-summed across the wave so every match lands at
//           its place in the order (state, rule, seed, depth-first) -- the order the oracle
//           enumerates, so event numbering is identical across engines and runs.
// ------------------------------------------------------------------------------------------

// Per-lane depth-first state: the edge bound at each depth, the bound-variable mask on entry to
// each depth, the next candidate and the end of the candidate range at each depth, and the
// variable bindings. Rules up to 16 LHS edges and variable ids up to 31 keep it per lane in
// private memory (JoinLaneBig); the common rules -- at most 4 LHS edges, variables below 16 --
// keep it in LDS, interleaved across the wave (element i of a lane at word i * 64 + lane), so
// every step of the depth-first search is an LDS access rather than a private-memory one.
template <int KM, int VM>
struct JoinLaneT {
    uint32_t matched[KM];
    uint32_t smask[KM + 1];
    uint32_t cidx[KM];
    uint32_t chi[KM];
    uint32_t bind[VM];
};
using JoinLaneBig = JoinLaneT<16, 32>;

struct LaneArr {
    uint32_t* p;            // this lane's element 0
    __device__ __forceinline__ uint32_t& operator[](uint32_t i) const { return p[i * WAVE]; }
};
struct JoinLaneLds {
    static constexpr uint32_t WORDS = 4 + 5 + 4 + 4 + 16;
    LaneArr matched, smask, cidx, chi, bind;
    __device__ __forceinline__ void at(uint32_t* block, uint32_t lane) {
        matched.p = block + lane;
        smask.p = matched.p + 4 * WAVE;
        cidx.p = smask.p + 5 * WAVE;
        chi.p = cidx.p + 4 * WAVE;
        bind.p = chi.p + 4 * WAVE;
    }
};

template <class Bind>
__device__ __forceinline__ bool bind_edge(const uint32_t* ev, uint32_t ea, const uint8_t* pv, uint32_t pa,
                                 Bind bind, uint32_t& mask) {   // match_core.hpp:32-49
    if (ea != pa) return false;
    for (uint32_t i = 0; i < ea; ++i) {
        const uint32_t var = pv[i], bit = 1u << var;
        if (mask & bit) { if (bind[var] != ev[i]) return false; }
        else { bind[var] = ev[i]; mask |= bit; }
    }
    return true;
}

// Candidates of depth d: the incidence list of the pivot vertex (a vertex the pattern edge
// shares with the bound prefix; the list holds every edge containing it, ascending), or every
// edge when the pattern edge shares nothing. Either way the candidates are tried in ascending
// edge order, so the matches come out in the same order as an all-edge scan.
template <class JoinLane>
__device__ __forceinline__ void cand_range(const DRule& R, uint32_t d, const JoinLane& L, const uint32_t* iend,
                                  uint32_t ne, uint32_t& lo, uint32_t& hi) {
    const uint32_t pv = R.piv[d];
    if (pv == 0xFFu) { lo = 0; hi = ne; return; }
    const uint32_t x = L.bind[R.lv[R.order[d]][pv]];
    lo = x ? iend[x - 1] : 0u;
    hi = iend[x];
}

// Depth-first completion of one seed; returns the number of matches. With q_out != nullptr
// every match is written at out_base + k (the k-th match of this seed).
template <class JoinLane>
__device__ __forceinline__ uint32_t join_seed(const DRule& R, const uint32_t* sv, const uint8_t* sa,
                                     uint32_t ne, uint32_t A, const uint32_t* iend, const uint16_t* ie,
                                     uint32_t seed, JoinLane& L,
                                     uint16_t* q_out, uint32_t K, uint32_t out_base,
                                     uint32_t* par_out, uint32_t* rule_out, uint32_t* cne_out,
                                     uint32_t parent, uint32_t rule_id, uint32_t child_ne) {
    const uint32_t n = R.n_lhs;
    uint32_t mask = 0;
    if (!bind_edge(sv + seed * A, sa[seed], R.lv[R.order[0]], R.la[R.order[0]], L.bind, mask))
        return 0;
    uint32_t found = 0;
    L.matched[0] = seed;
    if (n == 1) {
        if (q_out) {
            q_out[(size_t)out_base * K + R.order[0]] = (uint16_t)seed;
            par_out[out_base] = parent; rule_out[out_base] = rule_id; cne_out[out_base] = child_ne;
        }
        return 1;
    }
    L.smask[1] = mask;
    cand_range(R, 1, L, iend, ne, L.cidx[1], L.chi[1]);
    uint32_t d = 1;
    while (d >= 1) {
        const uint32_t p = R.order[d];
        const bool indexed = R.piv[d] != 0xFFu;
        bool advanced = false;
        while (L.cidx[d] < L.chi[d]) {
            const uint32_t j = L.cidx[d]++;
            const uint32_t e = indexed ? (uint32_t)ie[j] : j;
            bool taken = false;
            for (uint32_t t = 0; t < d; ++t) taken |= (L.matched[t] == e);
            if (taken) continue;
            uint32_t m = L.smask[d];
            if (!bind_edge(sv + e * A, sa[e], R.lv[p], R.la[p], L.bind, m)) continue;
            L.matched[d] = e;
            if (d + 1 == n) {
                if (q_out) {
                    const size_t o = (size_t)(out_base + found);
                    for (uint32_t t = 0; t < n; ++t) q_out[o * K + R.order[t]] = (uint16_t)L.matched[t];
                    par_out[o] = parent; rule_out[o] = rule_id; cne_out[o] = child_ne;
                }
                ++found;
                continue;
            }
            L.smask[d + 1] = m;
            cand_range(R, d + 1, L, iend, ne, L.cidx[d + 1], L.chi[d + 1]);
            ++d;
            advanced = true;
            break;
        }
        if (!advanced) --d;
    }
    return found;
}

// Order of two matches of one state, as a full join emits them: rule, then the edge
// positions in the rule's join order (q arrays are in LHS order).
__device__ __forceinline__ bool match_less(const DRule* rules, uint32_t ra, const uint16_t* qa, uint32_t rb,
                                  const uint16_t* qb) {
    if (ra != rb) return ra < rb;
    const DRule& R = rules[ra];
    for (uint32_t t = 0; t < R.n_lhs; ++t) {
        const uint32_t o = R.order[t];
        if (qa[o] != qb[o]) return qa[o] < qb[o];
    }
    return false;
}

// Delta matches (find_delta_matches, hypergraph/include/hypergraph/pattern_matcher.hpp:493-537):
// pattern edge p bound to child edge `anchor` (a produced edge), the rest completed depth-first
// in the anchored order aord[p]. LHS positions below p take kept edges only (positions below
// `kept`), so a match that uses produced edges is found exactly once -- from its lowest LHS
// position holding a produced edge. With d_out the k-th match is written at d_out + k * stride
// as (rule, edge positions in LHS order).
template <class JoinLane>
__device__ __forceinline__ uint32_t join_anchor(const DRule& R, uint32_t p, const uint32_t* sv, const uint8_t* sa,
                                       uint32_t ne, uint32_t A, const uint32_t* iend, const uint16_t* ie,
                                       uint32_t kept, uint32_t anchor, JoinLane& L, uint16_t* d_out,
                                       uint32_t stride, uint32_t rule_id) {
    const uint32_t n = R.n_lhs;
    const uint8_t* ord = R.aord[p];
    const uint8_t* piv = R.apiv[p];
    uint32_t mask = 0;
    if (!bind_edge(sv + anchor * A, sa[anchor], R.lv[p], R.la[p], L.bind, mask)) return 0;
    L.matched[0] = anchor;
    auto range = [&](uint32_t d, uint32_t& lo, uint32_t& hi) {
        const uint32_t pv = piv[d];
        if (pv == 0xFFu) { lo = 0; hi = ne; return; }
        const uint32_t x = L.bind[R.lv[ord[d]][pv]];
        lo = x ? iend[x - 1] : 0u;
        hi = iend[x];
    };
    uint32_t found = 0;
    if (n == 1) {
        if (d_out) { d_out[0] = (uint16_t)rule_id; d_out[1 + p] = (uint16_t)anchor; }
        return 1;
    }
    L.smask[1] = mask;
    range(1, L.cidx[1], L.chi[1]);
    uint32_t d = 1;
    while (d >= 1) {
        const uint32_t pe = ord[d];
        const bool indexed = piv[d] != 0xFFu;
        const uint32_t lim = pe < p ? kept : ne;      // candidates ascend: stop at the limit
        bool advanced = false;
        while (L.cidx[d] < L.chi[d]) {
            const uint32_t j = L.cidx[d]++;
            const uint32_t e = indexed ? (uint32_t)ie[j] : j;
            if (e >= lim) { L.cidx[d] = L.chi[d]; break; }
            bool taken = false;
            for (uint32_t t = 0; t < d; ++t) taken |= (L.matched[t] == e);
            if (taken) continue;
            uint32_t m = L.smask[d];
            if (!bind_edge(sv + e * A, sa[e], R.lv[pe], R.la[pe], L.bind, m)) continue;
            L.matched[d] = e;
            if (d + 1 == n) {
                if (d_out) {
                    uint16_t* o = d_out + (size_t)found * stride;
                    o[0] = (uint16_t)rule_id;
                    for (uint32_t t = 0; t < n; ++t) o[1 + ord[t]] = (uint16_t)L.matched[t];
                }
                ++found;
                continue;
            }
            L.smask[d + 1] = m;
            range(d + 1, L.cidx[d + 1], L.chi[d + 1]);
            ++d;
            advanced = true;
            break;
        }
        if (!advanced) --d;
    }
    return found;
}

// Depth-first completion of one seed with the search state in registers, for rules of at most
// N = 4 LHS edges: every depth is its own template level, so the edges bound so far are a
// register array indexed by constants, and a variable's value is read from the vertex slot
// where it was first bound (DRule::jsrc) instead of from a binding table. Same candidates in
// the same order as join_seed, so the same matches in the same order.
template <int N>
__device__ __forceinline__ uint32_t pick(const uint32_t (&m)[N], uint32_t d) {
    uint32_t r = m[0];
#pragma unroll
    for (int k = 1; k < N; ++k) r = (d == (uint32_t)k) ? m[k] : r;
    return r;
}

struct RegJoinOut {
    uint16_t* q;
    uint32_t K, base, *parent, *rule, *cne;
    uint32_t state, rule_id, child_ne;
};

template <int N, int D, bool OUT>
__device__ __forceinline__ uint32_t reg_join(const DRule& R, const uint32_t* sv, const uint8_t* sa, uint32_t ne,
                                             uint32_t A, const uint32_t* iend, const uint16_t* ie,
                                             uint32_t (&m)[N], const RegJoinOut& o, uint32_t found) {
    if constexpr (D == N) {
        if constexpr (OUT) {
            const size_t at = (size_t)(o.base + found);
#pragma unroll
            for (int t = 0; t < N; ++t) o.q[at * o.K + R.order[t]] = (uint16_t)m[t];
            o.parent[at] = o.state; o.rule[at] = o.rule_id; o.cne[at] = o.child_ne;
        }
        return found + 1;
    } else {
        const uint32_t pe = R.order[D], pa = R.la[pe];
        uint32_t lo = 0, hi = ne;
        const uint32_t pv = R.piv[D];
        const bool indexed = pv != 0xFFu;
        if (indexed) {
            const uint32_t src = R.jsrc[D][pv];
            const uint32_t x = sv[pick<N>(m, src >> 4) * A + (src & 15u)];
            lo = x ? iend[x - 1] : 0u;
            hi = iend[x];
        }
        for (uint32_t j = lo; j < hi; ++j) {
            const uint32_t e = indexed ? (uint32_t)ie[j] : j;
            bool ok = sa[e] == pa;
#pragma unroll
            for (int t = 0; t < D; ++t) ok &= (m[t] != e);
            if (!ok) continue;
            for (uint32_t i = 0; i < pa && ok; ++i) {
                const uint32_t src = R.jsrc[D][i];
                if (src == 0xFFu) continue;
                const uint32_t sd = src >> 4, si = src & 15u;
                const uint32_t want = sd == (uint32_t)D ? sv[e * A + si] : sv[pick<N>(m, sd) * A + si];
                ok = sv[e * A + i] == want;
            }
            if (!ok) continue;
            m[D] = e;
            found = reg_join<N, D + 1, OUT>(R, sv, sa, ne, A, iend, ie, m, o, found);
        }
        return found;
    }
}

template <int N, bool OUT>
__device__ __forceinline__ uint32_t reg_join_seed(const DRule& R, const uint32_t* sv, const uint8_t* sa,
                                                  uint32_t ne, uint32_t A, const uint32_t* iend,
                                                  const uint16_t* ie, uint32_t seed, const RegJoinOut& o) {
    uint32_t m[N];
#pragma unroll
    for (int t = 0; t < N; ++t) m[t] = 0xFFFFFFFFu;
    m[0] = seed;
    return reg_join<N, 1, OUT>(R, sv, sa, ne, A, iend, ie, m, o, 0u);
}

// seeds already satisfy position order[0]'s pattern (the SCAN checks it)
template <bool OUT>
__device__ __forceinline__ uint32_t reg_join_any(const DRule& R, const uint32_t* sv, const uint8_t* sa,
                                                 uint32_t ne, uint32_t A, const uint32_t* iend,
                                                 const uint16_t* ie, uint32_t seed, const RegJoinOut& o) {
    switch (R.n_lhs) {
        case 1: return reg_join_seed<1, OUT>(R, sv, sa, ne, A, iend, ie, seed, o);
        case 2: return reg_join_seed<2, OUT>(R, sv, sa, ne, A, iend, ie, seed, o);
        case 3: return reg_join_seed<3, OUT>(R, sv, sa, ne, A, iend, ie, seed, o);
        default: return reg_join_seed<4, OUT>(R, sv, sa, ne, A, iend, ie, seed, o);
    }
}

uint32_t match_lds_per_wave(uint32_t max_ne, uint32_t max_nextv, uint32_t A, uint32_t KL, uint32_t dmax,
                            bool small, uint32_t team, bool lanes) {
    auto al = [](uint32_t b) { return (b + 7u) & ~7u; };
    // + 64: carving slack; + 64: the team's reduction words at the end
    uint32_t b = al(max_ne * A * 4) + al(max_ne) + al(max_ne * 2) + al((max_nextv + 1) * 4) +
                 al(max_ne * A * 2) * 2 + al(max_ne * A * 4) + 128;
    if (small) b += (lanes ? JoinLaneLds::WORDS * WAVE * 4 * team : 0u) + 2 * (uint32_t)sizeof(DRule);
    if (dmax) b += al(dmax * (KL + 1) * 2) + al(dmax * 2) + al((dmax + 1) * 4);
    return b;
}

// W is the team: one wave per state (W = 1, up to four per workgroup, each with its own
// workspace) or the W waves of a workgroup sharing one state and one workspace (large states,
// whose workspace would otherwise leave a CU one or two waves). The one-wave passes are held to
// four waves per SIMD (128 registers): at five (96) the emit pass spilled 19 registers and 96
// bytes per lane to scratch -- 25 MB of writes per launch instead of 3 MB -- and config 2 ran 3%
// slower, while the batched pass times alike under both bounds.
// GWS: the workspace is a per-workgroup slice of global memory (a.scratch) -- the same code for
// states too large for LDS, with W = 8. G < 64 (with W = 1): sub-wave teams, 64 / G states per
// wave, each team with its own workspace, for small states whose seeds would leave most of a
// wave's lanes idle.
template <bool EMIT, class JoinLane, uint32_t W, bool GWS = false, uint32_t G = 64>
__global__ __launch_bounds__(W == 1 ? 256 : 64 * W, W == 1 ? 4 : 1) void k_match(MatchArgs a) {
    using TM = typename TeamSel<W, G>::type;
    extern __shared__ uint8_t lds[];
    constexpr uint32_t N = TM::N;
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    const uint32_t ti = W == 1 ? threadIdx.x / G : 0;         // team within the workgroup
    uint8_t* p;
    if constexpr (GWS) p = a.scratch + (size_t)blockIdx.x * a.lds_per_wave;
    else p = lds + (W == 1 ? (size_t)ti * a.lds_per_wave : 0);
    const TM tm((uint64_t*)(p + a.lds_per_wave - 64));         // the workspace's last 64 bytes
    const uint32_t A = a.P.A;
    uint32_t* sv = (uint32_t*)carve(p, a.max_ne * A * 4);      // vertices
    uint8_t* sa = carve(p, a.max_ne);                           // arity
    uint16_t* seeds = (uint16_t*)carve(p, a.max_ne * 2);
    uint32_t* iend = (uint32_t*)carve(p, (a.max_nextv + 1) * 4); // incidence list ends per vertex
    uint16_t* ie = (uint16_t*)carve(p, a.max_ne * A * 2);       // incidence lists
    uint16_t* ie2 = (uint16_t*)carve(p, a.max_ne * A * 2);      // the same, before ordering
    uint32_t* iown = (uint32_t*)carve(p, a.max_ne * A * 4);     // vertex owning each list slot
    const uint32_t dstride = a.KL + 1;                          // delta match: rule + positions
    uint16_t* dlist = (uint16_t*)carve(p, a.dmax * dstride * 2);   // empty when dmax = 0
    uint16_t* dperm = (uint16_t*)carve(p, a.dmax * 2);
    uint32_t* dhist = (uint32_t*)carve(p, a.dmax ? (a.dmax + 1) * 4 : 0);
    JoinLane L;
    // Small rules (<= 4 LHS edges, <= 2 rules per request): the request's rules staged in LDS per
    // state, so every pattern read in the depth-first search is an LDS read.
    constexpr bool SMALL = std::is_same<JoinLane, JoinLaneLds>::value;
    DRule* lr = nullptr;
    if constexpr (SMALL) {
        // one lane-interleaved block per wave; only the forwarding path's anchored joins use it
        // (the full join keeps its search in registers and checks seeds by first-binding slot)
        const uint32_t lane_words = a.fwd_n ? JoinLaneLds::WORDS * WAVE * W : 0u;
        L.at((uint32_t*)carve(p, lane_words * 4) + (W == 1 || !lane_words ? 0 : w * JoinLaneLds::WORDS * WAVE), l);
        lr = (DRule*)carve(p, 2 * sizeof(DRule));
    }

    // States strided over the teams: neighbouring states (often one request's, of similar size)
    // go to different teams, which balances the launch better than contiguous chunks do.
    const uint32_t tpb = W == 1 ? (blockDim.x / WAVE) * (WAVE / G) : 1;   // teams per workgroup
    const uint32_t teams = gridDim.x * tpb;
    const uint32_t n_items = a.nrange ? a.n_items : a.P.n;
    // The count pass visits the states scattered over the launch, so each request's states are
    // counted at an even pace from start to end: a request over its cap passes it with most of
    // its states still ahead, and those are skipped at the tally check below (in launch order a
    // request's states are counted at once, and the check stops few). Blocks of
    // kCountSpreadBlock consecutive states stay together: neighbouring small states share cache
    // lines, which one state at a time would fetch again for each.
    const bool spread = !EMIT && a.spread > 1;
    const uint32_t nblk = (n_items + kCountSpreadBlock - 1) / kCountSpreadBlock;
    const uint32_t n_loop = spread ? nblk * kCountSpreadBlock : n_items;
    for (uint32_t ls = blockIdx.x * tpb + (W == 1 ? ti : 0); ls < n_loop; ls += teams) {
        uint32_t li = ls;
        if (spread) {
            li = (uint32_t)((uint64_t)(ls / kCountSpreadBlock) * a.spread % nblk) * kCountSpreadBlock +
                 ls % kCountSpreadBlock;
            if (li >= n_items) continue;                        // the last block's tail
        }
        uint32_t s = li;
        if (a.nrange) {                                         // subset: logical -> state
            uint32_t lo = 0, hi = a.nrange;
            while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (a.rpre[mid] <= li) lo = mid; else hi = mid; }
            s = a.rb[lo] + (li - a.rpre[lo]);
        }
        // the state's record and its request's flags in two rounds of independent loads
        const uint32_t run = a.P.run[s];
        const uint32_t ne = a.P.ne[s];
        const uint64_t eo = a.P.eoff[s];
        const uint32_t nvid = a.P.nextv[s];                     // labels are < nvid
        const uint8_t active = a.active[run] && (!a.smask || a.smask[s]);
        const uint64_t budget = EMIT ? 0 : (uint64_t)a.budget[run];
        // the run's tally moves under other teams: one read, shared, so the team agrees
        const uint64_t tally = EMIT ? 0 : tm.bcast(tm.leader() ? (uint64_t)a.run_acc[run] : 0ull);
        if (!active || (!EMIT && tally > budget)) {
            if (!EMIT && tm.leader()) { a.cnt[s] = 0; a.state_edges[s] = 0; }
            continue;
        }
        for (uint32_t i = tm.r; i < ne * A; i += N) sv[i] = a.P.ev[eo * A + i];
        for (uint32_t i = tm.r; i < ne; i += N) sa[i] = a.P.ar[eo + i];
        for (uint32_t v = tm.r; v <= nvid; v += N) iend[v] = 0;
        tm.sync();
        // Incidence index (vertex -> ascending edges, one entry per (vertex, edge)): counting
        // sort with LDS atomics, then each list put in edge order.
        for (uint32_t e = tm.r; e < ne; e += N)
            for (uint32_t k = 0; k < sa[e]; ++k) {
                const uint32_t x = sv[e * A + k];
                bool dup = false;
                for (uint32_t k2 = 0; k2 < k; ++k2) dup |= (sv[e * A + k2] == x);
                if (!dup) atomicAdd(&iend[x], 1u);
            }
        tm.sync();
        const uint32_t nslot = tm.excl_scan_lds(iend, nvid + 1);    // iend[x] = start of x's list
        // The emit pass writes matches in the full join's order, and the delta joins of the
        // forwarding path (join_anchor) stop a list at the first edge past their limit: both need
        // each list in ascending edge order. The count pass's full join enumerates every
        // candidate of a list, whatever its order, and joins over the lists as filled.
        const bool sorted = EMIT || s < a.fwd_n;
        if (sorted)
            for (uint32_t x = tm.r; x < nvid; x += N)
                for (uint32_t i = iend[x]; i < iend[x + 1]; ++i) iown[i] = x;
        tm.sync();
        for (uint32_t e = tm.r; e < ne; e += N)
            for (uint32_t k = 0; k < sa[e]; ++k) {
                const uint32_t x = sv[e * A + k];
                bool dup = false;
                for (uint32_t k2 = 0; k2 < k; ++k2) dup |= (sv[e * A + k2] == x);
                if (!dup) ie2[atomicAdd(&iend[x], 1u)] = (uint16_t)e;   // -> iend[x] = end
            }
        tm.sync();
        // each list into ascending edge order: an entry's place is the number of smaller entries
        // of its list, so a hub vertex's list is placed by as many lanes as it has entries
        if (sorted) {
            for (uint32_t j = tm.r; j < nslot; j += N) {
                const uint32_t x = iown[j];
                const uint32_t b = x ? iend[x - 1] : 0u, n = iend[x] - b;
                const uint16_t key = ie2[j];
                uint32_t r = 0;
#pragma unroll 4
                for (uint32_t i = 0; i < n; ++i) r += ie2[b + i] < key ? 1u : 0u;
                ie[b + r] = key;
            }
            tm.sync();
        }
        const uint16_t* il = sorted ? ie : ie2;

        const DRun& RN = a.runs[run];
        const DRule* rules;
        if constexpr (SMALL) {
            // the whole rules when forwarding (anchored orders), else the prefix the full join
            // reads (DRule puts those fields first)
            const uint32_t pw = a.fwd_n ? (uint32_t)(sizeof(DRule) / 4) : (uint32_t)(offsetof(DRule, lv) / 4);
            const uint32_t words = RN.n_rules * pw;
            for (uint32_t i = tm.r; i < words; i += N) {
                const uint32_t r = i / pw, k = i - r * pw;
                ((uint32_t*)&lr[r])[k] = ((const uint32_t*)&RN.rules[r])[k];
            }
            tm.sync();
            rules = lr;
        } else {
            rules = RN.rules;
        }
        uint32_t total = 0;
        uint64_t edges_sum = 0;                                // count pass: edges of the children
        bool stop = false;
        uint32_t out = EMIT ? (uint32_t)a.moff[s] : 0u;
        bool full = true;
        if (s < a.fwd_n) {
            // ---- incremental forwarding (parallel_evolution.cpp:591-756): the parent's
            //      matches that avoid the consumed edges, plus delta joins at the produced
            //      edges, merged into the full join's order ----
            const DRule& CR = rules[a.pm_rule[s]];
            const uint32_t nc = CR.n_lhs, nprod = CR.n_rhs;
            const uint32_t kept = ne - nprod;                   // kept edges come first
            const uint16_t* cq = a.pm_q + (size_t)s * a.pm_K;  // consumed, parent positions
            const uint32_t par = a.pm_parent[s];
            const uint64_t lb = a.pm_moff[par], le = a.pm_moff[par + 1];
            uint32_t ncombo = 0;                                // (rule, anchor, produced edge)
            for (uint32_t r = 0; r < RN.n_rules; ++r) ncombo += rules[r].n_lhs * nprod;
            auto combo = [&](uint32_t i, uint32_t& r, uint32_t& pp, uint32_t& j) {
                r = 0;
                while (i >= rules[r].n_lhs * nprod) { i -= rules[r].n_lhs * nprod; ++r; }
                pp = i / nprod;
                j = i % nprod;
            };
            uint32_t nd = 0;
            uint64_t de = 0;                                    // child edges of the delta matches
            for (uint32_t base = 0; base < ncombo; base += N) {
                const uint32_t i = base + tm.r;
                uint32_t c = 0;
                if (i < ncombo) {
                    uint32_t r, pp, j;
                    combo(i, r, pp, j);
                    c = join_anchor(rules[r], pp, sv, sa, ne, A, iend, il, kept, kept + j, L, nullptr, 0, 0);
                    de += (uint64_t)c * (ne - rules[r].n_lhs + rules[r].n_rhs);
                }
                nd += tm.sum(c);
            }
            if (!EMIT || (a.dmax && nd <= a.dmax)) {
                full = false;
                if (EMIT) {
                    uint32_t o = 0;
                    for (uint32_t base = 0; base < ncombo; base += N) {
                        const uint32_t i = base + tm.r;
                        uint32_t c = 0, r = 0, pp = 0, j = 0;
                        if (i < ncombo) {
                            combo(i, r, pp, j);
                            c = join_anchor(rules[r], pp, sv, sa, ne, A, iend, il, kept, kept + j, L, nullptr, 0, 0);
                        }
                        uint32_t round;
                        const uint32_t inc = tm.incl_scan(c, round);
                        if (c)
                            join_anchor(rules[r], pp, sv, sa, ne, A, iend, il, kept, kept + j, L,
                                        dlist + (size_t)(o + inc - c) * dstride, dstride, r);
                        o += round;
                    }
                    tm.sync();
                    for (uint32_t i = tm.r; i < nd; i += N) {        // rank = place in order
                        const uint16_t* di = dlist + (size_t)i * dstride;
                        uint32_t rank = 0;
                        for (uint32_t j = 0; j < nd; ++j) {
                            const uint16_t* dj = dlist + (size_t)j * dstride;
                            rank += match_less(rules, dj[0], dj + 1, di[0], di + 1) ? 1u : 0u;
                        }
                        dperm[rank] = (uint16_t)i;
                    }
                    for (uint32_t i = tm.r; i <= nd; i += N) dhist[i] = 0;
                    tm.sync();
                }
                uint32_t nf = 0;
                for (uint64_t base = lb; base < le; base += N) {
                    const uint64_t m = base + tm.r;
                    bool fw = false;
                    uint32_t rm = 0, nm = 0;
                    uint16_t qm[16];
                    if (m < le) {
                        rm = a.pm_rule[m];
                        nm = rules[rm].n_lhs;
                        fw = true;
                        for (uint32_t t = 0; t < nm; ++t) {
                            const uint32_t x = a.pm_q[m * a.pm_K + t];
                            uint32_t below = 0;
                            for (uint32_t u = 0; u < nc; ++u) {
                                const uint32_t y = cq[u];
                                fw &= (x != y);
                                below += (y < x) ? 1u : 0u;
                            }
                            qm[t] = (uint16_t)(x - below);   // position among the kept edges
                        }
                    }
                    uint32_t nround;
                    const uint32_t at = tm.compact(fw, nround);
                    if (!EMIT && fw) de += ne - nm + rules[rm].n_rhs;
                    if (EMIT && fw) {
                        uint32_t lo = 0, hi = nd;                   // delta matches before it
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            const uint16_t* dm = dlist + (size_t)dperm[mid] * dstride;
                            if (match_less(rules, dm[0], dm + 1, rm, qm)) lo = mid + 1; else hi = mid;
                        }
                        const uint32_t o2 = out + nf + at + lo;
                        for (uint32_t t = 0; t < nm; ++t) a.M.q[(size_t)o2 * a.M.K + t] = qm[t];
                        a.M.parent[o2] = s;
                        a.M.rule[o2] = rm;
                        a.M.cne[o2] = ne - nm + rules[rm].n_rhs;
                        atomicAdd(&dhist[lo], 1u);
                    }
                    nf += nround;
                }
                if (EMIT) {
                    tm.sync();
                    tm.excl_scan_lds(dhist, nd + 1);          // dhist[k + 1] = forwarded before delta k
                    for (uint32_t k = tm.r; k < nd; k += N) {
                        const uint16_t* dm = dlist + (size_t)dperm[k] * dstride;
                        const uint32_t rr = dm[0], nm = rules[rr].n_lhs;
                        const uint32_t o2 = out + dhist[k + 1] + k;
                        for (uint32_t t = 0; t < nm; ++t) a.M.q[(size_t)o2 * a.M.K + t] = dm[1 + t];
                        a.M.parent[o2] = s;
                        a.M.rule[o2] = rr;
                        a.M.cne[o2] = ne - nm + rules[rr].n_rhs;
                    }
                    tm.sync();
                } else {
                    total = nf + nd;
                    edges_sum = tm.sum64(de);
                    if (tm.leader() && total) atomicAdd(&a.run_acc[run], (unsigned long long)total);
                }
            }
        }
        for (uint32_t r = 0; full && r < RN.n_rules && !stop; ++r) {
            const DRule& R = rules[r];
            const uint32_t p0 = R.order[0];
            const uint32_t child_ne = ne - R.n_lhs + R.n_rhs;
            // SCAN: compact the seeds of position order[0] in edge order.
            uint32_t nseed = 0;
            for (uint32_t base = 0; base < ne; base += N) {
                const uint32_t e = base + tm.r;
                bool ok = false;
                if (e < ne && sa[e] == R.la[p0]) {
                    if constexpr (SMALL) {
                        // repeated pattern variables: each position equals its first binding's
                        ok = true;
                        for (uint32_t i = 0; i < R.la[p0]; ++i) {
                            const uint32_t src = R.jsrc[0][i];
                            if (src != 0xFFu) ok &= sv[e * A + i] == sv[e * A + (src & 15u)];
                        }
                    } else {
                        uint32_t m = 0;
                        ok = bind_edge(sv + e * A, sa[e], R.lv[p0], R.la[p0], L.bind, m);
                    }
                }
                uint32_t nround;
                const uint32_t at = tm.compact(ok, nround);
                if (ok) seeds[nseed + at] = (uint16_t)e;
                nseed += nround;
            }
            tm.sync();
            // EXPAND + SINK, 64 seeds per round.
            uint16_t* sc = (SMALL && a.seedcnt && s >= a.fwd_n) ? a.seedcnt + 2 * eo + (size_t)r * ne : nullptr;
            for (uint32_t base = 0; base < nseed; base += N) {
                const uint32_t i = base + tm.r;
                uint32_t c = 0;
                if (i < nseed) {
                    if constexpr (SMALL) {
                        const uint32_t kept = (EMIT && sc) ? sc[i] : 0xFFFFu;
                        c = kept != 0xFFFFu ? kept
                                            : reg_join_any<false>(R, sv, sa, ne, A, iend, il, seeds[i], RegJoinOut{});
                        if (!EMIT && sc) sc[i] = (uint16_t)(c < 0xFFFFu ? c : 0xFFFFu);
                    } else {
                        c = join_seed(R, sv, sa, ne, A, iend, il, seeds[i], L, nullptr, 0, 0, nullptr, nullptr,
                                      nullptr, 0, 0, 0);
                    }
                }
                if (EMIT) {
                    uint32_t round;
                    const uint32_t inc = tm.incl_scan(c, round);
                    if (i < nseed && c) {
                        if constexpr (SMALL) {
                            const RegJoinOut ro{a.M.q, a.M.K, out + inc - c, a.M.parent, a.M.rule, a.M.cne,
                                                s, r, child_ne};
                            reg_join_any<true>(R, sv, sa, ne, A, iend, il, seeds[i], ro);
                        } else {
                            join_seed(R, sv, sa, ne, A, iend, il, seeds[i], L, a.M.q, a.M.K, out + inc - c,
                                      a.M.parent, a.M.rule, a.M.cne, s, r, child_ne);
                        }
                    }
                    out += round;
                } else {
                    // The run's shared tally: once it passes what the run may still take, the
                    // generation is over the cap whatever the other states hold, so stop. Every
                    // state's count equals what it added, so the totals stay over the cap.
                    const uint32_t round = tm.sum(c);
                    total += round;
                    edges_sum += (uint64_t)round * child_ne;
                    if (base + N >= nseed && r + 1 >= RN.n_rules) {
                        // the state's last round: nothing is left to stop, so add without waiting
                        if (tm.leader() && round) atomicAdd(&a.run_acc[run], (unsigned long long)round);
                    } else {
                        unsigned long long before = 0;
                        if (tm.leader() && round) before = atomicAdd(&a.run_acc[run], (unsigned long long)round);
                        before = tm.bcast(before);
                        if (before + round > budget) { stop = true; break; }
                    }
                }
            }
            tm.sync();
        }
        if (!EMIT && tm.leader()) {
            a.cnt[s] = total;
            a.state_edges[s] = edges_sum;
        }
    }
}


// ------------------------------------------------------------------------------------------
// Rewrite + hash + dedup + event + causal, one wave per child (Rewriter::apply,
// hypergraph/src/rewriter.cpp:11-205, with hgcommon rewrite semantics rewrite_core.hpp:53-87).
// ------------------------------------------------------------------------------------------

// A slot is read before it is claimed: most children repeat a hash already in the map, and a
// plain load that finds it (or another key: keys are only ever written once, 0 -> key) settles
// the probe without an atomic; only an empty-looking slot is claimed by compare-and-swap, whose
// result is authoritative if the load was stale.
__device__ inline bool map_insert(DRun* run, uint64_t h) {
    if (h == 0) return atomicExch(&run->zero_seen, 1u) == 0u;
    const uint32_t mask = run->map_mask;
    uint32_t i = (uint32_t)mix64(h) & mask;
    while (true) {
        // device scope is enough: only this device reads and claims the map
        uint64_t old = __hip_atomic_load(&run->map_keys[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 0)
            old = atomicCAS((unsigned long long*)&run->map_keys[i], 0ull, (unsigned long long)h);
        if (old == 0) return true;
        if (old == h) return false;
        i = (i + 1) & mask;
    }
}

// GWS selects the workspace at compile time (global scratch for states too large for LDS), so
// the LDS variant addresses its workspace with LDS instructions rather than flat ones. W is the
// team: W = 1, one wave per child (up to four per workgroup, each with its own workspace);
// W > 1, the workgroup's W waves share one child and one workspace, for large children whose
// workspace would otherwise leave a CU one or two waves. One-wave teams are held to four waves
// per SIMD (128 registers, a few spilled): on the batched pass the expansion took 9.8 ms so,
// 15.3 ms at three (no spill) and 11.1 ms at five (36 registers spilled).
template <bool GWS, uint32_t W, uint32_t G>
__global__ __launch_bounds__(W == 1 ? 256 : 64 * W, W == 1 ? 4 : 1) void k_expand(ExpandArgs a) {
    using TM = typename TeamSel<W, G>::type;
    extern __shared__ uint8_t lds[];
    const uint32_t tpb = W == 1 ? (blockDim.x / WAVE) * (WAVE / G) : 1;   // teams per workgroup
    const uint32_t w = W == 1 ? threadIdx.x / G : 0;                       // team within the workgroup
    uint8_t* my;
    if constexpr (GWS) my = a.scratch + ((size_t)blockIdx.x * tpb + w) * a.lds_per_wave;
    else my = lds + (size_t)w * a.lds_per_wave;
    const TM tm((uint64_t*)(my + a.lds_per_wave - 64));   // a wave team's reduction words: the last 64 bytes
    const uint32_t l = tm.lane();
    const uint32_t PA = a.P.A, A = a.C.A;                  // parent / child vertex strides
    const WLSizes z = expand_sizes(a.max_child_ne, a.max_nv, a.max_nextv, A, a.lean != 0);
    WLSpace s;
    uint8_t* p = my;
    wl_bind(s, p, z, A);
    // Child vertex labels and the label -> local id map live in the hash's srt and nbr regions,
    // which the hash only writes after they are read.
    uint32_t* gv = (uint32_t*)s.srt;                      // [NE * A]
    uint32_t* vmap = (uint32_t*)s.nbr;                    // [max_nextv + 1]

    // Children in chunks of kChunk consecutive ones, dealt round-robin to the teams (2 measured
    // best of 1, 2, 4 and contiguous per-team blocks on the batched workload; 1 and 2 measure
    // alike since the run counters left this kernel): a chunk keeps neighbouring children (one
    // parent) on one team, and the dealing spreads large and small requests' children over all
    // teams.
    constexpr uint32_t kChunk = 2;
    const uint32_t gw = xcd_block() * tpb + w, nw = gridDim.x * tpb;
    uint32_t rr = 0;                                       // range of logical index c (subsets)


    for (uint32_t lc = gw * kChunk; lc < a.n_items; lc += (lc % kChunk == kChunk - 1) ? (nw - 1) * kChunk + 1 : 1) {
        uint32_t c = lc;
        if (a.nrange) {
            if (lc % kChunk == 0 || lc == gw * kChunk) {       // a new chunk: find its range
                uint32_t lo = 0, hi = a.nrange;
                while (hi - lo > 1) { const uint32_t mid = (lo + hi) >> 1; if (a.rpre[mid] <= lc) lo = mid; else hi = mid; }
                rr = lo;
            }
            while (lc >= a.rpre[rr + 1]) ++rr;
            c = a.rb[rr] + (lc - a.rpre[rr]);
        }
        const uint32_t par = a.M.parent[c];
        const uint32_t rid = a.M.rule[c];
        const uint32_t run = a.P.run[par];
        DRun* RN = &a.runs[run];
        const DRule& R = RN->rules[rid];
        const uint32_t nl = R.n_lhs, nr = R.n_rhs;
        const uint32_t pne = a.P.ne[par];
        const uint64_t peo = a.P.eoff[par];
        const uint64_t ceo = a.C.eoff[c];
        const uint32_t cne = pne - nl + nr;
        const uint32_t pdep = a.P.depth[par];
        const uint32_t step = pdep + 1;                       // generation of the child
        const uint16_t* q = a.M.q + (size_t)c * a.M.K;
        // Causal inputs, fetched now so their latency hides under the rewrite and the hash:
        // lane t < nl holds the producer generation of consumed edge t (0: none) and that
        // producer's causal-ancestor mask from the parent's lineage table.
        uint32_t ps_t = 0;
        uint64_t anc_t = 0;
        if (l < nl) {
            ps_t = a.P.ps[peo + q[l]];
            if (ps_t) anc_t = a.P.lanc[(size_t)par * a.P.lstride + ps_t - 1];
        }

        if (a.E.cps && l < nl) a.E.cps[(size_t)c * a.E.cK + l] = (uint8_t)ps_t;

        // Parent edges minus the consumed ones, order kept (SparseBitset::derive).
        uint32_t kept = 0;
        for (uint32_t base = 0; base < pne; base += TM::N) {
            const uint32_t i = base + tm.r;
            bool keep = i < pne;
            if (keep) for (uint32_t t = 0; t < nl; ++t) keep &= (q[t] != i);
            uint32_t n_round;
            const uint32_t at = tm.compact(keep, n_round);
            if (keep) {
                const uint32_t pos = kept + at;
                const uint8_t ar = a.P.ar[peo + i];
                for (uint32_t k = 0; k < A; ++k) {
                    const uint32_t v = k < PA ? a.P.ev[(peo + i) * PA + k] : 0u;
                    a.C.ev[(ceo + pos) * A + k] = v;
                    gv[pos * A + k] = v;
                }
                a.C.ar[ceo + pos] = ar;
                a.C.ps[ceo + pos] = a.P.ps[peo + i];
                s.e_ar[pos] = ar;
            }
            kept += n_round;
        }
        // Produced edges in RHS order; fresh vertices ascending by variable
        // (assign_fresh_consecutive / resolve_rhs_vertices, rewrite_core.hpp:53-87).
        const uint32_t fresh_base = a.P.nextv[par];
        if (tm.r < nr) {
            const uint32_t pos = kept + tm.r;
            const uint32_t ra = R.ra[tm.r];
            for (uint32_t k = 0; k < A; ++k) {
                uint32_t v = 0;
                if (k < ra) {
                    const uint32_t var = R.rv[tm.r][k];
                    if (R.lhs_mask & (1u << var)) {
                        const uint32_t se = q[R.var_src_edge[var]];
                        v = a.P.ev[(peo + se) * PA + R.var_src_pos[var]];
                    } else {
                        v = fresh_base + __popc(R.new_mask & ((1u << var) - 1u));
                    }
                }
                a.C.ev[(ceo + pos) * A + k] = v;
                gv[pos * A + k] = v;
            }
            a.C.ar[ceo + pos] = (uint8_t)ra;
            a.C.ps[ceo + pos] = (uint8_t)step;
            s.e_ar[pos] = (uint8_t)ra;
        }
        tm.sync();
        // Local vertex numbering: ascending label, as the reference's sorted vertex list.
        const uint32_t nextv = fresh_base + R.n_new;
        for (uint32_t v = tm.r; v <= nextv; v += TM::N) vmap[v] = 0;
        tm.sync();
        for (uint32_t e = tm.r; e < cne; e += TM::N)
            for (uint32_t k = 0; k < s.e_ar[e]; ++k) vmap[gv[e * A + k]] = 1;
        tm.sync();
        const uint32_t nv = tm.excl_scan_lds(vmap, nextv + 1);
        for (uint32_t e = tm.r; e < cne; e += TM::N)
            for (uint32_t k = 0; k < A; ++k) s.e_lv[e * A + k] = (uint16_t)(k < s.e_ar[e] ? vmap[gv[e * A + k]] : 0);
        s.nE = cne;
        s.nV = nv;
        tm.sync();
        const uint64_t h = wl_hash_team(s, tm);

        // Causal: one triple per consumed edge with a producer; the producers are events of
        // this lineage, named by their generation (rewriter.cpp:121-191,
        // causal_graph.cpp:187-234). Transitive reduction: (P, C) is redundant exactly when P
        // is a causal ancestor of another producer of C. Every lane computes the same masks
        // from the nl lanes holding the consumed edges' producers.
        uint64_t prod = 0, anc_c = 0, others = 0;
        for (uint32_t u = 0; u < nl; ++u) {
            const uint32_t pu = tm.shfl(ps_t, u);
            const uint64_t au = ((uint64_t)tm.shfl((uint32_t)(anc_t >> 32), u) << 32) |
                                tm.shfl((uint32_t)anc_t, u);
            if (pu) {
                prod |= 1ull << pu;
                anc_c |= (1ull << pu) | au;
                if (pu != ps_t) others |= au;
            }
        }
        const bool rec_causal = (RN->flags & 0x1u) != 0;
        const bool drop = (RN->flags & 0x100u) && l < nl && ps_t && ((others >> ps_t) & 1ull);
        const uint64_t dropped = tm.ballot(drop);
        uint64_t tr = prod;
        for (uint32_t u = 0; u < nl; ++u) {
            const uint32_t pu = tm.shfl(ps_t, u);
            if ((dropped >> u) & 1ull) tr &= ~(1ull << pu);
        }
        const uint32_t triples = (uint32_t)__popcll(tm.ballot(l < nl && ps_t != 0));
        const uint32_t triples_tr = (uint32_t)__popcll(tm.ballot(l < nl && ps_t != 0 && ((tr >> ps_t) & 1ull)));
        if (!rec_causal) { prod = 0; tr = 0; anc_c = 0; }

        // Identity and bookkeeping (the team's first lane): the child's record and its event.
        const uint32_t ev_id = RN->ev_base + (c - RN->child_begin);
        const uint32_t child_sid = RN->n_roots + ev_id;
        if (tm.leader()) {
            a.C.run[c] = run;
            a.C.ne[c] = cne;
            a.C.nextv[c] = nextv;
            a.C.sid[c] = child_sid;
            a.C.hash[c] = h;
            a.C.depth[c] = (uint8_t)step;
            a.E.in_state[c] = a.P.sid[par];
            a.E.rule[c] = rid;
            a.E.prod[c] = prod;
            a.E.tr[c] = tr;
            // the run's counters are summed from these by k_run_events, not added here
            a.tcnt[c] = rec_causal ? (uint16_t)(triples | (triples_tr << 8)) : (uint16_t)0;
        }
        // Lineage table of the child: the parent's, then this event's ancestor mask.
        {
            const uint64_t* plan = a.P.lanc + (size_t)par * a.P.lstride;
            uint64_t* clan = a.C.lanc + (size_t)c * a.C.lstride;
            for (uint32_t t = tm.r; t < step; t += TM::N) clan[t] = (t + 1 == step) ? anc_c : plan[t];
        }
        tm.sync();
    }
}

// A generation's event counters per run (items of at most a few thousand of a run's children
// [cb, ce), one wave each): events, causal triples and pairs with and without transitive
// reduction, from the per-child records k_expand wrote -- one add per counter and item instead
// of adds from every team.
__global__ void k_run_events(const RunEvents* items, uint32_t n, const uint64_t* prod, const uint64_t* tr,
                             const uint16_t* tcnt, uint64_t* ctr) {
    const uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, l = lane_id();
    if (i >= n) return;
    const RunEvents it = items[i];
    uint64_t triples = 0, pairs = 0, pairs_tr = 0, triples_tr = 0;
    for (uint32_t c = it.cb + l; c < it.ce; c += WAVE) {
        const uint32_t t = tcnt[c];
        triples += t & 0xFFu;
        triples_tr += t >> 8;
        pairs += (uint64_t)__popcll(prod[c]);
        pairs_tr += (uint64_t)__popcll(tr[c]);
    }
    triples = wave_sum_u64(triples);
    triples_tr = wave_sum_u64(triples_tr);
    pairs = wave_sum_u64(pairs);
    pairs_tr = wave_sum_u64(pairs_tr);
    if (l == 0) {                                    // a large run is split over several items
        unsigned long long* row = (unsigned long long*)(ctr + (size_t)it.slot * CTR_N);
        atomicAdd(&row[CTR_EVENTS], (unsigned long long)(it.ce - it.cb));
        if (triples) atomicAdd(&row[CTR_TRIPLES], (unsigned long long)triples);
        if (pairs) atomicAdd(&row[CTR_PAIRS], (unsigned long long)pairs);
        if (pairs_tr) atomicAdd(&row[CTR_PAIRS_TR], (unsigned long long)pairs_tr);
        if (triples_tr) atomicAdd(&row[CTR_TRIPLES_TR], (unsigned long long)triples_tr);
    }
}

__global__ void k_fill(uint8_t* p, uint32_t word, uint64_t words, uint64_t bytes) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w = i; w < words; w += stride) ((uint32_t*)p)[w] = word;
    if (i < bytes - words * 4) p[words * 4 + i] = (uint8_t)word;          // the tail bytes
}

hipError_t launch_fill(void* p, uint8_t value, uint64_t bytes, hipStream_t st) {
    if (!bytes) return hipSuccess;
    const uint32_t word = 0x01010101u * value;
    // whole 32-bit words from a 4-byte aligned start; the rest bytewise
    uint8_t* b = (uint8_t*)p;
    const uint64_t lead = (4 - ((uintptr_t)b & 3)) & 3;
    if (lead) {
        const uint64_t k = lead < bytes ? lead : bytes;
        hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, st, b, word, (uint64_t)0, k);
        if (k == bytes) return hipGetLastError();
        b += k;
        bytes -= k;
    }
    const uint64_t words = bytes / 4;
    uint64_t blocks = (std::max<uint64_t>(words, 4) + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill, dim3((uint32_t)blocks), dim3(256), 0, st, b, word, words, bytes);
    return hipGetLastError();
}

// Calibration read: every 16-byte word of [p, p + words * 16) read once, one sum written, so the
// launch's HBM traffic is known (the bytes read, and 8 written).
__global__ void k_read_calibration(const uint4* p, uint64_t words, unsigned long long* sink) {
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 v = p[i];
        acc += (uint64_t)v.x + v.y + v.z + v.w;
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0) atomicAdd(sink, (unsigned long long)acc);
}

// Reads 16 bytes at the start of each of `lines` distinct 128-byte lines of the buffer: with n
// lines in the buffer (a power of two) line (i * P) mod n, P odd, is a bijection, so the lines
// read are distinct and spread over the whole buffer.
__global__ void k_read_gather(const uint8_t* p, uint32_t lines, uint32_t mask, unsigned long long* sink) {
    uint64_t acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < lines; i += gridDim.x * blockDim.x) {
        const uint32_t l = (uint32_t)(((uint64_t)i * 0x9E3779B1u) & mask);
        const uint4 v = *(const uint4*)(p + (uint64_t)l * 128);
        acc += (uint64_t)v.x + v.y + v.z + v.w;
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0) atomicAdd(sink, (unsigned long long)acc);
}

// One device-scope compare-and-swap on the first word of each of `lines` distinct lines (the
// buffer holds no zero word, so every one fails and nothing changes), then one device-scope add
// per thread into 4,096 words of 64-byte stride: the atomics' traffic, for the counters.
__global__ void k_atomic_calibration(uint8_t* p, uint32_t lines, uint32_t mask, unsigned long long* ctr,
                                     unsigned long long* sink) {
    uint64_t acc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < lines; i += gridDim.x * blockDim.x) {
        const uint32_t l = (uint32_t)(((uint64_t)i * 0x9E3779B1u) & mask);
        acc += atomicCAS((unsigned long long*)(p + (uint64_t)l * 128), 0ull, 1ull);
        atomicAdd(&ctr[(i & 4095u) * 8u], 1ull);
    }
    acc = wave_sum_u64(acc);
    if (lane_id() == 0) atomicAdd(sink, (unsigned long long)acc);
}

hipError_t launch_read_gather(void* p, uint64_t bytes, uint64_t* sink, uint64_t* ctr, hipStream_t st) {
    uint64_t n = 1;                                   // 128-byte lines in the buffer, a power of two
    while (n * 2 * 128 <= bytes) n *= 2;
    if (n < 4) return hipSuccess;
    hipLaunchKernelGGL(k_read_gather, dim3(2048), dim3(256), 0, st, (const uint8_t*)p, (uint32_t)(n / 4),
                       (uint32_t)(n - 1), (unsigned long long*)sink);
    if (hipGetLastError() != hipSuccess) return hipErrorLaunchFailure;
    hipLaunchKernelGGL(k_atomic_calibration, dim3(2048), dim3(256), 0, st, (uint8_t*)p, (uint32_t)(n / 4),
                       (uint32_t)(n - 1), (unsigned long long*)ctr, (unsigned long long*)sink);
    return hipGetLastError();
}

// HGX_OPTION_COUNTER_FLUSH: one streaming read of the flush buffer, whose lines displace every
// line the L2 holds (the sum goes nowhere the host reads; it is kept live by a store into the
// buffer's first word only when it is impossible, i.e. never).
__global__ void k_flush(const uint4* p, uint64_t words) {
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint4 v = p[i];
        acc += (uint64_t)v.x + v.y + v.z + v.w;
    }
    if (acc == 0xFFFFFFFFFFFFFFFFull) ((uint4*)p)[0] = make_uint4(1, 1, 1, 1);
}

hipError_t launch_flush(const void* buf, uint64_t bytes, hipStream_t st) {
    hipLaunchKernelGGL(k_flush, dim3(2048), dim3(256), 0, st, (const uint4*)buf, bytes / 16);
    return hipGetLastError();
}

hipError_t launch_read_calibration(const void* p, uint64_t bytes, uint64_t* sink, hipStream_t st) {
    const uint64_t words = bytes / 16;
    hipLaunchKernelGGL(k_read_calibration, dim3(2048), dim3(256), 0, st, (const uint4*)p, words,
                       (unsigned long long*)sink);
    return hipGetLastError();
}

// A run over its budget stops being counted once the count pass is sure of it, so its total is
// any value above the budget: the decision is what must agree, and below the budget the total.
__global__ void k_plan_check(const uint64_t* tot, const uint64_t* want, const uint32_t* budget, uint32_t n,
                             uint32_t* flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t b = budget[i];
    const bool over = tot[i] > b, want_over = want[i] > b;
    if (over != want_over || (!over && tot[i] != want[i])) atomicOr(flag, 1u);
}

hipError_t launch_plan_check(const uint64_t* tot, const uint64_t* want, const uint32_t* budget, uint32_t n,
                             uint32_t* flag, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_plan_check, dim3((n + 255) / 256), dim3(256), 0, st, tot, want, budget, n, flag);
    return hipGetLastError();
}

hipError_t launch_run_events(const RunEvents* items, uint32_t n, const EventView& E, const uint16_t* tcnt,
                             uint64_t* ctr, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_run_events, dim3((n + 3) / 4), dim3(256), 0, st, items, n, E.prod, E.tr, tcnt, ctr);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Branchial pairs of parents above BR_MAX_EDGES edges. Every (consumed edge slot, match) of such
// a parent becomes one 64-bit key (slot << 32 | match); sorted, the keys of one slot are its consuming
// matches in ascending order. Match i meets each partner j > i through the first of its own
// edges that j also consumes: from its key for that edge, the keys after it with the same slot.
// The work is the entries plus the sum over edges of (matches consuming it)^2 -- not the
// generation's edge count, nor (matches per parent)^2. One thread per match or entry.
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ bool brl_large(const BrListArgs& a, uint32_t m, uint32_t& par) {
    par = a.M.parent[m];
    return a.P.ne[par] > BR_MAX_EDGES && a.cnt[par] >= 2;
}

// Keys and payloads of match m at entries [m*K, m*K + K): (slot << 32 | m, m*K + k) for its
// consumed edges, the sentinel (past every slot) for the rest.
__global__ void k_brl_keys(BrListArgs a) {
    const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= a.M.n) return;
    const uint32_t K = a.M.K;
    uint32_t par, nl = 0;
    if (brl_large(a, m, par)) nl = a.runs[a.P.run[par]].rules[a.M.rule[m]].n_lhs;
    const uint16_t* q = a.M.q + (size_t)m * K;
    for (uint32_t k = 0; k < K; ++k) {
        const uint64_t x = (uint64_t)m * K + k;
        a.key_in[x] = k < nl ? ((a.P.eoff[par] + q[k]) << 32) | m : a.sentinel;
        a.val_in[x] = (uint32_t)x;
    }
}

// Where each (match, k) entry landed in the sorted keys.
__global__ void k_brl_pos(BrListArgs a, uint64_t n) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n || a.key[x] == a.sentinel) return;
    a.pos[a.val[x]] = (uint32_t)x;
}

template <bool EMIT>
__global__ void k_brl_pairs(BrListArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t par;
    if (i >= a.M.n || !brl_large(a, i, par)) return;
    const uint32_t run = a.P.run[par];
    const DRun& RN = a.runs[run];
    const uint32_t K = a.M.K;
    const uint32_t nl = RN.rules[a.M.rule[i]].n_lhs;
    const uint16_t* qi = a.M.q + (size_t)i * K;
    uint32_t c = 0;
    uint64_t o = 0;
    if (EMIT) {
        const uint32_t mb = (uint32_t)a.moff[par];
        o = a.poff[par] + (a.bo[i] - a.bo[mb]);
    }
    for (uint32_t k = 0; k < nl; ++k) {
        const uint64_t x0 = a.pos[(uint64_t)i * K + k];
        const uint64_t slot = a.key[x0] >> 32;
        for (uint64_t x = x0 + 1; x < a.n_keys && (a.key[x] >> 32) == slot; ++x) {
            const uint32_t j = (uint32_t)a.key[x];
            // count j here only if this is the first of i's edges that j consumes
            const uint16_t* qj = a.M.q + (size_t)j * K;
            const uint32_t nj = RN.rules[a.M.rule[j]].n_lhs;
            bool earlier = false;
            for (uint32_t k2 = 0; k2 < k && !earlier; ++k2)
                for (uint32_t y = 0; y < nj; ++y) earlier |= (qj[y] == qi[k2]);
            if (earlier) continue;
            if (EMIT) {
                a.pairs[2 * o] = RN.ev_base + (i - RN.child_begin);
                a.pairs[2 * o + 1] = RN.ev_base + (j - RN.child_begin);
                ++o;
            }
            ++c;
        }
    }
    if (!EMIT) {
        a.bc[i] = c;
        if (c) {
            atomicAdd((unsigned long long*)&a.ctr[(size_t)run * CTR_N + CTR_BRANCHIAL], (unsigned long long)c);
            if (a.pcnt) atomicAdd(&a.pcnt[par], c);
        }
    }
}

hipError_t launch_brl(const BrListArgs& a, int phase, hipStream_t st) {
    const uint32_t n = a.M.n;
    if (!n) return hipSuccess;
    const dim3 g((n + 255) / 256), b(256);
    switch (phase) {
        case 0: hipLaunchKernelGGL(k_brl_keys, g, b, 0, st, a); break;
        case 1: hipLaunchKernelGGL(k_brl_pos, dim3((uint32_t)((a.n_keys + 255) / 256)), b, 0, st, a, a.n_keys); break;
        case 2: hipLaunchKernelGGL(k_brl_pairs<false>, g, b, 0, st, a); break;
        default: hipLaunchKernelGGL(k_brl_pairs<true>, g, b, 0, st, a); break;
    }
    return hipGetLastError();
}

// State identity (Hypergraph::create_or_get_canonical_state, hypergraph/src/hypergraph.cpp:154-254):
// every child's hash into its request's state map, one lane per child, after k_expand -- so no
// wave of k_expand waits on a global compare-and-swap chain. A key inserted first counts a
// canonical state and adds to the fingerprint.
//
// Device-scope atomics are performed beyond the L2 and each moves a whole request over the
// fabric (measured on gfx950: 64 bytes per compare-and-swap, taken or not, 32 per add), and the
// children of one generation that hash alike all find the slot empty before any of them claims
// it. So a lane whose (request, hash) equals a lower lane's in the wave leaves the insert to that
// lane, and when the wave's lanes belong to one request (the usual case: a request's children
// are contiguous) the canonical count and fingerprint are summed over the wave and added once.
__global__ void k_dedup(GenView C, uint32_t n, DRun* runs, uint64_t* ctr) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = c < n;
    const uint32_t run = valid ? C.run[c] : 0xFFFFFFFFu;
    const uint64_t h = valid ? C.hash[c] : 0;
    const uint32_t lane = lane_id();
    bool dup = false;
    for (int k = 0; k < WAVE; ++k) {
        const uint64_t hk = __shfl(h, k, WAVE);
        const uint32_t rk = __shfl(run, k, WAVE);
        dup |= k < (int)lane && hk == h && rk == run;
    }
    const bool fresh = valid && !dup && map_insert(&runs[run], h);
    const uint32_t r0 = __shfl(run, 0, WAVE);
    if (__ballot(valid && run != r0) == 0) {
        const uint64_t nfresh = (uint64_t)__popcll(__ballot(fresh));
        const uint64_t fp = wave_sum_u64(fresh ? mix64(h) : 0ull);
        if (lane == 0 && nfresh) {
            atomicAdd((unsigned long long*)&ctr[(size_t)r0 * CTR_N + CTR_CANON], (unsigned long long)nfresh);
            atomicAdd((unsigned long long*)&ctr[(size_t)r0 * CTR_N + CTR_FINGERPRINT], (unsigned long long)fp);
        }
    } else if (fresh) {
        atomicAdd((unsigned long long*)&ctr[(size_t)run * CTR_N + CTR_CANON], 1ull);
        atomicAdd((unsigned long long*)&ctr[(size_t)run * CTR_N + CTR_FINGERPRINT], (unsigned long long)mix64(h));
    }
}

hipError_t launch_dedup(const GenView& C, uint32_t n, DRun* runs, uint64_t* ctr, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_dedup, dim3((n + 255) / 256), dim3(256), 0, st, C, n, runs, ctr);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Branchial pairs: two events of the SAME input state whose consumed edge sets intersect,
// one pair each (causal_graph.cpp:324-351). One wave per parent.
//
// Parents of at most 128 edges: the parent's matches are taken in chunks; for each chunk an
// LDS bitset per edge position marks the chunk's matches that consume it. Match i then meets
// every later match j of the chunk through the OR of the bitsets of its own edges: one popcount
// per 64 candidates instead of one test per pair. A consumed set is NW 64-bit words (one up to
// 64 edges, two up to 128); the bitsets and the cached consumed sets take the same 16 KiB of
// LDS either way, so a two-word parent works in half-size chunks. Larger parents take the
// per-edge match lists above (k_brl_*).
// ------------------------------------------------------------------------------------------

constexpr uint32_t BR_LDS_WORDS = 1024;             // per table: 8 KiB of bitsets, 8 KiB of masks

uint32_t branchial_lds_per_wave() { return 2 * BR_LDS_WORDS * 8 + 64; }

template <uint32_t NW>
__device__ __forceinline__ void consumed_mask(const BranchialArgs& a, const DRun& RN, uint32_t m,
                                              uint64_t (&msk)[NW]) {
    const uint32_t nl = RN.rules[a.M.rule[m]].n_lhs;
    const uint16_t* q = a.M.q + (size_t)m * a.M.K;
#pragma unroll
    for (uint32_t w = 0; w < NW; ++w) msk[w] = 0;
    for (uint32_t t = 0; t < nl; ++t) {
        const uint32_t e = q[t];
#pragma unroll
        for (uint32_t w = 0; w < NW; ++w)
            if ((e >> 6) == w) msk[w] |= 1ull << (e & 63);
    }
}

// One parent (cnt >= 2 matches from m0): its pairs counted (emit 0) or written (emit 1).
template <uint32_t NW>
__device__ __forceinline__ uint32_t branchial_parent(const BranchialArgs& a, const DRun& RN, uint32_t m0,
                                                     uint32_t cnt, uint32_t ne, uint32_t ev0, uint64_t out,
                                                     uint8_t* my, uint32_t l) {
    constexpr uint32_t WORDS = 16 / NW;              // 64-bit words per edge bitset
    constexpr uint32_t CHUNK = WORDS * 64;           // matches per chunk
    constexpr uint32_t MASKS = BR_LDS_WORDS / NW;    // consumed sets cached in LDS
    uint64_t* bits = (uint64_t*)my;                  // [64 * NW edges][WORDS]
    uint64_t* mcache = bits + BR_LDS_WORDS;          // [MASKS][NW]
    uint32_t total = 0;
    const uint32_t ncache = cnt < MASKS ? cnt : MASKS;
    for (uint32_t i = l; i < ncache; i += WAVE) {
        uint64_t mk[NW];
        consumed_mask<NW>(a, RN, m0 + i, mk);
#pragma unroll
        for (uint32_t w = 0; w < NW; ++w) mcache[i * NW + w] = mk[w];
    }
    wave_sync();
    auto mask_of = [&](uint32_t i, uint64_t (&mk)[NW]) {
        if (i < MASKS) {
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w) mk[w] = mcache[i * NW + w];
        } else {
            consumed_mask<NW>(a, RN, m0 + i, mk);
        }
    };
    for (uint32_t cb = 0; cb < cnt; cb += CHUNK) {
        const uint32_t ce = min(cnt, cb + CHUNK);
        const uint32_t words = (ce - cb + 63) >> 6;
        // only the words this chunk reads: the parent's edges x the chunk's 64-match words
        for (uint32_t x = l; x < ne * words; x += WAVE) {
            const uint32_t e = x / words;
            bits[e * WORDS + (x - e * words)] = 0;
        }
        wave_sync();
        for (uint32_t j = cb + l; j < ce; j += WAVE) {
            const uint32_t r = j - cb;
            uint64_t mj[NW];
            mask_of(j, mj);
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w)
                for (uint64_t mk = mj[w]; mk; mk &= mk - 1)
                    atomicOr((unsigned long long*)&bits[(w * 64 + __builtin_ctzll(mk)) * WORDS + (r >> 6)],
                             1ull << (r & 63));
        }
        wave_sync();
        for (uint32_t base = 0; base < ce; base += WAVE) {
            const uint32_t i = base + l;
            uint32_t c = 0;
            uint64_t mi[NW];
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w) mi[w] = 0;
            auto partners = [&](uint32_t wd) {         // later matches of word wd sharing an edge with i
                const uint32_t jb = cb + wd * 64;
                uint64_t word = 0;
#pragma unroll
                for (uint32_t w = 0; w < NW; ++w)
                    for (uint64_t mk = mi[w]; mk; mk &= mk - 1)
                        word |= bits[(w * 64 + __builtin_ctzll(mk)) * WORDS + wd];
                if (jb <= i) word &= ~((2ull << (i - jb)) - 1ull);
                return word;
            };
            if (i < ce) {
                mask_of(i, mi);
                for (uint32_t wd = 0; wd < words; ++wd) {
                    if (cb + wd * 64 + 63 <= i) continue;
                    c += __popcll(partners(wd));
                }
            }
            if (a.emit) {
                const uint32_t inc = wave_incl_scan_u32(c);
                uint64_t o = out + inc - c;
                if (c) {
                    for (uint32_t wd = 0; wd < words; ++wd) {
                        const uint32_t jb = cb + wd * 64;
                        if (jb + 63 <= i) continue;
                        for (uint64_t word = partners(wd); word; word &= word - 1) {
                            a.pairs[2 * o] = ev0 + i;
                            a.pairs[2 * o + 1] = ev0 + jb + __builtin_ctzll(word);
                            ++o;
                        }
                    }
                }
                out += __shfl(inc, WAVE - 1, WAVE);
            } else {
                total += c;
            }
        }
        wave_sync();
    }
    return total;
}

__global__ __launch_bounds__(256) void k_branchial(BranchialArgs a) {
    extern __shared__ uint8_t lds[];
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    uint8_t* my = lds + (size_t)w * a.lds_per_wave;
    const uint32_t waves = gridDim.x * (blockDim.x / WAVE);
    for (uint32_t s = blockIdx.x * (blockDim.x / WAVE) + w; s < a.P.n; s += waves) {
        const uint32_t cnt = a.cnt[s];
        const uint32_t ne = a.P.ne[s];
        if (cnt < 2 || ne > BR_MAX_EDGES) { if (a.emit == 0 && a.pcnt && l == 0) a.pcnt[s] = 0; continue; }
        const uint32_t run = a.P.run[s];
        const uint32_t m0 = (uint32_t)a.moff[s];
        const DRun& RN = a.runs[run];
        const uint32_t ev0 = RN.ev_base + (m0 - RN.child_begin);
        const uint64_t out = a.emit ? a.poff[s] : 0;
        const uint32_t total = ne <= 64 ? branchial_parent<1>(a, RN, m0, cnt, ne, ev0, out, my, l)
                                        : branchial_parent<2>(a, RN, m0, cnt, ne, ev0, out, my, l);
        if (!a.emit) {
            const uint32_t tot = wave_sum_u32(total);
            if (l == 0) {
                if (tot) atomicAdd((unsigned long long*)&a.ctr[(size_t)run * CTR_N + CTR_BRANCHIAL],
                                   (unsigned long long)tot);
                if (a.pcnt) a.pcnt[s] = tot;
            }
        }
        wave_sync();
    }
}

// ------------------------------------------------------------------------------------------
// WL hash of explicitly given states (dense local vertex ids per state). Optionally inserts
// each into its run's state map (roots).
// ------------------------------------------------------------------------------------------

template <bool GWS, uint32_t W>
__global__ __launch_bounds__(W == 1 ? 256 : 64 * W) void k_wl_states(WLStatesArgs a) {
    extern __shared__ uint8_t lds[];
    const uint32_t tpb = W == 1 ? blockDim.x / WAVE : 1;    // teams per workgroup
    const uint32_t w = W == 1 ? threadIdx.x / WAVE : 0;     // team within the workgroup
    uint8_t* p;
    if constexpr (GWS) p = a.scratch + ((size_t)blockIdx.x * tpb + w) * a.lds_per_wave;
    else p = lds + (size_t)w * a.lds_per_wave;
    const Team<W> tm((uint64_t*)(p + a.lds_per_wave - 64));   // the workspace's last 64 bytes
    const uint32_t A = a.A;
    const WLSizes z = WLSizes::make(a.max_ne, a.max_nv, A, a.lean != 0);
    WLSpace s;
    wl_bind(s, p, z, A);
    const uint32_t teams = gridDim.x * tpb;
    for (uint32_t i = blockIdx.x * tpb + w; i < a.n; i += teams) {
        const uint64_t e0 = a.state_off[i];
        const uint32_t ne = (uint32_t)(a.state_off[i + 1] - e0);
        for (uint32_t e = tm.r; e < ne; e += Team<W>::N) {
            const uint8_t ar = a.ar[e0 + e];
            s.e_ar[e] = ar;
            for (uint32_t k = 0; k < A; ++k) s.e_lv[e * A + k] = (uint16_t)(k < ar ? a.ev[(e0 + e) * A + k] : 0u);
        }
        s.nE = ne;
        s.nV = a.nv[i];
        tm.sync();
        const uint64_t h = wl_hash_team(s, tm);
        if (tm.leader()) {
            a.out[i] = h;
            if (a.run_of) {
                DRun* RN = &a.runs[a.run_of[i]];
                if (map_insert(RN, h)) {
                    atomicAdd((unsigned long long*)&a.ctr[(size_t)a.run_of[i] * CTR_N + CTR_CANON], 1ull);
                    atomicAdd((unsigned long long*)&a.ctr[(size_t)a.run_of[i] * CTR_N + CTR_FINGERPRINT],
                              (unsigned long long)mix64(h));
                }
            }
        }
        tm.sync();
    }
}

// Admissions of one step: one block per request.
__global__ void k_admit(const AdmitItem* items, uint64_t* ctr) {
    const AdmitItem it = items[blockIdx.x];
    const uint32_t* src = (const uint32_t*)it.rule_src;
    uint32_t* dst = (uint32_t*)it.rule_dst;
    const uint32_t words = it.n_rules * (uint32_t)(sizeof(DRule) / 4);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
    for (uint64_t i = threadIdx.x; i < it.map_cap; i += blockDim.x) it.map_keys[i] = 0;
    if (threadIdx.x < CTR_N) ctr[(size_t)it.slot * CTR_N + threadIdx.x] = 0;
}

hipError_t launch_admit(const AdmitItem* items, uint32_t n, uint64_t* ctr, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_admit, dim3(n), dim3(256), 0, st, items, ctr);
    return hipGetLastError();
}

// One block per prepared request: its roots' records and edges into the generation.
__global__ void k_copy_roots(const RootCopy* items, GenView v) {
    const RootCopy it = items[blockIdx.x];
    const uint64_t e0 = it.eoff[0], e1 = it.eoff[it.n_roots];
    for (uint32_t k = threadIdx.x; k <= it.n_roots; k += blockDim.x) {
        v.eoff[it.dst_state + k] = it.dst_edge + (it.eoff[k] - e0);
        if (k < it.n_roots) {
            v.run[it.dst_state + k] = it.slot;
            v.ne[it.dst_state + k] = it.ne[k];
            v.nextv[it.dst_state + k] = it.nextv[k];
            v.sid[it.dst_state + k] = k;
            v.depth[it.dst_state + k] = 0;
        }
    }
    for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const uint64_t d = it.dst_edge + (e - e0);
        v.ar[d] = it.ar[e];
        v.ps[d] = 0;
        for (uint32_t x = 0; x < v.A; ++x) v.ev[d * v.A + x] = x < it.src_A ? it.ev[e * it.src_A + x] : 0u;
    }
}

hipError_t launch_copy_roots(const RootCopy* items, uint32_t n, const GenView& v, hipStream_t st) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_copy_roots, dim3(n), dim3(256), 0, st, items, v);
    return hipGetLastError();
}

// Per-run match totals of a generation: runs' states are contiguous ranges.
// One wave per run: its matches from the offsets; its children's edges and the edges of its
// states that have matches (the parents its expansion reads), summed over its states -- a
// per-state store and these sums instead of every state adding to shared counters.
__global__ void k_run_totals(const uint64_t* moff, const uint64_t* state_edges, const uint32_t* cnt,
                             const uint32_t* ne, const uint32_t* run_sbeg, const uint32_t* run_send,
                             uint32_t nruns, uint64_t* tot, uint64_t* edges, uint64_t* parent_edges) {
    const uint32_t r = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE, l = lane_id();
    if (r >= nruns) return;
    const uint32_t b = run_sbeg[r], e = run_send[r];
    uint64_t sum = 0, pe = 0;
    for (uint32_t s = b + l; s < e; s += WAVE) {
        sum += state_edges[s];
        pe += cnt[s] ? ne[s] : 0u;
    }
    sum = wave_sum_u64(sum);
    pe = wave_sum_u64(pe);
    if (l == 0) {
        tot[r] = moff[e] - moff[b];
        edges[r] = sum;
        parent_edges[r] = pe;
    }
}

// Zero the counts of states whose run does not advance this generation.
__global__ void k_mask_counts(const uint32_t* run, const uint8_t* active, uint32_t n, uint32_t* cnt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && !active[run[s]]) cnt[s] = 0;
}

// Move one run's state map into a larger table.
__global__ void k_rehash(const uint64_t* old_keys, uint32_t old_cap, uint64_t* keys, uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= old_cap) return;
    const uint64_t h = old_keys[i];
    if (h == 0) return;
    uint32_t j = (uint32_t)mix64(h) & mask;
    while (true) {
        const uint64_t old = atomicCAS((unsigned long long*)&keys[j], 0ull, (unsigned long long)h);
        if (old == 0) return;
        j = (j + 1) & mask;
    }
}

// ------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------

static uint32_t grid_for(uint32_t items, uint32_t per_block, uint32_t cap);
uint32_t launch_grid(uint32_t items, uint32_t wpb, uint32_t grid_cap) { return grid_for(items, wpb, grid_cap); }

static uint32_t grid_for(uint32_t items, uint32_t per_block, uint32_t cap) {
    uint32_t g = (items + per_block - 1) / per_block;
    if (g > cap) g = cap;
    if (g < 8) g = 8;
    return (g + 7u) & ~7u;
}

template <uint32_t G>
static void launch_match_lanes(const MatchArgs& a, bool emit, uint32_t grid, uint32_t wpb, hipStream_t st) {
    const uint32_t threads = 64 * wpb;
    const size_t lds = (size_t)a.lds_per_wave * wpb * (64 / G);
    if (a.small) {
        if (emit) hipLaunchKernelGGL((k_match<true, JoinLaneLds, 1, false, G>), dim3(grid), dim3(threads), lds, st, a);
        else hipLaunchKernelGGL((k_match<false, JoinLaneLds, 1, false, G>), dim3(grid), dim3(threads), lds, st, a);
    } else {
        if (emit) hipLaunchKernelGGL((k_match<true, JoinLaneBig, 1, false, G>), dim3(grid), dim3(threads), lds, st, a);
        else hipLaunchKernelGGL((k_match<false, JoinLaneBig, 1, false, G>), dim3(grid), dim3(threads), lds, st, a);
    }
}

template <uint32_t W>
static void launch_match_team(const MatchArgs& a, bool emit, uint32_t grid, uint32_t wpb, hipStream_t st) {
    const uint32_t threads = W == 1 ? 64 * wpb : 64 * W;
    if constexpr (W == 8) {
        if (a.scratch) {
            if (a.small) {
                if (emit) hipLaunchKernelGGL((k_match<true, JoinLaneLds, 8, true>), dim3(grid), dim3(threads), 0, st, a);
                else hipLaunchKernelGGL((k_match<false, JoinLaneLds, 8, true>), dim3(grid), dim3(threads), 0, st, a);
            } else {
                if (emit) hipLaunchKernelGGL((k_match<true, JoinLaneBig, 8, true>), dim3(grid), dim3(threads), 0, st, a);
                else hipLaunchKernelGGL((k_match<false, JoinLaneBig, 8, true>), dim3(grid), dim3(threads), 0, st, a);
            }
            return;
        }
    }
    const size_t lds = (size_t)a.lds_per_wave * (W == 1 ? wpb : 1);
    if (a.small) {
        if (emit) hipLaunchKernelGGL((k_match<true, JoinLaneLds, W>), dim3(grid), dim3(threads), lds, st, a);
        else hipLaunchKernelGGL((k_match<false, JoinLaneLds, W>), dim3(grid), dim3(threads), lds, st, a);
    } else {
        if (emit) hipLaunchKernelGGL((k_match<true, JoinLaneBig, W>), dim3(grid), dim3(threads), lds, st, a);
        else hipLaunchKernelGGL((k_match<false, JoinLaneBig, W>), dim3(grid), dim3(threads), lds, st, a);
    }
}

hipError_t launch_match(const MatchArgs& a, bool emit, uint32_t grid_cap, uint32_t wpb, uint32_t team,
                        hipStream_t st, uint32_t lanes) {
    const uint32_t items = a.nrange ? a.n_items : a.P.n;
    if (!items) return hipSuccess;
    if (a.scratch && team != 8) return hipErrorInvalidValue;     // global workspaces: 8-wave teams
    if (lanes != 64) {                                           // sub-wave teams: one wave per team set
        if (team != 1 || a.scratch || (lanes != 16 && lanes != 32)) return hipErrorInvalidValue;
        const uint32_t grid = grid_for(items, wpb * (64 / lanes), grid_cap);
        if (lanes == 16) launch_match_lanes<16>(a, emit, grid, wpb, st);
        else launch_match_lanes<32>(a, emit, grid, wpb, st);
        return hipGetLastError();
    }
    const uint32_t grid = grid_for(items, team == 1 ? wpb : 1, grid_cap);
    switch (team) {
        case 1: launch_match_team<1>(a, emit, grid, wpb, st); break;
        case 4: launch_match_team<4>(a, emit, grid, wpb, st); break;
        case 8: launch_match_team<8>(a, emit, grid, wpb, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

uint32_t expand_lds_per_wave(uint32_t max_child_ne, uint32_t max_nv, uint32_t max_nextv, uint32_t A, bool lean) {
    const WLSizes z = expand_sizes(max_child_ne, max_nv, max_nextv, A, lean);
    // + 64: carving slack; + 64: the team's reduction words at the end
    return ((z.bytes(A) + 128u) + 15u) & ~15u;
}

template <uint32_t W, uint32_t G>
static void launch_expand_team(const ExpandArgs& a, uint32_t grid, uint32_t wpb, hipStream_t st) {
    const uint32_t threads = W == 1 ? 64 * wpb : 64 * W;
    const size_t lds = a.scratch ? 0 : (size_t)a.lds_per_wave * (W == 1 ? wpb * (64 / G) : 1);
    if constexpr (G == 64) {
        if (a.scratch) { hipLaunchKernelGGL((k_expand<true, W, 64>), dim3(grid), dim3(threads), lds, st, a); return; }
    }
    hipLaunchKernelGGL((k_expand<false, W, G>), dim3(grid), dim3(threads), lds, st, a);
}

uint32_t expand_grid(uint32_t n_items, uint32_t team, uint32_t wpb, uint32_t grid_cap, uint32_t lanes) {
    return grid_for(n_items, team == 1 ? wpb * (64 / lanes) : 1, grid_cap);
}

hipError_t launch_expand(const ExpandArgs& a, uint32_t grid_cap, uint32_t wpb, uint32_t team, hipStream_t st,
                         uint32_t lanes) {
    if (!a.n_items) return hipSuccess;
    if (lanes != 64 && (team != 1 || a.scratch || (lanes != 16 && lanes != 32))) return hipErrorInvalidValue;
    const uint32_t grid = expand_grid(a.n_items, team, wpb, grid_cap, lanes);
    switch (team * 100 + lanes) {
        case 164: launch_expand_team<1, 64>(a, grid, wpb, st); break;
        case 132: launch_expand_team<1, 32>(a, grid, wpb, st); break;
        case 116: launch_expand_team<1, 16>(a, grid, wpb, st); break;
        case 264: launch_expand_team<2, 64>(a, grid, wpb, st); break;
        case 464: launch_expand_team<4, 64>(a, grid, wpb, st); break;
        case 864: launch_expand_team<8, 64>(a, grid, wpb, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_branchial(const BranchialArgs& a, uint32_t grid_cap, uint32_t wpb, hipStream_t st) {
    const uint32_t grid = grid_for(a.P.n, wpb, grid_cap);
    const size_t lds = (size_t)a.lds_per_wave * wpb;
    hipLaunchKernelGGL(k_branchial, dim3(grid), dim3(64 * wpb), lds, st, a);
    return hipGetLastError();
}

uint32_t wl_states_lds_per_wave(uint32_t max_ne, uint32_t max_nv, uint32_t A, bool lean) {
    const WLSizes z = WLSizes::make(max_ne, max_nv, A, lean);
    // + 64: carving slack; + 64: the team's reduction words at the end
    return (z.bytes(A) + 128u + 15u) & ~15u;
}

template <uint32_t W>
static void launch_wl_states_team(const WLStatesArgs& a, uint32_t grid, uint32_t wpb, hipStream_t st) {
    const uint32_t threads = W == 1 ? 64 * wpb : 64 * W;
    const size_t lds = a.scratch ? 0 : (size_t)a.lds_per_wave * (W == 1 ? wpb : 1);
    if (a.scratch) hipLaunchKernelGGL((k_wl_states<true, W>), dim3(grid), dim3(threads), lds, st, a);
    else hipLaunchKernelGGL((k_wl_states<false, W>), dim3(grid), dim3(threads), lds, st, a);
}

uint32_t wl_states_grid(uint32_t n, uint32_t team, uint32_t wpb, uint32_t grid_cap) {
    return grid_for(n, team == 1 ? wpb : 1, grid_cap);
}

hipError_t launch_wl_states(const WLStatesArgs& a, uint32_t grid_cap, uint32_t wpb, uint32_t team, hipStream_t st) {
    if (!a.n) return hipSuccess;
    const uint32_t grid = wl_states_grid(a.n, team, wpb, grid_cap);
    switch (team) {
        case 1: launch_wl_states_team<1>(a, grid, wpb, st); break;
        case 2: launch_wl_states_team<2>(a, grid, wpb, st); break;
        case 4: launch_wl_states_team<4>(a, grid, wpb, st); break;
        case 8: launch_wl_states_team<8>(a, grid, wpb, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_run_totals(const uint64_t* moff, const uint64_t* state_edges, const uint32_t* cnt,
                             const uint32_t* ne, const uint32_t* b, const uint32_t* e, uint32_t nruns, uint64_t* tot,
                             uint64_t* edges, uint64_t* parent_edges, hipStream_t st) {
    if (nruns == 0) return hipSuccess;
    hipLaunchKernelGGL(k_run_totals, dim3((nruns + 3) / 4), dim3(256), 0, st, moff, state_edges, cnt, ne, b, e, nruns,
                       tot, edges, parent_edges);
    return hipGetLastError();
}

hipError_t launch_mask_counts(const uint32_t* run, const uint8_t* active, uint32_t n, uint32_t* cnt,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_mask_counts, dim3((n + 255) / 256), dim3(256), 0, st, run, active, n, cnt);
    return hipGetLastError();
}

hipError_t launch_rehash(const uint64_t* ok, uint32_t old_cap, uint64_t* keys, uint32_t mask, hipStream_t st) {
    if (!old_cap) return hipSuccess;
    hipLaunchKernelGGL(k_rehash, dim3((old_cap + 255) / 256), dim3(256), 0, st, ok, old_cap, keys, mask);
    return hipGetLastError();
}

}  // namespace hgx
