// hgx_wl.hpp -- the WL canonical hash of one state, computed by one wave64 over LDS.
//
// Bit-identical to hgcommon::wl_canonical_hash (common/include/hgcommon/wl_core.hpp:67-202)
// called the way WLHash::compute_state_hash_impl calls it
// (hypergraph/include/hypergraph/wl_hash.hpp:109-239). The reference runs the algorithm on
// one thread; here the 64 lanes share it:
//   occurrence CSR       counting sort with LDS atomics, then every vertex bucket is put in
//                        (edge, position) order so the CSR is the reference's exactly;
//   initial colours      one lane per vertex (wl_core.hpp:93-107), nbr_cap truncation kept;
//   distinct count       LDS open-addressing set with 64-bit CAS, exact;
//   refinement           one lane per vertex, neighbour multiset sorted in an LDS segment
//                        and folded in order (wl_core.hpp:119-141);
//   components           min-label propagation with pointer jumping; the per-component fold
//                        only needs the partition, which is what union-find gives;
//   folds                wave-wide sums (commutative, wrap-around).
// The result does not depend on local vertex numbering (every fold is order-independent and
// the truncated initial-colour prefix follows edge order), so the caller may use any dense
// numbering; the kernels use ascending vertex label, the reference's own choice.
#pragma once

#include "hgx_device.hpp"

namespace hgx {

// LDS workspace of one WL computation. Edges use a fixed vertex stride A. Arrays whose lives do
// not overlap share memory: the neighbour values being sorted, the distinct-count set and the
// per-position counts of the uniform-arity front end (nbr = hset); the sorted values and the
// second colour buffer of the register refinement (srt = nxt). A lean workspace is for states
// of one arity 2..4 only (the front end without the occurrence CSR): it has no occ_ep, oown or
// nnv, and its components reuse occ_off.
struct WLSpace {
    uint32_t nE, nV, A;
    uint8_t* e_ar;       // [NE]
    uint16_t* e_lv;      // [NE * A] local vertex index
    uint32_t* occ_off;   // [NV + 1]
    uint32_t* occ_ep;    // [OCC] (edge << 8) | pos, in (edge, pos) order per vertex (general only)
    uint32_t* cursor;    // [NV + 1] scratch (write cursors, then segment offsets)
    uint32_t* comp;      // [NV]
    uint64_t* cur;       // [NV]
    uint64_t* nxt;       // [NV]  (= srt)
    uint64_t* nbr;       // [NN] per-vertex neighbour segments (values being sorted)
    uint32_t* nb;        // [NN] per-vertex neighbour lists: (local id << 4) | position
    uint64_t* hset;      // [HT]  (= nbr)
    uint32_t* nnv;       // [NV] neighbour-list length per vertex (general only)
    uint32_t* oown;      // [OCC] vertex owning each occurrence slot (general only)
    uint32_t* own;       // [NN] vertex owning each neighbour-segment slot
    uint64_t* srt;       // [NN] segments in sorted order
    uint32_t ht_mask;
};

// Sizes a caller must provision for states with at most ne edges, nv vertices, arity <= A.
// r1_min / r2_min: 64-bit words the caller wants in the nbr / srt regions before the hash
// starts (the rewrite keeps its labels and vertex map there).
struct WLSizes {
    uint32_t NE, NV, OCC, NN, HT, R1, R2;
    bool lean;
    __host__ __device__ static WLSizes make(uint32_t ne, uint32_t nv, uint32_t A, bool lean = false,
                                            uint32_t r1_min = 0, uint32_t r2_min = 0) {
        WLSizes s;
        s.lean = lean;
        s.NE = ne ? ne : 1;
        s.OCC = s.NE * A;
        s.NV = nv ? nv : 1;
        if (s.NV > s.OCC) s.NV = s.OCC;
        // per-vertex segments: sum over vertices of max(nn, min(deg, cap)) <= OCC * max(A-1, 1)
        s.NN = s.OCC * (A > 2 ? A - 1 : 1);
        uint32_t ht = 64;
        while (ht < s.NV + s.NV / 4 + 1) ht <<= 1;     // load factor <= 0.8
        s.HT = ht;
        s.R1 = s.NN > ht ? s.NN : ht;                  // nbr, hset, per-position counts (NN >= NV)
        if (s.R1 < r1_min) s.R1 = r1_min;
        s.R2 = s.NN > r2_min ? s.NN : r2_min;          // srt, nxt
        return s;
    }
    // bytes, each array 8-byte aligned
    __host__ __device__ uint32_t bytes(uint32_t A) const {
        auto al = [](uint32_t b) { return (b + 7u) & ~7u; };
        uint32_t b = al(NE) + al(NE * A * 2) + al((NV + 1) * 4) + al((NV + 1) * 4) + 8 * NV +
                     8 * R1 + al(NN * 4) + al(NN * 4) + 8 * R2;
        if (!lean) b += al(OCC * 4) + al(NV * 4) + al(NV * 4) + al(OCC * 4);
        return b;
    }
};

__host__ __device__ __forceinline__ uint8_t* carve(uint8_t*& p, uint32_t bytes) {
    uint8_t* r = p;
    p += (bytes + 7u) & ~7u;
    return r;
}

__host__ __device__ __forceinline__ void wl_bind(WLSpace& s, uint8_t*& p, const WLSizes& z, uint32_t A) {
    s.A = A;
    s.e_ar = carve(p, z.NE);
    s.e_lv = (uint16_t*)carve(p, z.NE * A * 2);
    s.occ_off = (uint32_t*)carve(p, (z.NV + 1) * 4);
    s.cursor = (uint32_t*)carve(p, (z.NV + 1) * 4);
    s.cur = (uint64_t*)carve(p, z.NV * 8);
    s.nbr = (uint64_t*)carve(p, z.R1 * 8);
    s.hset = s.nbr;
    s.nb = (uint32_t*)carve(p, z.NN * 4);
    s.own = (uint32_t*)carve(p, z.NN * 4);
    s.srt = (uint64_t*)carve(p, z.R2 * 8);
    s.nxt = s.srt;
    if (z.lean) {
        s.occ_ep = nullptr;
        s.oown = nullptr;
        s.nnv = nullptr;
        s.comp = s.occ_off;
    } else {
        s.occ_ep = (uint32_t*)carve(p, z.OCC * 4);
        s.comp = (uint32_t*)carve(p, z.NV * 4);
        s.nnv = (uint32_t*)carve(p, z.NV * 4);
        s.oown = (uint32_t*)carve(p, z.OCC * 4);
    }
    s.ht_mask = z.HT - 1;
}

// Batcher's odd-even merge network for 8 keys (19 comparators), in registers. Unused slots
// hold ~0 and sort last, so the first n entries are the n keys in ascending order -- the order
// isort_u64 (core.hpp:54-61) produces.
#define HGX_CE(a, b) { const uint64_t lo_ = x[a] < x[b] ? x[a] : x[b]; \
                       const uint64_t hi_ = x[a] < x[b] ? x[b] : x[a]; x[a] = lo_; x[b] = hi_; }
__device__ inline void sort8(uint64_t (&x)[8]) {
    HGX_CE(0, 1) HGX_CE(2, 3) HGX_CE(4, 5) HGX_CE(6, 7)
    HGX_CE(0, 2) HGX_CE(1, 3) HGX_CE(4, 6) HGX_CE(5, 7)
    HGX_CE(1, 2) HGX_CE(5, 6) HGX_CE(0, 4) HGX_CE(3, 7)
    HGX_CE(1, 5) HGX_CE(2, 6)
    HGX_CE(1, 4) HGX_CE(3, 6)
    HGX_CE(2, 4) HGX_CE(3, 5)
    HGX_CE(3, 4)
}
#undef HGX_CE

// Every segment [so[v], so[v] + cnt[v]) of vals, sorted ascending into the same places of out:
// an entry's place is the number of smaller entries of its segment (equal ones by position), so
// the lanes share a long segment instead of one lane sorting it (the order isort_u64,
// core.hpp:54-61, produces). own[j] names the segment slot j belongs to; total = slots.
template <class TM, class T>
__device__ __forceinline__ void seg_rank_sort(const TM& tm, const T* vals, T* out, const uint32_t* own,
                                              const uint32_t* so, const uint32_t* cnt, uint32_t total) {
    for (uint32_t j = tm.r; j < total; j += TM::N) {
        const uint32_t v = own[j];
        const uint32_t b = so[v], n = cnt ? cnt[v] : so[v + 1] - b;     // null cnt: so is a CSR
        if (j - b >= n) continue;
        const T x = vals[j];
        uint32_t r = 0;
#pragma unroll 4
        for (uint32_t i = 0; i < n; ++i) {
            const T y = vals[b + i];
            r += (y < x || (y == x && b + i < j)) ? 1u : 0u;
        }
        out[b + r] = x;
    }
    tm.sync();
}

// Exact number of distinct values in cur[0..nV).
template <class TM>
__device__ __forceinline__ uint32_t wl_ndistinct(const TM& tm, WLSpace& s, const uint64_t* cur) {
    // An LDS hash set, whatever the team: on the batched pass it beat one colour per lane
    // compared with every lower lane (k_expand 9.8 -> 8.8 ms), whose lane reads are two
    // cross-lane permutes per colour in a sub-wave team.
    for (uint32_t i = tm.r; i <= s.ht_mask; i += TM::N) s.hset[i] = 0;
    tm.sync();
    uint32_t cnt = 0, zero = 0;
    for (uint32_t v = tm.r; v < s.nV; v += TM::N) {
        const uint64_t x = cur[v];
        if (x == 0) { zero = 1; continue; }
        uint32_t slot = (uint32_t)((x * 0x9E3779B97F4A7C15ULL) >> 40) & s.ht_mask;
        while (true) {
            const uint64_t old = atomicCAS((unsigned long long*)&s.hset[slot], 0ull,
                                           (unsigned long long)x);
            if (old == 0) { ++cnt; break; }
            if (old == x) break;
            slot = (slot + 1) & s.ht_mask;
        }
    }
    const uint32_t total = tm.sum(cnt);
    return total + (tm.any(zero != 0) ? 1u : 0u);
}

// The hash. Inputs in s: nE, nV, A, e_ar[nE], e_lv[nE*A]. Every member of the team must call it.
template <class TM>
__device__ inline uint64_t wl_hash_team(WLSpace& s, const TM& tm) {
    const uint32_t l = tm.r;
    const uint32_t nE = s.nE, nV = s.nV, A = s.A;
    if (nE == 0) return EMPTY_STATE_HASH;     // hypergraph.cpp:743-745
    if (nV == 0) return 0;                    // wl_core.hpp:85

    // Edges of one arity 2..4 (the common case): a vertex's sorted
    // initial keys (arity, position) are its occurrence counts per position in position order,
    // and nothing truncates (nbr_cap = max (arity-1)*deg >= deg), so the occurrence CSR is not
    // needed: count per position, size each neighbour segment (arity-1)*deg, and let every
    // occurrence write its neighbours straight into its vertex's segment (the segment is sorted
    // before it is folded, so the order of writes does not matter).
    uint32_t amn = 0xFFu, amx = 0;
    for (uint32_t e = l; e < nE; e += TM::N) {
        const uint32_t a = s.e_ar[e];
        amn = a < amn ? a : amn;
        amx = a > amx ? a : amx;
    }
    amx = tm.max(amx);
    amn = 0xFFu - tm.max(0xFFu - amn);
    const bool fast = amn == amx && amx >= 2 && amx <= 4 && nE < 65536;
    uint32_t nbr_cap = 1, nslots = 0;
    if (fast) {
        const uint32_t ua = amx;
        uint32_t* deg = s.occ_off;                     // [NV]: occurrences, then write cursors
        // [NV]: occurrences per position, 16 bits each (nE < 2^16), in the nbr region (free
        // until the refinement)
        uint64_t* pcnt = s.nbr;
        for (uint32_t v = l; v < nV; v += TM::N) { deg[v] = 0; pcnt[v] = 0; }
        tm.sync();
        for (uint32_t e = l; e < nE; e += TM::N)
            for (uint32_t p = 0; p < ua; ++p) {
                const uint32_t x = s.e_lv[e * A + p];
                atomicAdd(&deg[x], 1u);
                atomicAdd((unsigned long long*)&pcnt[x], 1ull << (16 * p));
            }
        tm.sync();
        uint32_t nn_max = 0;
        for (uint32_t v = l; v < nV; v += TM::N) {
            const uint32_t nn = deg[v] * (ua - 1);
            s.cursor[v] = nn;
            nn_max = nn > nn_max ? nn : nn_max;
        }
        nbr_cap = tm.max(nn_max);
        if (l == 0) s.cursor[nV] = 0;
        tm.sync();
        tm.excl_scan_lds(s.cursor, nV + 1);          // cursor: CSR of the segments
        nslots = s.cursor[nV];
        for (uint32_t v = l; v < nV; v += TM::N) {
            const uint32_t b = s.cursor[v], d = deg[v];
            for (uint32_t k = b; k < s.cursor[v + 1]; ++k) s.own[k] = v;
            // initial colour (wl_core.hpp:93-107): fnv(deg), then (arity, position) per occurrence
            // in ascending key order -- position order, each position as often as it occurs
            uint64_t h = fnv(FNV_OFFSET, d);
            const uint64_t pc = pcnt[v];
            for (uint32_t p = 0; p < ua; ++p)
                for (uint32_t i = 0; i < (uint32_t)((pc >> (16 * p)) & 0xFFFFu); ++i) { h = fnv(h, ua); h = fnv(h, p); }
            s.cur[v] = h;
            deg[v] = b;                                // write cursor
        }
        tm.sync();
        for (uint32_t e = l; e < nE; e += TM::N)
            for (uint32_t p = 0; p < ua; ++p) {
                uint32_t o = atomicAdd(&deg[s.e_lv[e * A + p]], ua - 1);
                for (uint32_t k = 0; k < ua; ++k)
                    if (k != p) s.nb[o++] = ((uint32_t)s.e_lv[e * A + k] << 4) | k;
            }
        tm.sync();
    } else {
        // ---- occurrence CSR (wl_core.hpp:87-97), with each vertex's neighbour-list length ----
        for (uint32_t v = l; v <= nV; v += TM::N) {
            s.occ_off[v] = 0;
            if (v < nV) s.nnv[v] = 0;
        }
        tm.sync();
        uint32_t amin = 0xFFu, amax = 0;
        for (uint32_t e = l; e < nE; e += TM::N) {
            const uint32_t a = s.e_ar[e];
            amin = a < amin ? a : amin;
            amax = a > amax ? a : amax;
            for (uint32_t p = 0; p < a; ++p) {
                const uint32_t x = s.e_lv[e * A + p];
                atomicAdd(&s.occ_off[x], 1u);
                if (a > 1) atomicAdd(&s.nnv[x], a - 1);   // wl_hash.hpp:201-210
            }
        }
        const bool uniform = tm.max(amax) == (0xFFu - tm.max(0xFFu - amin));
        const uint32_t ua = tm.max(amax);
        tm.sync();
        const uint32_t nocc = tm.excl_scan_lds(s.occ_off, nV + 1);
        for (uint32_t v = l; v <= nV; v += TM::N) s.cursor[v] = s.occ_off[v];
        tm.sync();
        for (uint32_t e = l; e < nE; e += TM::N)
            for (uint32_t p = 0; p < s.e_ar[e]; ++p) {
                const uint32_t w = atomicAdd(&s.cursor[s.e_lv[e * A + p]], 1u);
                s.occ_ep[w] = (e << 8) | p;
            }
        for (uint32_t v = l; v < nV; v += TM::N)
            for (uint32_t i = s.occ_off[v]; i < s.occ_off[v + 1]; ++i) s.oown[i] = v;
        tm.sync();
        // Each bucket into (edge, pos) order -- the order the reference's counting sort leaves.
        // The order only matters where the initial colour truncates a vertex's occurrences
        // (deg > nbr_cap), which needs unary edges: with every arity >= 2, nn >= deg for each
        // vertex, and every later use of a bucket sorts what it reads.
        const bool unary = (0xFFu - tm.max(0xFFu - amin)) <= 1;
        if (unary) {
            seg_rank_sort(tm, s.occ_ep, s.nb, s.oown, s.occ_off, (const uint32_t*)nullptr, nocc);
            for (uint32_t i = l; i < nocc; i += TM::N) s.occ_ep[i] = s.nb[i];
        }
        uint32_t nn_max = 0;
        for (uint32_t v = l; v < nV; v += TM::N) nn_max = s.nnv[v] > nn_max ? s.nnv[v] : nn_max;
        nbr_cap = tm.max(nn_max);
        if (nbr_cap < 1) nbr_cap = 1;
        tm.sync();
        // Segment sizes: room for the refinement list (nn) and the initial list (min(deg, cap)).
        uint32_t cnt_max = 0;
        for (uint32_t v = l; v < nV; v += TM::N) {
            const uint32_t deg = s.occ_off[v + 1] - s.occ_off[v];
            const uint32_t init = deg < nbr_cap ? deg : nbr_cap;
            const uint32_t nn = s.nnv[v];
            s.cursor[v] = nn > init ? nn : init;
            cnt_max = init > cnt_max ? init : cnt_max;
        }
        cnt_max = tm.max(cnt_max);
        tm.sync();
        tm.excl_scan_lds(s.cursor, nV + 1);                       // cursor[v] = segment offset
        nslots = s.cursor[nV];                                      // (cursor[nV] held no size)
        for (uint32_t v = l; v < nV; v += TM::N)
            for (uint32_t k = s.cursor[v]; k < s.cursor[v + 1]; ++k) s.own[k] = v;
        tm.sync();

        // ---- initial colours (wl_core.hpp:93-107): fnv(deg), then the first min(deg, cap)
        //      occurrences' (arity, position) keys in ascending order ----
        if (cnt_max <= 8) {
            for (uint32_t v = l; v < nV; v += TM::N) {
                const uint32_t b = s.occ_off[v], deg = s.occ_off[v + 1] - b;
                const uint32_t cnt = deg < nbr_cap ? deg : nbr_cap;
                uint64_t h = fnv(FNV_OFFSET, deg);
                uint64_t x[8];
    #pragma unroll
                for (uint32_t i = 0; i < 8; ++i) {
                    uint64_t key = ~0ull;
                    if (i < cnt) {
                        const uint32_t ep = s.occ_ep[b + i];
                        key = ((uint64_t)s.e_ar[ep >> 8] << 8) | (uint64_t)(ep & 0xFFu);
                    }
                    x[i] = key;
                }
                sort8(x);
    #pragma unroll
                for (uint32_t i = 0; i < 8; ++i)
                    if (i < cnt) { h = fnv(h, x[i] >> 8); h = fnv(h, x[i] & 0xFFu); }
                s.cur[v] = h;
            }
        } else {
            for (uint32_t v = l; v < nV; v += TM::N) {       // init counts, kept in comp until the components
                const uint32_t deg = s.occ_off[v + 1] - s.occ_off[v];
                s.comp[v] = deg < nbr_cap ? deg : nbr_cap;
            }
            tm.sync();
            for (uint32_t j = l; j < nslots; j += TM::N) {
                const uint32_t v = s.own[j], i = j - s.cursor[v];
                if (i < s.comp[v]) {
                    const uint32_t ep = s.occ_ep[s.occ_off[v] + i];
                    s.nbr[j] = ((uint64_t)s.e_ar[ep >> 8] << 8) | (uint64_t)(ep & 0xFFu);
                }
            }
            tm.sync();
            seg_rank_sort(tm, s.nbr, s.srt, s.own, s.cursor, s.comp, nslots);
            for (uint32_t v = l; v < nV; v += TM::N) {
                const uint32_t so = s.cursor[v], cnt = s.comp[v];
                uint64_t h = fnv(FNV_OFFSET, s.occ_off[v + 1] - s.occ_off[v]);
                for (uint32_t i = 0; i < cnt; ++i) { const uint64_t x = s.srt[so + i]; h = fnv(h, x >> 8); h = fnv(h, x & 0xFFu); }
                s.cur[v] = h;
            }
        }

        // ---- each vertex's neighbour list, fixed for every refinement round: (neighbour, position)
        //      for each occurrence and other position (nbr_cap never truncates it: nbr_cap is the
        //      largest such count) ----
        if (uniform && ua > 1) {
            // one lane per occurrence: the i-th occurrence of v fills slots i*(a-1) .. of v's segment
            for (uint32_t j = l; j < nocc; j += TM::N) {
                const uint32_t v = s.oown[j], i = j - s.occ_off[v];
                const uint32_t ep = s.occ_ep[j], e = ep >> 8, mypos = ep & 0xFFu;
                uint32_t o = s.cursor[v] + i * (ua - 1);
                for (uint32_t k = 0; k < ua; ++k)
                    if (k != mypos) s.nb[o++] = ((uint32_t)s.e_lv[e * A + k] << 4) | k;
            }
        } else {
            for (uint32_t v = l; v < nV; v += TM::N) {
                const uint32_t b = s.occ_off[v], deg = s.occ_off[v + 1] - b;
                uint32_t o = s.cursor[v];
                for (uint32_t j = 0; j < deg; ++j) {
                    const uint32_t ep = s.occ_ep[b + j];
                    const uint32_t e = ep >> 8, mypos = ep & 0xFFu;
                    const uint32_t a = s.e_ar[e];
                    for (uint32_t k = 0; k < a; ++k)
                        if (k != mypos) s.nb[o++] = ((uint32_t)s.e_lv[e * A + k] << 4) | k;
                }
            }
        }
        tm.sync();
    }
    // Segment lengths: the uniform-arity front end leaves cursor a CSR of exactly the neighbour
    // lists; the general one sizes segments for the initial lists too and keeps nnv.
    const uint32_t* segn = fast ? nullptr : s.nnv;
    uint64_t* cur = s.cur;
    uint64_t* nxt = s.nxt;
    uint32_t prev_d = wl_ndistinct(tm, s, cur);

    // ---- refinement (wl_core.hpp:119-141): sorted multiset of fnv(colour, position) over the
    //      neighbour list, folded in order -- in registers when every list has at most 8
    //      entries (two colour buffers), else sorted segment by segment across the team (every
    //      neighbour value is gathered before any colour changes, so colours update in place) ----
    for (uint32_t it = 0; it < WL_MAX_ITERS; ++it) {
        if (nbr_cap <= 8) {
            for (uint32_t v = l; v < nV; v += TM::N) {
                const uint32_t so = s.cursor[v], nn = segn ? segn[v] : s.cursor[v + 1] - so;
                uint64_t h = cur[v];
                uint64_t x[8];
#pragma unroll
                for (uint32_t i = 0; i < 8; ++i) {
                    uint64_t val = ~0ull;
                    if (i < nn) { const uint32_t t = s.nb[so + i]; val = fnv(cur[t >> 4], t & 15u); }
                    x[i] = val;
                }
                sort8(x);
#pragma unroll
                for (uint32_t i = 0; i < 8; ++i)
                    if (i < nn) h = fnv(h, x[i]);
                nxt[v] = h;
            }
            tm.sync();
            uint64_t* t = cur; cur = nxt; nxt = t;
        } else {
            for (uint32_t j = l; j < nslots; j += TM::N) {
                const uint32_t v = s.own[j];
                if (!segn || j - s.cursor[v] < segn[v]) { const uint32_t t = s.nb[j]; s.nbr[j] = fnv(cur[t >> 4], t & 15u); }
            }
            tm.sync();
            seg_rank_sort(tm, s.nbr, s.srt, s.own, s.cursor, segn, nslots);
            for (uint32_t v = l; v < nV; v += TM::N) {
                const uint32_t so = s.cursor[v], nn = segn ? segn[v] : s.cursor[v + 1] - so;
                uint64_t h = cur[v];
#pragma unroll 4
                for (uint32_t i = 0; i < nn; ++i) h = fnv(h, s.srt[so + i]);
                cur[v] = h;
            }
            tm.sync();
            nxt = s.srt;                                // free again for the folds
        }
        const uint32_t d = wl_ndistinct(tm, s, cur);
        if (d == prev_d) break;
        prev_d = d;
    }

    // ---- connected components (wl_core.hpp:143-168) ----
    for (uint32_t v = l; v < nV; v += TM::N) s.comp[v] = v;
    tm.sync();
    while (true) {
        uint32_t changed = 0;
        for (uint32_t e = l; e < nE; e += TM::N) {
            const uint32_t a = s.e_ar[e];
            uint32_t m = 0xFFFFFFFFu;
            for (uint32_t p = 0; p < a; ++p) { const uint32_t c = s.comp[s.e_lv[e * A + p]]; m = c < m ? c : m; }
            for (uint32_t p = 0; p < a; ++p) {
                const uint32_t x = s.e_lv[e * A + p];
                if (s.comp[x] > m) { atomicMin(&s.comp[x], m); changed = 1; }
            }
        }
        tm.sync();
        for (uint32_t v = l; v < nV; v += TM::N) {
            const uint32_t c = s.comp[v], cc = s.comp[c];
            if (cc < c) { atomicMin(&s.comp[v], cc); changed = 1; }
        }
        tm.sync();
        if (!tm.any(changed != 0)) break;
    }

    // ---- folds (wl_core.hpp:170-201) ----
    uint64_t* acc_col = nxt;          // refinement double buffer, finished with
    uint64_t* acc_cnt = s.hset;       // distinct-count set, finished with (HT >= NV)
    for (uint32_t v = l; v < nV; v += TM::N) { acc_col[v] = 0; acc_cnt[v] = 0; }
    tm.sync();
    uint64_t vsum = 0, esum = 0, csum = 0;
    for (uint32_t v = l; v < nV; v += TM::N) {
        const uint64_t sm = splitmix(cur[v]);
        vsum += sm;
        atomicAdd((unsigned long long*)&acc_col[s.comp[v]], (unsigned long long)sm);
        atomicAdd((unsigned long long*)&acc_cnt[s.comp[v]], 1ull);
    }
    for (uint32_t e = l; e < nE; e += TM::N) {
        const uint32_t a = s.e_ar[e];
        uint64_t eh = fnv(FNV_OFFSET, a);
        for (uint32_t p = 0; p < a; ++p) eh = fnv(eh, cur[s.e_lv[e * A + p]]);
        esum += splitmix(eh);
        atomicAdd((unsigned long long*)&acc_cnt[s.comp[s.e_lv[e * A]]], 1ull << 32);
    }
    tm.sync();
    for (uint32_t v = l; v < nV; v += TM::N) {
        if (s.comp[v] != v) continue;
        uint64_t ch = fnv(FNV_OFFSET, acc_cnt[v]);
        ch = fnv(ch, acc_col[v]);
        csum += splitmix(ch);
    }
    vsum = tm.sum64(vsum);
    esum = tm.sum64(esum);
    csum = tm.sum64(csum);
    tm.sync();
    uint64_t hash = FNV_OFFSET;
    hash = fnv(hash, nV);
    hash = fnv(hash, nE);
    hash = fnv(hash, vsum);
    hash = fnv(hash, esum);
    hash = fnv(hash, csum);
    return hash;
}

// One wave hashes the state.
__device__ inline uint64_t wl_hash_wave(WLSpace& s) {
    const Team<1> tm(nullptr);
    return wl_hash_team(s, tm);
}

}  // namespace hgx
