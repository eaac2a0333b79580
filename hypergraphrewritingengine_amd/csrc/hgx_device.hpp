// hgx_device.hpp -- shared device-side definitions for the gfx950 kernels: integer hash
// primitives, the device rule/run/generation views, and wave64 primitives.
//
// The hash primitives are the reference's (common/include/hgcommon/core.hpp:47-87). They are
// restated here because the whole WL value must be bit-identical to the reference's.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace hgx {

constexpr uint64_t FNV_OFFSET = 0xcbf29ce484222325ULL;   // core.hpp:47
constexpr uint64_t FNV_PRIME = 0x100000001b3ULL;         // core.hpp:48
constexpr uint64_t EMPTY_STATE_HASH = 0x9E3779B97F4A7C15ULL;  // core.hpp:169
constexpr uint32_t WL_MAX_ITERS = 100;                   // wl_core.hpp:52
constexpr uint32_t INV = 0xFFFFFFFFu;
constexpr int WAVE = 64;

__host__ __device__ inline uint64_t fnv(uint64_t h, uint64_t v) { return (h ^ v) * FNV_PRIME; }

__host__ __device__ inline uint64_t mix64(uint64_t x) {   // core.hpp:64-71
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33; return x;
}

__host__ __device__ inline uint64_t splitmix(uint64_t z) {   // core.hpp:82-87
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// A rule as the kernels read it. LHS order is the reference's join order
// (hypergraph/src/pattern.cpp:43-75), computed on the host; var_src_* name the first LHS
// (edge, position) of each LHS variable so a rewrite resolves a bound vertex in O(1).
struct DRule {
    // The fields the small-rule join reads come first, so it can stage just this prefix
    // (through jsrc) in LDS.
    uint8_t n_lhs, n_rhs, n_new, pad0;
    uint8_t la[16];
    uint8_t order[16];
    uint8_t piv[16];          // per join depth: a position sharing a bound variable, or 0xFF
    // Binding sources in the join order: jsrc[d][i] = (d' << 4) | i' when the variable at slot i
    // of pattern edge order[d] first appeared at slot i' of depth d' (an earlier depth, or an
    // earlier slot of the same edge), 0xFF where it appears first.
    uint8_t jsrc[16][16];
    uint8_t lv[16][16];
    uint8_t ra[16];
    uint8_t rv[16][16];
    uint8_t var_src_edge[32];
    uint8_t var_src_pos[32];
    uint32_t lhs_mask, rhs_mask, new_mask, pad1;
    // Anchored join orders for delta joins: aord[p] starts at LHS position p and adds the
    // remaining positions greedily by shared variables; apiv[p] are its pivots.
    uint8_t aord[16][16];
    uint8_t apiv[16][16];
};
static_assert(offsetof(DRule, lv) % 4 == 0 && sizeof(DRule) % 4 == 0, "DRule is staged in 32-bit words");

// Per request (run) state the kernels read and update.
struct DRun {
    const DRule* rules;      // this request's rules (device copy)
    uint32_t n_rules;
    uint32_t n_roots, flags, pad0;
    uint64_t* map_keys;      // open-addressing WL-hash set, first writer wins
    uint32_t map_mask;       // capacity - 1 (power of two)
    uint32_t zero_seen;      // the hash 0 is kept beside the table (0 is EMPTY)
    uint32_t ev_base;        // events before the current generation
    uint32_t child_begin;    // first child index of this run in the current generation
};

// Per-run counters, accumulated on the device.
enum {
    CTR_EVENTS = 0, CTR_CANON, CTR_TRIPLES, CTR_PAIRS, CTR_PAIRS_TR, CTR_TRIPLES_TR,
    CTR_BRANCHIAL, CTR_FINGERPRINT, CTR_N
};

// One generation of states (all runs mixed; each run's states contiguous).
struct GenView {
    uint32_t n;               // states
    uint32_t lstride;         // lineage-table entries per state (>= every depth held)
    uint32_t A;               // vertex stride per edge (max arity among these states)
    uint8_t* depth;           // [n] generation of each state inside its own run
    uint32_t* run;            // [n]
    uint64_t* eoff;           // [n] first edge in the pool
    uint32_t* ne;             // [n]
    uint32_t* nextv;          // [n] next fresh vertex label on the lineage
    uint32_t* sid;            // [n] state id inside its run
    uint64_t* hash;           // [n] WL hash
    uint64_t* lanc;           // [n * lstride] causal-ancestor masks of the lineage's events
    uint32_t* ev;             // edge pool: vertices [edges * A]
    uint8_t* ar;              // edge pool: arity
    uint8_t* ps;              // edge pool: generation of the producing event (0 = none)
};

// The matches of one generation (one per event / child).
struct MatchView {
    uint32_t n;
    uint32_t K;               // positions stored per match (batch max n_lhs)
    uint32_t* parent;         // [n] parent index in its generation
    uint32_t* rule;           // [n] rule index inside the run
    uint16_t* q;              // [n * K] matched edge positions in the parent, pattern order
    uint32_t* cne;            // [n] child edge count
};

// Per-event outputs of one generation.
struct EventView {
    uint32_t* in_state;       // [n] state id of the input (in run)
    uint32_t* rule;           // [n] rule index inside the run
    uint64_t* prod;           // [n] generations of the distinct producer events
    uint64_t* tr;             // [n] the subset kept by transitive reduction
    uint8_t* cps;             // [n * cK] or null: producer generation of each consumed edge
    uint32_t cK;              //   (LHS order, 0 = an initial edge), for the triple readback
};

// ---------------- wave64 primitives ----------------

__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Number of set bits of `mask` below this lane.
__device__ inline uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    return v;
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) { uint32_t w = __shfl_xor(v, o, WAVE); v = w > v ? w : v; }
    return v;
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t lo = __shfl_xor((uint32_t)v, o, WAVE);
        uint32_t hi = __shfl_xor((uint32_t)(v >> 32), o, WAVE);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

// Inclusive prefix sum across the wave.
__device__ inline uint32_t wave_incl_scan_u32(uint32_t v) {
    const uint32_t l = lane_id();
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t w = __shfl_up(v, o, WAVE);
        if (l >= (uint32_t)o) v += w;
    }
    return v;
}

// Exclusive prefix sum of a[0..n) in LDS, in place, by one wave; returns the total.
__device__ __forceinline__ uint32_t wave_excl_scan_lds(uint32_t* a, uint32_t n) {
    const uint32_t l = lane_id();
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += WAVE) {
        const uint32_t i = base + l;
        const uint32_t v = i < n ? a[i] : 0u;
        const uint32_t inc = wave_incl_scan_u32(v);
        if (i < n) a[i] = carry + inc - v;
        carry += __shfl(inc, WAVE - 1, WAVE);
    }
    wave_sync();
    return carry;
}

// ---------------- teams: the lanes that share one state ----------------
// A team is one wave (W = 1: wave barriers and cross-lane reductions, nothing in LDS) or the
// W waves of a workgroup (block barriers; the cross-wave step of a reduction goes through
// `red`, W words of LDS). Every member must make every call: each one is a barrier. Results
// are returned to every member.
template <uint32_t W>
struct Team {
    static constexpr uint32_t N = WAVE * W;
    uint32_t r;             // rank in the team
    uint64_t* red;          // [W] LDS, used when W > 1

    static constexpr uint32_t LANES = WAVE;   // lanes of one wave that hold the same state
    __device__ explicit Team(uint64_t* red_) : r(W == 1 ? lane_id() : (uint32_t)threadIdx.x), red(red_) {}
    __device__ bool leader() const { return r == 0; }
    // Per-wave helpers (every wave of a multi-wave team computes the same from its own lanes):
    // this lane's place in the wave, lane u's value, and a ballot with bit u for lane u.
    __device__ uint32_t lane() const { return lane_id(); }
    __device__ uint32_t shfl(uint32_t v, uint32_t u) const { return __shfl(v, (int)u, WAVE); }
    __device__ uint32_t read_lane(uint32_t v, uint32_t u) const {
        return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)u);
    }
    __device__ uint64_t ballot(bool b) const { return __ballot(b); }
    __device__ void sync() const {
        if constexpr (W == 1) wave_sync();
        else __syncthreads();
    }
    // Sum of one value per wave (already reduced inside the wave) over the team's waves.
    __device__ uint64_t across(uint64_t wave_value) const {
        if (lane_id() == 0) red[threadIdx.x / WAVE] = wave_value;
        __syncthreads();
        uint64_t t = 0;
        for (uint32_t k = 0; k < W; ++k) t += red[k];
        __syncthreads();
        return t;
    }
    __device__ uint32_t sum(uint32_t v) const {
        if constexpr (W == 1) return wave_sum_u32(v);
        else return (uint32_t)across(wave_sum_u32(v));
    }
    __device__ uint64_t sum64(uint64_t v) const {
        if constexpr (W == 1) return wave_sum_u64(v);
        else return across(wave_sum_u64(v));
    }
    __device__ uint32_t max(uint32_t v) const {
        v = wave_max_u32(v);
        if constexpr (W == 1) return v;
        else {
            if (lane_id() == 0) red[threadIdx.x / WAVE] = v;
            __syncthreads();
            uint32_t m = 0;
            for (uint32_t k = 0; k < W; ++k) m = (uint32_t)red[k] > m ? (uint32_t)red[k] : m;
            __syncthreads();
            return m;
        }
    }
    __device__ bool any(bool b) const {
        const bool w = __ballot(b) != 0;
        if constexpr (W == 1) return w;
        else return across(w && lane_id() == 0 ? 1u : 0u) != 0;
    }
    // Inclusive prefix sum of one value per member in rank order; the team's total in `total`.
    __device__ uint32_t incl_scan(uint32_t v, uint32_t& total) const {
        const uint32_t inc = wave_incl_scan_u32(v);
        if constexpr (W == 1) {
            total = __shfl(inc, WAVE - 1, WAVE);
            return inc;
        } else {
            const uint32_t w = threadIdx.x / WAVE;
            if (lane_id() == WAVE - 1) red[w] = inc;
            __syncthreads();
            uint32_t pre = 0, t = 0;
            for (uint32_t k = 0; k < W; ++k) {
                const uint32_t c = (uint32_t)red[k];
                pre += k < w ? c : 0u;
                t += c;
            }
            __syncthreads();
            total = t;
            return pre + inc;
        }
    }
    // The first member's value, to every member.
    __device__ uint64_t bcast(uint64_t v) const {
        if constexpr (W == 1) {
            const uint32_t lo = __shfl((uint32_t)v, 0, WAVE), hi = __shfl((uint32_t)(v >> 32), 0, WAVE);
            return ((uint64_t)hi << 32) | lo;
        } else {
            if (r == 0) red[0] = v;
            __syncthreads();
            const uint64_t x = red[0];
            __syncthreads();
            return x;
        }
    }
    // Stream compaction of one round of N items, item r kept when `keep`: returns the item's
    // place among the round's kept items and their number in `total`.
    __device__ uint32_t compact(bool keep, uint32_t& total) const {
        const uint64_t bal = __ballot(keep);
        const uint32_t below = lanes_below(bal);
        if constexpr (W == 1) {
            total = (uint32_t)__popcll(bal);
            return below;
        } else {
            const uint32_t w = threadIdx.x / WAVE;
            if (lane_id() == 0) red[w] = (uint64_t)__popcll(bal);
            __syncthreads();
            uint32_t pre = 0, t = 0;
            for (uint32_t k = 0; k < W; ++k) {
                const uint32_t c = (uint32_t)red[k];
                pre += k < w ? c : 0u;
                t += c;
            }
            __syncthreads();
            total = t;
            return pre + below;
        }
    }
    // Exclusive prefix sum of a[0..n) in LDS, in place; returns the total.
    __device__ uint32_t excl_scan_lds(uint32_t* a, uint32_t n) const {
        if constexpr (W == 1) return wave_excl_scan_lds(a, n);
        else {
            const uint32_t w = threadIdx.x / WAVE;
            uint32_t carry = 0;
            for (uint32_t base = 0; base < n; base += N) {
                const uint32_t i = base + r;
                const uint32_t v = i < n ? a[i] : 0u;
                const uint32_t inc = wave_incl_scan_u32(v);
                if (lane_id() == WAVE - 1) red[w] = inc;
                __syncthreads();
                uint32_t pre = 0, t = 0;
                for (uint32_t k = 0; k < W; ++k) {
                    const uint32_t c = (uint32_t)red[k];
                    pre += k < w ? c : 0u;
                    t += c;
                }
                if (i < n) a[i] = carry + pre + inc - v;
                carry += t;
                __syncthreads();
            }
            return carry;
        }
    }
};

// A sub-wave team: G consecutive lanes of one wave (G = 16 or 32) share one state, so a wave
// works on 64 / G small states at once, each with its own workspace. The groups of a wave may
// run different trip counts: every cross-lane step reads only the group's own lanes (shuffles
// within width G, ballots masked to the group), and a group's lanes are all active at each such
// step because the group's control flow depends on team-uniform values only. Synchronisation is
// the wave's (the groups share its program order).
template <uint32_t G>
struct SubTeam {
    static_assert(G == 16 || G == 32, "sub-wave teams are 16 or 32 lanes");
    static constexpr uint32_t N = G;
    static constexpr uint32_t LANES = G;
    uint32_t r;             // rank in the team
    uint32_t base;          // the team's first lane in the wave

    __device__ SubTeam() : r(lane_id() & (G - 1)), base(lane_id() & ~(G - 1)) {}
    __device__ explicit SubTeam(uint64_t*) : SubTeam() {}
    __device__ bool leader() const { return r == 0; }
    __device__ void sync() const { wave_sync(); }
    __device__ uint32_t lane() const { return r; }
    __device__ uint32_t shfl(uint32_t v, uint32_t u) const { return __shfl(v, (int)u, (int)G); }
    __device__ uint32_t read_lane(uint32_t v, uint32_t u) const { return __shfl(v, (int)u, (int)G); }
    __device__ uint64_t ballot(bool b) const { return (__ballot(b) >> base) & ((1ull << G) - 1ull); }
    __device__ uint32_t sum(uint32_t v) const {
        for (uint32_t o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, (int)o, WAVE);
        return v;
    }
    __device__ uint64_t sum64(uint64_t v) const {
        for (uint32_t o = G / 2; o > 0; o >>= 1) {
            const uint32_t lo = __shfl_xor((uint32_t)v, (int)o, WAVE);
            const uint32_t hi = __shfl_xor((uint32_t)(v >> 32), (int)o, WAVE);
            v += ((uint64_t)hi << 32) | lo;
        }
        return v;
    }
    __device__ uint32_t max(uint32_t v) const {
        for (uint32_t o = G / 2; o > 0; o >>= 1) { const uint32_t w = __shfl_xor(v, (int)o, WAVE); v = w > v ? w : v; }
        return v;
    }
    __device__ bool any(bool b) const { return ballot(b) != 0; }
    __device__ uint32_t incl_scan(uint32_t v, uint32_t& total) const {
        for (uint32_t o = 1; o < G; o <<= 1) {
            const uint32_t w = __shfl_up(v, o, (int)G);
            if (r >= o) v += w;
        }
        total = __shfl(v, (int)(G - 1), (int)G);
        return v;
    }
    __device__ uint64_t bcast(uint64_t v) const {
        const uint32_t lo = __shfl((uint32_t)v, 0, (int)G), hi = __shfl((uint32_t)(v >> 32), 0, (int)G);
        return ((uint64_t)hi << 32) | lo;
    }
    __device__ uint32_t compact(bool keep, uint32_t& total) const {
        const uint64_t m = ballot(keep);
        total = (uint32_t)__popcll(m);
        return (uint32_t)__popcll(m & ((1ull << r) - 1ull));
    }
    __device__ uint32_t excl_scan_lds(uint32_t* a, uint32_t n) const {
        uint32_t carry = 0;
        for (uint32_t b = 0; b < n; b += G) {
            const uint32_t i = b + r;
            const uint32_t v = i < n ? a[i] : 0u;
            uint32_t t;
            const uint32_t inc = incl_scan(v, t);
            if (i < n) a[i] = carry + inc - v;
            carry += t;
        }
        wave_sync();
        return carry;
    }
};

// The team type of a launch: W waves sharing one state (G = 64), or one wave split into teams
// of G lanes (W = 1, G = 16 or 32).
template <uint32_t W, uint32_t G>
struct TeamSel { using type = SubTeam<G>; };
template <uint32_t W>
struct TeamSel<W, 64> { using type = Team<W>; };

}  // namespace hgx
