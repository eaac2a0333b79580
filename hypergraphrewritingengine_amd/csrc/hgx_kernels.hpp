// hgx_kernels.hpp -- kernel argument blocks and launchers (host <-> device contract).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hgx_device.hpp"
#include "hgx_wl.hpp"

namespace hgx {

constexpr uint32_t kCountSpreadBlock = 32;   // measured best of 8, 16, 32, 64 (count pass 4.23, 4.22, 4.08, 4.24 ms)

struct MatchArgs {
    GenView P;
    const DRun* runs;
    const uint8_t* active;     // [runs] does this run advance this generation
    uint32_t* cnt;             // count pass: [P.n] matches per state
    const uint64_t* moff;      // emit pass: [P.n + 1] exclusive offsets
    MatchView M;               // emit pass: outputs
    uint32_t max_ne;           // largest state in P (LDS staging)
    uint32_t lds_per_wave;
    const uint32_t* budget;    // count pass: [runs] matches a run may still take
    unsigned long long* run_acc;   // count pass: [runs] running tally (zeroed per generation)
    uint64_t* state_edges;     // count pass: [P.n] edges of each state's children (summed per run after)
    uint32_t max_nextv;        // bound on vertex labels in P (incidence index)
    // Incremental forwarding. States [0, fwd_n) of P are the children of the previous
    // generation's matches (state s was made by match s there); their matches are the
    // parent's matches that avoid the consumed edges, renumbered, plus delta joins anchored at
    // the produced edges -- the same list, in the same order, as a full join.
    uint32_t fwd_n;
    const uint32_t* pm_parent; // [fwd_n] previous-frontier index of each state's parent
    const uint32_t* pm_rule;   // [fwd_n] rule of each previous match
    const uint16_t* pm_q;      // [fwd_n * pm_K] edge positions of each previous match
    uint32_t pm_K;
    const uint64_t* pm_moff;   // [previous frontier + 1] each parent's match range
    uint32_t KL;               // largest LHS of the frontier's active runs
    uint32_t dmax;             // delta matches a state may stage (else it takes the full join)
    uint32_t small;            // every active request has <= 2 rules, each with <= 4 LHS edges
                               //   and variables < 16
    // Optional subset of P's states, as ExpandArgs' ranges (nrange = 0: all of P).
    const uint32_t* rb;
    const uint32_t* rpre;
    uint32_t nrange;
    uint32_t n_items;
    // Count pass: the launch's states in blocks of kCountSpreadBlock consecutive ones, block b
    // visited as (b * spread) mod the number of blocks (spread coprime to it; 0 or 1: in order).
    uint32_t spread;
    // Small rules: the count pass's matches per seed of each fully joined state, kept for the
    // emit pass (which then joins each seed once, not twice): [2 * P's edges], state s's seeds
    // of rule r at 2 * eoff[s] + r * ne[s]; 0xFFFF means "at least that many, join again".
    // Null: every seed is counted in the emit pass too.
    uint16_t* seedcnt;
    // Optional [P.n] per-state flag: 0 = the state is not expanded (a sampled request's
    // exploration decisions), so it has no matches. Null: every state of an active run is.
    const uint8_t* smask;
    // Global-memory workspaces, lds_per_wave bytes per workgroup (eight-wave teams), for states
    // whose workspace does not fit LDS. Null: the workspaces are in LDS.
    uint8_t* scratch;
};

struct ExpandArgs {
    GenView P;                 // parents
    GenView C;                 // children (eoff filled; the rest written here)
    MatchView M;
    EventView E;
    DRun* runs;
    uint64_t* ctr;             // [runs * CTR_N]
    uint32_t max_child_ne;
    uint32_t max_nv;           // bound on distinct vertices of a child
    uint32_t max_nextv;        // bound on vertex labels of a child
    uint32_t lds_per_wave;
    uint8_t* scratch;          // when set: per-team workspace in global memory instead of LDS
                               //   (states too large for LDS), lds_per_wave bytes per team
    // Optional subset of the children: ranges [rb[r], rb[r] + rpre[r+1] - rpre[r]) for
    // r < nrange (rpre = exclusive prefix of the lengths). nrange = 0: all of M.
    const uint32_t* rb;
    const uint32_t* rpre;
    uint32_t nrange;
    uint32_t n_items;          // children this launch covers
    uint32_t lean;             // every child has one arity 2..4 and < 2^16 edges: lean workspace
    uint16_t* tcnt;            // [M] per child: causal triples | triples kept by TR << 8
};

// One run's children in a generation, for k_run_events.
struct RunEvents {
    uint32_t slot, cb, ce;
};
hipError_t launch_run_events(const RunEvents* items, uint32_t n, const EventView& E, const uint16_t* tcnt,
                             uint64_t* ctr, hipStream_t st);

// Workspace of one child in k_expand: the hash's, with room in its nbr region for the
// label -> local id map and in its srt region for the child's vertex labels.
__host__ __device__ inline WLSizes expand_sizes(uint32_t max_child_ne, uint32_t max_nv, uint32_t max_nextv,
                                                uint32_t A, bool lean) {
    const uint32_t ne = max_child_ne ? max_child_ne : 1;
    return WLSizes::make(ne, max_nv, A, lean, (max_nextv + 2) / 2, (ne * A + 1) / 2);
}

struct BranchialArgs {
    GenView P;
    MatchView M;
    const DRun* runs;
    const uint32_t* cnt;       // [P.n] matches per parent
    const uint64_t* moff;      // [P.n + 1]
    uint64_t* ctr;
    uint32_t emit;             // 0: count (into ctr and pcnt), 1: write pairs at poff
    uint32_t* pcnt;            // [P.n] pairs per parent (may be null in count mode)
    const uint64_t* poff;      // [P.n + 1] exclusive offsets (emit)
    uint32_t* pairs;           // [2 * total] event ids within the run
    uint32_t lds_per_wave;     // branchial_lds_per_wave()
};

// Branchial pairs of parents above 64 edges (per-edge lists of consuming matches).
struct BrListArgs {
    GenView P;
    MatchView M;
    const DRun* runs;
    const uint32_t* cnt;       // [P.n] matches per parent (zero for runs without branchial)
    const uint64_t* moff;      // [P.n + 1]
    uint64_t* key_in;          // [M.n * K] (slot << 32 | match), or the sentinel
    uint32_t* val_in;          // [M.n * K] entry index (match * K + k)
    const uint64_t* key;       // [n_keys] the keys, sorted
    const uint32_t* val;       // [n_keys] their entry indices
    uint32_t* pos;             // [M.n * K] sorted position of each entry
    uint64_t n_keys;           // M.n * K
    uint64_t sentinel;         // above every key: (edge slots) << 32
    uint32_t* bc;              // [M.n] pairs found per match (count phase)
    const uint64_t* bo;        // [M.n + 1] exclusive scan of bc (emit phase)
    uint64_t* ctr;
    uint32_t* pcnt;            // [P.n] pairs per parent, added to (may be null)
    const uint64_t* poff;      // [P.n + 1] (emit)
    uint32_t* pairs;           // (emit)
};
// phase 0 keys, 1 sorted positions, 2 pairs (count), 3 pairs (emit); the sort between 0 and 1
// is the caller's (sort_keys_u64).
hipError_t launch_brl(const BrListArgs& a, int phase, hipStream_t st);

struct WLStatesArgs {
    uint32_t n;
    uint32_t A;
    const uint64_t* state_off; // [n + 1] in edges
    const uint8_t* ar;
    const uint32_t* ev;        // dense local vertex ids, stride A
    const uint32_t* nv;        // [n] distinct vertices per state
    uint64_t* out;             // [n]
    const uint32_t* run_of;    // optional: insert into run maps
    DRun* runs;
    uint64_t* ctr;
    uint32_t max_ne, max_nv;
    uint32_t lds_per_wave;
    uint8_t* scratch;          // as ExpandArgs::scratch
    uint32_t lean;             // every state has one arity 2..4 and < 2^16 edges: lean workspace
};

// lanes: the launch may forward (fwd_n != 0), so the small-rule path needs its per-lane LDS
uint32_t match_lds_per_wave(uint32_t max_ne, uint32_t max_nextv, uint32_t A, uint32_t KL, uint32_t dmax,
                            bool small, uint32_t team, bool lanes);
// One admitted request: its rules (copied from a packed staging array), its state map (zeroed)
// and its counter row (zeroed) -- every admission of a step in one launch.
struct AdmitItem {
    const DRule* rule_src;     // its rules in device memory (staging array or prepared batch)
    uint32_t n_rules, pad0;
    DRule* rule_dst;
    uint64_t* map_keys;
    uint64_t map_cap;
    uint32_t slot, pad;
};
hipError_t launch_admit(const AdmitItem* items, uint32_t n, uint64_t* ctr, hipStream_t st);
// Initial states of prepared requests, copied from the batch's device arrays into a generation.
struct RootCopy {
    const uint32_t* ne;
    const uint32_t* nextv;
    const uint64_t* eoff;
    const uint8_t* ar;
    const uint32_t* ev;
    uint32_t src_A, n_roots, slot, dst_state;
    uint64_t dst_edge;
};
hipError_t launch_copy_roots(const RootCopy* items, uint32_t n, const GenView& v, hipStream_t st);

// team: waves per state (1, 4 or 8); with team 1, wpb states per workgroup.
// team: waves per state (1, 4, 8); lanes (team 1 only): 16 or 32 for sub-wave teams (64 / lanes
// states per wave, each with its own lds_per_wave workspace), 64 for one wave per state.
hipError_t launch_match(const MatchArgs& a, bool emit, uint32_t grid_cap, uint32_t wpb, uint32_t team,
                        hipStream_t st, uint32_t lanes = 64);
uint32_t expand_lds_per_wave(uint32_t max_child_ne, uint32_t max_nv, uint32_t max_nextv, uint32_t A, bool lean);
// team: waves per child (1, 2, 4 or 8); with team 1, wpb children per workgroup.
// team: waves sharing one child (1, 2, 4, 8); lanes: with team 1, the lanes sharing one child
// (64, or 32 / 16 for sub-wave teams, several small children per wave, LDS workspaces only).
hipError_t launch_expand(const ExpandArgs& a, uint32_t grid_cap, uint32_t wpb, uint32_t team, hipStream_t st,
                         uint32_t lanes = 64);
uint32_t expand_grid(uint32_t n_items, uint32_t team, uint32_t wpb, uint32_t grid_cap, uint32_t lanes = 64);
hipError_t launch_dedup(const GenView& C, uint32_t n, DRun* runs, uint64_t* ctr, hipStream_t st);
// grid used by the launchers for `items` work items (for sizing global scratch)
uint32_t launch_grid(uint32_t items, uint32_t wpb, uint32_t grid_cap);
uint32_t branchial_lds_per_wave();
// Parents above this many edges have their branchial pairs found through sorted per-edge match
// lists (launch_brl); smaller ones through LDS bitsets (launch_branchial).
constexpr uint32_t BR_MAX_EDGES = 128;
hipError_t launch_branchial(const BranchialArgs& a, uint32_t grid_cap, uint32_t wpb, hipStream_t st);
uint32_t wl_states_lds_per_wave(uint32_t max_ne, uint32_t max_nv, uint32_t A, bool lean);
// team: waves per state (1, 2, 4 or 8); with team 1, wpb states per workgroup.
hipError_t launch_wl_states(const WLStatesArgs& a, uint32_t grid_cap, uint32_t wpb, uint32_t team, hipStream_t st);
uint32_t wl_states_grid(uint32_t n, uint32_t team, uint32_t wpb, uint32_t grid_cap);
// Per run (slot r, states [b[r], e[r])): tot[r] = its matches, edges[r] = its children's edges,
// parent_edges[r] = the edges of its states with at least one match.
hipError_t launch_run_totals(const uint64_t* moff, const uint64_t* state_edges, const uint32_t* cnt,
                             const uint32_t* ne, const uint32_t* b, const uint32_t* e, uint32_t nruns, uint64_t* tot,
                             uint64_t* edges, uint64_t* parent_edges, hipStream_t st);
hipError_t launch_mask_counts(const uint32_t* run, const uint8_t* active, uint32_t n, uint32_t* cnt,
                              hipStream_t st);
// Byte fill of [p, p + bytes) by a kernel: the engine clears its buffers with it (never
// hipMemsetAsync), so a pass it records into a graph holds kernels and copies only.
hipError_t launch_fill(void* p, uint8_t value, uint64_t bytes, hipStream_t st);
// A read-only launch of known traffic (bytes read, 8 written) for counter calibration.
hipError_t launch_read_calibration(const void* p, uint64_t bytes, uint64_t* sink, hipStream_t st);
hipError_t launch_read_gather(void* p, uint64_t bytes, uint64_t* sink, uint64_t* ctr, hipStream_t st);
// Replay check of a recorded pass: flag[0] |= 1 when a run's child total (tot[slot], slots < n)
// leaves the recording's -- a different cap decision against budget[slot], or, within the budget,
// a different total.
hipError_t launch_plan_check(const uint64_t* tot, const uint64_t* want, const uint32_t* budget, uint32_t n,
                             uint32_t* flag, hipStream_t st);
hipError_t launch_rehash(const uint64_t* ok, uint32_t old_cap, uint64_t* keys, uint32_t mask, hipStream_t st);

}  // namespace hgx
