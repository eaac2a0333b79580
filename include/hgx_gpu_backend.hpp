// hgx_gpu_backend.hpp -- the paclet's GPU backend entry point on this library:
//
//   std::vector<uint8_t> run_gpu_evolution(const GpuJob& job, const HostBridge& host)
//
// the counterpart of paclet_source/hg_gpu_backend.hpp:119 (GpuJob :15-110, HostBridge
// paclet_source/hg_core.hpp:14-16), which run_rewriting_core calls for a parsed job routed to the
// device (hypergraph_ffi.cpp:739-742, hg_gpu_backend.cpp:266). GpuJob keeps the reference's field
// names and defaults, so a caller fills it the same way; the result is the WXF association the
// reference's backend marshals (States, Events, CausalEdges, BranchialEdges, Num*, Warnings). One
// device context lives for the process and serves calls one at a time, as the reference's one
// process-lifetime evolver does (hg_gpu_backend.cpp:177-183).
//
// Errors throw std::exception (std::runtime_error), as the reference's does. Options that would
// change what is explored and that this backend does not serve (canonical-state exploration,
// quotient initial states, sampling, caps, graph properties, per-step branchial selection,
// genesis events, sessions) throw with the reason rather than run a different evolution.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <utility>
#include <vector>

namespace hgx {

struct HostBridge {
    std::function<void(const std::string&)> progress;
};

struct GpuJob {
    // name -> {lhs_edges, rhs_edges}; each edge a list of pattern-variable indices
    const std::vector<std::pair<std::string, std::vector<std::vector<std::vector<int64_t>>>>>& rules;
    // one entry per initial (root) state: a list of edges
    const std::vector<std::vector<std::vector<int64_t>>>& initial_states;

    struct EventCanonCode { static constexpr int kNone = 0, kFull = 1, kAutomatic = 2; };
    struct StateCanonCode { static constexpr int kNone = 0, kAutomatic = 1, kFull = 2; };

    int steps = 0;
    int event_canon_mode = EventCanonCode::kNone;
    int state_canon_mode = StateCanonCode::kFull;
    bool transitive_reduction = true;
    bool explore_from_canonical_states_only = false;
    bool quotient_initial_states = false;
    double exploration_probability = 1.0;
    uint64_t random_seed = 0;
    uint64_t max_device_memory_bytes = 0;      // the context's HBM budget (0: 80% of free HBM)

    double transition_rate = 1.0;
    std::vector<double> rule_weights;
    uint64_t max_states_per_step = 0;
    uint64_t max_successor_states_per_parent = 0;
    uint64_t matches_per_state_rule = 0;

    bool include_states = true;
    bool include_events = true;
    bool include_events_minimal = false;
    bool include_causal_edges = true;
    bool include_branchial_edges = true;
    bool include_canonical_hashes = false;
    bool include_num_states = true;
    bool include_num_events = true;
    bool include_num_causal_edges = true;
    bool include_num_branchial_edges = true;
    bool include_branchial_state_edges = false;
    bool include_branchial_state_edges_all_siblings = false;
    bool include_global_edges = false;
    bool include_state_bitvectors = false;

    std::vector<std::string> graph_properties;
    bool edge_deduplication = true;
    int branchial_step = 0;
    bool show_genesis_events = false;

    std::string session_op;
    uint64_t session_handle = 0;
    std::vector<int64_t> session_from;
};

std::vector<uint8_t> run_gpu_evolution(const GpuJob& job, const HostBridge& host);

}  // namespace hgx
