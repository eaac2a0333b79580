// Drives the C++ adapter (include/hgx_engine.hpp) the way a caller of the reference's
// ParallelEvolutionEngine would, and prints what it read back as one JSON object per case.
// Compared against the oracle by tests/test_gpu_cpp_adapter.py.
#include <cstdio>
#include <string>

#include "hgx_engine.hpp"

static void report(const char* name, hgx::ParallelEvolutionEngine& e) {
    unsigned long long hsum = 0;
    for (uint64_t h : e.state_hashes()) hsum += h;
    unsigned long long esum = 0;
    for (const auto& ev : e.events()) esum += ev.input_state * 31ull + ev.output_state * 7ull + ev.rule_index + ev.step;
    size_t kept = 0;
    for (const auto& c : e.causal_edges()) kept += c.in_reduction ? 1 : 0;
    std::printf("{\"name\": \"%s\", \"states\": %zu, \"canonical\": %zu, \"events\": %zu, \"causal_edges\": %zu, "
                "\"causal_pairs\": %zu, \"branchial\": %zu, \"kept\": %zu, \"hash_sum\": %llu, \"event_sum\": %llu}\n",
                name, e.num_states(), e.num_canonical_states(), e.num_events(), e.num_causal_edges(),
                e.num_causal_event_pairs(), e.num_branchial_edges(), kept, hsum, esum);
}

int main() {
    auto dev = std::make_shared<hgx::Device>(0);
    {
        hgx::ParallelEvolutionEngine e(dev);
        e.add_rule(hgx::make_rule(0).lhs({0, 1}).lhs({1, 2}).rhs({0, 1}).rhs({1, 2}).rhs({2, 0}).build());
        e.set_transitive_reduction(true);
        e.evolve({{1, 2}, {2, 3}, {3, 1}}, 4);
        report("config1-s4-tr", e);
    }
    {
        hgx::ParallelEvolutionEngine e(dev);
        e.add_rule(hgx::make_rule(0).lhs({0, 1}).rhs({0, 2}).rhs({1, 2}).build());
        e.add_rule(hgx::make_rule(1).lhs({0, 1}).lhs({1, 2}).rhs({0, 2}).build());
        e.evolve({{0, 1}, {1, 2}}, 2);
        report("multi-rule", e);
    }
    {
        hgx::EvolveInput in;
        in.rules.push_back(hgx::make_rule(0).lhs({0, 1}).lhs({0, 2}).rhs({0, 2}).rhs({0, 3}).rhs({1, 3}).rhs({2, 3}).build());
        in.initial_states = {{{1, 1}, {1, 1}}, {{0, 1}, {0, 2}}};
        in.num_steps = 3;
        in.materialize_relations = true;
        in.edge_identity = true;
        hgx::EvolveResult r = hgx::evolve(in, dev);
        size_t edges = 0;
        for (const auto& s : r.states) edges += s.edges.size();
        std::printf("{\"name\": \"evolve-input\", \"states\": %zu, \"events\": %zu, \"causal\": %zu, "
                    "\"branchial\": %zu, \"state_edges\": %zu}\n",
                    r.states.size(), r.events.size(), r.causal_edges.size(), r.branchial_edges.size(), edges);
    }
    try {
        hgx::make_rule(0).lhs({0, 40});
        std::printf("{\"name\": \"length-error\", \"thrown\": false}\n");
    } catch (const std::length_error&) {
        std::printf("{\"name\": \"length-error\", \"thrown\": true}\n");
    }
    return 0;
}
