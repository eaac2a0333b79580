// ref_driver.cpp -- TEST INFRASTRUCTURE. A small front end of our own over the REFERENCE
// CPU engine, compiled from the reference sources where they lie (recipe: oracle/Makefile,
// output: oracle/_ref/hgref, never committed). It is the checker for the C restatement in
// hgx_oracle.c, the writer of the golden fixtures under tests/golden/, and the "reference"
// CPU baseline timed by bench.py. Nothing in the product links or runs it.
//
// It drives exactly the public surface a user of the reference drives:
//   Hypergraph + ParallelEvolutionEngine(hg, threads), add_rule(make_rule(i).lhs().rhs()
//   .build()), set_transitive_reduction, evolve(roots, steps)
//   (hypergraph/include/hypergraph/parallel_evolution.hpp:741-1041,
//    hypergraph/include/hypergraph/pattern.hpp:180-213)
// in StateCanonicalizationMode::None with EVENT_SIG_NONE (the defaults), and reads the
// results back through Hypergraph / CausalGraph accessors.
//
// Job file (whitespace-separated integers):
//   n_rules, then the rule program of include/hgx.h in decimal,
//   n_roots, then per root: n_edges, then per edge: arity v0 .. v_{arity-1},
//   steps, tr(0/1)
// Usage:
//   hgref JOB [--threads N] [--reps K] [--dump OUT]
// Prints one JSON line of counts (+ median wall seconds over K reps). --dump writes the
// raw relation for the LAST rep as text: per state "S hash step", per event
// "E in out rule step", per causal triple "C producer consumer", per branchial "B e1 e2".

// Built a second way, with -DHGX_DROPIN against include/hgx_engine.hpp (tests/cpp/ref_driver_hgx,
// by __graft_entry__.build()): the same driving code over the device engine, with only the
// include and the namespace below changed -- the check that the C++ surface is a drop-in.
#ifdef HGX_DROPIN
#include "hgx_engine.hpp"
#else
#include "hypergraph/hypergraph.hpp"
#include "hypergraph/parallel_evolution.hpp"
#include "hypergraph/pattern.hpp"
#endif

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <set>
#include <string>
#include <vector>

#ifdef HGX_DROPIN
using namespace hgx;
#else
using namespace hypergraph;
#endif

namespace {

struct Job {
    std::vector<RewriteRule> rules;
    std::vector<std::vector<std::vector<VertexId>>> roots;
    size_t steps = 0;
    bool tr = false;
};

bool read_job(const char* path, Job& job) {
    std::ifstream in(path);
    if (!in) return false;
    long n_rules;
    if (!(in >> n_rules)) return false;
    for (long r = 0; r < n_rules; ++r) {
        long n_lhs, n_rhs;
        in >> n_lhs >> n_rhs;
        auto b = make_rule(static_cast<uint16_t>(r));
        for (long side = 0; side < 2; ++side) {
            const long ne = side ? n_rhs : n_lhs;
            for (long e = 0; e < ne; ++e) {
                long a;
                in >> a;
                std::vector<uint8_t> vars(a);
                for (long i = 0; i < a; ++i) { long v; in >> v; vars[i] = static_cast<uint8_t>(v); }
                if (side) b.rhs(vars); else b.lhs(vars);
            }
        }
        job.rules.push_back(b.build());
    }
    long n_roots;
    in >> n_roots;
    for (long r = 0; r < n_roots; ++r) {
        long ne;
        in >> ne;
        std::vector<std::vector<VertexId>> st;
        for (long e = 0; e < ne; ++e) {
            long a;
            in >> a;
            std::vector<VertexId> ed(a);
            for (long i = 0; i < a; ++i) { long v; in >> v; ed[i] = static_cast<VertexId>(v); }
            st.push_back(std::move(ed));
        }
        job.roots.push_back(std::move(st));
    }
    long steps, tr;
    in >> steps >> tr;
    job.steps = static_cast<size_t>(steps);
    job.tr = tr != 0;
    return static_cast<bool>(in);
}

uint64_t fp_mix64(uint64_t x) {
    x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33; return x;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: hgref JOB [--threads N] [--reps K] [--dump OUT]\n");
        return 2;
    }
    size_t threads = 1;
    int reps = 1;
    const char* dump = nullptr;
    for (int i = 2; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--threads") && i + 1 < argc) threads = std::strtoul(argv[++i], nullptr, 10);
        else if (!std::strcmp(argv[i], "--reps") && i + 1 < argc) reps = std::atoi(argv[++i]);
        else if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dump = argv[++i];
    }
    Job job;
    if (!read_job(argv[1], job)) { std::fprintf(stderr, "bad job file\n"); return 2; }

    std::vector<double> secs;
    for (int rep = 0; rep < reps; ++rep) {
        Hypergraph hg;
        ParallelEvolutionEngine engine(&hg, threads);
        for (const auto& r : job.rules) engine.add_rule(r);
        engine.set_transitive_reduction(job.tr);
        auto t0 = std::chrono::steady_clock::now();
        if (job.roots.size() == 1) engine.evolve(job.roots[0], job.steps);
        else engine.evolve(job.roots, job.steps);
        auto t1 = std::chrono::steady_clock::now();
        secs.push_back(std::chrono::duration<double>(t1 - t0).count());
        if (rep + 1 < reps) continue;

        std::set<uint64_t> distinct;
        const uint32_t ns = hg.num_published_states();
        for (uint32_t sid = 0; sid < ns; ++sid) {
            if (hg.get_state(sid).id == INVALID_ID) continue;
            distinct.insert(hg.get_or_compute_canonical_hash(sid));
        }
        uint64_t fp = 0;
        for (uint64_t h : distinct) fp += fp_mix64(h);
        std::sort(secs.begin(), secs.end());
        std::printf("{\"raw_states\": %u, \"events\": %u, \"canonical_states\": %zu, "
                    "\"causal_triples\": %zu, \"causal_pairs\": %zu, \"branchial_pairs\": %zu, "
                    "\"fingerprint\": %llu, \"threads\": %zu, \"reps\": %d, \"median_s\": %.6f, "
                    "\"min_s\": %.6f}\n",
                    hg.num_states(), hg.num_events(), distinct.size(),
                    hg.causal_graph().num_causal_edges(),
                    hg.causal_graph().num_causal_event_pairs(),
                    hg.causal_graph().num_branchial_edges(),
                    static_cast<unsigned long long>(fp), threads, reps,
                    secs[secs.size() / 2], secs.front());
        if (dump) {
            FILE* f = std::fopen(dump, "w");
            if (!f) return 3;
            for (uint32_t sid = 0; sid < ns; ++sid) {
                const State& s = hg.get_state(sid);
                if (s.id == INVALID_ID) continue;
                std::fprintf(f, "S %u %llu %u\n", sid,
                             static_cast<unsigned long long>(hg.get_or_compute_canonical_hash(sid)),
                             s.step);
            }
            const uint32_t ne = hg.num_published_events();
            for (uint32_t eid = 0; eid < ne; ++eid) {
                const Event& ev = hg.get_event(eid);
                if (ev.id == INVALID_ID) continue;
                std::fprintf(f, "E %u %u %u %u %u\n", eid, ev.input_state, ev.output_state,
                             static_cast<unsigned>(ev.rule_index),
                             hg.get_state(ev.output_state).step);
            }
            for (const auto& ce : hg.causal_graph().get_causal_edges())
                std::fprintf(f, "C %u %u\n", ce.producer, ce.consumer);
            for (const auto& be : hg.causal_graph().get_branchial_edges())
                std::fprintf(f, "B %u %u\n", be.event1, be.event2);
            std::fclose(f);
        }
    }
    return 0;
}
