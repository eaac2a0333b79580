// ref_ir.cpp -- TEST INFRASTRUCTURE. A small front end of our own over the REFERENCE exact
// canonicalizer (IRCanonicalizer, hypergraph/src/ir_canonicalization.cpp over
// common/include/hgcommon/ir_core.hpp), compiled from the reference sources where they lie
// (recipe: oracle/Makefile, output: oracle/_ref/irref, never committed). It writes the IR
// vectors under tests/golden/ that pin the host canonicalizer; nothing in the product links
// or runs it.
//
// Input (whitespace-separated integers): n_states, then per state: n_edges, then per edge:
// arity v0 .. v_{arity-1}. Output: one line per state,
//   <compute_canonical_hash> <vertex_count> <n_edges> then each canonical edge as
//   <arity> <labels...> (canonicalize_edges().canonical_form, in canonical order), then each
//   input edge's rank (compute_canonical_hash_with_edge_rank).

#include "hypergraph/ir_canonicalization.hpp"

#include <cstdio>
#include <fstream>
#include <vector>

using namespace hypergraph;

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: irref STATES\n"); return 2; }
    std::ifstream in(argv[1]);
    size_t ns = 0;
    in >> ns;
    IRCanonicalizer ir;
    for (size_t s = 0; s < ns; ++s) {
        size_t ne = 0;
        in >> ne;
        std::vector<std::vector<VertexId>> edges(ne);
        for (auto& e : edges) {
            size_t a = 0;
            in >> a;
            e.resize(a);
            for (auto& v : e) { unsigned long long x; in >> x; v = (VertexId)x; }
        }
        const uint64_t h = ir.compute_canonical_hash(edges);
        auto c = ir.canonicalize_edges(edges);
        std::printf("%llu %llu %zu", (unsigned long long)h, (unsigned long long)c.canonical_form.vertex_count,
                    c.canonical_form.edges.size());
        for (const auto& e : c.canonical_form.edges) {
            std::printf(" %zu", e.size());
            for (auto v : e) std::printf(" %llu", (unsigned long long)v);
        }
        std::vector<uint32_t> rank;
        {
            auto mk = worker_scratch().mark();
            SVec<SVec<VertexId>> se;
            se.reserve(edges.size());
            for (const auto& e : edges) se.emplace_back(e.begin(), e.end());
            ir.compute_canonical_hash_with_edge_rank(se, rank);
            worker_scratch().release(mk);
        }
        for (size_t e = 0; e < edges.size(); ++e) std::printf(" %u", e < rank.size() ? rank[e] : 0u);
        std::printf("\n");
    }
    return 0;
}
