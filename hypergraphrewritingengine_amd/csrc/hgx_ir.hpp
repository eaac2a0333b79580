// hgx_ir.hpp -- exact canonical hash and canonical form of a state by individualization-
// refinement, on the host.
//
// The value is the reference's IRCanonicalizer::compute_canonical_hash /
// canonicalize_edges (hypergraph/src/ir_canonicalization.cpp:30-105, :107-198), whose algorithm
// is common/include/hgcommon/ir_core.hpp:596-955: the smallest leaf form of the IR search tree,
// hashed with FNV-1a. The device evolves and hashes states with WL; this is the exact audit
// and the identity of CanonicalizeStates -> Full, computed where the reference computes it (the
// GPU backend's marshaller, paclet_source/hg_gpu_backend.cpp:269-310), on the host.
//
// What fixes the value, and is therefore kept exactly:
//   * local numbering: the state's distinct vertex labels in ascending order;
//   * the initial partition: cells of equal degree signature (the sorted multiset of
//     (arity, position) over a vertex's occurrences), in ascending signature order, runs
//     compared lexicographically with a proper prefix first;
//   * refinement: the lowest queued cell splits the cells of the vertices on its edges by the
//     sorted multiset of (arity, own position, positions held by the splitter) keys; a split
//     cell's pieces are [untouched, then signature groups ascending], the largest (earliest on
//     a tie) keeps the cell's id, the others take new ids in that order and are queued; every
//     refinement starts with all cells queued;
//   * search: individualize the lowest-id non-singleton cell's vertices in ascending order (the
//     chosen vertex's singleton takes the cell's place, the rest follow, later cells shift);
//     a leaf labels each vertex by its singleton's id; its form is the edges relabelled and
//     sorted (tuples lexicographic, a proper prefix first), flattened as [arity, labels...];
//     the smallest form over the leaves wins;
//   * hash: FNV-1a over the vertex count, then every label of every edge in form order, each
//     edge closed by 0xDEADBEEF; 0 for an empty state, and 0 when a search path needs more
//     than 63 individualizations (the reference gives up there).
// Branches automorphic to an explored one are skipped (automorphisms found as leaves equal to
// the first leaf); automorphic branches reach the same forms, so this changes time, not value.
#pragma once

#include <stdint.h>

#include <vector>

namespace hgx {
namespace ir {

// Canonical hash of one state given by global vertex labels: edge e holds
// verts[edge_off[e] .. edge_off[e+1]). When `form` is given it receives the canonical form,
// flat [arity, labels...] per edge in canonical order (empty when the hash is 0).
uint64_t state_hash(const uint32_t* edge_off, const uint32_t* verts, uint32_t n_edges,
                    std::vector<uint32_t>* form = nullptr, uint32_t* rank = nullptr);
// rank (when given, n_edges entries): each input edge's canonical RANK, its place in the
// winning leaf's edge order (canonical content, then input index) -- the per-edge identity
// events are canonicalized by (ir_core.hpp:589-595, :697-700); all zero when the hash is 0.

}  // namespace ir
}  // namespace hgx
