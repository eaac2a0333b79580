/*
 * hgx_oracle.h -- TEST INFRASTRUCTURE. A plain-C, single-threaded restatement of the
 * reference engine's multiway evolution (None state mode, EVENT_SIG_NONE), used only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker. The
 * product path (hypergraphrewritingengine_amd/) never links, loads or calls it.
 *
 * Every function in hgx_oracle.c cites the reference file:line it restates. Parity of
 * each piece is pinned by tests/test_oracle.py against reference/golden_matrix.txt,
 * the brute-force event counts of the 5-step binary rule (3/15/117/1,329/20,739), the
 * reference-built golden vectors under tests/golden/, and closed-form/brute-force checks.
 */
#ifndef HGX_ORACLE_H
#define HGX_ORACLE_H

#include <stdint.h>
#include "../include/hgx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hgxo_result hgxo_result;

/* Evolve one request on the CPU. Returns HGX_OK or HGX_ERR_INVALID. */
int  hgxo_evolve(const hgx_run_desc* run, hgxo_result** out);
void hgxo_free(hgxo_result* r);
int  hgxo_counts(const hgxo_result* r, hgx_counts* out);
int  hgxo_states(const hgxo_result* r, uint64_t* hash, uint32_t* step, uint32_t* creator_event);
int  hgxo_events(const hgxo_result* r, hgx_event* out);
/* Deduplicated causal pairs in (consumer, producer) creation order; in_tr per pair. */
uint64_t hgxo_num_causal_pairs(const hgxo_result* r);
int  hgxo_causal(const hgxo_result* r, hgx_pair* pairs, uint8_t* in_tr);
uint64_t hgxo_num_branchial_pairs(const hgxo_result* r);
int  hgxo_branchial(const hgxo_result* r, hgx_pair* pairs);
/* Edge lists of all raw states: same layout as hgx_result_state_edges. */
int  hgxo_state_edges(const hgxo_result* r, uint64_t* state_offsets, uint32_t* edge_offsets,
                      uint32_t* vertices, uint64_t* n_edges_total, uint64_t* n_vertices_total);

/* Consumed edges of every event as positions in its input state, pattern order: offsets
 * [n_events+1] into positions (either may be NULL; *total gets the position count). */
int  hgxo_consumed(const hgxo_result* r, uint32_t* offsets, uint16_t* positions, uint64_t* total);

/* WL canonical hash of one state (edge_offsets[n_edges+1] into vertices). */
uint64_t hgxo_wl_hash(const uint32_t* edge_offsets, const uint32_t* vertices, uint32_t n_edges);

/* All matches of one rule in one state; count returned, up to cap matches written
 * (n_lhs edge indices each, pattern order, lexicographic in join order). -1 on error. */
int64_t hgxo_match(const uint8_t* rules, uint32_t rules_len, uint32_t n_rules, uint32_t rule,
                   const uint32_t* edge_offsets, const uint32_t* vertices, uint32_t n_edges,
                   uint32_t* match_edges, uint64_t cap);

/* Join order of a rule (pattern.cpp:43-75), for tests. */
int hgxo_match_order(const uint8_t* rules, uint32_t rules_len, uint32_t n_rules, uint32_t rule,
                     uint8_t* order_out);

#ifdef __cplusplus
}
#endif

#endif
