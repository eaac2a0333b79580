/*
 * hgx.h -- C ABI of the MI355X multiway-evolution backend.
 *
 * Plain C: PODs, pointers and sizes, integer status codes, no exceptions and no
 * framework types cross this boundary. Host C++ (include/hgx_engine.hpp), Python
 * (ctypes) and a paclet FFI all bind to exactly these symbols.
 *
 * WHAT ONE REQUEST IS. One multiway evolution, with the semantics of the reference
 * CPU engine in StateCanonicalizationMode::None + EVENT_SIG_NONE: every raw state is
 * expanded, every match is an event, every event creates its own raw state, and
 * each raw state carries its WL canonical hash
 *   (hypergraph/src/parallel_evolution.cpp:74-180, hypergraph/src/rewriter.cpp:11-205,
 *    hypergraph/src/hypergraph.cpp:154-254, common/include/hgcommon/wl_core.hpp:67-202).
 * NumStates is the number of distinct WL hashes over the raw states; the raw count is
 * reported beside it.
 *
 * WHAT THIS REPLACES.
 *   hgx_evolve_batch / hgx_request_*  stand in for hg_gpu::PersistentEvolver::run
 *       (gpu/include/hg_gpu/evolve.hpp:466-497) as called by run_gpu_evolution
 *       (paclet_source/hg_gpu_backend.cpp:174-183), and sit under the C++ adapter
 *       hgx::ParallelEvolutionEngine that mirrors ParallelEvolutionEngine::add_rule /
 *       evolve (hypergraph/include/hypergraph/parallel_evolution.hpp:741-1041).
 *   hgx_wl_hash_batch               stands in for the per-state WL call
 *       Hypergraph::compute_wl_hash (hypergraph/src/hypergraph.cpp:740-755).
 *   hgx_match_count                 stands in for the all-matches join of one state
 *       (common/include/hgcommon/join_core.hpp:143-208, gpu/src/match.cu:343-398).
 */
#ifndef HGX_H
#define HGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Structural limits, the reference's (common/include/hgcommon/core.hpp:42-44). */
#define HGX_MAX_ARITY          16
#define HGX_MAX_PATTERN_EDGES  16
#define HGX_MAX_VARS           32
/* Causal lineage masks are 64-bit, one bit per generation. */
#define HGX_MAX_STEPS          63

/* Status codes. Every entry point returns one; hgx_last_error() gives the text. */
enum {
    HGX_OK = 0,
    HGX_ERR_INVALID = 1,      /* malformed rule / state / argument                    */
    HGX_ERR_NO_DEVICE = 2,    /* no HIP device or the kernels are missing for it       */
    HGX_ERR_DEVICE = 3,       /* a HIP runtime call failed                             */
    HGX_ERR_CAPACITY = 4,     /* the request does not fit the context's HBM budget     */
    HGX_ERR_NOT_READY = 5,    /* hgx_request_poll: still running                        */
    HGX_ERR_UNSUPPORTED = 6   /* a size beyond what this build handles                  */
};

/* Request flags (hgx_run_desc.flags). Recording flags mirror hgcommon::RecordSet
 * (common/include/hgcommon/core.hpp:128-155); the TR flag mirrors
 * ParallelEvolutionEngine::set_transitive_reduction (parallel_evolution.hpp:781). */
#define HGX_REC_CAUSAL      0x1u   /* count causal triples / pairs                   */
#define HGX_REC_BRANCHIAL   0x2u   /* count branchial pairs                          */
#define HGX_REC_RELATIONS   0x4u   /* also materialise the pair lists for readback   */
#define HGX_REC_STATES      0x8u   /* keep every state's edge list for readback      */
#define HGX_OPT_TR          0x100u /* causal pairs: keep the transitive-reduction tag */
/* Quotient exploration (ParallelEvolutionEngine::set_explore_from_canonical_states_only,
 * parallel_evolution.hpp:921-927; the device's state_survives_dedup with dedup on,
 * gpu/include/hg_gpu/exploration.hpp:66-122): a new state is expanded only when it is the
 * first state, in state-id order, of its exact (IR) isomorphism class; later states of the
 * class are kept as states and events but never matched. The exact hashes are computed on the
 * host between generations (a host round trip per generation, as for sampling). Every root is
 * expanded unless HGX_OPT_QUOTIENT_ROOTS is also set, when a root isomorphic to an earlier root
 * is not (EvolveInput::quotient_initial_states, gpu/include/hg_gpu/evolve.hpp:44-52).
 * hgx_result_unfold gives the raw unfolding beneath such a result. */
#define HGX_OPT_QUOTIENT        0x200u
#define HGX_OPT_QUOTIENT_ROOTS  0x400u
#define HGX_REC_DEFAULT     (HGX_REC_CAUSAL | HGX_REC_BRANCHIAL)

/*
 * RULE PROGRAM. A rule set is a byte string, rule after rule:
 *   n_lhs, n_rhs,
 *   n_lhs x ( arity, var[arity] ),
 *   n_rhs x ( arity, var[arity] )
 * Variables are small integers < HGX_MAX_VARS. A variable that occurs only on the RHS
 * is fresh: each application mints a new vertex for it, ascending by variable index
 * (common/include/hgcommon/rewrite_core.hpp:53-62). Rule i is reported as rule index i.
 * Limits: 1 <= n_lhs <= 16, 0 <= n_rhs <= 16, 1 <= arity <= 16.
 */

/* Sampling and capping of one request: the options TransitionRate, RuleWeights,
 * MatchesPerStateRule, ExplorationProbability, MaxStatesPerStep, MaxSuccessorStatesPerParent
 * and RandomSeed (paclet_source/ffi_job.hpp:48-72, as the GPU backend takes them,
 * hg_gpu_backend.cpp:152-160), with the GPU engine's decisions, taken generation by generation:
 *   - a transition (input state, rule, consumed edges) is keyed by the event signature over the
 *     input state's exact (IR) hash, the rule and the consumed edges' canonical ranks in pattern
 *     order (event_core.hpp:55-84); matches_per_state_rule keeps the k matches of each
 *     (state, rule) with the smallest transition rank (ties in event order), then each kept
 *     match survives with probability transition_rate x rule_weights[rule] (clamped to [0, 1];
 *     rules past the weights weigh 1), by a draw keyed on the transition and the seed
 *     (common/include/hgcommon/sampling_core.hpp). A transition that does not survive makes
 *     no event, no state and no relation;
 *   - a new state below the last step is expanded when the exploration coin passes
 *     (exploration_probability, keyed on the seed, its step and its state id; a seed of 0 with
 *     a coin to flip is replaced by a random one) and then, counted in event order, while its
 *     step has expanded fewer than max_states_per_step states and its parent fewer than
 *     max_successor_states_per_parent (0 = no bound) (gpu/include/hg_gpu/exploration.hpp:66-122).
 * Transition decisions need the exact hash and edge ranks of every expanded state: the host
 * computes them (the IR canonicalizer) between generations, so a sampled request costs a host
 * round trip per generation. Sampled requests are not recorded into graphs. */
typedef struct hgx_sampling {
    double          transition_rate;          /* 1: every transition                     */
    const double*   rule_weights;             /* per-rule multipliers, or NULL            */
    uint32_t        n_rule_weights;
    uint32_t        reserved;
    uint64_t        seed;
    double          exploration_probability;  /* 1: every new state is expanded          */
    uint64_t        max_states_per_step;      /* 0: no bound                             */
    uint64_t        max_successor_states_per_parent;
    uint64_t        matches_per_state_rule;
} hgx_sampling;

/* One evolution request. Vertex ids of the initial state(s) are arbitrary uint32; each
 * root is renumbered densely by first appearance, as the FFI does
 * (paclet_source/hypergraph_ffi.cpp:599-636). */
typedef struct hgx_run_desc {
    const uint8_t*  rules;         /* rule program                                    */
    uint32_t        rules_len;     /* bytes in the rule program                       */
    uint32_t        n_rules;
    const uint32_t* edge_offsets;  /* [n_edges+1] offsets of each edge in `vertices`   */
    const uint32_t* vertices;
    uint32_t        n_edges;
    const uint32_t* root_offsets;  /* [n_roots+1] edge-index ranges, or NULL: one root */
    uint32_t        n_roots;
    uint32_t        steps;         /* generations; 0 gives the roots only             */
    uint32_t        state_cap;     /* 0 = none. Otherwise the run stops after the
                                      largest step s <= steps whose cumulative raw
                                      state count is <= state_cap.                    */
    uint32_t        flags;         /* HGX_REC_* | HGX_OPT_TR                          */
    const hgx_sampling* sampling;  /* NULL: the full multiway evolution               */
} hgx_run_desc;

/* Per-request totals. Field meanings follow the reference's counters:
 *   raw_states        Hypergraph::num_states()
 *   events            Hypergraph::num_events()
 *   canonical_states  distinct WL hashes over the raw states
 *   causal_triples    CausalGraph::num_causal_edges() with TR off (one per consumed
 *                     edge that has a producer event)
 *   causal_pairs      CausalGraph::num_causal_event_pairs() with TR off
 *   causal_pairs_tr   the same relation after transitive reduction
 *   causal_triples_tr triples whose pair survives the reduction (num_causal_edges, TR on)
 *   branchial_pairs   CausalGraph::num_branchial_edges()
 *   fingerprint       sum of mix64(h) over distinct WL hashes
 *                     (reference/golden_matrix.hpp:101-107)
 * (hypergraph/src/causal_graph.cpp:187-234, :324-351, :356-386). */
typedef struct hgx_counts {
    uint64_t raw_states;
    uint64_t events;
    uint64_t canonical_states;
    uint64_t causal_triples;
    uint64_t causal_pairs;
    uint64_t causal_pairs_tr;
    uint64_t causal_triples_tr;
    uint64_t branchial_pairs;
    uint64_t fingerprint;
    uint32_t steps_done;       /* generations actually evolved                        */
    uint32_t truncated;        /* 1 when state_cap stopped the run before `steps`;
                                  2 when a generation did not fit the context's HBM
                                  budget: the counts are those of the generations
                                  completed before it (hgx_result_status says why)    */
} hgx_counts;

/* Event record. States are numbered roots first (0..n_roots-1), then the child of event
 * e is state n_roots + e. Events are numbered generation by generation; inside a
 * generation by input state, then rule, then match. */
typedef struct hgx_event {
    uint32_t input_state;
    uint32_t output_state;
    uint32_t rule;
    uint32_t step;             /* the output state's step                               */
} hgx_event;

typedef struct hgx_pair {
    uint32_t a;                /* causal: producer event; branchial: smaller event id   */
    uint32_t b;                /* causal: consumer event; branchial: larger event id    */
} hgx_pair;

typedef struct hgx_ctx hgx_ctx;
typedef struct hgx_result hgx_result;

/* ---- context ------------------------------------------------------------------------ */

/* budget_bytes = 0 takes 80% of the device's free HBM. */
int  hgx_ctx_create(int device, uint64_t budget_bytes, hgx_ctx** out);
void hgx_ctx_destroy(hgx_ctx* ctx);
/* Text of the last failure on this context (or of the last failed hgx_ctx_create when
 * ctx is NULL). Returns the full length; writes at most len-1 bytes plus NUL. */
size_t hgx_last_error(const hgx_ctx* ctx, char* buf, size_t len);
/* Kernel/device description string (arch, CU count, HBM), for reports. */
size_t hgx_device_info(const hgx_ctx* ctx, char* buf, size_t len);

/* Context options (none changes any result, only how it is computed):
 *   HGX_OPTION_FORWARD  incremental match forwarding for states made by the previous
 *                       generation (ParallelEvolutionEngine's forwarding, enabled there by
 *                       rule_analysis forwarding_pays, hypergraph/src/rule_analysis.cpp:134):
 *                       1 on, 0 off (full join for every state), -1 automatic (default).
 *   HGX_OPTION_DELTA_CAP delta matches one forwarded state may stage on chip, 1..4096
 *                       (default 256); a state with more takes the full join.
 *   HGX_OPTION_EXPAND_TEAM waves that rewrite and hash one child: 1, 2, 4 or 8 for every
 *                       child, 0 automatic (default: one wave for children whose workspace is
 *                       small, a workgroup of 4 or 8 waves sharing one workspace for large ones).
 *   HGX_OPTION_JOIN_TEAM waves that join one parent state: 1, 4 or 8 for every state, 0 automatic
 *                       (chosen the same way).
 *   HGX_OPTION_EXPAND_LANES lanes that rewrite and hash one child when one wave is the team:
 *                       64, or 32 / 16 (sub-wave teams: two or four children per wave, each in
 *                       its own workspace), for every size class whose wave workspaces fit a
 *                       workgroup's LDS; 0 automatic (default: 16 lanes for children of up to 32
 *                       edges, 32 up to 64, else 64).
 *   HGX_OPTION_JOIN_LANES lanes that join one parent state when one wave is the team: 64, or
 *                       32 / 16 (two or four states per wave, each in its own workspace); 0
 *                       automatic (the tuned or default shape per size class).
 *   HGX_OPTION_COUNTER_FLUSH for hardware-counter runs only: MiB (16..1024) of a buffer that a
 *                       read kernel (k_flush) streams through before and after every kernel the
 *                       engine times, so the L2 holds no dirty line of another kernel while one
 *                       runs and a kernel's write-backs land in it or in the flush after it;
 *                       0 off (default). Results are unchanged; passes take longer.
 * Returns HGX_ERR_INVALID for an unknown option or value. */
#define HGX_OPTION_FORWARD 1
#define HGX_OPTION_DELTA_CAP 2
#define HGX_OPTION_EXPAND_TEAM 3
#define HGX_OPTION_JOIN_TEAM 4
#define HGX_OPTION_EXPAND_LANES 5
#define HGX_OPTION_JOIN_LANES 6
#define HGX_OPTION_COUNTER_FLUSH 7
int hgx_ctx_set_option(hgx_ctx* ctx, uint32_t option, int64_t value);

/* ---- synchronous batch: the plain and the batched path -------------------------------- */

/* Evolve n independent requests together on the device, each with its own rules, roots,
 * steps and cap. results[i] receives a handle per request (release each). Results do not
 * depend on which other requests share the batch. */
int hgx_evolve_batch(hgx_ctx* ctx, const hgx_run_desc* runs, uint32_t n, hgx_result** results);

/* Prepared batches: the requests are parsed and their rules and initial states staged in
 * device memory once (hgx_batch_prepare); hgx_batch_evolve then runs them from device-resident
 * inputs, any number of times, with results identical to hgx_evolve_batch on the same
 * requests. No counterpart in the reference: it is the batched path of the service with its
 * inputs already on the device. A batch must be released before its context is destroyed. */
typedef struct hgx_batch hgx_batch;
int  hgx_batch_prepare(hgx_ctx* ctx, const hgx_run_desc* runs, uint32_t n, hgx_batch** out);
int  hgx_batch_evolve(hgx_ctx* ctx, const hgx_batch* batch, hgx_result** results);
void hgx_batch_release(hgx_batch* batch);

/* A prepared batch's whole pass as one HIP graph. hgx_batch_capture runs the batch once eagerly,
 * keeping every host decision a generation takes (the per-run child totals and sizes read after
 * the count pass), then runs it again under stream capture with those decisions, no host
 * synchronisation, every buffer kept and every host copy staged in pinned memory: one graph of
 * all generations' launches. hgx_replay_launch launches it (one graph launch per pass), reads the
 * counters and returns a result per request, identical to hgx_batch_evolve's; the graph also
 * checks on the device that every generation's per-run totals equal the recording's, and the
 * launch fails with HGX_ERR_DEVICE if they do not. A replay's results read the replay's buffers:
 * they are valid until the next launch or the release of that replay, which must come before the
 * batch's and the context's. A batch whose pass does not fit the budget is refused
 * (HGX_ERR_CAPACITY). No counterpart in the reference; it is the north star's hipGraph-captured
 * batched step, for batches evolved repeatedly from device-resident inputs. */
typedef struct hgx_replay hgx_replay;
int  hgx_batch_capture(hgx_ctx* ctx, const hgx_batch* batch, hgx_replay** out);
int  hgx_replay_launch(hgx_ctx* ctx, hgx_replay* replay, hgx_result** results);
void hgx_replay_release(hgx_replay* replay);

/* ---- the paclet's job envelope (WXF in, WXF out) -------------------------------------- */

/* run_rewriting_core (paclet_source/hg_core.hpp:21-22) on this library: job is a WXF
 * Association with the FFI's keys (InitialStates or InitialEdges, Rules as name ->
 * Rule[lhs, rhs], Steps, Options; hypergraph_ffi.cpp:136-376). On HGX_OK *out holds a WXF
 * Association shaped as the GPU backend's result (States, Events, CausalEdges, BranchialEdges,
 * NumStates, NumEvents, NumCausalEdges, NumBranchialEdges, Warnings; hg_gpu_backend.cpp:362-640)
 * of *out_n bytes, to be freed with hgx_buffer_free. Options this backend does not serve
 * (sessions, genesis events, graph properties, per-step branchial selection, canonical
 * exploration under content identity) give HGX_ERR_INVALID with the reason in hgx_last_error. */
int  hgx_run_wxf(hgx_ctx* ctx, const uint8_t* job, size_t n, uint8_t** out, size_t* out_n);
void hgx_buffer_free(uint8_t* p);
/* Parse and check a WXF job without a device: HGX_OK, or HGX_ERR_INVALID with the reason. */
int  hgx_wxf_validate(const uint8_t* job, size_t n, char* err, size_t err_len);

/* ---- served path: continuous batching ------------------------------------------------- */

/* Queue a request. It is admitted when its HBM footprint estimate fits what running
 * requests leave of the budget, joins the next generation launch, and completes on its
 * own. Submission is thread-safe. */
int hgx_request_submit(hgx_ctx* ctx, const hgx_run_desc* run, uint64_t* request_id);
/* Drive the device by one generation of every admitted request (called by the serving
 * loop; hgx_request_wait calls it too). Returns the number of requests still live. */
int hgx_serve_step(hgx_ctx* ctx, uint32_t* live);
/* HGX_OK and *out set when done; HGX_ERR_NOT_READY while running. */
int hgx_request_poll(hgx_ctx* ctx, uint64_t request_id, hgx_result** out);
int hgx_request_wait(hgx_ctx* ctx, uint64_t request_id, hgx_result** out);
/* Results of requests that finished and were not yet taken, up to cap of them; each is
 * removed from the finished set (ids[i] names results[i]). */
int hgx_serve_drain(hgx_ctx* ctx, uint64_t* ids, hgx_result** results, uint32_t cap, uint32_t* n);
/* Footprint estimate (bytes of HBM) the admission control charges for a request. */
uint64_t hgx_request_footprint(const hgx_run_desc* run);

/* ---- results ------------------------------------------------------------------------- */

int hgx_result_counts(const hgx_result* r, hgx_counts* out);
/* The same for n results at once (one call for a whole batch). */
int hgx_result_counts_many(const hgx_result* const* rs, uint32_t n, hgx_counts* out);
/* Per raw state: WL hash, step, and the event that created it (0xFFFFFFFF for roots).
 * Any pointer may be NULL. cap is the capacity in elements; returns HGX_ERR_INVALID if
 * cap < raw_states. */
int hgx_result_states(const hgx_result* r, uint64_t* hash, uint32_t* step,
                      uint32_t* creator_event, uint64_t cap);
int hgx_result_events(const hgx_result* r, hgx_event* out, uint64_t cap);
/* Consumed edges of every event (with HGX_REC_STATES): event e consumed the edges at
 * positions[offsets[e] .. offsets[e+1]) of its input state's edge list (hgx_result_state_edges
 * order), one per LHS edge in rule order; its produced edges are the last n_rhs edges of its
 * output state. offsets has events + 1 entries; positions may be NULL to learn the total
 * (offsets[events]). The device counterpart of Event::consumed_edges / produced_edges
 * (hypergraph/src/hypergraph.cpp:432-520), as positions rather than global edge ids. */
int hgx_result_consumed(const hgx_result* r, uint32_t* offsets, uint16_t* positions, uint64_t cap);

/* Edge lists of every raw state (needs HGX_REC_STATES): state_offsets[raw_states+1] in
 * edges, edge_offsets[n_edges_total+1] in vertices, vertices[]. Call with NULL arrays to
 * learn the totals. Vertex labels are per-lineage: each root renumbered from 0, fresh
 * vertices continue from the largest label on the lineage. */
int hgx_result_state_edges(const hgx_result* r, uint64_t* state_offsets, uint32_t* edge_offsets,
                           uint32_t* vertices, uint64_t* n_edges_total, uint64_t* n_vertices_total);
/* Causal pairs (needs HGX_REC_CAUSAL|HGX_REC_RELATIONS), deduplicated; in_tr[i] = 1 if the
 * pair survives transitive reduction. */
int hgx_result_causal(const hgx_result* r, hgx_pair* pairs, uint8_t* in_tr, uint64_t cap);
/* Causal triples (needs HGX_REC_CAUSAL|HGX_REC_RELATIONS): one per consumed edge that an event
 * produced -- the entries of CausalGraph::get_causal_edges() (hypergraph/src/causal_graph.cpp:
 * 187-234, CausalEdge{producer, consumer, edge}, hypergraph/include/hypergraph/types.hpp:239-248)
 * -- with the consumed edge named by its LHS index in the consumer's match (the reference names
 * it by its global EdgeId, kept there "for debugging/viz"). causal_triples of them, in consumer
 * order, then LHS order; in_tr[i] = 1 if the triple's pair survives transitive reduction (the
 * triples the reference keeps with reduction on). */
typedef struct hgx_triple {
    uint32_t producer;
    uint32_t consumer;
    uint32_t lhs_index;
} hgx_triple;
int hgx_result_causal_triples(const hgx_result* r, hgx_triple* triples, uint8_t* in_tr, uint64_t cap);
/* Branchial pairs (needs HGX_REC_BRANCHIAL|HGX_REC_RELATIONS), (smaller, larger) event ids. */
int hgx_result_branchial(const hgx_result* r, hgx_pair* pairs, uint64_t cap);

/* The raw unfolding beneath an evolution: the multiway system every raw state expanded, counted
 * without materialising it. The result's expansions are grouped into exact (IR) classes; each
 * class's expansion is recorded once, from the first state of the class that was expanded, in
 * slots (its edges' canonical ranks); then the device replays every class expansion against
 * every raw instance of the class, generation by generation from one instance per root, each
 * instance carrying the producing generation of every slot and the causal ancestry of the events
 * on its branch. The reference's reconstruction under quotient exploration
 * (common/include/hgcommon/quotient_replay_core.hpp:112-226 qr_apply, driven by
 * hypergraph/src/hypergraph.cpp:977-1195; counts hypergraph.cpp:1413-1445). Needs a complete
 * result made with HGX_REC_STATES; for a result made without HGX_OPT_QUOTIENT the counts equal the
 * result's own (hgx_counts). States whose exact hash the canonicalizer gave up on, or above 4,096
 * edges, give HGX_ERR_UNSUPPORTED. */
typedef struct hgx_unfolding {
    uint64_t raw_events;         /* applications: num_reconstructed_raw_events               */
    uint64_t causal_triples;     /* consumed edges with a producer: num_reconstructed_causal_edges */
    uint64_t causal_pairs;       /* distinct (producer, consumer) event pairs               */
    uint64_t causal_pairs_tr;    /* the pairs that survive transitive reduction             */
    uint64_t causal_triples_tr;  /* triples whose pair survives                             */
    uint64_t branchial_pairs;    /* sibling applications with overlapping consumed edges     */
    uint64_t instances;          /* raw states below the last step (roots included)        */
    uint64_t classes;            /* exact classes whose expansion was recorded              */
    uint64_t class_matches;      /* recorded (class, match) expansions                      */
} hgx_unfolding;
int hgx_result_unfold(hgx_ctx* ctx, const hgx_result* r, hgx_unfolding* out);
/* Device time spent on this request's generations, and its latency
 * (completion - submission) on the served path, in nanoseconds. */
int hgx_result_timing(const hgx_result* r, uint64_t* device_ns, uint64_t* latency_ns);
/* HGX_OK for a complete result (or one its state cap stopped), HGX_ERR_CAPACITY for a partial one:
 * a generation did not fit the context's HBM budget, and the result holds the generations before
 * it (none, when the request's initial states did not fit). The reason goes to msg (at most len-1
 * bytes plus NUL; msg may be NULL). The counterpart of EvolveResult::warnings and the partial
 * result the reference returns on capacity overflow (gpu/src/evolve.cu:573-700,
 * gpu/include/hg_gpu/errors.hpp:31-69). */
int hgx_result_status(const hgx_result* r, char* msg, size_t len);
void hgx_result_release(hgx_result* r);
/* The same for n results at once (a whole batch; null entries are skipped). */
void hgx_result_release_many(hgx_result* const* rs, uint32_t n);

/* ---- building blocks, exposed for tests and the match-join workload -------------------- */

/* WL canonical hash of n states given as edge lists (state_offsets in edges,
 * edge_offsets in vertices), computed on the device. Bit-identical to
 * hgcommon::wl_canonical_hash with the caller sizing of wl_hash.hpp:109-239; an empty
 * state hashes to EMPTY_STATE_CANONICAL_HASH (hypergraph/src/hypergraph.cpp:743-745). */
int hgx_wl_hash_batch(hgx_ctx* ctx, const uint64_t* state_offsets, uint32_t n_states,
                      const uint32_t* edge_offsets, const uint32_t* vertices, uint64_t* out_hash);

/* Exact canonical hash of n states given like hgx_wl_hash_batch's input: the reference's
 * IRCanonicalizer::compute_canonical_hash (hypergraph/src/ir_canonicalization.cpp:180-198,
 * common/include/hgcommon/ir_core.hpp:596-955), the identity of CanonicalizeStates -> Full.
 * Host computation on worker threads (as the reference's GPU backend does it after the
 * device run, paclet_source/hg_gpu_backend.cpp:269-310); ctx may be NULL. An empty state, and
 * one whose search would need more than 63 individualization levels, hash to 0 (the
 * reference's values). */
int hgx_ir_hash_batch(hgx_ctx* ctx, const uint64_t* state_offsets, uint32_t n_states,
                      const uint32_t* edge_offsets, const uint32_t* vertices, uint64_t* out_hash);
/* The same, and each edge's canonical rank: out_rank[e - state_offsets[0]] for every edge e of
 * the states, its place in its state's canonical edge order (canonical content, then input
 * index; the ranks of compute_canonical_hash_with_edge_rank, ir_canonicalization.cpp:227-243)
 * -- the per-edge identity of event canonicalization (common/include/hgcommon/event_core.hpp). */
int hgx_ir_rank_batch(hgx_ctx* ctx, const uint64_t* state_offsets, uint32_t n_states,
                      const uint32_t* edge_offsets, const uint32_t* vertices, uint64_t* out_hash,
                      uint32_t* out_rank);
/* The canonical form of one state (IRCanonicalizer::canonicalize_edges,
 * ir_canonicalization.cpp:107-170): its edges relabelled 0..n-1 and in canonical order,
 * out_edge_offsets[n_edges+1] and out_vertices[edge_offsets[n_edges] - edge_offsets[0]];
 * *out_hash its canonical hash. Host only. */
int hgx_ir_canonical_form(const uint32_t* edge_offsets, const uint32_t* vertices, uint32_t n_edges,
                          uint32_t* out_edge_offsets, uint32_t* out_vertices, uint64_t* out_hash);

/* Number of matches of rule `rule` (index into the program) in one state, computed on the
 * device with the large-state path (CSR incidence index + seeded join). If
 * match_edges is non-NULL it receives up to cap matches, n_lhs edge indices each in
 * pattern order. */
int hgx_match_count(hgx_ctx* ctx, const uint8_t* rules, uint32_t rules_len, uint32_t n_rules,
                    uint32_t rule, const uint32_t* edge_offsets, const uint32_t* vertices,
                    uint32_t n_edges, uint64_t* count, uint32_t* match_edges, uint64_t cap,
                    float* device_ms);

/* The same join for callers that cannot allocate device memory: no counterpart in the reference;
 * it exists for code that must not call hipMalloc (under HIP graph capture, in tight serving
 * loops, in harnesses that forbid allocation). The state and the output are the caller's device
 * memory: vertices [n_edges * stride] (edge e's vertices at e * stride, ids below n_vertices),
 * arity [n_edges] (1..stride), match_edges [cap * n_lhs] (may be NULL when cap is 0); *count is a
 * host pointer. Every other byte the join uses comes from a fixed device array of
 * HGX_NOALLOC_SCRATCH_MIB MiB per device, reserved at the first hgx_ctx_create on the device and
 * held until the process ends; a state or a frontier that does not fit it gives HGX_ERR_CAPACITY
 * with the reason, before anything is written. Matches and their order are hgx_match_count's. The
 * array is shared by every call on a device: calls through one context are serialised, calls
 * that overlap on one device through different contexts race. */
#define HGX_NOALLOC_SCRATCH_MIB 64
int hgx_match_count_noalloc(hgx_ctx* ctx, const uint8_t* rules, uint32_t rules_len, uint32_t n_rules,
                            uint32_t rule, const uint32_t* vertices, const uint8_t* arity, uint32_t n_edges,
                            uint32_t stride, uint32_t n_vertices, uint64_t* count, uint32_t* match_edges,
                            uint64_t cap);

/* ---- autotuning ------------------------------------------------------------------------- */

/* Launch shapes per size class -- for the expansion (children of up to 32, 64, 128, 512, 2048 edges,
 * more) the lanes per child (16, 32, 64) and workgroups per CU; for the join (parents of up to 128,
 * 512, 2048 edges, more) the waves per state (by size, 1, 4) and workgroups per CU -- measured on
 * the device over `batch` (each candidate run reps times after a warm-up pass, launches
 * event-timed per class) and the fastest kept per class. The choice takes effect in ctx at once
 * and, with cache_path, is written to a JSON cache under the key "<arch>:<CUs>" (other keys kept);
 * hgx_ctx_create loads its device's entry from HGX_TUNE_CACHE, else tune_cache.json beside the
 * library. report (may be NULL) receives the choice and every measurement as JSON. No result
 * depends on the shapes. No counterpart in the reference, whose GPU launch shapes are fixed by
 * hand (gpu/src/persistent.cu:917-923). */
int hgx_autotune(hgx_ctx* ctx, const hgx_batch* batch, uint32_t reps, const char* cache_path,
                 char* report, size_t report_len);
/* Load this device's tuned shapes from a cache written by hgx_autotune. */
int hgx_tune_load(hgx_ctx* ctx, const char* path);
/* The shapes in effect, as JSON (with their source); returns the full length. */
size_t hgx_tune_get(const hgx_ctx* ctx, char* buf, size_t len);

/* ---- kernel-level measurement ----------------------------------------------------------- */

/* With profiling on, every launch of the generation kernels is bracketed by HIP events on the
 * context's stream and its algorithmic HBM bytes (the minimum the launch must move: inputs read
 * once, outputs written once) are recorded. No counterpart in the reference; it feeds the
 * roofline report. */
typedef struct hgx_kernel_stat {
    char     name[32];
    uint64_t launches;
    double   total_ms;        /* sum of event-timed launch durations                     */
    uint64_t alg_bytes;       /* sum of algorithmic bytes over those launches            */
    uint64_t items;           /* states / children processed                              */
} hgx_kernel_stat;
int hgx_profile_enable(hgx_ctx* ctx, int on);
/* Counter calibration: one fill kernel writes `bytes` (a multiple of 16), then one read-only
 * kernel (k_read_calibration) reads them once and writes 8 bytes, then a gather kernel
 * (k_read_gather) reads 16 bytes from each of n/4 distinct 128-byte lines scattered over the
 * buffer (n = the buffer's 128-byte lines rounded down to a power of two), then one kernel
 * (k_atomic_calibration) makes one failing compare-and-swap on each of those lines and one add
 * per line into 4,096 counters -- launches of known traffic for scaling FETCH_SIZE / WRITE_SIZE
 * readings: streamed, scattered, and device-scope atomics. No counterpart in the reference. */
int hgx_calibration_read(hgx_ctx* ctx, uint64_t bytes);
/* Copies up to cap entries, sets *n, and clears the accumulators. */
int hgx_profile_read(hgx_ctx* ctx, hgx_kernel_stat* out, uint32_t cap, uint32_t* n);

#ifdef __cplusplus
}
#endif

#endif /* HGX_H */
