// hgx_engine.hpp -- header-only C++ face of the C ABI (include/hgx.h), shaped like the two C++
// surfaces the reference's callers use for this path:
//
//   hgx::Hypergraph + hgx::ParallelEvolutionEngine(Hypergraph*, num_threads, ExecutionMode)
//       add_rule() / evolve(), mirroring hypergraph::ParallelEvolutionEngine
//       (hypergraph/include/hypergraph/parallel_evolution.hpp:741-1041; the constructor :741-742,
//       add_rule :756, evolve :1009 and the multi-root overload :1041, setters :758-935). The
//       results land in the caller-owned Hypergraph and are read through its accessors
//       (hypergraph/include/hypergraph/hypergraph.hpp:555-812: get_state, num_published_states,
//       get_or_compute_canonical_hash, num_canonical_states, get_event, num_events,
//       causal_graph()), so code that drives the reference engine compiles against this header
//       with the include and the namespace changed. The engine also keeps the direct accessors of
//       earlier versions (num_states(), state_hashes(), events(), ...).
//   hgx::EvolveInput / EvolveResult / evolve()   mirroring hg_gpu::EvolveInput / EvolveResult
//       (gpu/include/hg_gpu/evolve.hpp:26-259) for the fields of a full-multiway run, the
//       struct run_gpu_evolution (paclet_source/hg_gpu_backend.cpp:174) builds and reads.
//
// Errors: the reference throws std::runtime_error from evolve() on worker failures
// (parallel_evolution.cpp:243-250) and std::length_error for over-limit rules; this adapter
// throws hgx::Error (a std::runtime_error) carrying the C status code, and length_error the same
// way for rules. A state cap that stops a run early is not an error: truncated() says so, as the
// reference's capacity warning does.
#pragma once

#include <cstdint>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hgx.h"

namespace hgx {

using VertexId = uint32_t;
using StateId = uint32_t;
using EventId = uint32_t;
using EdgeId = uint32_t;
using RuleIndex = uint16_t;
constexpr uint32_t INVALID_ID = 0xFFFFFFFFu;      // hypergraph/include/hypergraph/types.hpp

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// hypergraph::RewriteRule / RuleBuilder / make_rule (pattern.hpp:72-213): LHS and RHS pattern
// edges over small variable indices; RHS-only variables are fresh.
struct RewriteRule {
    uint16_t index = 0;
    std::vector<std::vector<uint8_t>> lhs, rhs;
};

class RuleBuilder {
public:
    explicit RuleBuilder(uint16_t index) { r_.index = index; }
    RuleBuilder& lhs(std::initializer_list<uint8_t> vars) { return add(r_.lhs, std::vector<uint8_t>(vars)); }
    RuleBuilder& rhs(std::initializer_list<uint8_t> vars) { return add(r_.rhs, std::vector<uint8_t>(vars)); }
    RuleBuilder& lhs(const std::vector<uint8_t>& vars) { return add(r_.lhs, vars); }
    RuleBuilder& rhs(const std::vector<uint8_t>& vars) { return add(r_.rhs, vars); }
    RewriteRule build() const { return r_; }

private:
    RuleBuilder& add(std::vector<std::vector<uint8_t>>& side, const std::vector<uint8_t>& vars) {
        if (side.size() >= HGX_MAX_PATTERN_EDGES) throw std::length_error("RuleBuilder: exceeds MAX_PATTERN_EDGES");
        if (vars.empty() || vars.size() > HGX_MAX_ARITY) throw std::length_error("RuleBuilder: arity must be 1..16");
        for (uint8_t v : vars)
            if (v >= HGX_MAX_VARS) throw std::length_error("RuleBuilder: variable index exceeds MAX_VARS");
        side.push_back(vars);
        return *this;
    }
    RewriteRule r_;
};

inline RuleBuilder make_rule(uint16_t index = 0) { return RuleBuilder(index); }

inline std::vector<uint8_t> encode_rules(const std::vector<RewriteRule>& rules) {
    std::vector<uint8_t> p;
    for (const auto& r : rules) {
        p.push_back((uint8_t)r.lhs.size());
        p.push_back((uint8_t)r.rhs.size());
        for (const auto* side : {&r.lhs, &r.rhs})
            for (const auto& e : *side) {
                p.push_back((uint8_t)e.size());
                p.insert(p.end(), e.begin(), e.end());
            }
    }
    return p;
}

namespace detail {
inline void check(int rc, hgx_ctx* ctx) {
    if (rc == HGX_OK) return;
    char buf[512] = {0};
    hgx_last_error(ctx, buf, sizeof(buf));
    throw Error(rc, std::string("hgx: ") + buf);
}
struct CtxDeleter { void operator()(hgx_ctx* c) const { hgx_ctx_destroy(c); } };
struct ResDeleter { void operator()(hgx_result* r) const { hgx_result_release(r); } };
}  // namespace detail

// A device context: create one per process and GPU, share it between engines.
class Device {
public:
    explicit Device(int device = 0, uint64_t budget_bytes = 0) {
        hgx_ctx* c = nullptr;
        detail::check(hgx_ctx_create(device, budget_bytes, &c), nullptr);
        ctx_.reset(c);
    }
    hgx_ctx* get() const { return ctx_.get(); }
    // Incremental match forwarding: 1 on, 0 off, -1 automatic (results never depend on it).
    void set_forwarding(int mode) { detail::check(hgx_ctx_set_option(ctx_.get(), HGX_OPTION_FORWARD, mode), ctx_.get()); }
    std::string info() const {
        char buf[512] = {0};
        hgx_device_info(ctx_.get(), buf, sizeof(buf));
        return buf;
    }

private:
    std::unique_ptr<hgx_ctx, detail::CtxDeleter> ctx_;
};

// hypergraph::State / Event / CausalEdge / BranchialEdge (types.hpp:131-258) for this path.
struct State {
    StateId id = INVALID_ID;
    uint32_t step = 0;
    uint64_t canonical_hash = 0;            // the WL hash
    EventId parent_event = INVALID_ID;      // the event that created it, INVALID_ID for a root
};
struct Event {
    EventId id = INVALID_ID;
    StateId input_state = INVALID_ID, output_state = INVALID_ID;
    uint32_t rule_index = 0, step = 0;
};
// A causal edge: the reference lists one per (producer, consumer, consumed edge) and names the
// edge by its global EdgeId; here the edge is its LHS index in the consumer's match. The
// engine's causal_edges() lists deduplicated event pairs (edge = INVALID_ID there).
struct CausalEdge { EventId producer, consumer; EdgeId edge; bool in_reduction; };
struct BranchialEdge { EventId event1, event2; EdgeId shared_edge = INVALID_ID; };

// The process-wide device context the Hypergraph-shaped engines share (the reference's GPU
// backend likewise keeps one process-lifetime evolver, paclet_source/hg_gpu_backend.cpp:177-183).
inline std::shared_ptr<Device> default_device() {
    static std::shared_ptr<Device> dev = std::make_shared<Device>(0);
    return dev;
}

// hypergraph::CausalGraph's counters and edge lists (causal_graph.cpp:187-386) over one result.
class CausalGraph {
public:
    size_t num_causal_edges() const { return (size_t)(tr_ ? c_.causal_triples_tr : c_.causal_triples); }
    size_t num_causal_event_pairs() const { return (size_t)(tr_ ? c_.causal_pairs_tr : c_.causal_pairs); }
    size_t num_branchial_edges() const { return (size_t)c_.branchial_pairs; }
    // One entry per causal triple (with reduction on: the triples whose pair it keeps). Needs the
    // relations materialised (the engine's default).
    std::vector<CausalEdge> get_causal_edges() const {
        std::vector<CausalEdge> out;
        if (!res_ || !c_.causal_triples) return out;
        std::vector<hgx_triple> t(c_.causal_triples);
        std::vector<uint8_t> k(c_.causal_triples);
        detail::check(hgx_result_causal_triples(res_.get(), t.data(), k.data(), t.size()), nullptr);
        out.reserve(t.size());
        for (size_t i = 0; i < t.size(); ++i)
            if (!tr_ || k[i]) out.push_back(CausalEdge{t[i].producer, t[i].consumer, t[i].lhs_index, k[i] != 0});
        return out;
    }
    std::vector<BranchialEdge> get_branchial_edges() const {
        std::vector<BranchialEdge> out;
        if (!res_ || !c_.branchial_pairs) return out;
        std::vector<hgx_pair> p(c_.branchial_pairs);
        detail::check(hgx_result_branchial(res_.get(), p.data(), p.size()), nullptr);
        out.reserve(p.size());
        for (const auto& x : p) out.push_back(BranchialEdge{x.a, x.b});
        return out;
    }

private:
    friend class Hypergraph;
    std::shared_ptr<hgx_result> res_;
    hgx_counts c_{};
    bool tr_ = false;
};

// hypergraph::Hypergraph as the results of an evolution are read from it (hypergraph.hpp:555-812).
// Owned by the caller; an engine constructed over it fills it on evolve(). Raw states are numbered
// roots first, then the output state of event e is n_roots + e; every raw state is published.
class Hypergraph {
public:
    uint32_t num_states() const { return (uint32_t)c_.raw_states; }
    uint32_t num_published_states() const { return (uint32_t)c_.raw_states; }
    uint32_t num_events() const { return (uint32_t)c_.events; }
    uint32_t num_published_events() const { return (uint32_t)c_.events; }
    size_t num_canonical_states() const { return (size_t)c_.canonical_states; }
    const State& get_state(StateId sid) const {
        load_states();
        static const State invalid{};
        return sid < states_.size() ? states_[sid] : invalid;
    }
    uint64_t get_or_compute_canonical_hash(StateId sid) const { return get_state(sid).canonical_hash; }
    const Event& get_event(EventId eid) const {
        load_events();
        static const Event invalid{};
        return eid < events_.size() ? events_[eid] : invalid;
    }
    const CausalGraph& causal_graph() const { return cg_; }
    const hgx_counts& counts() const { return c_; }

    // Filled by ParallelEvolutionEngine::evolve.
    void assign(std::shared_ptr<hgx_result> res, const hgx_counts& c, bool tr) {
        res_ = res;
        c_ = c;
        cg_.res_ = res;
        cg_.c_ = c;
        cg_.tr_ = tr;
        states_.clear();
        events_.clear();
    }

private:
    void load_states() const {
        if (!res_ || states_.size() == c_.raw_states) return;
        std::vector<uint64_t> h(c_.raw_states);
        std::vector<uint32_t> st(c_.raw_states), cr(c_.raw_states);
        detail::check(hgx_result_states(res_.get(), h.data(), st.data(), cr.data(), h.size()), nullptr);
        states_.resize(h.size());
        for (size_t s = 0; s < h.size(); ++s) states_[s] = State{(StateId)s, st[s], h[s], cr[s]};
    }
    void load_events() const {
        if (!res_ || events_.size() == c_.events || !c_.events) return;
        std::vector<hgx_event> raw(c_.events);
        detail::check(hgx_result_events(res_.get(), raw.data(), raw.size()), nullptr);
        events_.resize(raw.size());
        for (size_t i = 0; i < raw.size(); ++i)
            events_[i] = Event{(EventId)i, raw[i].input_state, raw[i].output_state, raw[i].rule, raw[i].step};
    }
    std::shared_ptr<hgx_result> res_;
    hgx_counts c_{};
    CausalGraph cg_;
    mutable std::vector<State> states_;
    mutable std::vector<Event> events_;
};

// hypergraph::ParallelEvolutionEngine for the full-multiway, WL-hashed path.
class ParallelEvolutionEngine {
public:
    enum class ExecutionMode { Parallel, Serial };   // both run on the device; kept for signature parity

    // The reference's constructor (parallel_evolution.hpp:741-742): results go to *hg.
    explicit ParallelEvolutionEngine(Hypergraph* hg, size_t num_threads = 0,
                                     ExecutionMode mode = ExecutionMode::Parallel)
        : dev_(default_device()), hg_(hg) {
        (void)num_threads;
        (void)mode;
    }
    // On a given device context (several engines may share one); hg may be null.
    explicit ParallelEvolutionEngine(std::shared_ptr<Device> dev, Hypergraph* hg = nullptr)
        : dev_(dev ? dev : default_device()), hg_(hg) {}
    ParallelEvolutionEngine() : dev_(default_device()) {}
    ParallelEvolutionEngine(const ParallelEvolutionEngine&) = delete;
    ParallelEvolutionEngine& operator=(const ParallelEvolutionEngine&) = delete;

    void add_rule(const RewriteRule& rule) {
        if (rule.lhs.empty()) throw std::length_error("add_rule: a rule needs at least one LHS edge");
        rules_.push_back(rule);
    }
    void set_transitive_reduction(bool on) { tr_ = on; }
    // Deterministic cap: stop after the largest step whose cumulative raw-state count fits.
    void set_max_states(size_t max) { cap_ = (uint32_t)max; }
    void set_record_causal(bool on) { rec_causal_ = on; }
    void set_record_branchial(bool on) { rec_branchial_ = on; }
    void set_materialize_relations(bool on) { relations_ = on; }
    void set_keep_state_edges(bool on) { keep_edges_ = on; }

    void evolve(const std::vector<std::vector<VertexId>>& initial_edges, size_t steps) {
        run({initial_edges}, steps, false);
    }
    void evolve(const std::vector<std::vector<std::vector<VertexId>>>& initial_states, size_t steps) {
        run(initial_states, steps, true);
    }

    size_t num_states() const { return (size_t)counts_.raw_states; }
    size_t num_canonical_states() const { return (size_t)counts_.canonical_states; }
    size_t num_events() const { return (size_t)counts_.events; }
    size_t num_causal_edges() const { return (size_t)(tr_ ? counts_.causal_triples_tr : counts_.causal_triples); }
    size_t num_causal_event_pairs() const { return (size_t)(tr_ ? counts_.causal_pairs_tr : counts_.causal_pairs); }
    size_t num_branchial_edges() const { return (size_t)counts_.branchial_pairs; }
    bool truncated() const { return counts_.truncated != 0; }
    uint32_t steps_done() const { return counts_.steps_done; }
    const hgx_counts& counts() const { return counts_; }

    // Per raw state: WL hash (get_or_compute_canonical_hash) and step.
    std::vector<uint64_t> state_hashes() const {
        std::vector<uint64_t> h(counts_.raw_states);
        detail::check(hgx_result_states(res_.get(), h.data(), nullptr, nullptr, h.size()), dev_->get());
        return h;
    }
    std::vector<uint32_t> state_steps() const {
        std::vector<uint32_t> s(counts_.raw_states);
        detail::check(hgx_result_states(res_.get(), nullptr, s.data(), nullptr, s.size()), dev_->get());
        return s;
    }
    std::vector<Event> events() const {
        std::vector<hgx_event> raw(counts_.events);
        if (!raw.empty()) detail::check(hgx_result_events(res_.get(), raw.data(), raw.size()), dev_->get());
        std::vector<Event> out(raw.size());
        for (size_t i = 0; i < raw.size(); ++i)
            out[i] = Event{(EventId)i, raw[i].input_state, raw[i].output_state, raw[i].rule, raw[i].step};
        return out;
    }
    // Deduplicated (producer, consumer) pairs with their transitive-reduction tag.
    std::vector<CausalEdge> causal_edges() const {
        std::vector<hgx_pair> p(counts_.causal_pairs);
        std::vector<uint8_t> t(counts_.causal_pairs);
        detail::check(hgx_result_causal(res_.get(), p.data(), t.data(), p.size()), dev_->get());
        std::vector<CausalEdge> out(p.size());
        for (size_t i = 0; i < p.size(); ++i) out[i] = CausalEdge{p[i].a, p[i].b, INVALID_ID, t[i] != 0};
        return out;
    }
    std::vector<BranchialEdge> branchial_edges() const {
        std::vector<hgx_pair> p(counts_.branchial_pairs);
        if (!p.empty()) detail::check(hgx_result_branchial(res_.get(), p.data(), p.size()), dev_->get());
        std::vector<BranchialEdge> out(p.size());
        for (size_t i = 0; i < p.size(); ++i) out[i] = BranchialEdge{p[i].a, p[i].b};
        return out;
    }
    // Edge lists of every raw state (needs set_keep_state_edges(true) before evolve).
    std::vector<std::vector<std::vector<VertexId>>> state_edges() const {
        uint64_t ne = 0, nv = 0;
        detail::check(hgx_result_state_edges(res_.get(), nullptr, nullptr, nullptr, &ne, &nv), dev_->get());
        std::vector<uint64_t> so(counts_.raw_states + 1);
        std::vector<uint32_t> eo(ne + 1), vv(nv ? nv : 1);
        detail::check(hgx_result_state_edges(res_.get(), so.data(), eo.data(), vv.data(), nullptr, nullptr),
                      dev_->get());
        std::vector<std::vector<std::vector<VertexId>>> out(counts_.raw_states);
        for (size_t s = 0; s < out.size(); ++s)
            for (uint64_t e = so[s]; e < so[s + 1]; ++e)
                out[s].emplace_back(vv.begin() + eo[e], vv.begin() + eo[e + 1]);
        return out;
    }

private:
    void run(const std::vector<std::vector<std::vector<VertexId>>>& roots, size_t steps, bool multi) {
        if (rules_.empty()) return;                    // the reference returns without work
        if (steps > HGX_MAX_STEPS) throw Error(HGX_ERR_UNSUPPORTED, "steps above HGX_MAX_STEPS");
        const std::vector<uint8_t> prog = encode_rules(rules_);
        std::vector<uint32_t> eoff{0}, verts, roff{0};
        for (const auto& st : roots) {
            for (const auto& e : st) {
                verts.insert(verts.end(), e.begin(), e.end());
                eoff.push_back((uint32_t)verts.size());
            }
            roff.push_back((uint32_t)(eoff.size() - 1));
        }
        if (verts.empty()) verts.push_back(0);
        hgx_run_desc d{};
        d.rules = prog.data();
        d.rules_len = (uint32_t)prog.size();
        d.n_rules = (uint32_t)rules_.size();
        d.edge_offsets = eoff.data();
        d.vertices = verts.data();
        d.n_edges = (uint32_t)(eoff.size() - 1);
        d.root_offsets = multi ? roff.data() : nullptr;
        d.n_roots = multi ? (uint32_t)roots.size() : 1;
        d.steps = (uint32_t)steps;
        d.state_cap = cap_;
        d.flags = (rec_causal_ ? HGX_REC_CAUSAL : 0u) | (rec_branchial_ ? HGX_REC_BRANCHIAL : 0u) |
                  (relations_ ? HGX_REC_RELATIONS : 0u) | (keep_edges_ ? HGX_REC_STATES : 0u) |
                  (tr_ ? HGX_OPT_TR : 0u);
        hgx_result* r = nullptr;
        detail::check(hgx_evolve_batch(dev_->get(), &d, 1, &r), dev_->get());
        res_.reset(r, detail::ResDeleter());
        detail::check(hgx_result_counts(r, &counts_), dev_->get());
        if (hg_) hg_->assign(res_, counts_, tr_);
    }

    std::shared_ptr<Device> dev_;
    Hypergraph* hg_ = nullptr;
    std::vector<RewriteRule> rules_;
    bool tr_ = false, rec_causal_ = true, rec_branchial_ = true, relations_ = true, keep_edges_ = false;
    uint32_t cap_ = 0;
    hgx_counts counts_{};
    std::shared_ptr<hgx_result> res_;
};

// hg_gpu::EvolveInput / EvolveResult for a full-multiway run (the fields run_gpu_evolution sets
// and reads on this path). Defaults follow the GPU surface: transitive reduction on.
struct EvolveInput {
    std::vector<RewriteRule> rules;
    std::vector<std::vector<VertexId>> initial_state;
    std::vector<std::vector<std::vector<VertexId>>> initial_states;   // multi-root when non-empty
    uint32_t num_steps = 0;
    bool transitive_reduction = true;
    bool materialize_relations = false;
    bool record_causal = true, record_branchial = true;
    bool edge_identity = false;          // keep states' edges for readback
    uint32_t max_states = 0;             // deterministic step-prefix cap
};

struct CanonicalState {
    StateId id;
    uint64_t canonical_hash;
    std::vector<std::vector<VertexId>> edges;   // filled when edge_identity
};

struct EvolveResult {
    std::vector<CanonicalState> states;
    std::vector<Event> events;
    std::vector<CausalEdge> causal_edges;
    std::vector<BranchialEdge> branchial_edges;
    hgx_counts counts{};
    std::vector<std::string> warnings;   // e.g. the state cap stopped the run
};

inline EvolveResult evolve(const EvolveInput& in, std::shared_ptr<Device> dev = nullptr) {
    ParallelEvolutionEngine eng(dev);
    for (const auto& r : in.rules) eng.add_rule(r);
    eng.set_transitive_reduction(in.transitive_reduction);
    eng.set_record_causal(in.record_causal);
    eng.set_record_branchial(in.record_branchial);
    eng.set_materialize_relations(in.materialize_relations);
    eng.set_keep_state_edges(in.edge_identity);
    eng.set_max_states(in.max_states);
    if (!in.initial_states.empty()) eng.evolve(in.initial_states, in.num_steps);
    else eng.evolve(in.initial_state, in.num_steps);
    EvolveResult out;
    out.counts = eng.counts();
    const auto h = eng.state_hashes();
    std::vector<std::vector<std::vector<VertexId>>> edges;
    if (in.edge_identity) edges = eng.state_edges();
    out.states.resize(h.size());
    for (size_t s = 0; s < h.size(); ++s) {
        out.states[s].id = (StateId)s;
        out.states[s].canonical_hash = h[s];
        if (in.edge_identity) out.states[s].edges = std::move(edges[s]);
    }
    out.events = eng.events();
    if (in.materialize_relations) {
        if (in.record_causal) out.causal_edges = eng.causal_edges();
        if (in.record_branchial) out.branchial_edges = eng.branchial_edges();
    }
    if (eng.truncated())
        out.warnings.push_back("state cap reached: the evolution stopped after step " +
                               std::to_string(eng.steps_done()));
    return out;
}

}  // namespace hgx
