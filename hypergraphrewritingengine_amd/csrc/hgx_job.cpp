// hgx_job.cpp -- the paclet's job envelope: a WXF job in, a WXF result out, evolved on the device.
//
// The counterpart of run_rewriting_core (paclet_source/hg_core.hpp:21-22) on the multiway path
// this library serves: the job keys and option semantics of hypergraph_ffi.cpp:136-376 (parse)
// and the result layout of hg_gpu_backend.cpp:362-640 (States, Events, CausalEdges,
// BranchialEdges, Num*). It is written over the library's own C ABI (hgx_evolve_batch and the
// hgx_result_* readbacks).
//
// Served: multiway evolution with CanonicalizeStates -> None, Automatic or Full (the GPU
// backend's grouping, hg_gpu_backend.cpp:269-310: None gives every state its own class, Automatic
// groups states of equal content, Full groups by the exact IR hash computed on the host and emits
// each class's representative relabelled to its canonical form; CanonicalHash is the IR hash in
// every mode), CanonicalizeEvents -> None, Full, Automatic, Positional or a list of signature
// keys (events identified by the signature of common/include/hgcommon/event_core.hpp:60-84 over
// exact state hashes and canonical edge ranks, as Hypergraph::create_event does,
// hypergraph/src/hypergraph.cpp:448-499), ExploreFromCanonicalStatesOnly and
// QuotientInitialStates (quotient exploration on the device, HGX_OPT_QUOTIENT), any rule set, one
// or more initial states, causal edges with or without transitive reduction, branchial edges,
// RequestedData selection, and the sampling and capping options (TransitionRate, RuleWeights,
// MatchesPerStateRule, ExplorationProbability, MaxStatesPerStep, MaxSuccessorStatesPerParent,
// RandomSeed; decided by the engine, include/hgx.h hgx_sampling). With exact states and quotient
// exploration or Automatic event identity the job takes the reference's reconstruction route:
// events are identified by the recorded class expansion they replay, and with quotient
// exploration NumEvents / NumCausalEdges / NumBranchialEdges are the raw unfolding's
// (hgx_result_unfold). Options that change what is explored and are not served (sessions,
// genesis events, graph properties, quotient exploration under content identity) are refused
// with an error rather than ignored, so a caller never receives a different evolution than it
// asked for; options with no effect on the result (ShowProgress, ...) are accepted and reported
// under "Warnings" as skipped, with the reference's Kind/Count/Context schema.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hgx_gpu_backend.hpp"
#include "hgx_runtime.hpp"
#include "hgx_wxf.hpp"

namespace hgx {
namespace job {

using wxf::Value;

// The sampling and capping options (ffi_job.hpp:48-72, as the GPU backend takes them,
// hg_gpu_backend.cpp:152-160): per-transition thinning, per-rule weights, a per-(state, rule) cap,
// a per-state exploration coin and the two arrival-order bounds on the states expanded. The
// engine takes the decisions (hgx_sampling).
struct Sampling {
    double explore_p = 1.0;                 // ExplorationProbability
    double rate = 1.0;                      // TransitionRate
    std::vector<double> weights;            // RuleWeights
    uint64_t seed = 0;                      // RandomSeed
    uint64_t per_step = 0;                  // MaxStatesPerStep (or MatchesPerStep with UniformRandom)
    uint64_t per_parent = 0;                // MaxSuccessorStatesPerParent
    uint64_t per_state_rule = 0;            // MatchesPerStateRule
    // the transition draw is taken when the rate is below 1 or weights are given (rewrite.cu:270)
    bool draws() const { return rate < 1.0 || !weights.empty(); }
    bool transitions() const { return draws() || per_state_rule != 0; }
    bool active() const { return transitions() || explore_p < 1.0 || per_step || per_parent; }
};

struct Job {
    std::vector<std::vector<std::vector<int64_t>>> states;          // initial states as given
    std::vector<std::pair<std::vector<std::vector<int64_t>>, std::vector<std::vector<int64_t>>>> rules;
    int64_t steps = 1;
    bool tr = true;                          // CausalTransitiveReduction (FFI default True)
    bool inc_states = true, inc_events = true, inc_events_min = false, inc_causal = true,
         inc_branchial = true, inc_nstates = true, inc_nevents = true, inc_ncausal = true,
         inc_nbranchial = true, inc_hashes = false;
    bool full = false;                       // CanonicalizeStates -> Full
    bool automatic = false;                  // CanonicalizeStates -> Automatic (content identity)
    uint32_t ev_keys = 0;                    // CanonicalizeEvents: event signature keys
    bool positional = false;                 // CanonicalizeEvents -> Positional (raw labelling)
    bool quotient = false;                   // ExploreFromCanonicalStatesOnly
    bool quotient_roots = false;             // QuotientInitialStates
    Sampling smp;
    std::vector<std::pair<std::string, std::string>> warnings;   // (kind, context)
};

struct JobError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

static int64_t as_int(const Value& v, const char* what) {
    if (v.kind != Value::Int) throw JobError(std::string(what) + " must be an integer");
    return v.i;
}

static double as_real(const Value& v, const char* what) {
    if (v.kind == Value::Real) return v.r;
    if (v.kind == Value::Int) return (double)v.i;
    throw JobError(std::string(what) + " must be a number");
}

static uint64_t as_count(const Value& v, const char* what) {
    const int64_t x = as_int(v, what);
    if (x < 0) throw JobError(std::string(what) + " must be non-negative");
    return (uint64_t)x;
}

static bool as_bool(const Value& v, const char* what) {
    if (v.is_symbol("True")) return true;
    if (v.is_symbol("False")) return false;
    throw JobError(std::string(what) + " must be True or False");
}

static std::string as_name(const Value& v) {
    if (v.kind == Value::String || v.kind == Value::Symbol) return v.s;
    throw JobError("expected a string or symbol");
}

static const Value& need_list(const Value& v, const char* what) {
    if (!v.is_list()) throw JobError(std::string(what) + " must be a list");
    return v;
}

// {{v...}, ...}
static std::vector<std::vector<int64_t>> edges_of(const Value& v, const char* what) {
    need_list(v, what);
    std::vector<std::vector<int64_t>> out;
    for (size_t k = 0; k < v.nargs(); ++k) {
        const Value& e = need_list(v.arg(k), what);
        std::vector<int64_t> ed;
        for (size_t j = 0; j < e.nargs(); ++j) ed.push_back(as_int(e.arg(j), what));
        out.push_back(std::move(ed));
    }
    return out;
}

static Job parse(const Value& root) {
    if (root.kind != Value::Assoc) throw JobError("the job must be an Association");
    Job j;
    bool have_rules = false;
    uint64_t max_per_step = 0, matches_per_step = 0;
    bool uniform = false;
    for (const auto& kv : root.rules) {
        const std::string key = as_name(kv.first);
        const Value& val = kv.second;
        if (key == "InitialStates") {
            need_list(val, "InitialStates");
            for (size_t k = 0; k < val.nargs(); ++k) j.states.push_back(edges_of(val.arg(k), "InitialStates"));
        } else if (key == "InitialEdges") {
            j.states.push_back(edges_of(val, "InitialEdges"));
        } else if (key == "Rules") {
            if (val.kind != Value::Assoc) throw JobError("Rules must be an Association of name -> Rule[lhs, rhs]");
            for (const auto& r : val.rules) {
                const Value& rv = r.second;
                if (!rv.is_head("Rule") || rv.nargs() != 2) throw JobError("Expected Rule[lhs, rhs]");
                j.rules.emplace_back(edges_of(rv.arg(0), "rule LHS"), edges_of(rv.arg(1), "rule RHS"));
            }
            have_rules = true;
        } else if (key == "Steps") {
            j.steps = as_int(val, "Steps");
            if (j.steps < 0) throw JobError("Steps must be non-negative, got " + std::to_string(j.steps));
        } else if (key == "Options") {
            if (val.kind != Value::Assoc) throw JobError("Options must be an Association");
            for (const auto& o : val.rules) {
                const std::string ok = as_name(o.first);
                const Value& ov = o.second;
                auto refuse = [&](const std::string& why) {
                    throw JobError("option " + ok + ": " + why + " (this backend evolves the full multiway system)");
                };
                if (ok == "CanonicalizeStates") {
                    const std::string m = as_name(ov);
                    j.full = m == "Full" || m == "True";
                    j.automatic = m == "Automatic";
                    if (!j.full && !j.automatic && m != "None" && m != "False")
                        refuse("unknown state canonicalization " + m);
                } else if (ok == "CanonicalizeEvents") {
                    // event_core.hpp:26-45; the FFI's names (hypergraph_ffi.cpp:236-262)
                    constexpr uint32_t In = 1, Out = 2, Step = 4, Rule = 8, Cons = 16, Prod = 32;
                    if (ov.is_head("List")) {
                        j.ev_keys = 0;
                        for (size_t k = 0; k < ov.nargs(); ++k) {
                            const std::string key = as_name(ov.arg(k));
                            if (key == "InputState") j.ev_keys |= In;
                            else if (key == "OutputState") j.ev_keys |= Out;
                            else if (key == "Step") j.ev_keys |= Step;
                            else if (key == "Rule") j.ev_keys |= Rule;
                            else if (key == "ConsumedEdges") j.ev_keys |= Cons;
                            else if (key == "ProducedEdges") j.ev_keys |= Prod;
                        }
                    } else {
                        const std::string m = as_name(ov);
                        if (m == "None") j.ev_keys = 0;
                        else if (m == "Full") j.ev_keys = In | Out;
                        else if (m == "Automatic" || m == "Positional") j.ev_keys = In | Out | Step | Cons | Prod;
                        else refuse("unknown event canonicalization " + m);
                        j.positional = m == "Positional";
                    }
                } else if (ok == "CausalTransitiveReduction") {
                    j.tr = as_bool(ov, ok.c_str());
                } else if (ok == "IncludeCanonicalHashes") {
                    j.inc_hashes = as_bool(ov, ok.c_str());
                } else if (ok == "ExploreFromCanonicalStatesOnly") {
                    j.quotient = as_bool(ov, ok.c_str());
                } else if (ok == "QuotientInitialStates") {
                    j.quotient_roots = as_bool(ov, ok.c_str());
                } else if (ok == "ShowGenesisEvents") {
                    if (as_bool(ov, ok.c_str())) refuse("True is not served");
                } else if (ok == "MaxSuccessorStatesPerParent") {
                    j.smp.per_parent = as_count(ov, ok.c_str());
                } else if (ok == "MaxStatesPerStep") {
                    max_per_step = as_count(ov, ok.c_str());
                } else if (ok == "MatchesPerStep") {
                    matches_per_step = as_count(ov, ok.c_str());
                } else if (ok == "UniformRandom") {
                    uniform = as_bool(ov, ok.c_str());
                } else if (ok == "MatchesPerStateRule") {
                    j.smp.per_state_rule = as_count(ov, ok.c_str());
                } else if (ok == "RandomSeed") {
                    j.smp.seed = (uint64_t)as_int(ov, ok.c_str());
                } else if (ok == "ExplorationProbability") {
                    j.smp.explore_p = as_real(ov, ok.c_str());
                } else if (ok == "TransitionRate") {
                    j.smp.rate = as_real(ov, ok.c_str());
                } else if (ok == "RuleWeights") {
                    need_list(ov, "RuleWeights");
                    j.smp.weights.clear();
                    for (size_t k = 0; k < ov.nargs(); ++k) j.smp.weights.push_back(as_real(ov.arg(k), "RuleWeights"));
                } else if (ok == "GraphProperties") {
                    if (!ov.is_list() || ov.nargs() != 0) refuse("graph properties are not served");
                } else if (ok == "BranchialStep") {
                    if (as_int(ov, ok.c_str()) != 0) refuse("per-step branchial selection is not served");
                } else if (ok == "RequestedData") {
                    need_list(ov, "RequestedData");
                    j.inc_states = j.inc_events = j.inc_events_min = j.inc_causal = j.inc_branchial = false;
                    j.inc_nstates = j.inc_nevents = j.inc_ncausal = j.inc_nbranchial = false;
                    for (size_t k = 0; k < ov.nargs(); ++k) {
                        const std::string c = as_name(ov.arg(k));
                        if (c == "States") j.inc_states = true;
                        else if (c == "Events") j.inc_events = true;
                        else if (c == "EventsMinimal") j.inc_events_min = true;
                        else if (c == "CausalEdges") j.inc_causal = true;
                        else if (c == "BranchialEdges") j.inc_branchial = true;
                        else if (c == "NumStates") j.inc_nstates = true;
                        else if (c == "NumEvents") j.inc_nevents = true;
                        else if (c == "NumCausalEdges") j.inc_ncausal = true;
                        else if (c == "NumBranchialEdges") j.inc_nbranchial = true;
                        else refuse("component " + c + " is not served");
                    }
                } else {
                    // ShowProgress, EdgeDeduplication and unknown keys: nothing here depends on them
                    j.warnings.emplace_back("OptionSkipped", ok);
                }
            }
        } else if (key == "Op") {
            const std::string op = as_name(val);
            if (!op.empty() && op != "Evolve") throw JobError("Op " + op + ": sessions are not served");
        } else if (key == "Session" || key == "From" || key == "Delivery") {
            throw JobError(key + ": sessions are not served");
        } else {
            j.warnings.emplace_back("KeySkipped", key);
        }
    }
    if (!have_rules || j.rules.empty()) throw JobError("the job has no Rules");
    // MatchesPerStep is MaxStatesPerStep under another name when UniformRandom selects it
    // (hypergraph_ffi.cpp:378-385)
    j.smp.per_step = uniform && matches_per_step ? matches_per_step : max_per_step;
    if (j.quotient && j.automatic)
        throw JobError("option ExploreFromCanonicalStatesOnly: quotient exploration under content identity "
                       "(CanonicalizeStates -> Automatic) is not served");
    return j;
}

// Rule program bytes (per rule: n_lhs, n_rhs, then (arity, vars...) per LHS then RHS edge), with
// the FFI's checks: variables 0..31, arity <= 16, at most 16 edges per side, empty edges dropped,
// a non-empty LHS (hypergraph_ffi.cpp:520-590).
static std::vector<uint8_t> program(const Job& j) {
    std::vector<uint8_t> out;
    for (size_t r = 0; r < j.rules.size(); ++r) {
        auto side = [&](const std::vector<std::vector<int64_t>>& edges, const char* name,
                        std::vector<std::vector<uint8_t>>& dst) {
            for (size_t e = 0; e < edges.size(); ++e) {
                if (edges[e].empty()) continue;
                if (dst.size() >= HGX_MAX_PATTERN_EDGES)
                    throw JobError("rule " + std::to_string(r) + " " + name + " has more than 16 edges");
                if (edges[e].size() > HGX_MAX_ARITY)
                    throw JobError("rule " + std::to_string(r) + " " + name + " edge " + std::to_string(e) +
                                   " has arity above 16");
                std::vector<uint8_t> ed;
                for (int64_t v : edges[e]) {
                    if (v < 0) throw JobError("rule " + std::to_string(r) + " " + name + " edge " +
                                              std::to_string(e) + " has a negative pattern variable");
                    if (v >= HGX_MAX_VARS) throw JobError("rule " + std::to_string(r) + " " + name +
                                                          " uses pattern variable " + std::to_string(v) +
                                                          ", but the maximum is 31");
                    ed.push_back((uint8_t)v);
                }
                dst.push_back(std::move(ed));
            }
        };
        std::vector<std::vector<uint8_t>> lhs, rhs;
        side(j.rules[r].first, "LHS", lhs);
        side(j.rules[r].second, "RHS", rhs);
        if (lhs.empty()) throw JobError("rule " + std::to_string(r) + " has an empty left-hand side");
        out.push_back((uint8_t)lhs.size());
        out.push_back((uint8_t)rhs.size());
        for (auto* sd : {&lhs, &rhs})
            for (auto& ed : *sd) {
                out.push_back((uint8_t)ed.size());
                out.insert(out.end(), ed.begin(), ed.end());
            }
    }
    return out;
}

// Initial states as the FFI takes them: negative vertices and empty edges dropped, empty states
// dropped (hypergraph_ffi.cpp:598-636). Each state's int64 labels are renumbered densely by
// first appearance here, as the FFI does, so any non-negative int64 label is taken; the library
// renumbers the same way, so the dense labels reach it unchanged.
// The device numbers each initial state's vertices 0.. by first appearance; `orig` (when given)
// receives, per kept initial state, the job's label of each device label.
static void roots(const Job& j, std::vector<uint32_t>& eoff, std::vector<uint32_t>& verts,
                  std::vector<uint32_t>& roff, std::vector<std::vector<int64_t>>* orig = nullptr) {
    eoff.assign(1, 0);
    roff.assign(1, 0);
    std::unordered_map<int64_t, uint32_t> label;
    std::vector<int64_t> names;
    for (const auto& st : j.states) {
        uint32_t kept = 0;
        label.clear();
        names.clear();
        for (const auto& e : st) {
            uint32_t n = 0;
            for (int64_t v : e) {
                if (v < 0) continue;
                const auto it = label.emplace(v, (uint32_t)label.size());
                if (it.second) names.push_back(v);
                verts.push_back(it.first->second);
                ++n;
            }
            if (!n) continue;
            if (n > HGX_MAX_ARITY) throw JobError("initial edge with arity above 16");
            eoff.push_back((uint32_t)verts.size());
            ++kept;
        }
        if (kept) {
            roff.push_back((uint32_t)(eoff.size() - 1));
            if (orig) orig->push_back(names);
        }
    }
}

static Value ilist(const std::vector<int64_t>& xs) {
    Value v = Value::list();
    v.args.reserve(xs.size() + 1);
    for (int64_t x : xs) v.push(Value::integer(x));
    return v;
}

// Edges a rule's right-hand side produces (empty pattern edges make none).
static size_t rhs_edges(const Job& j, uint32_t rule) {
    const auto& rhs = j.rules[rule].second;
    return rhs.size() - (size_t)std::count_if(rhs.begin(), rhs.end(), [](const std::vector<int64_t>& ed) { return ed.empty(); });
}

struct Run {
    hgx_counts c{};
    std::vector<uint64_t> hash;
    std::vector<uint32_t> step, creator;
    std::vector<hgx_event> events;
    std::vector<hgx_pair> causal, branchial;
    std::vector<uint8_t> causal_tr;
    std::vector<uint64_t> soff;
    std::vector<uint32_t> eoff, verts, coff;
    std::vector<uint16_t> cpos;
};

static void check(int rc, hgx_ctx* ctx, const char* what) {
    if (rc == HGX_OK) return;
    char buf[512] = {0};
    hgx_last_error(ctx, buf, sizeof(buf));
    throw JobError(std::string(what) + " failed (status " + std::to_string(rc) + "): " + buf);
}

// CanonicalizeStates -> Automatic: states are the same when their edges, in order, hold the
// same vertices (hg_gpu_backend.cpp:269-310 groups by the content key the evolution deduplicated
// with, common/include/hgcommon/wl_core.hpp:20-48; reference/oracle_corpus.hpp:85-96 states the
// same identity as a string). Content compares vertex NAMES, so it depends on how fresh vertices
// are named: the reference takes every fresh vertex from one global counter, so a vertex a rule
// creates is a name no other event ever creates. The device names fresh vertices per lineage
// (siblings may reuse a number), so the comparison here runs on global names instead: an initial
// vertex is named by the job's own label, and a vertex that first appears in state s (created by
// the event that produced s) gets a name of its own; a kept vertex keeps its parent's name.
// Any naming with those equalities gives the reference's classes. The representative of a class
// is its first state in id order.
static void content_classes(const Run& R, const std::vector<std::vector<int64_t>>& orig, uint32_t n_roots,
                            std::vector<int64_t>& rep, std::vector<uint64_t>& reps) {
    const uint64_t ns = R.soff.size() - 1;
    std::vector<uint64_t> name(R.verts.size());           // global name of every vertex slot
    uint64_t fresh = 0;
    std::unordered_map<uint32_t, uint64_t> of_parent;
    for (uint64_t s = 0; s < ns; ++s) {
        const uint64_t x0 = R.eoff[R.soff[s]], x1 = R.eoff[R.soff[s + 1]];
        if (s < n_roots) {
            for (uint64_t x = x0; x < x1; ++x) name[x] = (uint64_t)orig[s][R.verts[x]] << 1;
            continue;
        }
        const uint32_t e = R.creator[s];
        if (e >= R.events.size()) throw JobError("a state without its creating event");
        const uint32_t p = R.events[e].input_state;
        of_parent.clear();
        for (uint64_t x = R.eoff[R.soff[p]]; x < R.eoff[R.soff[p + 1]]; ++x) of_parent.emplace(R.verts[x], name[x]);
        std::unordered_map<uint32_t, uint64_t> made;
        for (uint64_t x = x0; x < x1; ++x) {
            const auto it = of_parent.find(R.verts[x]);
            if (it != of_parent.end()) { name[x] = it->second; continue; }
            const auto m = made.emplace(R.verts[x], (fresh << 1) | 1u);
            if (m.second) ++fresh;
            name[x] = m.first->second;
        }
    }
    // class key: edge count, then per edge its arity and names; hashed, then compared exactly
    auto key_of = [&](uint64_t s, std::vector<uint64_t>& k) {
        k.clear();
        k.push_back(R.soff[s + 1] - R.soff[s]);
        for (uint64_t ed = R.soff[s]; ed < R.soff[s + 1]; ++ed) {
            k.push_back(R.eoff[ed + 1] - R.eoff[ed]);
            for (uint64_t x = R.eoff[ed]; x < R.eoff[ed + 1]; ++x) k.push_back(name[x]);
        }
    };
    auto hash_of = [](const std::vector<uint64_t>& k) {
        uint64_t h = 0x9ae16a3b2f90404fULL;
        for (uint64_t v : k) {
            h ^= v + 0x632be59bd9b4e019ULL + (h << 6) + (h >> 2);
            h *= 0xff51afd7ed558ccdULL;
        }
        return h;
    };
    std::unordered_multimap<uint64_t, uint32_t> seen;
    std::vector<uint64_t> k, k2;
    for (uint64_t s = 0; s < ns; ++s) {
        key_of(s, k);
        const uint64_t h = hash_of(k);
        int64_t found = -1;
        for (auto [it, end] = seen.equal_range(h); it != end; ++it) {
            key_of(it->second, k2);
            if (k2 == k) { found = it->second; break; }
        }
        if (found < 0) {
            seen.emplace(h, (uint32_t)s);
            rep[s] = (int64_t)s;
            reps.push_back(s);
        } else {
            rep[s] = found;
        }
    }
}

static Value evolve(hgx_ctx* ctx, const Job& j) {
    const std::vector<uint8_t> prog = program(j);
    std::vector<uint32_t> eoff, verts, roff;
    std::vector<std::vector<int64_t>> orig;
    roots(j, eoff, verts, roff, &orig);
    const uint32_t n_roots = (uint32_t)roff.size() - 1;
    if (n_roots == 0) throw JobError("the job has no non-empty initial state");
    if (j.steps > HGX_MAX_STEPS) throw JobError("Steps above " + std::to_string(HGX_MAX_STEPS));
    uint32_t flags = HGX_REC_CAUSAL | HGX_REC_BRANCHIAL | HGX_REC_RELATIONS;
    constexpr uint32_t In = 1, Out = 2, Step = 4, Cons = 16, Prod = 32;
    // The reconstruction route (ParallelEvolutionEngine::configure_identity_and_quotient,
    // parallel_evolution.cpp:1446-1545; the device's qc_route, gpu/src/evolve.cu:183-198): exact
    // states, and quotient exploration or Automatic event identity. There an event is a recorded
    // (class, match) expansion and the relations are the raw unfolding's.
    const bool auto_keys = !j.positional && j.ev_keys == (In | Out | Step | Cons | Prod);
    const bool wants_route = j.quotient || auto_keys;
    const bool route = wants_route && j.full && j.steps > 0;
    std::vector<std::pair<std::string, std::string>> warns = j.warnings;
    if (wants_route && !j.full)
        warns.emplace_back("Engine", "quotient exploration and Automatic event identity are defined over exact "
                                     "classes, which only CanonicalizeStates -> Full computes: every state is "
                                     "explored and events are identified by their own states' labelling");
    // a key set whose signature would depend on how a class's slots are numbered
    if (route && (((j.ev_keys & Cons) && !(j.ev_keys & In)) || ((j.ev_keys & (Prod | Step)) && !(j.ev_keys & Out))))
        throw JobError("option CanonicalizeEvents: on the reconstruction route (exact states with quotient "
                       "exploration or Automatic identity) ConsumedEdges needs InputState, and ProducedEdges "
                       "and Step need OutputState");
    // state edges: for the output, and for the IR hash (Full's grouping, CanonicalHash)
    const bool need_ir = j.full || (j.inc_hashes && j.inc_states) || j.ev_keys;
    const bool need_rank = (j.ev_keys & (Cons | Prod)) != 0 || route;
    if (j.inc_states || j.inc_events || need_ir || j.automatic || route) flags |= HGX_REC_STATES;
    if (j.tr) flags |= HGX_OPT_TR;
    if (j.quotient && j.full) flags |= HGX_OPT_QUOTIENT | (j.quotient_roots ? HGX_OPT_QUOTIENT_ROOTS : 0u);
    hgx_run_desc d{};
    d.rules = prog.data(); d.rules_len = (uint32_t)prog.size(); d.n_rules = (uint32_t)j.rules.size();
    d.edge_offsets = eoff.data(); d.vertices = verts.empty() ? nullptr : verts.data();
    d.n_edges = (uint32_t)(eoff.size() - 1);
    d.root_offsets = roff.data(); d.n_roots = n_roots;
    d.steps = (uint32_t)j.steps; d.state_cap = 0; d.flags = flags;
    // sampling and caps: decided by the engine generation by generation (include/hgx.h)
    hgx_sampling sp{};
    if (j.smp.active()) {
        sp.transition_rate = j.smp.rate;
        sp.rule_weights = j.smp.weights.empty() ? nullptr : j.smp.weights.data();
        sp.n_rule_weights = (uint32_t)j.smp.weights.size();
        sp.seed = j.smp.seed;
        sp.exploration_probability = j.smp.explore_p;
        sp.max_states_per_step = j.smp.per_step;
        sp.max_successor_states_per_parent = j.smp.per_parent;
        sp.matches_per_state_rule = j.smp.per_state_rule;
        d.sampling = &sp;
    }
    hgx_result* res = nullptr;
    check(hgx_evolve_batch(ctx, &d, 1, &res), ctx, "evolution");
    Run R;
    struct Release { hgx_result* r; ~Release() { hgx_result_release(r); } } rel{res};
    check(hgx_result_counts(res, &R.c), ctx, "counts");
    // a generation beyond the HBM budget: the result holds the ones before it, and says so
    {
        char why[512] = {0};
        if (hgx_result_status(res, why, sizeof(why)) == HGX_ERR_CAPACITY) warns.emplace_back("CapacityExceeded", why);
    }
    const uint64_t ns = R.c.raw_states, ne = R.c.events;
    R.hash.resize(ns); R.step.resize(ns); R.creator.resize(ns);
    check(hgx_result_states(res, R.hash.data(), R.step.data(), R.creator.data(), ns), ctx, "states");
    R.events.resize(ne);
    if (ne) check(hgx_result_events(res, R.events.data(), ne), ctx, "events");
    R.causal.resize(R.c.causal_pairs); R.causal_tr.resize(R.c.causal_pairs);
    if (R.c.causal_pairs) check(hgx_result_causal(res, R.causal.data(), R.causal_tr.data(), R.c.causal_pairs), ctx, "causal");
    R.branchial.resize(R.c.branchial_pairs);
    if (R.c.branchial_pairs) check(hgx_result_branchial(res, R.branchial.data(), R.c.branchial_pairs), ctx, "branchial");
    if (flags & HGX_REC_STATES) {
        uint64_t te = 0, tv = 0;
        check(hgx_result_state_edges(res, nullptr, nullptr, nullptr, &te, &tv), ctx, "state edges");
        R.soff.resize(ns + 1); R.eoff.resize(te + 1); R.verts.resize(tv);
        check(hgx_result_state_edges(res, R.soff.data(), R.eoff.data(), R.verts.data(), nullptr, nullptr), ctx,
              "state edges");
        R.coff.resize(ne + 1);
        check(hgx_result_consumed(res, R.coff.data(), nullptr, 0), ctx, "consumed edges");
        R.cpos.resize(R.coff[ne]);
        check(hgx_result_consumed(res, R.coff.data(), R.cpos.data(), R.cpos.size()), ctx, "consumed edges");
    }

    // Exact (IR) hash of every state, on the host, when Full groups by it or CanonicalHash is
    // asked for. Classes: None -> every state its own; Full -> equal IR hash, represented by the
    // first state in id order (hg_gpu_backend.cpp:278-310).
    std::vector<uint64_t> ir(need_ir ? ns : 0);
    std::vector<uint32_t> rank(need_rank ? R.soff[ns] : 0);
    if (need_ir && ns)
        check(hgx_ir_rank_batch(ctx, R.soff.data(), (uint32_t)ns, R.eoff.data(), R.verts.empty() ? nullptr : R.verts.data(),
                                ir.data(), need_rank ? rank.data() : nullptr), ctx, "canonical hashes");
    std::vector<int64_t> rep(ns);
    std::vector<uint64_t> reps;                               // class representatives, in id order
    if (j.automatic) {
        content_classes(R, orig, n_roots, rep, reps);
    } else if (j.full) {
        std::unordered_map<uint64_t, uint32_t> rep_of_hash;
        for (uint64_t s = 0; s < ns; ++s) {
            auto it = rep_of_hash.emplace(ir[s], (uint32_t)s);
            rep[s] = it.first->second;
            if (it.second) reps.push_back(s);
        }
    } else {
        for (uint64_t s = 0; s < ns; ++s) { rep[s] = (int64_t)s; reps.push_back(s); }
    }

    // Event identity (Hypergraph::create_event, hypergraph.cpp:448-499): the signature of
    // event_core.hpp:60-84 over the exact hashes of the input and output states, the step of
    // the output state's class (its first state), the rule, and the canonical ranks of the
    // consumed edges (pattern order) and the produced edges (RHS order); the first event in id
    // order with a signature is its class's id, and NumEvents counts the classes.
    // On the reconstruction route an event is signed by the recorded expansion it replays
    // (quotient_replay_core.hpp:133-144): each class's expansion is recorded from the first of
    // its states that was expanded, and an event of any other state of the class is that
    // recorded event with the same rule and consumed ranks (canonical ranks are one frame for
    // the whole class); NumEvents counts the recorded expansions' signatures.
    std::vector<int64_t> ev_rep(ne);
    uint64_t n_ev_classes = ne;
    std::vector<uint32_t> rec_of;                              // route: the recorded event of each event
    if (route) {
        std::unordered_map<uint64_t, uint32_t> rec_state;     // exact hash -> recording state
        std::unordered_map<uint64_t, uint32_t> rec_by_key;    // (class, rule, consumed ranks) -> event
        rec_of.assign(ne, 0);
        auto tkey = [&](uint64_t e) {
            const hgx_event& x = R.events[e];
            uint64_t k = fnv(fnv(FNV_OFFSET, ir[x.input_state]), (uint64_t)x.rule);
            for (uint32_t q = R.coff[e]; q < R.coff[e + 1]; ++q) k = fnv(k, rank[R.soff[x.input_state] + R.cpos[q]]);
            return k;
        };
        for (uint64_t e = 0; e < ne; ++e) {
            const uint32_t in = R.events[e].input_state;
            const uint32_t rs = rec_state.emplace(ir[in], in).first->second;
            if (rs == in) rec_by_key.emplace(tkey(e), (uint32_t)e);
        }
        for (uint64_t e = 0; e < ne; ++e) {
            auto it = rec_by_key.find(tkey(e));
            if (it == rec_by_key.end())
                throw JobError("an event has no recorded expansion of its class (an exact hash gave up?)");
            rec_of[e] = it->second;
        }
    }
    if (j.ev_keys) {
        std::unordered_map<uint64_t, uint32_t> first_of_class;     // exact hash -> first state
        for (uint64_t s = 0; s < ns; ++s) first_of_class.emplace(ir[s], (uint32_t)s);
        std::unordered_map<uint64_t, uint32_t> first_of_sig;
        std::vector<uint64_t> sig_of(route ? ne : 0);
        auto sign = [&](uint64_t e) {
            const hgx_event& x = R.events[e];
            uint64_t sig = FNV_OFFSET;
            if (j.ev_keys & 1) sig = fnv(sig, ir[x.input_state]);
            if (j.ev_keys & 2) sig = fnv(sig, ir[x.output_state]);
            if (j.ev_keys & 4) sig = fnv(sig, R.step[first_of_class[ir[x.output_state]]]);
            if (j.ev_keys & 8) sig = fnv(sig, (uint64_t)(uint16_t)x.rule);
            if (j.ev_keys & 16)
                for (uint32_t k = R.coff[e]; k < R.coff[e + 1]; ++k)
                    sig = fnv(sig, rank[R.soff[x.input_state] + R.cpos[k]]);
            if (j.ev_keys & 32) {
                const uint64_t o1 = R.soff[x.output_state + 1];
                const uint64_t nrhs = rhs_edges(j, x.rule);
                for (uint64_t k = o1 - nrhs; k < o1; ++k) sig = fnv(sig, rank[k]);
            }
            if (sig == 0 || sig == FNV_OFFSET) sig = 1;
            return sig;
        };
        for (uint64_t e = 0; e < ne; ++e) {
            if (route && rec_of[e] != e) continue;            // signed through its recorded event below
            const uint64_t sig = sign(e);
            if (route) sig_of[e] = sig;
            ev_rep[e] = first_of_sig.emplace(sig, (uint32_t)e).first->second;
        }
        if (route)
            for (uint64_t e = 0; e < ne; ++e) ev_rep[e] = first_of_sig[sig_of[rec_of[e]]];
        n_ev_classes = first_of_sig.size();
    } else {
        for (uint64_t e = 0; e < ne; ++e) ev_rep[e] = (int64_t)e;
    }

    // On the reconstruction route with quotient exploration the counts a caller is told are the
    // raw unfolding's (hypergraph_ffi.cpp:1672-1720): every class expansion replayed on every raw
    // instance, on the device. Full capture is the raw unfolding already.
    uint64_t n_branchial = R.c.branchial_pairs;
    int64_t unfold_causal = -1;
    if (route && j.quotient) {
        hgx_unfolding u{};
        check(hgx_result_unfold(ctx, res, &u), ctx, "raw unfolding");
        if (!j.ev_keys) n_ev_classes = u.raw_events;
        unfold_causal = (int64_t)(j.tr ? u.causal_pairs_tr : u.causal_pairs);
        n_branchial = u.branchial_pairs;
    }

    Value out = Value::assoc();
    if (j.inc_states) {
        Value sts = Value::assoc();
        sts.rules.reserve(reps.size());
        std::vector<uint32_t> coff, cv;
        for (uint64_t s : reps) {
            Value sa = Value::assoc();
            sa.put("Id", Value::integer((int64_t)s));
            sa.put("CanonicalId", Value::integer(rep[s]));
            sa.put("ContentStateId", Value::integer(rep[s]));
            sa.put("Step", Value::integer(R.step[s]));
            // Full relabels to the canonical form; None keeps the state's own labels
            const uint32_t e0 = (uint32_t)R.soff[s], ne_s = (uint32_t)(R.soff[s + 1] - R.soff[s]);
            const uint32_t* eo = R.eoff.data() + e0;
            const uint32_t* vv = R.verts.data();
            if (j.full && ne_s) {
                coff.assign(ne_s + 1, 0);
                cv.assign(R.eoff[e0 + ne_s] - R.eoff[e0], 0);
                uint64_t h = 0;
                check(hgx_ir_canonical_form(eo, vv, ne_s, coff.data(), cv.data(), &h), ctx, "canonical form");
                eo = coff.data();
                vv = cv.data();
            }
            Value edges = Value::list();
            for (uint32_t k = 0; k < ne_s; ++k) {
                Value ed = Value::list();
                ed.push(Value::integer((int64_t)k));
                for (uint32_t x = eo[k]; x < eo[k + 1]; ++x) ed.push(Value::integer(vv[x]));
                edges.push(std::move(ed));
            }
            sa.put("Edges", std::move(edges));
            sa.put("IsInitial", Value::boolean(s < n_roots));
            if (j.inc_hashes) sa.put("CanonicalHash", Value::integer((int64_t)ir[s]));
            sts.put((int64_t)s, std::move(sa));
        }
        out.put("States", std::move(sts));
    }
    if (j.inc_events || j.inc_events_min) {
        Value evs = Value::assoc();
        evs.rules.reserve(ne);
        for (uint64_t e = 0; e < ne; ++e) {
            const hgx_event& x = R.events[e];
            Value ea = Value::assoc();
            ea.put("Id", Value::integer((int64_t)e));
            if (j.inc_events) {
                ea.put("CanonicalId", Value::integer(ev_rep[e]));
                ea.put("RuleIndex", Value::integer(x.rule));
            }
            ea.put("InputState", Value::integer(x.input_state));
            ea.put("OutputState", Value::integer(x.output_state));
            if (j.inc_events) {
                ea.put("CanonicalInputState", Value::integer(rep[x.input_state]));
                ea.put("CanonicalOutputState", Value::integer(rep[x.output_state]));
                std::vector<int64_t> cons, prod;
                for (uint32_t k = R.coff[e]; k < R.coff[e + 1]; ++k) cons.push_back(R.cpos[k]);
                const uint64_t out_ne = R.soff[x.output_state + 1] - R.soff[x.output_state];
                const size_t nrhs = rhs_edges(j, x.rule);
                for (uint64_t k = out_ne - nrhs; k < out_ne; ++k) prod.push_back((int64_t)k);
                ea.put("ConsumedEdges", ilist(cons));
                ea.put("ProducedEdges", ilist(prod));
            }
            evs.put((int64_t)e, std::move(ea));
        }
        out.put("Events", std::move(evs));
    }
    int64_t n_causal = 0;
    {
        Value cl = Value::list();
        for (size_t k = 0; k < R.causal.size(); ++k) {
            if (j.tr && !R.causal_tr[k]) continue;
            ++n_causal;
            if (!j.inc_causal) continue;
            Value ed = Value::assoc();
            ed.put("From", Value::integer(R.causal[k].a));
            ed.put("To", Value::integer(R.causal[k].b));
            ed.put("RawFrom", Value::integer(R.causal[k].a));
            ed.put("RawTo", Value::integer(R.causal[k].b));
            cl.push(std::move(ed));
        }
        if (j.inc_causal) out.put("CausalEdges", std::move(cl));
    }
    if (j.inc_branchial) {
        Value bl = Value::list();
        for (const auto& p : R.branchial) {
            Value ed = Value::assoc();
            ed.put("From", Value::integer(p.a));
            ed.put("To", Value::integer(p.b));
            bl.push(std::move(ed));
        }
        out.put("BranchialEdges", std::move(bl));
    }
    if (j.inc_nstates) out.put("NumStates", Value::integer((int64_t)reps.size()));
    if (j.inc_nevents) out.put("NumEvents", Value::integer((int64_t)n_ev_classes));
    if (j.inc_ncausal) out.put("NumCausalEdges", Value::integer(unfold_causal >= 0 ? unfold_causal : n_causal));
    if (j.inc_nbranchial) out.put("NumBranchialEdges", Value::integer((int64_t)n_branchial));
    if (!warns.empty()) {
        Value wl = Value::list();
        for (auto& w : warns) {
            Value wa = Value::assoc();
            wa.put("Kind", Value::string(w.first));
            wa.put("Count", Value::integer(1));
            wa.put("Context", Value::string(w.second));
            wl.push(std::move(wa));
        }
        out.put("Warnings", std::move(wl));
    }
    return out;
}

static uint8_t* to_buffer(const std::vector<uint8_t>& bytes, size_t* n) {
    uint8_t* p = (uint8_t*)std::malloc(bytes.size() ? bytes.size() : 1);
    if (!p) throw JobError("out of host memory");
    if (!bytes.empty()) std::memcpy(p, bytes.data(), bytes.size());
    *n = bytes.size();
    return p;
}

}  // namespace job
}  // namespace hgx

using namespace hgx::job;

extern "C" int hgx_wxf_validate(const uint8_t* job, size_t n, char* err, size_t err_len) {
    try {
        if (!job) throw JobError("null job");
        const Job j = parse(hgx::wxf::parse(job, n));
        (void)program(j);
        std::vector<uint32_t> eoff, verts, roff;
        roots(j, eoff, verts, roff);
        if (roff.size() < 2) throw JobError("the job has no non-empty initial state");
        return HGX_OK;
    } catch (const std::exception& e) {
        if (err && err_len) std::snprintf(err, err_len, "%s", e.what());
        return HGX_ERR_INVALID;
    }
}

extern "C" int hgx_run_wxf(hgx_ctx* ctx, const uint8_t* job, size_t n, uint8_t** out, size_t* out_n) {
    if (!ctx || !job || !out || !out_n) return HGX_ERR_INVALID;
    *out = nullptr;
    *out_n = 0;
    try {
        const Job j = parse(hgx::wxf::parse(job, n));
        const Value res = evolve(ctx, j);
        *out = to_buffer(hgx::wxf::serialize(res), out_n);
        return HGX_OK;
    } catch (const std::exception& e) {
        std::lock_guard<std::recursive_mutex> g(ctx->mu);
        ctx->err = e.what();
        return HGX_ERR_INVALID;
    }
}

extern "C" void hgx_buffer_free(uint8_t* p) { std::free(p); }

// ---- run_gpu_evolution (paclet_source/hg_gpu_backend.hpp:119) ------------------------------

namespace hgx {

std::vector<uint8_t> run_gpu_evolution(const GpuJob& g, const HostBridge& host) {
    auto refuse = [](const std::string& what) {
        throw std::runtime_error(what + " is not served (this backend evolves the full multiway system)");
    };
    if (!g.graph_properties.empty()) refuse("GraphProperties");
    if (g.branchial_step != 0) refuse("BranchialStep");
    if (g.show_genesis_events) refuse("ShowGenesisEvents");
    if (g.include_branchial_state_edges || g.include_branchial_state_edges_all_siblings ||
        g.include_global_edges || g.include_state_bitvectors)
        refuse("that RequestedData component");
    if (!g.session_op.empty() && g.session_op != "Evolve") refuse("a session (" + g.session_op + ")");
    if (g.steps < 0) throw std::runtime_error("Steps must be non-negative");

    Job j;
    for (const auto& r : g.rules) {
        if (r.second.size() != 2) throw std::runtime_error("rule " + r.first + " must be {lhs, rhs}");
        j.rules.emplace_back(r.second[0], r.second[1]);
    }
    if (j.rules.empty()) throw std::runtime_error("the job has no Rules");
    j.states = g.initial_states;
    j.steps = g.steps;
    j.tr = g.transitive_reduction;
    j.full = g.state_canon_mode == GpuJob::StateCanonCode::kFull;
    j.automatic = g.state_canon_mode == GpuJob::StateCanonCode::kAutomatic;
    j.quotient = g.explore_from_canonical_states_only;
    j.quotient_roots = g.quotient_initial_states;
    j.smp.explore_p = g.exploration_probability;
    j.smp.rate = g.transition_rate;
    j.smp.weights = g.rule_weights;
    j.smp.seed = g.random_seed;
    j.smp.per_step = g.max_states_per_step;
    j.smp.per_parent = g.max_successor_states_per_parent;
    j.smp.per_state_rule = g.matches_per_state_rule;
    constexpr uint32_t In = 1, Out = 2, Step = 4, Cons = 16, Prod = 32;
    j.ev_keys = g.event_canon_mode == GpuJob::EventCanonCode::kFull ? (In | Out)
              : g.event_canon_mode == GpuJob::EventCanonCode::kAutomatic ? (In | Out | Step | Cons | Prod) : 0u;
    j.inc_states = g.include_states;
    j.inc_events = g.include_events;
    j.inc_events_min = g.include_events_minimal;
    j.inc_causal = g.include_causal_edges;
    j.inc_branchial = g.include_branchial_edges;
    j.inc_hashes = g.include_canonical_hashes;
    j.inc_nstates = g.include_num_states;
    j.inc_nevents = g.include_num_events;
    j.inc_ncausal = g.include_num_causal_edges;
    j.inc_nbranchial = g.include_num_branchial_edges;

    // one context for the process, created on first use with the first job's budget
    static std::mutex mu;
    static hgx_ctx* ctx = nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ctx) {
        const int rc = hgx_ctx_create(0, g.max_device_memory_bytes, &ctx);
        if (rc != HGX_OK) {
            char buf[512] = {0};
            hgx_last_error(nullptr, buf, sizeof(buf));
            ctx = nullptr;
            throw std::runtime_error(std::string("no device context: ") + buf);
        }
    }
    if (host.progress) host.progress("hgx: evolving " + std::to_string(g.steps) + " steps on the device");
    const Value res = job::evolve(ctx, j);
    return wxf::serialize(res);
}

}  // namespace hgx
