// hgx_engine.hpp -- the generation engine behind both the batch and the served path.
//
// The engine holds a FRONTIER: one device generation of states drawn from every admitted
// request, each at its own depth. step() advances all of them by one generation -- join,
// cap decision, rewrite + WL + dedup + causal, branchial -- and admits queued requests by
// appending their roots to the generation it builds, so a request that arrives while others
// run joins the very next launch. A request that finishes leaves its counters and a view of
// its generations behind; its slot, rules and state map are recycled.
#pragma once

#include <chrono>
#include <deque>
#include <unordered_map>
#include <unordered_set>
#include <memory>
#include <string>
#include <vector>

#include "hgx_runtime.hpp"

namespace hgx {

// One generation on the device. `meta` holds per-state and per-event arrays (kept while any
// result may read them); `edges` holds the edge pool (released once the next generation has
// been built from it, unless a request asked to keep its states' edges).
struct DevGen {
    std::shared_ptr<ChunkPool> pool;
    Arena meta;
    Arena edges;
    GenView v{};
    EventView ev{};
    uint64_t n_edges = 0;
    uint32_t max_ne = 0, max_nextv = 0;
    uint32_t* br_pairs = nullptr;    // branchial pairs of the events that created these states
    uint16_t* cq = nullptr;          // with kept edges: consumed edge positions of each event
    uint32_t cK = 0;                 //   (stride)
    bool keep_edges = false;
    explicit DevGen(std::shared_ptr<ChunkPool> p)
        : pool(p), meta(p.get(), (size_t)1 << 20), edges(p.get(), (size_t)4 << 20) {}
};

struct Segment {
    std::shared_ptr<DevGen> g;       // generation holding this run's states
    uint32_t b, e;                   // their index range in it
    uint64_t br_b, br_e;             // this run's branchial pairs in g->br_pairs
};

using Clock = std::chrono::steady_clock;

struct RunRec {
    uint64_t id = 0;
    std::shared_ptr<const RunHost> hp;   // parsed request (shared with a prepared batch)
    uint32_t slot = 0xFFFFFFFFu;
    DRule* d_rules = nullptr;
    uint64_t* map_keys = nullptr;
    uint32_t map_cap = 0;
    uint64_t states = 0, events = 0;
    uint32_t depth = 0, steps_done = 0;
    uint32_t ne_bound = 0, nv_bound = 0;   // bounds on edges / vertex labels of its frontier states
    bool truncated = false, done = false, admitted = false;
    // Why the run stopped early: HGX_OK (it did not, or its state cap stopped it) or
    // HGX_ERR_CAPACITY (a generation did not fit the context's HBM budget: the result holds the
    // generations completed before it, with `warning` saying so).
    int status = 0;
    std::string warning;
    bool finishing = false;          // listed for finalize in the current step
    uint64_t footprint = 0;
    std::vector<Segment> segs;       // every state of the run, in state-id order
    // A sampled request's frontier states: exact hash and canonical edge ranks (edge position q
    // of frontier state k has rank smp_rank[smp_roff[k] + q]), for its transition keys.
    std::vector<uint64_t> smp_ir;
    std::vector<uint32_t> smp_roff, smp_rank;
    const void* smp_gen = nullptr;   // the generation smp_* were computed for
    // Quotient exploration: the exact classes met so far (hash of each).
    std::unordered_set<uint64_t> q_seen;
    hgx_counts counts{};
    hipStream_t st = nullptr;        // the context's stream (contexts outlive results)
    Clock::time_point t_submit, t_admit, t_done;
    double device_ms = 0;
};

// The new states of a request whose frontier the host decides (sampling, quotient exploration)
// in a new generation: [cb, ce), its events before them.
struct SampledKids {
    RunRec* r;
    uint32_t cb, ce;
    uint64_t ev_base;
    bool roots = false;              // [cb, ce) are the request's roots
};

// The host decisions of one eager pass over a prepared batch, per generation: the per-slot child
// edges, child totals and parent edges read after the count pass, and the branchial pair offsets
// read before the pairs are written. A pass recorded into a graph takes them from here instead of
// reading them from the device, and every replay checks the totals on the device.
struct PassPlan {
    std::vector<std::vector<uint64_t>> rb;
    std::vector<std::vector<uint64_t>> poff;
};

enum class EngineMode {
    Normal,     // every decision from the device, synchronising where a size is needed
    Record,     // the same, keeping the decisions in a PassPlan
    Capture     // under stream capture: decisions from the PassPlan, no synchronisation, every
                // host-to-device copy from pinned staging the engine keeps, no buffer reused
};

class Engine {
public:
    explicit Engine(hgx_ctx* ctx);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    uint64_t submit(std::shared_ptr<const RunHost> h, uint64_t id);
    void reserve_slots(uint32_t n) { grow_slots(n); live_.reserve(n); finished.reserve(n); }
    // One generation of every admitted request, plus admissions. Returns live + queued runs.
    // A generation that does not fit the HBM budget does not fail the engine: the requests in
    // the frontier finish with the generations they completed (status HGX_ERR_CAPACITY), the
    // requests admitted into that generation go back to the queue, and a request that fails
    // alone finishes with no states -- so every step makes progress.
    uint32_t step();
    uint32_t live_or_queued() const { return (uint32_t)(live_.size() + pending_.size()); }
    // A finished run (removed from the finished set), or null.
    std::shared_ptr<RunRec> take(uint64_t id);
    bool known(uint64_t id) const;
    std::unordered_map<uint64_t, std::shared_ptr<RunRec>> finished;
    double soft_budget_frac = 0.6;

    // Record / capture of a batched pass (hgx_batch_capture). check: a device word the replay
    // check sets when the device's totals leave the plan.
    void set_mode(EngineMode m, std::shared_ptr<PassPlan> plan, uint32_t* check);
    // After a replay of the captured pass: every finished run's counters read from the device.
    void resolve_counts();

private:
    hgx_ctx* ctx_;
    hipStream_t st_;
    std::shared_ptr<ChunkPool> pool_;
    SlabAlloc slab_;
    std::unique_ptr<Arena> slot_arena_;
    uint32_t cap_slots_ = 0;
    uint32_t hw_slots_ = 0;                   // one past the highest slot ever taken
    DRun* d_runs_ = nullptr;
    uint64_t* d_ctr_ = nullptr;
    uint64_t* d_ctr_snap_ = nullptr;          // the counters at the start of the step
    uint8_t* d_active_ = nullptr;
    uint32_t* d_budget_ = nullptr;
    uint64_t* d_run_acc_ = nullptr;
    uint64_t* d_run_edges_ = nullptr;
    uint32_t* d_sbeg_ = nullptr;
    uint32_t* d_send_ = nullptr;
    uint64_t* d_tot_ = nullptr;               // [cap_slots + 1]
    uint64_t* d_run_pe_ = nullptr;            // [cap_slots] edges of each run's states with matches
    std::vector<DRun> h_runs_;
    std::vector<uint8_t> h_tab_;              // host side of budget, sbeg, send, active
    std::vector<uint32_t> free_slots_;
    std::vector<uint32_t> release_next_;      // slots whose states are still in the frontier
    std::unordered_map<uint64_t, std::shared_ptr<RunRec>> live_;
    std::deque<std::shared_ptr<RunRec>> pending_;
    std::shared_ptr<DevGen> F_;               // current frontier
    std::vector<std::shared_ptr<RunRec>> F_runs_;   // runs with states in F, in F order
    std::vector<uint32_t> F_b_, F_e_;         // their ranges, parallel to F_runs_
    uint64_t live_footprint_ = 0;
    uint32_t admit_limit_ = 0xFFFFFFFFu;       // admissions per step (1 after a failed joint start)
    // The previous generation's match list, kept one generation longer: state s < prev_.n of
    // the frontier was made by match s, and its parent's matches seed its own (forwarding).
    struct PrevMatches {
        std::unique_ptr<Arena> ar;
        uint32_t n = 0, K = 1;
        uint32_t* parent = nullptr;
        uint32_t* rule = nullptr;
        uint16_t* q = nullptr;
        uint64_t* moff = nullptr;
    } prev_;

    // Sampled requests: which frontier states are expanded ([F_->v.n]; empty: every one), and
    // whether the last generation's match list was thinned (then its matches do not forward:
    // a child's matches come from its parent's full list).
    std::vector<uint8_t> F_expand_;
    bool thinned_ = false;
    // After the emit pass: the transition decisions of the sampled requests' matches (their
    // frontier's exact hashes and ranks from frontier_ir), and the match list, the per-state
    // counts and offsets, the per-run totals and child edges compacted to the survivors.
    void sample_transitions(uint32_t n, uint32_t K, MatchView& mv, uint32_t* cnt, uint64_t* moff,
                            std::vector<uint64_t>& tot, std::vector<uint64_t>& h_edges, uint64_t& M,
                            std::vector<uint32_t>& h_parent, hipStream_t st);
    // After a generation is built: which of the sampled requests' new states are expanded
    // (exploration coin, per-step and per-parent bounds, in event order), into F_expand_.
    void sample_frontier(const std::vector<SampledKids>& kids, uint32_t nC, const MatchView& mv, uint64_t M,
                         std::vector<uint32_t>& h_parent, hipStream_t st);

    EngineMode mode_ = EngineMode::Normal;
    std::shared_ptr<PassPlan> plan_;
    size_t plan_rb_ = 0, plan_poff_ = 0;      // next plan entries (capture)
    uint32_t* d_check_ = nullptr;
    std::vector<void*> pinned_;               // pinned staging of captured copies

    // Host-to-device copy of n elements: straight from src, or (capture) through pinned staging
    // that lives as long as the engine, since a replay copies from it again.
    void put(void* dst, const void* src, size_t bytes, hipStream_t st);
    template <class T>
    void up(T* dst, const T* src, size_t n, hipStream_t st) { put(dst, src, sizeof(T) * n, st); }
    void fill_counts(RunRec& r, const uint64_t* x);

    uint32_t step_impl(std::vector<std::shared_ptr<RunRec>>& admitted);
    void recover(const std::string& why, std::vector<std::shared_ptr<RunRec>>& admitted, bool frontier_live);
    void grow_slots(uint32_t need);
    uint32_t take_slot();
    void admit(std::shared_ptr<RunRec>& r);
    void finalize(std::vector<std::shared_ptr<RunRec>>& runs);
};

void engine_destroy(std::unique_ptr<Engine>& e);

}  // namespace hgx

struct hgx_result {
    std::shared_ptr<hgx::RunRec> run;
    uint64_t device_ns = 0, latency_ns = 0;
};
