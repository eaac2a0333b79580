// hgx_engine.cpp -- the generation engine (see hgx_engine.hpp), the batch and served entry
// points, and result readback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <unordered_map>

#include "hgx_engine.hpp"
#include "hgx_kernels.hpp"

namespace hgx {

// HGX_TRACE=1 prints a host-side phase breakdown of every generation step to stderr (the role of
// the reference's HG_GPU_DBG_TIME, gpu/src/evolve.cu:358).
static bool trace_on() {
    static const bool on = [] { const char* e = std::getenv("HGX_TRACE"); return e && *e && *e != '0'; }();
    return on;
}
static double ms_since(Clock::time_point t) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t).count();
}

Engine::Engine(hgx_ctx* ctx) : ctx_(ctx), st_(ctx->stream), pool_(ctx->pool), slab_(ctx->pool.get()) {
    grow_slots(256);
}

Engine::~Engine() {
    (void)hipStreamSynchronize(st_);
    for (void* p : pinned_) (void)hipHostFree(p);
}

void Engine::set_mode(EngineMode m, std::shared_ptr<PassPlan> plan, uint32_t* check) {
    mode_ = m;
    plan_ = std::move(plan);
    plan_rb_ = plan_poff_ = 0;
    d_check_ = check;
}

void Engine::put(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (!bytes) return;
    if (mode_ != EngineMode::Capture) {
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
        return;
    }
    void* h = nullptr;
    HIP_TRY(hipHostMalloc(&h, bytes, hipHostMallocDefault));
    pinned_.push_back(h);
    std::memcpy(h, src, bytes);
    HIP_TRY(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, st));
}

void Engine::fill_counts(RunRec& r, const uint64_t* x) {
    hgx_counts& c = r.counts;
    std::memset(&c, 0, sizeof(c));
    c.raw_states = r.states;
    c.events = r.events;
    c.canonical_states = x[CTR_CANON];
    c.fingerprint = x[CTR_FINGERPRINT];
    const bool rec_causal = r.hp->flags & HGX_REC_CAUSAL;
    c.causal_triples = rec_causal ? x[CTR_TRIPLES] : 0;
    c.causal_pairs = rec_causal ? x[CTR_PAIRS] : 0;
    const bool tr = r.hp->flags & HGX_OPT_TR;
    c.causal_pairs_tr = tr ? x[CTR_PAIRS_TR] : c.causal_pairs;
    c.causal_triples_tr = tr ? x[CTR_TRIPLES_TR] : c.causal_triples;
    c.branchial_pairs = (r.hp->flags & HGX_REC_BRANCHIAL) ? x[CTR_BRANCHIAL] : 0;
    c.steps_done = r.steps_done;
    c.truncated = r.status == HGX_ERR_CAPACITY ? 2u : (r.truncated ? 1u : 0u);
}

void Engine::resolve_counts() {
    std::vector<uint64_t> ctr((size_t)hw_slots_ * CTR_N);
    d2h(ctr.data(), d_ctr_, ctr.size(), st_);
    HIP_TRY(hipStreamSynchronize(st_));
    for (auto& kv : finished) fill_counts(*kv.second, &ctr[(size_t)kv.second->slot * CTR_N]);
}

void engine_destroy(std::unique_ptr<Engine>& e) { e.reset(); }

void Engine::grow_slots(uint32_t need) {
    if (need <= cap_slots_) return;
    uint32_t cap = std::max<uint32_t>(256, cap_slots_);
    while (cap < need) cap *= 2;
    auto na = std::make_unique<Arena>(pool_.get());
    DRun* runs = dalloc<DRun>(*na, cap);
    uint64_t* ctr = dalloc<uint64_t>(*na, (size_t)cap * CTR_N);
    uint64_t* snap = dalloc<uint64_t>(*na, (size_t)cap * CTR_N);
    // per-slot tables uploaded together each generation: budget, range begin, range end (u32),
    // active (u8); and the ones zeroed together: accumulated matches, child edges, totals, the
    // edges of the states with matches
    uint8_t* tab = dalloc<uint8_t>(*na, 13ull * cap);
    uint32_t* budget = (uint32_t*)tab;
    uint32_t* sb = budget + cap;
    uint32_t* se = sb + cap;
    uint8_t* active = (uint8_t*)(se + cap);
    uint64_t* acc = dalloc<uint64_t>(*na, 4ull * cap + 1);
    uint64_t* redges = acc + cap;
    uint64_t* tot = redges + cap;                     // [cap + 1]
    uint64_t* rpe = tot + cap + 1;
    HIP_TRY(hipMemsetAsync(ctr, 0, sizeof(uint64_t) * cap * CTR_N, st_));
    HIP_TRY(hipMemsetAsync(snap, 0, sizeof(uint64_t) * cap * CTR_N, st_));
    if (cap_slots_) {
        // slots grow only at admission, before any kernel of the step has changed a counter, so
        // the old counters are the step's snapshot as well
        HIP_TRY(hipMemcpyAsync(ctr, d_ctr_, sizeof(uint64_t) * cap_slots_ * CTR_N, hipMemcpyDeviceToDevice, st_));
        HIP_TRY(hipMemcpyAsync(snap, d_ctr_, sizeof(uint64_t) * cap_slots_ * CTR_N, hipMemcpyDeviceToDevice, st_));
    }
    HIP_TRY(hipStreamSynchronize(st_));
    h_runs_.resize(cap);
    h_tab_.assign(13ull * cap, 0);
    for (uint32_t s = cap; s-- > cap_slots_;) {
        std::memset(&h_runs_[s], 0, sizeof(DRun));
        free_slots_.push_back(s);
    }
    slot_arena_ = std::move(na);
    d_runs_ = runs; d_ctr_ = ctr; d_ctr_snap_ = snap; d_active_ = active; d_budget_ = budget; d_run_acc_ = acc;
    d_run_edges_ = redges;
    d_sbeg_ = sb; d_send_ = se; d_tot_ = tot; d_run_pe_ = rpe;
    cap_slots_ = cap;
}

uint32_t Engine::take_slot() {
    if (free_slots_.empty()) grow_slots(cap_slots_ * 2);
    const uint32_t s = free_slots_.back();
    hw_slots_ = std::max(hw_slots_, s + 1);
    free_slots_.pop_back();
    return s;
}

uint64_t Engine::submit(std::shared_ptr<const RunHost> h, uint64_t id) {
    auto r = std::make_shared<RunRec>();
    r->id = id;
    r->hp = std::move(h);
    r->footprint = footprint_bytes(*r->hp);
    r->st = st_;
    r->t_submit = Clock::now();
    pending_.push_back(r);
    return id;
}

bool Engine::known(uint64_t id) const {
    if (finished.count(id) || live_.count(id)) return true;
    for (auto& r : pending_) if (r->id == id) return true;
    return false;
}

std::shared_ptr<RunRec> Engine::take(uint64_t id) {
    auto it = finished.find(id);
    if (it == finished.end()) return nullptr;
    auto r = it->second;
    finished.erase(it);
    return r;
}

// Host side of an admission; the device side (rules copied, map and counters zeroed) is one
// k_admit launch for all of a step's admissions.
void Engine::admit(std::shared_ptr<RunRec>& r) {
    const uint32_t slot = take_slot();
    r->slot = slot;
    uint64_t want = r->hp->cap ? (uint64_t)r->hp->cap + r->hp->n_roots() : std::max<uint64_t>(512, r->hp->n_roots());
    uint64_t mc = 64;
    while (mc < 2 * want) mc <<= 1;
    r->map_cap = (uint32_t)mc;
    r->d_rules = (DRule*)slab_.alloc(sizeof(DRule) * r->hp->rules.size());
    r->map_keys = (uint64_t*)slab_.alloc(8 * mc);
    DRun& d = h_runs_[slot];
    std::memset(&d, 0, sizeof(d));
    d.rules = r->d_rules;
    d.n_rules = (uint32_t)r->hp->rules.size();
    d.n_roots = r->hp->n_roots();
    d.flags = r->hp->flags;
    d.map_keys = r->map_keys;
    d.map_mask = (uint32_t)(mc - 1);
    r->ne_bound = r->hp->max_root_ne;
    r->nv_bound = r->hp->max_root_nv;
    r->q_seen.clear();
    r->smp_gen = nullptr;
    r->admitted = true;
    r->t_admit = Clock::now();
    live_[r->id] = r;
    live_footprint_ += r->footprint;
}

void Engine::finalize(std::vector<std::shared_ptr<RunRec>>& runs) {
    if (runs.empty()) return;
    const auto now = Clock::now();
    if (mode_ == EngineMode::Capture) {
        // counters are read after each replay (resolve_counts); the run's slot, rules and map
        // stay with it, since the graph writes them again on every replay
        for (auto& r : runs) {
            r->t_done = now;
            finished[r->id] = r;
            live_.erase(r->id);
            live_footprint_ -= r->footprint;
        }
        return;
    }
    std::vector<uint64_t> ctr((size_t)hw_slots_ * CTR_N);       // slots ever taken
    d2h(ctr.data(), d_ctr_, ctr.size(), st_);
    HIP_TRY(hipStreamSynchronize(st_));
    for (auto& r : runs) {
        fill_counts(*r, &ctr[(size_t)r->slot * CTR_N]);
        r->t_done = now;
        r->device_ms = std::chrono::duration<double, std::milli>(now - r->t_admit).count();
        finished[r->id] = r;
        live_.erase(r->id);
        live_footprint_ -= r->footprint;
        release_next_.push_back(r->slot);
        slab_.free(r->d_rules, sizeof(DRule) * r->hp->rules.size());
        slab_.free(r->map_keys, 8ull * r->map_cap);
        r->d_rules = nullptr; r->map_keys = nullptr;
    }
}

namespace {
// Children of the requests whose child-edge bound falls in one size class.
struct ExpandClass {
    uint32_t ne = 0, nextv = 0;
    bool uniform = true;                      // every request of the class has one arity 2..4
    uint64_t n = 0;
    std::vector<uint32_t> rb, rpre{0};
};
constexpr uint32_t kClassEdges[6] = {32, 64, 128, 512, 2048, 0xFFFFFFFFu};
// The expansion's size classes are finer at the small end, where a child's workspace is a few KiB
// and several children share one wave (sub-wave teams, by default: a wave hashes four children of
// up to 32 edges, or two of up to 64, each in its own workspace; the autotuner may choose others).
constexpr int kExpandClasses = 6;
constexpr uint32_t kExpandEdges[kExpandClasses] = {32, 64, 128, 512, 2048, 0xFFFFFFFFu};
static_assert(kExpandClasses == kTuneExpandClasses, "one tuned shape per expansion size class");
static_assert(sizeof(kClassEdges) / sizeof(kClassEdges[0]) == kTuneJoinClasses, "one tuned shape per join size class");

// A stride near n / golden ratio and coprime to n, so i -> i * stride mod n visits 0..n-1 once
// each, consecutive i far apart (1 for n <= 2: in order).
uint32_t coprime_stride(uint32_t n) {
    if (n <= 2) return 1;
    uint32_t s = (uint32_t)((double)n * 0.6180339887498949);
    if (s < 1) s = 1;
    auto gcd = [](uint32_t a, uint32_t b) { while (b) { const uint32_t t = a % b; a = b; b = t; } return a; };
    while (gcd(s, n) != 1) ++s;
    return s;
}
}  // namespace

uint32_t Engine::step() {
    // runs of the frontier still evolving, and the counters before this generation touches them
    bool frontier_live = false;
    for (auto& r : F_runs_) frontier_live |= !r->done && live_.count(r->id) != 0;
    if (mode_ == EngineMode::Capture) {             // no recovery inside a graph being recorded
        std::vector<std::shared_ptr<RunRec>> admitted;
        return step_impl(admitted);
    }
    if (frontier_live && hw_slots_)
        HIP_TRY(hipMemcpyAsync(d_ctr_snap_, d_ctr_, sizeof(uint64_t) * hw_slots_ * CTR_N,
                               hipMemcpyDeviceToDevice, st_));
    std::vector<std::shared_ptr<RunRec>> admitted;
    try {
        const uint32_t left = step_impl(admitted);
        admit_limit_ = 0xFFFFFFFFu;
        return left;
    } catch (const Error& e) {
        if (e.code != HGX_ERR_CAPACITY) throw;
        recover(e.msg, admitted, frontier_live);
        return live_or_queued();
    }
}

// A generation ran out of the HBM budget (the reference's capacity warning and partial result,
// gpu/src/evolve.cu:573-700): the counters go back to the snapshot taken before the generation,
// the frontier's requests finish with the generations they completed, and the requests admitted
// into the failed generation return to the front of the queue. With no frontier, the admitted
// requests failed by themselves: one alone finishes with no states; several are re-queued and
// started one at a time.
void Engine::recover(const std::string& why, std::vector<std::shared_ptr<RunRec>>& admitted, bool frontier_live) {
    (void)hipStreamSynchronize(st_);
    (void)hipGetLastError();
    if (frontier_live && hw_slots_)
        HIP_TRY(hipMemcpyAsync(d_ctr_, d_ctr_snap_, sizeof(uint64_t) * hw_slots_ * CTR_N,
                               hipMemcpyDeviceToDevice, st_));
    std::vector<std::shared_ptr<RunRec>> done_now, failed_alone;
    if (!frontier_live && admitted.size() == 1) failed_alone.push_back(admitted[0]);
    else if (!frontier_live) admit_limit_ = 1;
    for (auto it = admitted.rbegin(); it != admitted.rend(); ++it) {
        auto& r = *it;
        if (!failed_alone.empty() && r == failed_alone[0]) continue;
        if (r->slot != 0xFFFFFFFFu) free_slots_.push_back(r->slot);
        slab_.free(r->d_rules, sizeof(DRule) * r->hp->rules.size());
        slab_.free(r->map_keys, 8ull * r->map_cap);
        r->d_rules = nullptr; r->map_keys = nullptr;
        r->slot = 0xFFFFFFFFu;
        if (r->admitted) {
            live_.erase(r->id);
            live_footprint_ -= r->footprint;
        }
        r->admitted = false;
        pending_.push_front(r);
    }
    for (auto& r : failed_alone) {
        // no generation of it completed: it finishes with no states, the capacity status and
        // its reason, and whatever its admission took is given back
        if (r->slot != 0xFFFFFFFFu) free_slots_.push_back(r->slot);
        slab_.free(r->d_rules, sizeof(DRule) * r->hp->rules.size());
        slab_.free(r->map_keys, 8ull * r->map_cap);
        r->d_rules = nullptr; r->map_keys = nullptr;
        if (r->admitted) {
            live_.erase(r->id);
            live_footprint_ -= r->footprint;
        }
        r->states = 0; r->events = 0; r->steps_done = 0;
        r->segs.clear();
        r->status = HGX_ERR_CAPACITY;
        r->warning = "the request's initial states do not fit the context's HBM budget: " + why;
        std::memset(&r->counts, 0, sizeof(r->counts));
        r->counts.truncated = 2;
        r->t_done = Clock::now();
        r->finishing = true;
        finished[r->id] = r;
    }
    for (auto& r : F_runs_) {
        if (!live_.count(r->id) || r->finishing) continue;
        if (!r->done) {
            r->done = true;
            r->truncated = true;
            r->status = HGX_ERR_CAPACITY;
            r->warning = "generation " + std::to_string(r->depth + 1) +
                         " does not fit the context's HBM budget; the result holds generations 0.." +
                         std::to_string(r->depth) + ": " + why;
        }
        r->finishing = true;
        done_now.push_back(r);
    }
    finalize(done_now);
    // the frontier is gone with its requests: every slot it held is free now
    if (F_ && !F_->keep_edges) F_->edges.release();
    F_.reset();
    prev_ = PrevMatches{};
    F_expand_.clear();
    F_runs_.clear(); F_b_.clear(); F_e_.clear();
    for (uint32_t s2 : release_next_) free_slots_.push_back(s2);
    release_next_.clear();
}

void frontier_ir(const DevGen& G, uint32_t b, uint32_t e, hipStream_t st, RunRec& r);

uint32_t Engine::step_impl(std::vector<std::shared_ptr<RunRec>>& admitted) {
    const auto t_step = Clock::now();
    double t_count = 0, t_emit = 0, t_child = 0, t_launch = 0, t_fin = 0;
    hipStream_t st = st_;
    Arena tmp(pool_.get(), (size_t)4 << 20);
    auto keep = std::make_unique<Arena>(pool_.get(), (size_t)4 << 20);   // this step's match list
    std::vector<ExpandClass> mcls_count, mcls_emit;   // host range lists of the join launches
    std::vector<uint64_t> h_edges;                      // per-slot child edges of this generation
    const uint32_t n = F_ ? F_->v.n : 0;
    const uint32_t nF = (uint32_t)F_runs_.size();
    std::vector<uint64_t> tot(nF, 0);
    uint64_t M = 0, parent_edges = 0;
    uint32_t K = 1;
    MatchView mv{};
    uint32_t* cnt = nullptr;
    uint64_t* moff = nullptr;
    std::vector<uint8_t> h_active(cap_slots_, 0);
    std::vector<uint32_t> h_parent;        // sampled generations: each match's parent, on the host
    std::vector<SampledKids> kids;         // children of requests that sample their frontier
    thinned_ = false;

    // ---- 1. the join over the frontier, with the state-cap decision ----
    if (n) {
        std::fill(h_tab_.begin(), h_tab_.end(), 0);
        uint32_t* h_budget = (uint32_t*)h_tab_.data();
        uint32_t* h_b = h_budget + cap_slots_;
        uint32_t* h_e = h_b + cap_slots_;
        bool any = false;
        uint32_t KL = 1;
        bool small = true;
        for (uint32_t i = 0; i < nF; ++i) {
            RunRec& r = *F_runs_[i];
            h_b[r.slot] = F_b_[i];
            h_e[r.slot] = F_e_[i];
            if (r.done || r.depth >= r.hp->steps) { r.done = true; continue; }
            h_active[r.slot] = 1;
            h_budget[r.slot] = !r.hp->cap ? 0xFFFFFFFFu : (r.states >= r.hp->cap ? 0u : (uint32_t)(r.hp->cap - r.states));
            small &= r.hp->rules.size() <= 2;
            for (auto& R : r.hp->rules) {
                KL = std::max<uint32_t>(KL, R.n_lhs);
                small &= R.n_lhs <= 4 && (R.lhs_mask >> 16) == 0;
            }
            any = true;
        }
        if (any) {
            // sampled requests: the exact hash and edge ranks of their frontier states, for the
            // transition keys of this generation's matches
            for (uint32_t i = 0; i < nF; ++i) {
                RunRec& r = *F_runs_[i];
                if (h_active[r.slot] && r.hp->smp.transitions() && r.smp_gen != F_.get())
                    frontier_ir(*F_, F_b_[i], F_e_[i], st, r);
            }
            std::memcpy(h_tab_.data() + 12ull * cap_slots_, h_active.data(), cap_slots_);
            up((uint8_t*)d_budget_, h_tab_.data(), h_tab_.size(), st);    // budget, sbeg, send, active
            HIP_TRY(launch_fill(d_run_acc_, 0, sizeof(uint64_t) * (4ull * cap_slots_ + 1), st));
            cnt = dalloc<uint32_t>(tmp, n + 1);
            moff = dalloc<uint64_t>(*keep, n + 1);
            if (F_->max_ne > 65535)
                throw Error{HGX_ERR_UNSUPPORTED, "states above 65535 edges need the large-state join"};
            MatchArgs ma{};
            ma.P = F_->v; ma.runs = d_runs_; ma.active = d_active_; ma.cnt = cnt; ma.moff = moff;
            if (!F_expand_.empty()) {                 // sampled requests' unexpanded states
                if (F_expand_.size() != n) throw Error{HGX_ERR_DEVICE, "expansion flags out of step"};
                uint8_t* d_smask = dalloc<uint8_t>(tmp, n);
                up(d_smask, F_expand_.data(), n, st);
                ma.smask = d_smask;
            }
            ma.max_ne = std::max(1u, F_->max_ne);
            ma.max_nextv = std::max(1u, F_->max_nextv);
            // forwarding: on when asked; automatic forwards the states of the size classes above
            // 512 edges (count and emit measured up to 2x cheaper than their full join there) and
            // joins smaller states in full (cheaper up to a few hundred edges)
            const bool fwd = prev_.n > 0 && prev_.n <= n && ctx_->forward != 0;
            ma.fwd_n = fwd ? prev_.n : 0;
            ma.pm_parent = prev_.parent; ma.pm_rule = prev_.rule; ma.pm_q = prev_.q; ma.pm_K = prev_.K;
            ma.pm_moff = prev_.moff;
            ma.KL = KL;
            ma.small = small ? 1u : 0u;
            // small rules: the count pass keeps each seed's match count for the emit pass
            ma.seedcnt = small ? dalloc<uint16_t>(tmp, 2 * std::max<uint64_t>(1, F_->n_edges)) : nullptr;
            ma.dmax = fwd ? std::min(ctx_->delta_cap, 16384u / (2u * KL + 8u)) : 0u;   // <= 16 KiB LDS
            ma.budget = d_budget_;
            ma.run_acc = (unsigned long long*)d_run_acc_;
            ma.state_edges = dalloc<uint64_t>(tmp, n);
            const uint64_t match_bytes = F_->n_edges * (4ull * F_->v.A + 1) + (uint64_t)n * 24;
            // One launch per size class of the requests taking part (by their bound on state
            // edges), each with LDS staging sized for its class; states outside the launches
            // (requests not advancing) keep the zero count set here.
            auto launch_classes = [&](bool emit, std::vector<ExpandClass>& cls, const char* name,
                                      uint64_t bytes) {
                cls.assign(6, ExpandClass{});
                for (uint32_t i = 0; i < nF; ++i) {
                    const RunRec& r = *F_runs_[i];
                    if (emit ? !tot[i] : !h_active[r.slot]) continue;
                    if (F_e_[i] == F_b_[i]) continue;
                    int k = 0;
                    while (r.ne_bound > kClassEdges[k]) ++k;
                    ExpandClass& x = cls[k];
                    x.ne = std::max(x.ne, r.ne_bound);
                    x.nextv = std::max(x.nextv, r.nv_bound);
                    x.rb.push_back(F_b_[i]);
                    x.n += F_e_[i] - F_b_[i];
                    x.rpre.push_back((uint32_t)x.n);
                }
                for (auto& x : cls) {
                    if (!x.n) continue;
                    MatchArgs mc = ma;
                    if (ctx_->forward == -1 && x.ne <= 512) { mc.fwd_n = 0; mc.dmax = 0; }
                    mc.max_ne = std::max(1u, std::min(x.ne, F_->max_ne));
                    mc.max_nextv = std::max(1u, std::min(x.nextv, F_->max_nextv));
                    // Team: as for the expansion, by the one-wave workspace
                    auto lpw = [&](uint32_t t) {
                        return (match_lds_per_wave(mc.max_ne, mc.max_nextv, F_->v.A, mc.KL, mc.dmax, mc.small != 0,
                                                   t, mc.fwd_n != 0) + 15u) & ~15u;
                    };
                    const int kc = (int)(&x - cls.data());
                    uint32_t team = pick_team(lpw(1), ctx_->join_team ? ctx_->join_team : ctx_->tune.join_team[kc]);
                    mc.lds_per_wave = lpw(team);
                    uint32_t gws_grid = 0;
                    if (lpw(8) > ctx_->max_lds) {
                        // states whose workspace does not fit LDS: eight-wave teams over
                        // per-workgroup slices of global memory (at most 2 per CU, 4 GiB in all)
                        team = 8;
                        mc.lds_per_wave = lpw(8);
                        uint32_t cap = std::max<uint32_t>(8u, std::min<uint64_t>(2ull * ctx_->num_cu,
                                                                                 (4ull << 30) / mc.lds_per_wave));
                        gws_grid = launch_grid((uint32_t)x.n, 1, cap);
                        mc.scratch = (uint8_t*)tmp.alloc((size_t)gws_grid * mc.lds_per_wave);
                    }
                    if (x.n != n) {
                        uint32_t* d_r = dalloc<uint32_t>(tmp, x.rb.size() + x.rpre.size());
                        up(d_r, x.rb.data(), x.rb.size(), st);
                        up(d_r + x.rb.size(), x.rpre.data(), x.rpre.size(), st);
                        mc.rb = d_r; mc.rpre = d_r + x.rb.size(); mc.nrange = (uint32_t)x.rb.size();
                        mc.n_items = (uint32_t)x.n;
                    }
                    if (!emit)
                        mc.spread = coprime_stride((uint32_t)((x.n + kCountSpreadBlock - 1) / kCountSpreadBlock));
                    // sub-wave teams for small states, when the wave's workspaces fit a workgroup
                    uint32_t lanes = 64;
                    if (team == 1 && !mc.scratch) {
                        const uint32_t want = ctx_->join_lanes ? ctx_->join_lanes : ctx_->tune.join_lanes[kc];
                        if (want < 64 && (uint64_t)mc.lds_per_wave * (64 / want) <= ctx_->max_lds) lanes = want;
                    }
                    const uint32_t wpb_m = mc.scratch ? 1u : waves_per_block(mc.lds_per_wave * (64 / lanes), ctx_->max_lds);
                    // up to 16 workgroups per CU (about three times what stays resident): measured
                    // best of 5, 8, 16, 32 and 64 on the batched workload (0.8 ms per pass over 8)
                    const uint32_t gcap = mc.scratch ? gws_grid : ctx_->num_cu * ctx_->tune.join_grid[kc];
                    timed(ctx_, st, name, (uint64_t)((double)bytes * x.n / n), x.n,
                          [&] { return launch_match(mc, emit, gcap, team == 1 ? wpb_m : 1, team, st, lanes); }, kc);
                }
            };
            HIP_TRY(launch_fill(cnt, 0, sizeof(uint32_t) * (n + 1), st));
            launch_classes(false, mcls_count, "k_match_count", match_bytes + n * 4ull);
            scan_u32_to_u64(tmp, cnt, moff, n, st);
            HIP_TRY(launch_run_totals(moff, ma.state_edges, cnt, F_->v.ne, d_sbeg_, d_send_, cap_slots_, d_tot_,
                                      d_run_edges_, d_run_pe_, st));
            // per-slot child edges, totals and parent edges are adjacent
            // ([edges cap][totals cap + 1][parent edges cap]): one read
            std::vector<uint64_t> rb(3ull * cap_slots_ + 1);
            t_launch = ms_since(t_step);
            if (mode_ == EngineMode::Capture) {
                // the recorded decision; each replay checks the device's totals against it
                if (plan_rb_ >= plan_->rb.size() || plan_->rb[plan_rb_].size() != rb.size())
                    throw Error{HGX_ERR_INVALID, "the pass being captured left its recording"};
                rb = plan_->rb[plan_rb_++];
                uint64_t* want = dalloc<uint64_t>(tmp, cap_slots_);
                up(want, rb.data() + cap_slots_, cap_slots_, st);
                HIP_TRY(launch_plan_check(d_tot_, want, d_budget_, cap_slots_, d_check_, st));
            } else {
                d2h(rb.data(), d_run_edges_, rb.size(), st);
                HIP_TRY(hipStreamSynchronize(st));
                if (mode_ == EngineMode::Record) plan_->rb.push_back(rb);
            }
            t_count = ms_since(t_step);
            const uint64_t* htot = rb.data() + cap_slots_;
            const uint64_t* hpe = htot + cap_slots_ + 1;
            h_edges.assign(rb.begin(), rb.begin() + cap_slots_);
            bool masked = false;
            for (uint32_t i = 0; i < nF; ++i) {
                RunRec& r = *F_runs_[i];
                if (!h_active[r.slot]) continue;
                const uint64_t t = htot[r.slot];
                if (r.hp->cap && r.states + t > r.hp->cap) {
                    r.truncated = true; r.done = true; h_active[r.slot] = 0; masked = true;
                } else if (t == 0) {
                    r.done = true; h_active[r.slot] = 0;
                } else {
                    tot[i] = t;
                    parent_edges += hpe[r.slot];          // parents the expansion reads
                    for (auto& R : r.hp->rules) K = std::max<uint32_t>(K, R.n_lhs);
                    if (!r.hp->cap) {
                        // an uncapped request is charged what it holds once this generation
                        // is built, when that exceeds its admission estimate, so a request that
                        // grows keeps the queue from being admitted past the budget
                        const uint64_t f = footprint_bytes(*r.hp, r.states + t);
                        if (f > r.footprint) { live_footprint_ += f - r.footprint; r.footprint = f; }
                    }
                }
            }
            // HBM check before anything of the generation is written: what this generation will
            // take (match list, child state and event records, child edges, the step's scratch)
            // against what the budget still hands out. When it does not fit, the requests with
            // the largest children stop here -- they finish with the generations they completed,
            // status HGX_ERR_CAPACITY (the reference's capacity warning with a partial result,
            // gpu/src/evolve.cu:573-700) -- and the others build this generation unchanged, so
            // their results are what they would be alone.
            if (mode_ != EngineMode::Capture) {
                // per run: (bytes of its children's records in new arenas, frontier index); the
                // match list and scratch grow arenas that already hold a chunk
                std::vector<std::pair<uint64_t, uint32_t>> need;
                uint64_t fresh = 0, grow = 0;
                for (uint32_t i = 0; i < nF; ++i) {
                    if (!tot[i]) continue;
                    const RunHost& h = *F_runs_[i]->hp;
                    uint32_t k = 1;
                    for (auto& R : h.rules) k = std::max<uint32_t>(k, R.n_lhs);
                    const uint64_t rec = tot[i] * (57ull + 8ull * h.steps) +
                                         h_edges[F_runs_[i]->slot] * (4ull * h.max_arity + 2ull);
                    const uint64_t scr = tot[i] * (8ull + 2ull * k + 16ull);
                    need.push_back({rec + scr, i});
                    fresh += rec;
                    grow += scr;
                }
                // arenas take chunks of at least 64 MiB: the child generation opens two (state
                // and event records, edges); one more chunk covers what the rest adds
                const uint64_t chunk = (uint64_t)64 << 20;
                auto fits = [&] {
                    const uint64_t want = std::max(fresh, 2 * chunk) + grow + chunk;
                    return pool_->in_use() + want <= pool_->budget();
                };
                if (!fits()) {
                    std::sort(need.begin(), need.end(), [](auto& x, auto& y) { return x.first > y.first; });
                    for (auto& [b, i] : need) {
                        if (fits()) break;
                        RunRec& r = *F_runs_[i];
                        r.done = true;
                        r.truncated = true;
                        r.status = HGX_ERR_CAPACITY;
                        r.warning = "generation " + std::to_string(r.depth + 1) + " (" + std::to_string(tot[i]) +
                                    " states) does not fit the context's HBM budget; the result holds generations 0.." +
                                    std::to_string(r.depth);
                        h_active[r.slot] = 0;
                        masked = true;
                        parent_edges -= hpe[r.slot];
                        const RunHost& h = *r.hp;
                        uint32_t k = 1;
                        for (auto& R : h.rules) k = std::max<uint32_t>(k, R.n_lhs);
                        fresh -= tot[i] * (57ull + 8ull * h.steps) + h_edges[r.slot] * (4ull * h.max_arity + 2ull);
                        grow -= tot[i] * (8ull + 2ull * k + 16ull);
                        tot[i] = 0;
                    }
                }
            }
            if (masked) {
                up(d_active_, h_active.data(), cap_slots_, st);
                HIP_TRY(launch_mask_counts(F_->v.run, d_active_, n, cnt, st));
                scan_u32_to_u64(tmp, cnt, moff, n, st);
            }
            for (uint32_t i = 0; i < nF; ++i) M += tot[i];
            if (M > 0xFFFFFFF0ull) throw Error{HGX_ERR_UNSUPPORTED, "more than 2^32 events in one generation"};
            if (M) {
                // uncapped maps grow ahead of the inserts
                for (uint32_t i = 0; i < nF; ++i) {
                    if (!tot[i]) continue;
                    RunRec& r = *F_runs_[i];
                    const uint64_t need = 2 * (r.states + tot[i]);
                    if (need <= r.map_cap) continue;
                    uint64_t c = r.map_cap;
                    while (c < need) c <<= 1;
                    if (c > 0x80000000ull) throw Error{HGX_ERR_UNSUPPORTED, "state map above 2^31 slots"};
                    uint64_t* keys = (uint64_t*)slab_.alloc(8 * c);
                    HIP_TRY(launch_fill(keys, 0, 8 * c, st));
                    HIP_TRY(launch_rehash(r.map_keys, r.map_cap, keys, (uint32_t)(c - 1), st));
                    slab_.free(r.map_keys, 8ull * r.map_cap);
                    r.map_keys = keys; r.map_cap = (uint32_t)c;
                    DRun& d = h_runs_[r.slot];
                    d.map_keys = keys; d.map_mask = (uint32_t)(c - 1);
                }
                mv.n = (uint32_t)M; mv.K = K;
                mv.parent = dalloc<uint32_t>(*keep, M);
                mv.rule = dalloc<uint32_t>(*keep, M);
                mv.q = dalloc<uint16_t>(*keep, M * K);
                mv.cne = dalloc<uint32_t>(tmp, M + 1);
                HIP_TRY(launch_fill(mv.cne + M, 0, sizeof(uint32_t), st));
                ma.M = mv;
                launch_classes(true, mcls_emit, "k_match_emit", match_bytes + M * (12ull + 2ull * K));
                bool thin = false;
                for (uint32_t i = 0; i < nF; ++i) thin |= tot[i] && F_runs_[i]->hp->smp.transitions();
                if (thin) sample_transitions(n, K, mv, cnt, moff, tot, h_edges, M, h_parent, st);
            }
        }
    }

    t_emit = ms_since(t_step);
    // ---- 2. admissions: queued requests join the generation built now ----
    const uint64_t soft = (uint64_t)(pool_->budget() * soft_budget_frac);
    while (!pending_.empty() && admitted.size() < admit_limit_) {
        auto r = pending_.front();
        if (!live_.empty() && live_footprint_ + r->footprint > soft) break;
        pending_.pop_front();
        admitted.push_back(r);                  // listed first: a failed admission is rolled back
        admit(r);
    }
    // requests with host inputs first, then prepared ones (device-resident inputs)
    std::stable_partition(admitted.begin(), admitted.end(),
                          [](const std::shared_ptr<RunRec>& r) { return !r->hp->staged.rules; });
    uint32_t NR = 0, NRh = 0;
    uint64_t NER = 0, NERh = 0;
    for (auto& r : admitted) {
        NR += r->hp->n_roots(); NER += r->hp->root_edges;
        if (!r->hp->staged.rules) { NRh += r->hp->n_roots(); NERh += r->hp->root_edges; }
    }
    std::vector<DRule> a_rules;                 // staging; outlives the synchronisation below
    std::vector<AdmitItem> a_items;
    if (!admitted.empty()) {
        size_t nh = 0;
        for (auto& r : admitted) if (!r->hp->staged.rules) nh += r->hp->rules.size();
        DRule* d_stage = nh ? dalloc<DRule>(tmp, nh) : nullptr;
        for (auto& r : admitted) {
            AdmitItem it{};
            if (r->hp->staged.rules) {
                it.rule_src = r->hp->staged.rules;
            } else {
                it.rule_src = d_stage + a_rules.size();
                a_rules.insert(a_rules.end(), r->hp->rules.begin(), r->hp->rules.end());
            }
            it.n_rules = (uint32_t)r->hp->rules.size();
            it.rule_dst = r->d_rules;
            it.map_keys = r->map_keys;
            it.map_cap = r->map_cap;
            it.slot = r->slot;
            a_items.push_back(it);
        }
        AdmitItem* d_items = dalloc<AdmitItem>(tmp, a_items.size());
        if (nh) up(d_stage, a_rules.data(), a_rules.size(), st);
        up(d_items, a_items.data(), a_items.size(), st);
        HIP_TRY(launch_admit(d_items, (uint32_t)a_items.size(), d_ctr_, st));
    }

    // ---- 3. the child generation: children of the frontier, then the new roots ----
    std::vector<std::shared_ptr<RunRec>> next_runs;
    std::vector<uint32_t> next_b, next_e;
    std::vector<std::shared_ptr<RunRec>> done_now;
    std::shared_ptr<DevGen> C;
    // host staging for the roots; must outlive the stream synchronisation below
    std::vector<uint32_t> r_run, r_ne, r_nextv, r_sid, r_ev;
    std::vector<uint64_t> r_eoff;
    std::vector<uint8_t> r_depth, r_ar;
    std::vector<RootCopy> r_copy;
    ExpandClass cls[kExpandClasses];            // host range lists of the expand launches
    std::vector<RunEvents> rev;                 // per-run child ranges for k_run_events
    const uint32_t nC = (uint32_t)M + NR;
    if (nC) {
        C = std::make_shared<DevGen>(pool_);
        uint32_t A = 1, lstride = 1, max_ne = 0, max_nextv = 0;
        bool keep = false;
        for (uint32_t i = 0; i < nF; ++i) {
            if (!tot[i]) continue;
            const RunRec& r = *F_runs_[i];
            A = std::max(A, r.hp->max_arity);
            lstride = std::max(lstride, r.hp->steps);
            keep |= (r.hp->flags & HGX_REC_STATES) != 0;
        }
        if (M) {
            uint32_t gain = 0, nw = 0;
            for (uint32_t i = 0; i < nF; ++i)
                if (tot[i]) { gain = std::max(gain, F_runs_[i]->hp->max_gain); nw = std::max(nw, F_runs_[i]->hp->max_new); }
            max_ne = F_->max_ne + gain;
            max_nextv = F_->max_nextv + nw;
        }
        for (auto& r : admitted) {
            A = std::max(A, r->hp->max_arity);
            lstride = std::max(lstride, r->hp->steps);
            keep |= (r->hp->flags & HGX_REC_STATES) != 0;
            max_ne = std::max(max_ne, r->hp->max_root_ne);
            max_nextv = std::max(max_nextv, r->hp->max_root_nv);
        }
        C->keep_edges = keep;
        C->max_ne = max_ne;
        C->max_nextv = max_nextv;
        GenView& v = C->v;
        v.n = nC; v.A = A; v.lstride = lstride;
        v.run = dalloc<uint32_t>(C->meta, nC);
        v.eoff = dalloc<uint64_t>(C->meta, nC + 1);
        v.ne = dalloc<uint32_t>(C->meta, nC);
        v.nextv = dalloc<uint32_t>(C->meta, nC);
        v.sid = dalloc<uint32_t>(C->meta, nC);
        v.hash = dalloc<uint64_t>(C->meta, nC);
        v.lanc = dalloc<uint64_t>(C->meta, (size_t)nC * lstride);
        v.depth = dalloc<uint8_t>(C->meta, nC);
        if (M) {
            C->ev.in_state = dalloc<uint32_t>(C->meta, M);
            C->ev.rule = dalloc<uint32_t>(C->meta, M);
            C->ev.prod = dalloc<uint64_t>(C->meta, M);
            C->ev.tr = dalloc<uint64_t>(C->meta, M);
            // with causal relations materialised: the producer generation of every consumed
            // edge, for the causal triple readback
            bool triples = false;
            for (uint32_t i = 0; i < nF; ++i)
                if (tot[i] && (F_runs_[i]->hp->flags & (HGX_REC_CAUSAL | HGX_REC_RELATIONS)) ==
                                  (HGX_REC_CAUSAL | HGX_REC_RELATIONS))
                    triples = true;
            C->ev.cps = triples ? dalloc<uint8_t>(C->meta, M * K) : nullptr;
            C->ev.cK = K;
            scan_u32_to_u64(tmp, mv.cne, v.eoff, (uint32_t)M, st);
            if (keep) {                          // consumed positions, for event edge readback
                C->cq = dalloc<uint16_t>(C->meta, M * K);
                C->cK = K;
                HIP_TRY(hipMemcpyAsync(C->cq, mv.q, sizeof(uint16_t) * M * K, hipMemcpyDeviceToDevice, st));
            }
        }
        uint64_t EC = 0;                         // child edges, from the count pass's per-run sums
        for (uint32_t i = 0; i < nF; ++i)
            if (tot[i]) EC += h_edges[F_runs_[i]->slot];
        C->n_edges = EC + NER;
        v.ev = dalloc<uint32_t>(C->edges, (EC + NER) * A);
        v.ar = dalloc<uint8_t>(C->edges, EC + NER);
        v.ps = dalloc<uint8_t>(C->edges, EC + NER);

        // per-run child ranges and event numbering
        uint32_t o = 0;
        std::vector<uint32_t> cb(nF, 0), ce(nF, 0);
        for (uint32_t i = 0; i < nF; ++i) {
            if (!tot[i]) continue;
            RunRec& r = *F_runs_[i];
            cb[i] = o; o += (uint32_t)tot[i]; ce[i] = o;
            h_runs_[r.slot].ev_base = (uint32_t)r.events;
            h_runs_[r.slot].child_begin = cb[i];
        }
        up(d_runs_, h_runs_.data(), hw_slots_, st);              // slots ever taken

        if (M) {
            // Size classes: children grouped by their request's bound on child edges, each class
            // launched with a workspace sized for it, so one large request in the mix does not
            // push every child onto the large-state workspace.
            for (uint32_t i = 0; i < nF; ++i) {
                if (!tot[i]) continue;
                const RunRec& r = *F_runs_[i];
                const uint32_t cne = r.ne_bound + r.hp->max_gain, cnv = r.nv_bound + r.hp->max_new;
                int k = 0;
                while (cne > kExpandEdges[k]) ++k;
                ExpandClass& x = cls[k];
                x.ne = std::max(x.ne, cne);
                x.nextv = std::max(x.nextv, cnv);
                x.uniform = x.uniform && r.hp->uni_arity != 0;
                x.rb.push_back(cb[i]);
                x.n += ce[i] - cb[i];
                x.rpre.push_back((uint32_t)x.n);
            }
            const uint64_t bytes = M * (8ull + 2ull * K) + parent_edges * (4ull * F_->v.A + 2) +
                                   EC * (4ull * A + 2) + M * (28ull + 8ull * lstride) + M * 24ull;
            uint16_t* tcnt = dalloc<uint16_t>(tmp, M);   // per child: causal triples, for k_run_events
            for (int k = 0; k < kExpandClasses; ++k) {
                ExpandClass& x = cls[k];
                if (!x.n) continue;
                ExpandArgs ea{};
                ea.P = F_->v; ea.C = v; ea.M = mv; ea.E = C->ev; ea.runs = d_runs_; ea.ctr = d_ctr_;
                ea.tcnt = tcnt;
                ea.max_child_ne = std::max(1u, std::min(x.ne, max_ne));
                ea.max_nextv = std::max(1u, std::min(x.nextv, max_nextv));
                ea.max_nv = std::max(1u, std::min(ea.max_nextv, ea.max_child_ne * A));
                ea.lean = x.uniform && ea.max_child_ne < 65536 ? 1u : 0u;
                ea.lds_per_wave = expand_lds_per_wave(ea.max_child_ne, ea.max_nv, ea.max_nextv, A, ea.lean != 0);
                ea.n_items = (uint32_t)x.n;
                if (x.n != M) {
                    uint32_t* d_r = dalloc<uint32_t>(tmp, x.rb.size() + x.rpre.size());
                    up(d_r, x.rb.data(), x.rb.size(), st);
                    up(d_r + x.rb.size(), x.rpre.data(), x.rpre.size(), st);
                    ea.rb = d_r; ea.rpre = d_r + x.rb.size(); ea.nrange = (uint32_t)x.rb.size();
                }
                // Team: one wave per child while a CU still holds several workspaces, else the
                // workgroup's waves share one child (a CU holds 160 KiB of LDS).
                const uint32_t team = pick_team(ea.lds_per_wave, ctx_->expand_team);
                const uint32_t gcap = ctx_->num_cu * ctx_->tune.expand_grid[k];
                uint32_t wpb_e = 1, lanes = 64;
                if (team == 1) {
                    // sub-wave teams for small children, when the wave's workspaces fit a workgroup
                    const uint32_t want = ctx_->expand_lanes ? ctx_->expand_lanes : ctx_->tune.expand_lanes[k];
                    if (want < 64 && (uint64_t)ea.lds_per_wave * (64 / want) <= ctx_->max_lds) {
                        lanes = want;
                        wpb_e = waves_per_block(ea.lds_per_wave * (64 / lanes), ctx_->max_lds);
                    } else {
                        wpb_e = place_workspace(tmp, ea.lds_per_wave, ea.n_items, ctx_->max_lds,
                                                gcap, &ea.scratch);
                    }
                } else if (ea.lds_per_wave > ctx_->max_lds) {
                    const uint32_t grid = expand_grid(ea.n_items, team, 1, gcap);
                    ea.scratch = (uint8_t*)tmp.alloc((size_t)grid * ea.lds_per_wave);
                }
                timed(ctx_, st, "k_expand", (uint64_t)((double)bytes * x.n / M), x.n,
                      [&] { return launch_expand(ea, gcap, wpb_e, team, st, lanes); }, k);
            }
            timed(ctx_, st, "k_dedup", M * 12ull + M * 8ull, M,
                  [&] { return launch_dedup(v, (uint32_t)M, d_runs_, d_ctr_, st); });
            // each run's event counters from its children's records: one add per run
            for (uint32_t i = 0; i < nF; ++i)                // items of at most 4,096 children
                for (uint64_t b = cb[i]; b < ce[i]; b += 4096)
                    rev.push_back(RunEvents{F_runs_[i]->slot, (uint32_t)b,
                                            (uint32_t)std::min<uint64_t>(ce[i], b + 4096)});
            RunEvents* d_rev = dalloc<RunEvents>(tmp, rev.size());
            up(d_rev, rev.data(), rev.size(), st);
            HIP_TRY(launch_run_events(d_rev, (uint32_t)rev.size(), C->ev, tcnt, d_ctr_, st));
        }

        // new roots: uploaded (host inputs) or copied on the device (prepared inputs), then
        // hashed and inserted into their maps
        if (NR) {
            if (NRh) {
                r_run.resize(NRh); r_ne.resize(NRh); r_nextv.resize(NRh); r_sid.resize(NRh);
                r_eoff.resize(NRh + 1); r_depth.assign(NRh, 0); r_ar.resize(NERh); r_ev.assign(NERh * A, 0);
                uint32_t s = 0;
                uint64_t e = 0;
                for (auto& r : admitted) {
                    if (r->hp->staged.rules) continue;
                    const RunHost& h = *r->hp;
                    for (uint32_t k = 0; k < h.n_roots(); ++k, ++s) {
                        r_run[s] = r->slot;
                        r_ne[s] = h.root_ne(k);
                        r_nextv[s] = h.root_nv[k];
                        r_sid[s] = k;
                        r_eoff[s] = EC + e;
                        for (uint32_t ed = h.root_eb[k]; ed < h.root_eb[k + 1]; ++ed, ++e) {
                            const uint32_t a = h.arity(ed);
                            r_ar[e] = (uint8_t)a;
                            std::memcpy(&r_ev[e * A], &h.verts[h.edge_vb[ed]], 4 * a);
                        }
                    }
                }
                r_eoff[NRh] = EC + e;
                up(v.run + M, r_run.data(), NRh, st);
                up(v.ne + M, r_ne.data(), NRh, st);
                up(v.nextv + M, r_nextv.data(), NRh, st);
                up(v.sid + M, r_sid.data(), NRh, st);
                up(v.depth + M, r_depth.data(), NRh, st);
                up(v.eoff + M, r_eoff.data(), NRh + 1, st);
                if (NERh) {
                    up(v.ar + EC, r_ar.data(), NERh, st);
                    up(v.ev + EC * A, r_ev.data(), NERh * A, st);
                    HIP_TRY(launch_fill(v.ps + EC, 0, NERh, st));
                }
            }
            if (NR > NRh) {
                uint32_t s = (uint32_t)M + NRh;
                uint64_t e = EC + NERh;
                for (auto& r : admitted) {
                    const StagedSrc& g = r->hp->staged;
                    if (!g.rules) continue;
                    RootCopy c{g.ne, g.nextv, g.eoff, g.ar, g.ev, g.A, r->hp->n_roots(), r->slot, s, e};
                    r_copy.push_back(c);
                    s += r->hp->n_roots();
                    e += r->hp->root_edges;
                }
                RootCopy* d_copy = dalloc<RootCopy>(tmp, r_copy.size());
                up(d_copy, r_copy.data(), r_copy.size(), st);
                HIP_TRY(launch_copy_roots(d_copy, (uint32_t)r_copy.size(), v, st));
            }
            uint32_t rmax_ne = 1, rmax_nv = 1;
            for (auto& r : admitted) {
                rmax_ne = std::max(rmax_ne, r->hp->max_root_ne);
                rmax_nv = std::max(rmax_nv, r->hp->max_root_nv);
            }
            WLStatesArgs w{};
            w.n = NR; w.A = A; w.state_off = v.eoff + M; w.ar = v.ar; w.ev = v.ev; w.nv = v.nextv + M;
            w.out = v.hash + M; w.run_of = v.run + M; w.runs = d_runs_; w.ctr = d_ctr_;
            w.max_ne = rmax_ne; w.max_nv = rmax_nv;
            bool roots_uniform = true;
            for (auto& r : admitted) roots_uniform = roots_uniform && r->hp->root_uni;
            uint32_t wpb = 1;
            const uint32_t team = place_wl_states(tmp, w, roots_uniform, ctx_->max_lds, ctx_->num_cu * 8,
                                                  ctx_->expand_team, &wpb);
            timed(ctx_, st, "k_wl_states", NER * (4ull * A + 1) + NR * 36ull, NR,
                  [&] { return launch_wl_states(w, ctx_->num_cu * 8, wpb, team, st); });
        }

        // branchial relation of this generation's events
        std::vector<uint64_t> h_poff;
        if (M) {
            bool any_b = false, all_b = true, any_rel = false;
            std::vector<uint8_t> bflag(cap_slots_, 0);
            for (uint32_t i = 0; i < nF; ++i) {
                if (!tot[i]) continue;
                const RunRec& r = *F_runs_[i];
                const bool b = (r.hp->flags & HGX_REC_BRANCHIAL) != 0;
                bflag[r.slot] = b ? 1 : 0;
                any_b |= b; all_b &= b;
                any_rel |= b && (r.hp->flags & HGX_REC_RELATIONS);
            }
            if (any_b) {
                uint32_t* bcnt = cnt;
                if (!all_b) {
                    bcnt = dalloc<uint32_t>(tmp, n + 1);
                    HIP_TRY(hipMemcpyAsync(bcnt, cnt, sizeof(uint32_t) * (n + 1), hipMemcpyDeviceToDevice, st));
                    uint8_t* d_bf = dalloc<uint8_t>(tmp, cap_slots_);
                    up(d_bf, bflag.data(), cap_slots_, st);
                    HIP_TRY(launch_mask_counts(F_->v.run, d_bf, n, bcnt, st));
                    if (mode_ != EngineMode::Capture) HIP_TRY(hipStreamSynchronize(st));
                }
                BranchialArgs ba{};
                ba.P = F_->v; ba.M = mv; ba.runs = d_runs_; ba.cnt = bcnt; ba.moff = moff; ba.ctr = d_ctr_;
                ba.emit = 0; ba.lds_per_wave = branchial_lds_per_wave();
                uint32_t* pcnt = nullptr;
                if (any_rel) {
                    pcnt = dalloc<uint32_t>(tmp, n + 1);
                    HIP_TRY(launch_fill(pcnt, 0, sizeof(uint32_t) * (n + 1), st));
                }
                ba.pcnt = pcnt;
                timed(ctx_, st, "k_branchial", M * (6ull + 2ull * K) + n * 12ull, n,
                      [&] { return launch_branchial(ba, ctx_->num_cu * 8, 2, st); });
                // parents above BR_MAX_EDGES edges: (edge slot, match) keys sorted into per-edge lists
                BrListArgs bl{};
                const bool large = F_->max_ne > BR_MAX_EDGES;
                if (large) {
                    const uint64_t E = F_->n_edges, NK = M * K;
                    if (E >= 0xFFFFFFFFull) throw Error{HGX_ERR_UNSUPPORTED, "branchial lists above 2^32 edges"};
                    bl.P = F_->v; bl.M = mv; bl.runs = d_runs_; bl.cnt = bcnt; bl.moff = moff;
                    bl.ctr = d_ctr_; bl.pcnt = pcnt;
                    bl.n_keys = NK;
                    bl.sentinel = E << 32;
                    bl.key_in = dalloc<uint64_t>(tmp, std::max<uint64_t>(1, NK));
                    bl.val_in = dalloc<uint32_t>(tmp, std::max<uint64_t>(1, NK));
                    uint64_t* key = dalloc<uint64_t>(tmp, std::max<uint64_t>(1, NK));
                    uint32_t* val = dalloc<uint32_t>(tmp, std::max<uint64_t>(1, NK));
                    bl.key = key; bl.val = val;
                    bl.pos = dalloc<uint32_t>(tmp, std::max<uint64_t>(1, NK));
                    bl.bc = dalloc<uint32_t>(tmp, M + 1);
                    HIP_TRY(launch_fill(bl.bc, 0, sizeof(uint32_t) * (M + 1), st));
                    int end_bit = 32;
                    while (end_bit < 64 && (E >> (end_bit - 32)) != 0) ++end_bit;   // keys <= E << 32
                    const uint64_t lbytes = NK * (8ull + 4ull) * 4 + NK * 4ull + M * (12ull + 2ull * K);
                    timed(ctx_, st, "k_branchial_lists", lbytes, M, [&] {
                        hipError_t e = launch_brl(bl, 0, st);
                        if (e == hipSuccess) {
                            sort_pairs_u64(tmp, bl.key_in, key, bl.val_in, val, NK, end_bit, st);
                            e = hipGetLastError();
                        }
                        if (e == hipSuccess) e = launch_brl(bl, 1, st);
                        if (e == hipSuccess) e = launch_brl(bl, 2, st);
                        return e;
                    });
                }
                if (any_rel) {
                    uint64_t* poff = dalloc<uint64_t>(tmp, n + 1);
                    scan_u32_to_u64(tmp, pcnt, poff, n, st);
                    h_poff.resize(n + 1);
                    if (mode_ == EngineMode::Capture) {
                        if (plan_poff_ >= plan_->poff.size() || plan_->poff[plan_poff_].size() != h_poff.size())
                            throw Error{HGX_ERR_INVALID, "the pass being captured left its recording"};
                        h_poff = plan_->poff[plan_poff_++];
                    } else {
                        d2h(h_poff.data(), poff, n + 1, st);
                        HIP_TRY(hipStreamSynchronize(st));
                        if (mode_ == EngineMode::Record) plan_->poff.push_back(h_poff);
                    }
                    const uint64_t NB = h_poff[n];
                    if (NB) {
                        C->br_pairs = dalloc<uint32_t>(C->meta, 2 * NB);
                        ba.emit = 1; ba.poff = poff; ba.pairs = C->br_pairs;
                        HIP_TRY(launch_branchial(ba, ctx_->num_cu * 8, 2, st));
                        if (large) {
                            uint64_t* bo = dalloc<uint64_t>(tmp, M + 1);
                            scan_u32_to_u64(tmp, bl.bc, bo, (uint32_t)M, st);
                            bl.bo = bo; bl.poff = poff; bl.pairs = C->br_pairs;
                            HIP_TRY(launch_brl(bl, 3, st));
                        }
                    }
                }
            }
        }

        // bookkeeping
        for (uint32_t i = 0; i < nF; ++i) {
            if (!tot[i]) continue;
            auto& r = F_runs_[i];
            if (r->hp->host_frontier()) kids.push_back(SampledKids{r.get(), cb[i], ce[i], r->events});
            Segment sg{C, cb[i], ce[i], 0, 0};
            if (!h_poff.empty()) { sg.br_b = h_poff[F_b_[i]]; sg.br_e = h_poff[F_e_[i]]; }
            r->segs.push_back(sg);
            r->states += tot[i];
            r->events += tot[i];
            r->depth += 1;
            r->steps_done = r->depth;
            r->ne_bound += r->hp->max_gain;
            r->nv_bound += r->hp->max_new;
            if (r->depth >= r->hp->steps) r->done = true;
            next_runs.push_back(r);
            next_b.push_back(cb[i]);
            next_e.push_back(ce[i]);
        }
        uint32_t s = (uint32_t)M;
        for (auto& r : admitted) {
            const uint32_t nr = r->hp->n_roots();
            if (r->hp->quotient()) kids.push_back(SampledKids{r.get(), s, s + nr, 0, true});
            r->segs.push_back(Segment{C, s, s + nr, 0, 0});
            r->states = nr;
            if (r->hp->steps == 0) r->done = true;
            next_runs.push_back(r);
            next_b.push_back(s);
            next_e.push_back(s + nr);
            s += nr;
        }
    }

    t_child = ms_since(t_step);
    // ---- 4. finished runs, slot recycling, the new frontier ----
    for (auto* list : {&F_runs_, &next_runs})
        for (auto& r : *list)
            if (r->done && !r->finishing && live_.count(r->id)) {
                r->finishing = true;
                done_now.push_back(r);
            }
    std::vector<uint32_t> releasable;
    releasable.swap(release_next_);
    finalize(done_now);                        // synchronises the stream (not when capturing)
    if (done_now.empty() && mode_ != EngineMode::Capture) HIP_TRY(hipStreamSynchronize(st));
    t_fin = ms_since(t_step);
    prof_resolve(ctx_);
    for (uint32_t s : releasable) free_slots_.push_back(s);
    if (F_ && !F_->keep_edges) F_->edges.release();
    F_ = C;
    sample_frontier(kids, C ? C->v.n : 0, mv, M, h_parent, st);
    if (M && !thinned_) {
        prev_.ar = std::move(keep);
        prev_.n = (uint32_t)M; prev_.K = K;
        prev_.parent = mv.parent; prev_.rule = mv.rule; prev_.q = mv.q; prev_.moff = moff;
    } else {
        prev_ = PrevMatches{};
    }
    F_runs_.swap(next_runs);
    F_b_.swap(next_b);
    F_e_.swap(next_e);
    // a frontier holding only finished runs is dead
    bool live_in_F = false;
    for (auto& r : F_runs_) live_in_F |= !r->done;
    if (!live_in_F) {
        if (F_ && !F_->keep_edges) F_->edges.release();
        F_.reset();
        prev_ = PrevMatches{};
        F_expand_.clear();
        F_runs_.clear(); F_b_.clear(); F_e_.clear();
        for (uint32_t s2 : release_next_) free_slots_.push_back(s2);
        release_next_.clear();
    }
    if (trace_on())
        std::fprintf(stderr, "[hgx] step: frontier %u states, %llu children, %u admitted, %zu finished | "
                             "count launched %.3f ms, +sync %.3f, +emit %.3f, +child gen %.3f, +finish %.3f, "
                             "total %.3f ms\n",
                     n, (unsigned long long)M, (unsigned)admitted.size(), done_now.size(), t_launch, t_count,
                     t_emit, t_child, t_fin, ms_since(t_step));
    return live_or_queued();
}

// ------------------------------------------------------------------------------------------
// sampling (hgx_sampling): the GPU engine's decisions, generation by generation
// ------------------------------------------------------------------------------------------

// The exact (IR) hash and canonical edge ranks of a request's frontier states [b, e), computed
// on the host from the generation's edges (the reference's device computes the same with its IR
// canonicalizer when sampling, gpu/src/ir_canon.cu:152-280).
void frontier_ir(const DevGen& G, uint32_t b, uint32_t e, hipStream_t st, RunRec& r) {
    const uint32_t n = e - b, A = G.v.A;
    std::vector<uint64_t> eo(n);
    std::vector<uint32_t> ne(n);
    d2h(eo.data(), G.v.eoff + b, n, st);
    d2h(ne.data(), G.v.ne + b, n, st);
    HIP_TRY(hipStreamSynchronize(st));
    uint64_t E = 0;
    for (uint32_t k = 0; k < n; ++k) E += ne[k];
    std::vector<uint8_t> ar(E);
    std::vector<uint32_t> ev(E * A);
    // states of one request are contiguous in the generation, their edges too
    if (E) {
        const uint64_t e0 = eo[0];
        d2h(ar.data(), G.v.ar + e0, E, st);
        d2h(ev.data(), G.v.ev + e0 * A, E * A, st);
        HIP_TRY(hipStreamSynchronize(st));
    }
    std::vector<uint64_t> soff(n + 1, 0);
    std::vector<uint32_t> eoff(E + 1, 0), verts;
    verts.reserve(E * 2);
    r.smp_roff.assign(n + 1, 0);
    uint64_t x = 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (ne[k] && eo[k] != eo[0] + x) throw Error{HGX_ERR_DEVICE, "a request's states are not contiguous"};
        soff[k] = x;
        r.smp_roff[k] = (uint32_t)x;
        for (uint32_t j = 0; j < ne[k]; ++j, ++x) {
            for (uint32_t t = 0; t < ar[x]; ++t) verts.push_back(ev[x * A + t]);
            eoff[x + 1] = (uint32_t)verts.size();
        }
    }
    soff[n] = x;
    r.smp_roff[n] = (uint32_t)x;
    r.smp_gen = &G;
    r.smp_ir.assign(n, 0);
    r.smp_rank.assign(std::max<uint64_t>(E, 1), 0);
    if (n && hgx_ir_rank_batch(nullptr, soff.data(), n, eoff.data(), verts.empty() ? nullptr : verts.data(),
                               r.smp_ir.data(), r.smp_rank.data()) != HGX_OK)
        throw Error{HGX_ERR_DEVICE, "exact hashes of a sampled frontier failed"};
}

void Engine::sample_transitions(uint32_t n, uint32_t K, MatchView& mv, uint32_t* cnt, uint64_t* moff,
                                std::vector<uint64_t>& tot, std::vector<uint64_t>& h_edges, uint64_t& M,
                                std::vector<uint32_t>& h_parent, hipStream_t st) {
    const uint64_t M0 = M;
    h_parent.resize(M0);
    std::vector<uint32_t> rule(M0), cne(M0);
    std::vector<uint16_t> q(M0 * K);
    std::vector<uint64_t> off(n + 1);
    d2h(h_parent.data(), mv.parent, M0, st);
    d2h(rule.data(), mv.rule, M0, st);
    d2h(q.data(), mv.q, M0 * K, st);
    d2h(cne.data(), mv.cne, M0, st);
    d2h(off.data(), moff, n + 1, st);
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<uint8_t> keep(M0, 1);
    bool dropped = false;
    const uint32_t nF = (uint32_t)F_runs_.size();
    std::vector<std::pair<uint64_t, uint64_t>> grp;
    for (uint32_t i = 0; i < nF; ++i) {
        if (!tot[i]) continue;
        RunRec& r = *F_runs_[i];
        const SampleHost& S = r.hp->smp;
        if (!S.transitions()) continue;
        const uint64_t c0 = off[F_b_[i]], c1 = off[F_e_[i]];
        std::vector<uint64_t> key(c1 - c0);
        uint32_t ranks[HGX_MAX_PATTERN_EDGES];
        for (uint64_t c = c0; c < c1; ++c) {
            const uint32_t k = h_parent[c] - F_b_[i];
            if (k >= r.smp_ir.size() || rule[c] >= r.hp->rules.size())
                throw Error{HGX_ERR_DEVICE, "a sampled match outside its request"};
            const uint32_t nl = r.hp->rules[rule[c]].n_lhs;
            for (uint32_t t = 0; t < nl; ++t) ranks[t] = r.smp_rank[r.smp_roff[k] + q[c * K + t]];
            key[c - c0] = transition_key(r.smp_ir[k], rule[c], ranks, nl);
        }
        if (S.per_state_rule) {
            // a state's matches are contiguous: per (state, rule), the k smallest ranks survive
            // (ties in event order; gpu/src/match.cu:229-256)
            for (uint64_t a = c0; a < c1;) {
                uint64_t z = a;
                while (z < c1 && h_parent[z] == h_parent[a]) ++z;
                for (uint32_t ru = 0; ru < r.hp->rules.size(); ++ru) {
                    grp.clear();
                    for (uint64_t c = a; c < z; ++c)
                        if (rule[c] == ru) grp.push_back({transition_rank(key[c - c0], S.seed), c});
                    if (grp.size() <= S.per_state_rule) continue;
                    std::sort(grp.begin(), grp.end());
                    for (size_t x = S.per_state_rule; x < grp.size(); ++x) keep[grp[x].second] = 0;
                }
                a = z;
            }
        }
        if (S.draws())                          // the rate draw (gpu/src/rewrite.cu:255-274)
            for (uint64_t c = c0; c < c1; ++c)
                if (keep[c] && !transition_survives(key[c - c0], S.seed, S.rate_for(rule[c]))) keep[c] = 0;
        for (uint64_t c = c0; c < c1; ++c) dropped |= !keep[c];
    }
    if (!dropped) return;
    // the survivors, in order: match list, per-state counts and offsets, per-run totals
    std::vector<uint32_t> p2, r2, cne2, cnt2(n + 1, 0);
    std::vector<uint16_t> q2;
    std::vector<uint64_t> off2(n + 1, 0);
    p2.reserve(M0); r2.reserve(M0); cne2.reserve(M0 + 1); q2.reserve(M0 * K);
    for (uint32_t st_ = 0; st_ < n; ++st_) {
        off2[st_] = p2.size();
        for (uint64_t c = off[st_]; c < off[st_ + 1]; ++c) {
            if (!keep[c]) continue;
            p2.push_back(h_parent[c]);
            r2.push_back(rule[c]);
            cne2.push_back(cne[c]);
            q2.insert(q2.end(), q.begin() + c * K, q.begin() + (c + 1) * K);
            ++cnt2[st_];
        }
    }
    off2[n] = p2.size();
    const uint64_t M1 = p2.size();
    for (uint32_t i = 0; i < nF; ++i) {
        if (!tot[i]) continue;
        RunRec& r = *F_runs_[i];
        const uint64_t a = off2[F_b_[i]], z = off2[F_e_[i]];
        tot[i] = z - a;
        uint64_t edges = 0;
        for (uint64_t c = a; c < z; ++c) edges += cne2[c];
        h_edges[r.slot] = edges;
        if (!tot[i]) r.done = true;             // no transition of the generation survived
    }
    cne2.push_back(0);
    if (M1) {
        up(mv.parent, p2.data(), M1, st);
        up(mv.rule, r2.data(), M1, st);
        up(mv.q, q2.data(), M1 * K, st);
    }
    up(mv.cne, cne2.data(), M1 + 1, st);
    up(cnt, cnt2.data(), n + 1, st);
    up(moff, off2.data(), n + 1, st);
    M = M1;
    mv.n = (uint32_t)M1;
    h_parent.swap(p2);
    thinned_ = true;
}

void Engine::sample_frontier(const std::vector<SampledKids>& kids, uint32_t nC, const MatchView& mv, uint64_t M,
                             std::vector<uint32_t>& h_parent, hipStream_t st) {
    F_expand_.clear();
    bool any = false;
    for (const SampledKids& k : kids) any |= k.r->depth < k.r->hp->steps;
    if (!any || !F_) return;
    F_expand_.assign(nC, 1);
    bool masked = false;
    for (const SampledKids& k : kids) {
        RunRec& r = *k.r;
        const SampleHost& S = r.hp->smp;
        const uint32_t step = r.depth;              // the new states' step
        if (step >= r.hp->steps) continue;          // the last step: nothing expands them
        const bool q = r.hp->quotient();
        if (q) frontier_ir(*F_, k.cb, k.ce, st, r); // exact hashes (and ranks, for the next join)
        const bool merge_roots = (r.hp->flags & HGX_OPT_QUOTIENT_ROOTS) != 0;
        const bool smp = !k.roots && S.frontier();
        if (smp && S.per_parent && h_parent.size() != M) {
            h_parent.resize(M);
            d2h(h_parent.data(), mv.parent, M, st);
            HIP_TRY(hipStreamSynchronize(st));
        }
        uint64_t at_step = 0;
        std::unordered_map<uint32_t, uint64_t> at_parent;
        const uint64_t n_roots = r.hp->n_roots();
        for (uint32_t c = k.cb; c < k.ce; ++c) {
            const uint64_t sid = k.roots ? (c - k.cb) : n_roots + k.ev_base + (c - k.cb);
            bool go = true;
            // exploration.hpp:66-92: the first state of an exact class is expanded, later ones
            // are not (a hash of 0 -- the canonicalizer gave up -- is no class: kept); every root
            // is expanded unless isomorphic roots merge (persistent.cu:122-135)
            if (q) {
                const uint64_t h = r.smp_ir[c - k.cb];
                if (h && !r.q_seen.insert(h).second && (!k.roots || merge_roots)) go = false;
            }
            // exploration.hpp:93-121: the coin, then the step's bound, then the parent's
            if (go && smp && S.thr != 0xFFFFFFFFu)
                go = S.thr != 0u && (uint32_t)splitmix(S.coin_seed ^ ((uint64_t)step << 32) ^ sid) < S.thr;
            if (go && smp && S.per_step) go = at_step++ < S.per_step;
            if (go && smp && S.per_parent) go = at_parent[h_parent[c]]++ < S.per_parent;
            F_expand_[c] = go ? 1 : 0;
            masked |= !go;
        }
    }
    if (!masked) F_expand_.clear();
}

// ------------------------------------------------------------------------------------------
// result readback
// ------------------------------------------------------------------------------------------

struct RunTables {
    std::vector<uint64_t> hash;
    std::vector<uint32_t> step, creator;
    std::vector<uint32_t> ev_in, ev_rule, ev_step;
    std::vector<uint64_t> ev_prod, ev_tr;
};

static RunTables run_tables(const RunRec& R) {
    RunTables T;
    const uint64_t ns = R.states;
    const uint32_t nroots = R.hp->n_roots();
    const uint64_t ne = ns >= nroots ? ns - nroots : 0;
    T.hash.resize(ns); T.step.resize(ns); T.creator.assign(ns, 0xFFFFFFFFu);
    T.ev_in.resize(ne); T.ev_rule.resize(ne); T.ev_step.resize(ne); T.ev_prod.resize(ne); T.ev_tr.resize(ne);
    for (size_t k = 0; k < R.segs.size(); ++k) {
        const Segment& sg = R.segs[k];
        const uint32_t m = sg.e - sg.b;
        if (!m) continue;
        std::vector<uint64_t> h(m);
        std::vector<uint32_t> sid(m);
        std::vector<uint8_t> dep(m);
        d2h(h.data(), sg.g->v.hash + sg.b, m, R.st);
        d2h(sid.data(), sg.g->v.sid + sg.b, m, R.st);
        d2h(dep.data(), sg.g->v.depth + sg.b, m, R.st);
        std::vector<uint32_t> in, rule;
        std::vector<uint64_t> prod, tr;
        const bool events = k > 0;                  // the first segment holds the roots
        if (events) {
            in.resize(m); rule.resize(m); prod.resize(m); tr.resize(m);
            d2h(in.data(), sg.g->ev.in_state + sg.b, m, R.st);
            d2h(rule.data(), sg.g->ev.rule + sg.b, m, R.st);
            d2h(prod.data(), sg.g->ev.prod + sg.b, m, R.st);
            d2h(tr.data(), sg.g->ev.tr + sg.b, m, R.st);
        }
        HIP_TRY(hipStreamSynchronize(R.st));
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t s = sid[i];
            if (s >= ns) throw Error{HGX_ERR_DEVICE, "state id out of range in readback"};
            T.hash[s] = h[i];
            T.step[s] = dep[i];
            if (events) {
                const uint32_t e = s - nroots;
                T.creator[s] = e;
                T.ev_in[e] = in[i]; T.ev_rule[e] = rule[i]; T.ev_step[e] = dep[i];
                T.ev_prod[e] = prod[i]; T.ev_tr[e] = tr[i];
            }
        }
    }
    return T;
}

}  // namespace hgx

using namespace hgx;

static hgx_result* make_result(std::shared_ptr<RunRec> r) {
    auto* out = new hgx_result();
    out->run = r;
    out->device_ns = (uint64_t)(r->device_ms * 1e6);
    out->latency_ns = (uint64_t)std::chrono::duration<double, std::nano>(r->t_done - r->t_submit).count();
    return out;
}

extern "C" int hgx_evolve_batch(hgx_ctx* ctx, const hgx_run_desc* runs, uint32_t n, hgx_result** results) {
    if (!ctx || (!runs && n) || (!results && n)) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        const auto t0 = Clock::now();
        std::vector<RunHost> hs;
        hs.reserve(n);
        for (uint32_t i = 0; i < n; ++i) hs.push_back(parse_run(&runs[i]));
        if (trace_on()) std::fprintf(stderr, "[hgx] batch: parsed %u requests in %.3f ms\n", n, ms_since(t0));
        Engine E(ctx);
        E.soft_budget_frac = 0.8;
        E.reserve_slots(n + 1);
        for (uint32_t i = 0; i < n; ++i) E.submit(std::make_shared<const RunHost>(std::move(hs[i])), i + 1);
        while (E.live_or_queued()) E.step();
        for (uint32_t i = 0; i < n; ++i) results[i] = make_result(E.take(i + 1));
        if (trace_on()) std::fprintf(stderr, "[hgx] batch: total %.3f ms\n", ms_since(t0));
        return HGX_OK;
    });
}

// ---- prepared batches: inputs parsed and staged in device memory once --------------------

struct hgx_batch {
    hgx_ctx* ctx = nullptr;
    std::vector<std::shared_ptr<const RunHost>> runs;
    std::unique_ptr<Arena> mem;                // device copies of rules and initial states
};

extern "C" int hgx_batch_prepare(hgx_ctx* ctx, const hgx_run_desc* runs, uint32_t n, hgx_batch** out) {
    if (!ctx || (!runs && n) || !out) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        std::vector<RunHost> hs;
        hs.reserve(n);
        for (uint32_t i = 0; i < n; ++i) hs.push_back(parse_run(&runs[i]));
        uint64_t nrules = 0, nroots = 0, nedges = 0;
        uint32_t A = 1;
        for (auto& h : hs) {
            nrules += h.rules.size(); nroots += h.n_roots(); nedges += h.root_edges;
            for (uint64_t e = 0; e < h.root_edges; ++e) A = std::max(A, h.arity(e));
        }
        std::vector<DRule> rules;
        rules.reserve(nrules);
        std::vector<uint32_t> ne, nv, ev(std::max<uint64_t>(1, nedges * A), 0);
        std::vector<uint64_t> eoff;
        std::vector<uint8_t> ar(std::max<uint64_t>(1, nedges), 0);
        std::vector<uint64_t> rule_b, root_b;
        uint64_t e = 0;
        for (auto& h : hs) {
            rule_b.push_back(rules.size());
            root_b.push_back(ne.size());
            rules.insert(rules.end(), h.rules.begin(), h.rules.end());
            for (uint32_t k = 0; k < h.n_roots(); ++k) {
                ne.push_back(h.root_ne(k));
                nv.push_back(h.root_nv[k]);
                eoff.push_back(e);
                for (uint32_t ed = h.root_eb[k]; ed < h.root_eb[k + 1]; ++ed, ++e) {
                    const uint32_t a = h.arity(ed);
                    ar[e] = (uint8_t)a;
                    std::memcpy(&ev[e * A], &h.verts[h.edge_vb[ed]], 4 * a);
                }
            }
        }
        eoff.push_back(e);
        auto b = std::make_unique<hgx_batch>();
        b->ctx = ctx;
        b->mem = std::make_unique<Arena>(ctx->pool.get(), (size_t)1 << 20);
        const hipStream_t st = ctx->stream;
        DRule* d_rules = dalloc<DRule>(*b->mem, std::max<uint64_t>(1, nrules));
        uint32_t* d_ne = dalloc<uint32_t>(*b->mem, std::max<uint64_t>(1, nroots));
        uint32_t* d_nv = dalloc<uint32_t>(*b->mem, std::max<uint64_t>(1, nroots));
        uint64_t* d_eoff = dalloc<uint64_t>(*b->mem, nroots + 1);
        uint8_t* d_ar = dalloc<uint8_t>(*b->mem, ar.size());
        uint32_t* d_ev = dalloc<uint32_t>(*b->mem, ev.size());
        if (nrules) h2d(d_rules, rules.data(), nrules, st);
        if (nroots) { h2d(d_ne, ne.data(), nroots, st); h2d(d_nv, nv.data(), nroots, st); }
        h2d(d_eoff, eoff.data(), nroots + 1, st);
        h2d(d_ar, ar.data(), ar.size(), st);
        h2d(d_ev, ev.data(), ev.size(), st);
        HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < n; ++i) {
            StagedSrc& sg = hs[i].staged;
            sg.rules = d_rules + rule_b[i];
            sg.ne = d_ne + root_b[i];
            sg.nextv = d_nv + root_b[i];
            sg.eoff = d_eoff + root_b[i];
            sg.ar = d_ar;
            sg.ev = d_ev;
            sg.A = A;
            b->runs.push_back(std::make_shared<const RunHost>(std::move(hs[i])));
        }
        *out = b.release();
        return HGX_OK;
    });
}

extern "C" int hgx_batch_evolve(hgx_ctx* ctx, const hgx_batch* b, hgx_result** results) {
    if (!ctx || !b || b->ctx != ctx || (!results && !b->runs.empty())) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        const uint32_t n = (uint32_t)b->runs.size();
        const auto t0 = Clock::now();
        double t_setup = 0, t_steps = 0, t_results = 0;
        {
            Engine E(ctx);
            E.soft_budget_frac = 0.8;
            E.reserve_slots(n + 1);
            for (uint32_t i = 0; i < n; ++i) E.submit(b->runs[i], i + 1);
            t_setup = ms_since(t0);
            while (E.live_or_queued()) E.step();
            t_steps = ms_since(t0);
            for (uint32_t i = 0; i < n; ++i) results[i] = make_result(E.take(i + 1));
            t_results = ms_since(t0);
        }
        if (trace_on())
            std::fprintf(stderr, "[hgx] batch: %u requests | setup %.3f ms, +steps %.3f, +results %.3f, "
                                 "+engine release %.3f\n", n, t_setup, t_steps, t_results, ms_since(t0));
        return HGX_OK;
    });
}

uint32_t hgx_batch_size(const hgx_batch* b) { return b ? (uint32_t)b->runs.size() : 0u; }

// ---- a prepared batch's pass recorded into one graph --------------------------------------

struct hgx_replay {
    hgx_ctx* ctx = nullptr;
    const hgx_batch* batch = nullptr;
    std::unique_ptr<Arena> mem;               // the replay check word
    uint32_t* d_check = nullptr;
    std::unique_ptr<Engine> E;                // the captured engine: its runs, generations, maps
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<Chunk> held;                  // every buffer the captured pass touched
    uint32_t n = 0;
    ~hgx_replay() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        E.reset();
        for (auto& c : held) ctx->pool->put(c);
        mem.reset();
    }
};

extern "C" int hgx_batch_capture(hgx_ctx* ctx, const hgx_batch* b, hgx_replay** out) {
    if (!ctx || !b || b->ctx != ctx || !out) return HGX_ERR_INVALID;
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        const uint32_t n = (uint32_t)b->runs.size();
        if (!n) throw Error{HGX_ERR_INVALID, "the batch is empty"};
        for (const auto& h : b->runs)
            if (h->smp.on || h->quotient())
                throw Error{HGX_ERR_UNSUPPORTED, "sampled requests and quotient exploration take host decisions "
                                                 "every generation; they are not recorded into graphs"};
        hipStream_t st = ctx->stream;
        auto plan = std::make_shared<PassPlan>();
        {   // 1. an eager pass keeps every host decision
            Engine E(ctx);
            E.set_mode(EngineMode::Record, plan, nullptr);
            E.soft_budget_frac = 0.8;
            E.reserve_slots(n + 1);
            for (uint32_t i = 0; i < n; ++i) E.submit(b->runs[i], i + 1);
            while (E.live_or_queued()) E.step();
            for (uint32_t i = 0; i < n; ++i) {
                auto r = E.take(i + 1);
                if (!r || r->status != HGX_OK)
                    throw Error{HGX_ERR_CAPACITY, "the batch does not fit the budget in one pass; it cannot be captured"};
            }
        }
        // 2. the same pass recorded into a graph, its decisions taken from the plan
        auto rp = std::make_unique<hgx_replay>();
        rp->ctx = ctx;
        rp->batch = b;
        rp->n = n;
        rp->mem = std::make_unique<Arena>(ctx->pool.get());
        rp->d_check = dalloc<uint32_t>(*rp->mem, 64);
        rp->E = std::make_unique<Engine>(ctx);
        rp->E->set_mode(EngineMode::Capture, plan, rp->d_check);
        rp->E->soft_budget_frac = 0.8;
        rp->E->reserve_slots(n + 1);
        for (uint32_t i = 0; i < n; ++i) rp->E->submit(b->runs[i], i + 1);
        HIP_TRY(hipStreamSynchronize(st));
        struct ProfOff {                       // no event timing inside the graph
            hgx_ctx* c; bool was;
            ~ProfOff() { c->prof.on = was; }
        } prof_off{ctx, ctx->prof.on};
        ctx->prof.on = false;
        ctx->pool->hold(true);
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        try {
            HIP_TRY(launch_fill(rp->d_check, 0, sizeof(uint32_t), st));
            while (rp->E->live_or_queued()) rp->E->step();
        } catch (...) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(st, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            ctx->pool->hold(false);
            rp->held = ctx->pool->take_held();
            throw;
        }
        const hipError_t e = hipStreamEndCapture(st, &rp->graph);
        ctx->pool->hold(false);
        rp->held = ctx->pool->take_held();
        if (e != hipSuccess) throw Error{HGX_ERR_DEVICE, std::string("stream capture: ") + hipGetErrorString(e)};
        HIP_TRY(hipGraphInstantiate(&rp->exec, rp->graph, nullptr, nullptr, 0));
        if (rp->E->live_or_queued() || rp->E->finished.size() != n)
            throw Error{HGX_ERR_DEVICE, "the captured pass did not finish every request"};
        *out = rp.release();
        return HGX_OK;
    });
}

extern "C" int hgx_replay_launch(hgx_ctx* ctx, hgx_replay* rp, hgx_result** results) {
    if (!ctx || !rp || rp->ctx != ctx || !results) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        const auto t0 = Clock::now();
        HIP_TRY(hipGraphLaunch(rp->exec, ctx->stream));
        uint32_t check = 0;
        HIP_TRY(hipMemcpyAsync(&check, rp->d_check, sizeof(check), hipMemcpyDeviceToHost, ctx->stream));
        rp->E->resolve_counts();                      // synchronises the stream
        if (check) throw Error{HGX_ERR_DEVICE, "a replay's totals left its recording"};
        const auto t1 = Clock::now();
        for (uint32_t i = 0; i < rp->n; ++i) {
            auto r = rp->E->finished.at(i + 1);
            r->t_submit = t0;
            r->t_admit = t0;
            r->t_done = t1;
            r->device_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            results[i] = make_result(r);
        }
        return HGX_OK;
    });
}

extern "C" void hgx_replay_release(hgx_replay* rp) {
    if (!rp) return;
    hgx_ctx* ctx = rp->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    delete rp;
}

extern "C" void hgx_batch_release(hgx_batch* b) {
    if (!b) return;
    hgx_ctx* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    delete b;
}

extern "C" int hgx_request_submit(hgx_ctx* ctx, const hgx_run_desc* run, uint64_t* request_id) {
    if (!ctx || !run || !request_id) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        RunHost h = parse_run(run);
        if (!ctx->serve) { HIP_TRY(hipSetDevice(ctx->device)); ctx->serve = std::make_unique<Engine>(ctx); }
        const uint64_t id = ctx->next_request_id++;
        ctx->serve->submit(std::make_shared<const RunHost>(std::move(h)), id);
        *request_id = id;
        return HGX_OK;
    });
}

extern "C" int hgx_serve_step(hgx_ctx* ctx, uint32_t* live) {
    if (!ctx) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        uint32_t l = 0;
        if (ctx->serve) { HIP_TRY(hipSetDevice(ctx->device)); l = ctx->serve->step(); }
        if (live) *live = l;
        return HGX_OK;
    });
}

extern "C" int hgx_request_poll(hgx_ctx* ctx, uint64_t id, hgx_result** out) {
    if (!ctx || !out) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    *out = nullptr;
    if (!ctx->serve || !ctx->serve->known(id)) { ctx->err = "unknown request id"; return HGX_ERR_INVALID; }
    auto r = ctx->serve->take(id);
    if (!r) return HGX_ERR_NOT_READY;
    *out = make_result(r);
    return HGX_OK;
}

extern "C" int hgx_request_wait(hgx_ctx* ctx, uint64_t id, hgx_result** out) {
    if (!ctx || !out) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        *out = nullptr;
        if (!ctx->serve || !ctx->serve->known(id)) { ctx->err = "unknown request id"; return HGX_ERR_INVALID; }
        HIP_TRY(hipSetDevice(ctx->device));
        // every step finishes or advances something (a generation that does not fit the budget
        // finishes the requests it held), so this ends; the bound guards against a lost id
        while (true) {
            auto r = ctx->serve->take(id);
            if (r) { *out = make_result(r); return HGX_OK; }
            if (!ctx->serve->known(id) || ctx->serve->step() == 0) {
                auto late = ctx->serve->take(id);
                if (late) { *out = make_result(late); return HGX_OK; }
                ctx->err = "request " + std::to_string(id) + " is no longer queued, running or finished";
                return HGX_ERR_INVALID;
            }
        }
    });
}

extern "C" int hgx_serve_drain(hgx_ctx* ctx, uint64_t* ids, hgx_result** results, uint32_t cap, uint32_t* n) {
    if (!ctx || !n || (cap && (!ids || !results))) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    uint32_t k = 0;
    if (ctx->serve) {
        auto& fin = ctx->serve->finished;
        for (auto it = fin.begin(); it != fin.end() && k < cap;) {
            ids[k] = it->first;
            results[k] = make_result(it->second);
            ++k;
            it = fin.erase(it);
        }
    }
    *n = k;
    return HGX_OK;
}

extern "C" uint64_t hgx_request_footprint(const hgx_run_desc* run) {
    try {
        return footprint_bytes(parse_run(run));
    } catch (...) {
        return 0;
    }
}

extern "C" int hgx_result_counts(const hgx_result* r, hgx_counts* out) {
    if (!r || !out) return HGX_ERR_INVALID;
    *out = r->run->counts;
    return HGX_OK;
}

extern "C" int hgx_result_counts_many(const hgx_result* const* rs, uint32_t n, hgx_counts* out) {
    if ((!rs || !out) && n) return HGX_ERR_INVALID;
    for (uint32_t i = 0; i < n; ++i) {
        if (!rs[i]) return HGX_ERR_INVALID;
        out[i] = rs[i]->run->counts;
    }
    return HGX_OK;
}

extern "C" int hgx_result_timing(const hgx_result* r, uint64_t* device_ns, uint64_t* latency_ns) {
    if (!r) return HGX_ERR_INVALID;
    if (device_ns) *device_ns = r->device_ns;
    if (latency_ns) *latency_ns = r->latency_ns;
    return HGX_OK;
}

extern "C" int hgx_result_status(const hgx_result* r, char* msg, size_t len) {
    if (!r) return HGX_ERR_INVALID;
    const std::string& w = r->run->warning;
    if (msg && len) {
        const size_t k = std::min(len - 1, w.size());
        std::memcpy(msg, w.data(), k);
        msg[k] = 0;
    }
    return r->run->status;
}

extern "C" void hgx_result_release(hgx_result* r) { delete r; }

extern "C" void hgx_result_release_many(hgx_result* const* rs, uint32_t n) {
    if (!rs) return;
    for (uint32_t i = 0; i < n; ++i) delete rs[i];
}

extern "C" int hgx_result_states(const hgx_result* r, uint64_t* hash, uint32_t* step, uint32_t* creator,
                                 uint64_t cap) {
    if (!r || cap < r->run->counts.raw_states) return HGX_ERR_INVALID;
    try {
        RunTables T = run_tables(*r->run);
        for (uint64_t s = 0; s < r->run->states; ++s) {
            if (hash) hash[s] = T.hash[s];
            if (step) step[s] = T.step[s];
            if (creator) creator[s] = T.creator[s];
        }
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_events(const hgx_result* r, hgx_event* out, uint64_t cap) {
    if (!r || (!out && r->run->events) || cap < r->run->events) return HGX_ERR_INVALID;
    try {
        RunTables T = run_tables(*r->run);
        const uint32_t nroots = r->run->hp->n_roots();
        for (uint64_t e = 0; e < r->run->events; ++e)
            out[e] = hgx_event{T.ev_in[e], (uint32_t)(e + nroots), T.ev_rule[e], T.ev_step[e]};
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_causal(const hgx_result* r, hgx_pair* pairs, uint8_t* in_tr, uint64_t cap) {
    if (!r) return HGX_ERR_INVALID;
    const RunRec& R = *r->run;
    if (!(R.hp->flags & HGX_REC_CAUSAL) || cap < R.counts.causal_pairs) return HGX_ERR_INVALID;
    try {
        RunTables T = run_tables(R);
        uint64_t o = 0;
        for (uint64_t c = 0; c < R.events; ++c) {
            for (uint64_t m = T.ev_prod[c]; m; m &= m - 1) {
                const uint32_t s = (uint32_t)__builtin_ctzll(m);
                uint32_t x = T.ev_in[c];                 // walk up the lineage to generation s
                while (T.step[x] > s) x = T.ev_in[T.creator[x]];
                if (o >= cap) return HGX_ERR_INVALID;
                if (pairs) pairs[o] = hgx_pair{T.creator[x], (uint32_t)c};
                if (in_tr) in_tr[o] = (uint8_t)((T.ev_tr[c] >> s) & 1u);
                ++o;
            }
        }
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_causal_triples(const hgx_result* r, hgx_triple* triples, uint8_t* in_tr, uint64_t cap) {
    if (!r) return HGX_ERR_INVALID;
    const RunRec& R = *r->run;
    if ((R.hp->flags & (HGX_REC_CAUSAL | HGX_REC_RELATIONS)) != (HGX_REC_CAUSAL | HGX_REC_RELATIONS) ||
        cap < R.counts.causal_triples)
        return HGX_ERR_INVALID;
    try {
        RunTables T = run_tables(R);
        const uint32_t nroots = R.hp->n_roots();
        // producer generation of every consumed edge, per event
        std::vector<uint8_t> cps;
        uint32_t K = 0;
        for (size_t k = 1; k < R.segs.size(); ++k) {          // segment 0 holds the roots
            const Segment& sg = R.segs[k];
            const uint32_t m = sg.e - sg.b;
            if (!m) continue;
            if (!sg.g->ev.cps) throw Error{HGX_ERR_INVALID, "causal triples were not recorded"};
            const uint32_t gk = sg.g->ev.cK;
            if (!K) { K = 16; cps.assign((size_t)R.events * K, 0); }
            std::vector<uint8_t> rows((size_t)m * gk);
            std::vector<uint32_t> sid(m);
            d2h(rows.data(), sg.g->ev.cps + (size_t)sg.b * gk, rows.size(), R.st);
            d2h(sid.data(), sg.g->v.sid + sg.b, m, R.st);
            HIP_TRY(hipStreamSynchronize(R.st));
            for (uint32_t i = 0; i < m; ++i) {
                const uint64_t e = (uint64_t)sid[i] - nroots;
                if (e >= R.events) throw Error{HGX_ERR_DEVICE, "event id out of range in readback"};
                std::memcpy(&cps[e * K], &rows[(size_t)i * gk], std::min<uint32_t>(gk, K));
            }
        }
        uint64_t o = 0;
        for (uint64_t c = 0; c < R.events; ++c) {
            const uint32_t nl = R.hp->rules.at(T.ev_rule[c]).n_lhs;
            for (uint32_t t = 0; t < nl; ++t) {
                const uint32_t ps = cps[c * K + t];
                if (!ps) continue;
                uint32_t x = T.ev_in[c];                 // up the lineage to generation ps
                while (T.step[x] > ps) x = T.ev_in[T.creator[x]];
                if (o >= cap) return HGX_ERR_INVALID;
                if (triples) triples[o] = hgx_triple{T.creator[x], (uint32_t)c, t};
                if (in_tr) in_tr[o] = (uint8_t)((T.ev_tr[c] >> ps) & 1u);
                ++o;
            }
        }
        return o == R.counts.causal_triples ? HGX_OK : HGX_ERR_DEVICE;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_branchial(const hgx_result* r, hgx_pair* pairs, uint64_t cap) {
    if (!r) return HGX_ERR_INVALID;
    const RunRec& R = *r->run;
    if (!(R.hp->flags & HGX_REC_BRANCHIAL) || !(R.hp->flags & HGX_REC_RELATIONS)) return HGX_ERR_INVALID;
    if (cap < R.counts.branchial_pairs) return HGX_ERR_INVALID;
    try {
        uint64_t o = 0;
        for (const Segment& sg : R.segs) {
            if (sg.br_e <= sg.br_b || !sg.g->br_pairs) continue;
            std::vector<uint32_t> tmp(2 * (sg.br_e - sg.br_b));
            d2h(tmp.data(), sg.g->br_pairs + 2 * sg.br_b, tmp.size(), R.st);
            HIP_TRY(hipStreamSynchronize(R.st));
            for (uint64_t i = 0; i < sg.br_e - sg.br_b; ++i) {
                if (o >= cap) return HGX_ERR_INVALID;
                pairs[o++] = hgx_pair{tmp[2 * i], tmp[2 * i + 1]};
            }
        }
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_consumed(const hgx_result* r, uint32_t* offsets, uint16_t* positions, uint64_t cap) {
    if (!r) return HGX_ERR_INVALID;
    const RunRec& R = *r->run;
    if (!(R.hp->flags & HGX_REC_STATES)) return HGX_ERR_INVALID;
    try {
        uint64_t o = 0, ev = 0;
        for (size_t k = 1; k < R.segs.size(); ++k) {            // segment 0 holds the roots
            const Segment& sg = R.segs[k];
            const uint32_t m = sg.e - sg.b;
            if (!m) continue;
            const uint32_t K = sg.g->cK;
            if (!sg.g->cq || !K) throw Error{HGX_ERR_INVALID, "consumed positions were not kept"};
            std::vector<uint16_t> q((size_t)m * K);
            std::vector<uint32_t> rule(m);
            d2h(q.data(), sg.g->cq + (size_t)sg.b * K, q.size(), R.st);
            d2h(rule.data(), sg.g->ev.rule + sg.b, m, R.st);
            HIP_TRY(hipStreamSynchronize(R.st));
            for (uint32_t i = 0; i < m; ++i, ++ev) {
                const uint32_t nl = R.hp->rules.at(rule[i]).n_lhs;
                if (offsets) offsets[ev] = (uint32_t)o;
                for (uint32_t t = 0; t < nl; ++t) {
                    if (positions) {
                        if (o >= cap) return HGX_ERR_INVALID;
                        positions[o] = q[(size_t)i * K + t];
                    }
                    ++o;
                }
            }
        }
        if (offsets) offsets[ev] = (uint32_t)o;
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_state_edges(const hgx_result* r, uint64_t* state_offsets, uint32_t* edge_offsets,
                                      uint32_t* vertices, uint64_t* n_edges_total, uint64_t* n_vertices_total) {
    if (!r) return HGX_ERR_INVALID;
    const RunRec& R = *r->run;
    if (!(R.hp->flags & HGX_REC_STATES)) return HGX_ERR_INVALID;
    try {
        uint64_t ne_tot = 0, nv_tot = 0, s_idx = 0;
        for (const Segment& sg : R.segs) {
            const uint32_t m = sg.e - sg.b;
            if (!m) continue;
            const uint32_t A = sg.g->v.A;
            std::vector<uint64_t> eo(m + 1);
            d2h(eo.data(), sg.g->v.eoff + sg.b, m + 1, R.st);
            HIP_TRY(hipStreamSynchronize(R.st));
            const uint64_t E0 = eo[0], E1 = eo[m];
            std::vector<uint8_t> ar(E1 - E0);
            std::vector<uint32_t> ev((E1 - E0) * A);
            d2h(ar.data(), sg.g->v.ar + E0, ar.size(), R.st);
            d2h(ev.data(), sg.g->v.ev + E0 * A, ev.size(), R.st);
            HIP_TRY(hipStreamSynchronize(R.st));
            for (uint32_t i = 0; i < m; ++i) {
                if (state_offsets) state_offsets[s_idx] = ne_tot;
                ++s_idx;
                for (uint64_t e = eo[i]; e < eo[i + 1]; ++e) {
                    const uint32_t a = ar[e - E0];
                    if (edge_offsets) edge_offsets[ne_tot] = (uint32_t)nv_tot;
                    if (vertices) for (uint32_t k = 0; k < a; ++k) vertices[nv_tot + k] = ev[(e - E0) * A + k];
                    nv_tot += a;
                    ++ne_tot;
                }
            }
        }
        if (state_offsets) state_offsets[s_idx] = ne_tot;
        if (edge_offsets) edge_offsets[ne_tot] = (uint32_t)nv_tot;
        if (n_edges_total) *n_edges_total = ne_tot;
        if (n_vertices_total) *n_vertices_total = nv_tot;
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

