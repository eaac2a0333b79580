// hgx_runtime.cpp -- host side of the C ABI (include/hgx.h): device memory, request parsing,
// the generation loop, results. Compiled by hipcc for gfx950 together with the kernels.
//
// One generation of every live request is one sequence of launches on the context's stream:
//   k_match<count> -> scan -> per-run totals (+ state-cap decision) -> k_match<emit> -> scan ->
//   k_expand -> k_branchial
// Requests of different rules, sizes and depths share each launch; each keeps its own state
// map, counters and numbering, so its results do not depend on what else is in the batch.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hgx.h"
#include "hgx_device.hpp"
#include "hgx_kernels.hpp"
#include "hgx_runtime.hpp"

namespace hgx {

// ------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------

struct Error {
    int code;
    std::string msg;
};

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            throw Error{HGX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)};    \
    } while (0)

static std::string g_create_error;

// ------------------------------------------------------------------------------------------
// device memory: a pool of large chunks (allocated once, reused), bump arenas on top
// ------------------------------------------------------------------------------------------

ChunkPool::~ChunkPool() {
    for (auto& c : free_) (void)hipFree(c.p);
}

Chunk ChunkPool::get(size_t min_bytes) {
    std::lock_guard<std::mutex> g(mu_);
    size_t best = SIZE_MAX;
    for (size_t i = 0; i < free_.size(); ++i)
        if (free_[i].size >= min_bytes && (best == SIZE_MAX || free_[i].size < free_[best].size)) best = i;
    if (best != SIZE_MAX) {
        Chunk c = free_[best];
        free_.erase(free_.begin() + best);
        return c;
    }
    size_t sz = std::max(min_bytes, chunk_bytes_);
    sz = (sz + (2u << 20) - 1) & ~((size_t)(2u << 20) - 1);
    if (total_ + sz > budget_) {
        // give back idle chunks before refusing
        for (auto& c : free_) { (void)hipFree(c.p); total_ -= c.size; }
        free_.clear();
        if (total_ + sz > budget_)
            throw Error{HGX_ERR_CAPACITY, "HBM budget exhausted: need " + std::to_string(sz >> 20) +
                                              " MiB more, budget " + std::to_string(budget_ >> 20) + " MiB"};
    }
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, sz));
    total_ += sz;
    return Chunk{p, sz};
}

void ChunkPool::put(const Chunk& c) {
    std::lock_guard<std::mutex> g(mu_);
    free_.push_back(c);
}

Arena::~Arena() { release(); }

void Arena::release() {
    if (pool_) for (auto& c : chunks_) pool_->put(c);
    chunks_.clear();
    used_ = 0;
}

void* Arena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (bytes == 0) bytes = 256;
    if (chunks_.empty() || used_ + bytes > chunks_.back().size) {
        chunks_.push_back(pool_->get(bytes));
        used_ = 0;
    }
    void* p = (uint8_t*)chunks_.back().p + used_;
    used_ += bytes;
    bytes_ += bytes;
    return p;
}

// ------------------------------------------------------------------------------------------
// request parsing (host): rule program -> DRule with the reference's join order
// ------------------------------------------------------------------------------------------

static uint32_t var_mask(const uint8_t* v, uint32_t a) {
    uint32_t m = 0;
    for (uint32_t i = 0; i < a; ++i) m |= 1u << v[i];
    return m;
}

// RewriteRule::edge_constraint_score / compute_match_order (hypergraph/src/pattern.cpp:31-75).
static void join_order(DRule& R) {
    for (int i = 0; i < 16; ++i) R.order[i] = (uint8_t)i;
    if (R.n_lhs <= 1) return;
    auto score = [&](int e) {
        int s = 0;
        for (int i = 0; i < R.la[e]; ++i)
            for (int j = i + 1; j < R.la[e]; ++j)
                if (R.lv[e][i] == R.lv[e][j]) s += 100;
        for (int o = 0; o < R.n_lhs; ++o)
            if (o != e && (var_mask(R.lv[e], R.la[e]) & var_mask(R.lv[o], R.la[o]))) s += 1;
        return s;
    };
    bool used[16] = {};
    int first = 0, best = -1;
    for (int e = 0; e < R.n_lhs; ++e) { int s = score(e); if (s > best) { best = s; first = e; } }
    R.order[0] = (uint8_t)first; used[first] = true;
    uint32_t bound = var_mask(R.lv[first], R.la[first]);
    for (int pos = 1; pos < R.n_lhs; ++pos) {
        int pick = 0, ps = -1, pself = -1;
        for (int e = 0; e < R.n_lhs; ++e) {
            if (used[e]) continue;
            int sh = __builtin_popcount(var_mask(R.lv[e], R.la[e]) & bound), self = score(e);
            if (sh > ps || (sh == ps && self > pself)) { pick = e; ps = sh; pself = self; }
        }
        R.order[pos] = (uint8_t)pick; used[pick] = true;
        bound |= var_mask(R.lv[pick], R.la[pick]);
    }
}

static void parse_rules(const uint8_t* p, uint32_t len, uint32_t n, std::vector<DRule>& out) {
    uint32_t k = 0;
    for (uint32_t r = 0; r < n; ++r) {
        DRule R;
        std::memset(&R, 0, sizeof(R));
        if (k + 2 > len) throw Error{HGX_ERR_INVALID, "rule program truncated"};
        R.n_lhs = p[k++]; R.n_rhs = p[k++];
        if (R.n_lhs < 1 || R.n_lhs > 16 || R.n_rhs > 16)
            throw Error{HGX_ERR_INVALID, "rule needs 1..16 LHS and 0..16 RHS edges"};
        for (int side = 0; side < 2; ++side) {
            const int ne = side ? R.n_rhs : R.n_lhs;
            for (int e = 0; e < ne; ++e) {
                if (k >= len) throw Error{HGX_ERR_INVALID, "rule program truncated"};
                const uint8_t a = p[k++];
                if (a < 1 || a > 16 || k + a > len) throw Error{HGX_ERR_INVALID, "pattern edge arity must be 1..16"};
                for (int i = 0; i < a; ++i) {
                    const uint8_t v = p[k++];
                    if (v >= 32) throw Error{HGX_ERR_INVALID, "pattern variable must be < 32"};
                    if (side) R.rv[e][i] = v; else R.lv[e][i] = v;
                }
                if (side) R.ra[e] = a; else R.la[e] = a;
            }
        }
        for (int e = 0; e < R.n_lhs; ++e) R.lhs_mask |= var_mask(R.lv[e], R.la[e]);
        for (int e = 0; e < R.n_rhs; ++e) R.rhs_mask |= var_mask(R.rv[e], R.ra[e]);
        R.new_mask = R.rhs_mask & ~R.lhs_mask;
        R.n_new = (uint8_t)__builtin_popcount(R.new_mask);
        std::memset(R.var_src_edge, 0, sizeof(R.var_src_edge));
        for (int v = 0; v < 32; ++v) {
            if (!(R.lhs_mask & (1u << v))) continue;
            bool done = false;
            for (int e = 0; e < R.n_lhs && !done; ++e)
                for (int i = 0; i < R.la[e] && !done; ++i)
                    if (R.lv[e][i] == v) { R.var_src_edge[v] = (uint8_t)e; R.var_src_pos[v] = (uint8_t)i; done = true; }
        }
        join_order(R);
        out.push_back(R);
    }
    if (k != len) throw Error{HGX_ERR_INVALID, "rule program has trailing bytes"};
}

struct RunHost {
    std::vector<DRule> rules;
    std::vector<std::vector<std::vector<uint32_t>>> roots;   // renumbered per root
    std::vector<uint32_t> root_nv;
    uint32_t steps = 0, cap = 0, flags = 0;
    uint32_t max_arity = 1, max_lhs = 1;
};

static RunHost parse_run(const hgx_run_desc* d) {
    if (!d) throw Error{HGX_ERR_INVALID, "null request"};
    if (!d->rules || d->n_rules == 0) throw Error{HGX_ERR_INVALID, "request has no rules"};
    if (d->steps > HGX_MAX_STEPS) throw Error{HGX_ERR_UNSUPPORTED, "steps above HGX_MAX_STEPS"};
    RunHost h;
    parse_rules(d->rules, d->rules_len, d->n_rules, h.rules);
    h.steps = d->steps; h.cap = d->state_cap; h.flags = d->flags;
    if (d->n_edges && (!d->edge_offsets || !d->vertices)) throw Error{HGX_ERR_INVALID, "null state arrays"};
    const uint32_t n_roots = d->root_offsets ? d->n_roots : 1;
    if (d->root_offsets && (d->n_roots == 0 || d->root_offsets[0] != 0 || d->root_offsets[d->n_roots] != d->n_edges))
        throw Error{HGX_ERR_INVALID, "root_offsets must span the edges"};
    for (uint32_t r = 0; r < n_roots; ++r) {
        const uint32_t e0 = d->root_offsets ? d->root_offsets[r] : 0;
        const uint32_t e1 = d->root_offsets ? d->root_offsets[r + 1] : d->n_edges;
        if (e1 < e0) throw Error{HGX_ERR_INVALID, "root_offsets not ascending"};
        std::unordered_map<uint32_t, uint32_t> lab;   // dense by first appearance (FFI :599-636)
        std::vector<std::vector<uint32_t>> st;
        for (uint32_t e = e0; e < e1; ++e) {
            const uint32_t a = d->edge_offsets[e + 1] - d->edge_offsets[e];
            if (d->edge_offsets[e + 1] < d->edge_offsets[e] || a < 1 || a > 16)
                throw Error{HGX_ERR_INVALID, "state edge arity must be 1..16"};
            std::vector<uint32_t> ed(a);
            for (uint32_t i = 0; i < a; ++i) {
                auto it = lab.emplace(d->vertices[d->edge_offsets[e] + i], (uint32_t)lab.size()).first;
                ed[i] = it->second;
            }
            h.max_arity = std::max(h.max_arity, a);
            st.push_back(std::move(ed));
        }
        h.roots.push_back(std::move(st));
        h.root_nv.push_back((uint32_t)lab.size());
    }
    for (auto& R : h.rules) {
        for (int e = 0; e < R.n_rhs; ++e) h.max_arity = std::max<uint32_t>(h.max_arity, R.ra[e]);
        h.max_lhs = std::max<uint32_t>(h.max_lhs, R.n_lhs);
    }
    return h;
}

// ------------------------------------------------------------------------------------------
// the batch: all device data of one set of requests evolved together
// ------------------------------------------------------------------------------------------

struct Gen {
    GenView v{};
    uint64_t n_edges = 0;
    uint32_t max_ne = 0, max_nextv = 0;
    // per run: state range in this generation
    std::vector<uint32_t> sbeg, send;
    // generation >= 1: the events that created these states
    EventView ev{};
    uint32_t* br_pairs = nullptr;            // materialised branchial pairs of events whose
    uint64_t n_br = 0;                       //   input state is in THIS generation's parents
    std::vector<uint64_t> br_run_off;        // per run offset into br_pairs (host)
};

struct BatchStore {
    std::shared_ptr<ChunkPool> pool;
    Arena arena;
    hipStream_t st = nullptr;         // the owning context's stream (contexts outlive results)
    std::vector<RunHost> runs;
    std::vector<Gen> gens;
    std::vector<hgx_counts> counts;
    uint32_t A = 1, K = 1;
    double device_ms = 0;
    // host caches for readback
    std::mutex mu;
    bool host_ready = false;
    std::vector<std::vector<uint64_t>> h_hash;      // per gen
    std::vector<std::vector<uint32_t>> h_sid;
    std::vector<std::vector<uint32_t>> h_in, h_rule;
    std::vector<std::vector<uint64_t>> h_prod, h_tr;
    explicit BatchStore(std::shared_ptr<ChunkPool> p) : pool(p), arena(p.get()) {}
};

}  // namespace hgx

struct hgx_result {
    std::shared_ptr<hgx::BatchStore> store;
    uint32_t run = 0;
    hgx_counts counts{};
    uint64_t device_ns = 0, latency_ns = 0;
};

namespace hgx {
struct ProfPending {
    std::string name;
    hipEvent_t a, b;
    uint64_t bytes, items;
};
struct Profiler {
    bool on = false;
    std::vector<ProfPending> pend;
    std::map<std::string, hgx_kernel_stat> acc;
};
}  // namespace hgx

struct hgx_ctx {
    hgx::Profiler prof;
    int device = 0;
    hipStream_t stream = nullptr;
    std::shared_ptr<hgx::ChunkPool> pool;
    std::string err;
    std::string info;
    uint32_t num_cu = 256;
    uint32_t max_lds = 65536;
    std::mutex mu;
    hgx::ServeState* serve = nullptr;
};

namespace hgx {

template <class T>
static T* dalloc(Arena& a, size_t n) { return (T*)a.alloc(sizeof(T) * (n ? n : 1)); }

template <class T>
static void h2d(T* dst, const T* src, size_t n, hipStream_t st) {
    if (n) HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, st));
}

template <class T>
static void d2h(T* dst, const T* src, size_t n, hipStream_t st) {
    if (n) HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, st));
}

struct CastU64 {
    __host__ __device__ uint64_t operator()(uint32_t x) const { return x; }
};

// Exclusive prefix sum of n uint32 values into n+1 uint64 offsets (last = total).
static void scan_u32_to_u64(Arena& ar, const uint32_t* in, uint64_t* out, uint32_t n, hipStream_t st) {
    // in must hold n+1 readable values with in[n] == 0
    hipcub::TransformInputIterator<uint64_t, CastU64, const uint32_t*> it(in, CastU64());
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, it, out, n + 1, st));
    void* t = ar.alloc(tmp);
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(t, tmp, it, out, n + 1, st));
}

static uint32_t waves_per_block(uint32_t lds_per_wave, uint32_t max_lds) {
    uint32_t w = 4;
    while (w > 1 && (uint64_t)lds_per_wave * w > max_lds) --w;
    if ((uint64_t)lds_per_wave * w > max_lds)
        throw Error{HGX_ERR_UNSUPPORTED, "a state is too large for the LDS-resident path (" +
                                             std::to_string(lds_per_wave) + " bytes per wave)"};
    return w;
}

// Launch with optional event timing; `bytes` are the launch's algorithmic HBM bytes.
template <class F>
static void timed(hgx_ctx* ctx, hipStream_t st, const char* name, uint64_t bytes, uint64_t items, F&& launch) {
    if (!ctx->prof.on) { HIP_TRY(launch()); return; }
    ProfPending p{name, nullptr, nullptr, bytes, items};
    HIP_TRY(hipEventCreate(&p.a));
    HIP_TRY(hipEventCreate(&p.b));
    HIP_TRY(hipEventRecord(p.a, st));
    HIP_TRY(launch());
    HIP_TRY(hipEventRecord(p.b, st));
    ctx->prof.pend.push_back(p);
}

// After a stream synchronisation: fold finished launches into the per-kernel totals.
static void prof_resolve(hgx_ctx* ctx) {
    for (auto& p : ctx->prof.pend) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
        hgx_kernel_stat& s = ctx->prof.acc[p.name];
        std::snprintf(s.name, sizeof(s.name), "%s", p.name.c_str());
        s.launches += 1;
        s.total_ms += ms;
        s.alg_bytes += p.bytes;
        s.items += p.items;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    ctx->prof.pend.clear();
}

// Evolve a set of requests to completion. Returns the store; counts filled.
std::shared_ptr<BatchStore> run_batch(hgx_ctx* ctx, std::vector<RunHost>&& runs) {
    hipStream_t st = ctx->stream;
    auto S = std::make_shared<BatchStore>(ctx->pool);
    S->st = st;
    S->runs = std::move(runs);
    const uint32_t R = (uint32_t)S->runs.size();
    Arena& ar = S->arena;
    uint32_t A = 1, K = 1, max_rhs_gain = 0, max_new = 0;
    bool any_branch = false, any_rel = false;
    for (auto& h : S->runs) {
        A = std::max(A, h.max_arity);
        K = std::max(K, h.max_lhs);
        for (auto& r : h.rules) {
            max_rhs_gain = std::max<uint32_t>(max_rhs_gain, r.n_rhs > r.n_lhs ? r.n_rhs - r.n_lhs : 0);
            max_new = std::max<uint32_t>(max_new, r.n_new);
        }
        any_branch |= (h.flags & HGX_REC_BRANCHIAL) != 0;
        any_rel |= (h.flags & HGX_REC_RELATIONS) != 0;
    }
    S->A = A; S->K = K;

    hipEvent_t t0, t1;
    HIP_TRY(hipEventCreate(&t0));
    HIP_TRY(hipEventCreate(&t1));
    HIP_TRY(hipEventRecord(t0, st));

    // ---- rules, runs, counters ----
    std::vector<DRule> all_rules;
    std::vector<DRun> hruns(R);
    uint32_t total_roots = 0, max_root_ne = 0, max_root_nv = 0;
    uint64_t root_edges = 0;
    for (uint32_t r = 0; r < R; ++r) {
        auto& h = S->runs[r];
        DRun& d = hruns[r];
        std::memset(&d, 0, sizeof(d));
        d.rule_begin = (uint32_t)all_rules.size();
        d.n_rules = (uint32_t)h.rules.size();
        d.n_roots = (uint32_t)h.roots.size();
        d.flags = h.flags;
        for (auto& x : h.rules) all_rules.push_back(x);
        total_roots += d.n_roots;
        for (size_t i = 0; i < h.roots.size(); ++i) {
            max_root_ne = std::max<uint32_t>(max_root_ne, (uint32_t)h.roots[i].size());
            max_root_nv = std::max(max_root_nv, h.root_nv[i]);
            root_edges += h.roots[i].size();
        }
    }
    DRule* d_rules = dalloc<DRule>(ar, all_rules.size());
    h2d(d_rules, all_rules.data(), all_rules.size(), st);
    // state maps: capacity >= 2x the states they will hold
    std::vector<uint32_t> map_cap(R);
    for (uint32_t r = 0; r < R; ++r) {
        const uint64_t want = S->runs[r].cap ? (uint64_t)S->runs[r].cap + hruns[r].n_roots : 64;
        uint64_t c = 64;
        while (c < 2 * want) c <<= 1;
        map_cap[r] = (uint32_t)c;
    }
    {
        uint64_t tot = 0;
        for (uint32_t r = 0; r < R; ++r) tot += map_cap[r];
        uint64_t* keys = dalloc<uint64_t>(ar, tot);
        uint32_t* vals = dalloc<uint32_t>(ar, tot);
        HIP_TRY(hipMemsetAsync(keys, 0, sizeof(uint64_t) * tot, st));
        uint64_t o = 0;
        for (uint32_t r = 0; r < R; ++r) {
            hruns[r].map_keys = keys + o;
            hruns[r].map_vals = vals + o;
            hruns[r].map_mask = map_cap[r] - 1;
            o += map_cap[r];
        }
    }
    DRun* d_runs = dalloc<DRun>(ar, R);
    uint64_t* d_ctr = dalloc<uint64_t>(ar, (size_t)R * CTR_N);
    HIP_TRY(hipMemsetAsync(d_ctr, 0, sizeof(uint64_t) * R * CTR_N, st));
    uint8_t* d_active = dalloc<uint8_t>(ar, R);
    h2d(d_runs, hruns.data(), R, st);

    // ---- generation 0: the roots ----
    {
        Gen g;
        g.v.n = total_roots; g.v.step = 0; g.v.A = A;
        std::vector<uint32_t> run(total_roots), ne(total_roots), nextv(total_roots), sid(total_roots);
        std::vector<uint64_t> eoff(total_roots + 1);
        std::vector<uint32_t> ev((size_t)root_edges * A, 0);
        std::vector<uint8_t> ea(root_edges);
        g.sbeg.resize(R); g.send.resize(R);
        uint32_t s = 0; uint64_t e = 0;
        for (uint32_t r = 0; r < R; ++r) {
            g.sbeg[r] = s;
            auto& h = S->runs[r];
            for (size_t i = 0; i < h.roots.size(); ++i, ++s) {
                run[s] = r; ne[s] = (uint32_t)h.roots[i].size(); nextv[s] = h.root_nv[i];
                sid[s] = (uint32_t)i; eoff[s] = e;
                for (auto& ed : h.roots[i]) {
                    ea[e] = (uint8_t)ed.size();
                    for (size_t k = 0; k < ed.size(); ++k) ev[e * A + k] = ed[k];
                    ++e;
                }
            }
            g.send[r] = s;
        }
        eoff[total_roots] = e;
        g.v.run = dalloc<uint32_t>(ar, total_roots);
        g.v.eoff = dalloc<uint64_t>(ar, total_roots + 1);
        g.v.ne = dalloc<uint32_t>(ar, total_roots);
        g.v.nextv = dalloc<uint32_t>(ar, total_roots);
        g.v.sid = dalloc<uint32_t>(ar, total_roots);
        g.v.hash = dalloc<uint64_t>(ar, total_roots);
        g.v.lanc = dalloc<uint64_t>(ar, 1);
        g.v.ev = dalloc<uint32_t>(ar, root_edges * A);
        g.v.ar = dalloc<uint8_t>(ar, root_edges);
        g.v.ps = dalloc<uint8_t>(ar, root_edges);
        h2d(g.v.run, run.data(), total_roots, st);
        h2d(g.v.eoff, eoff.data(), total_roots + 1, st);
        h2d(g.v.ne, ne.data(), total_roots, st);
        h2d(g.v.nextv, nextv.data(), total_roots, st);
        h2d(g.v.sid, sid.data(), total_roots, st);
        h2d(g.v.ev, ev.data(), ev.size(), st);
        h2d(g.v.ar, ea.data(), ea.size(), st);
        if (root_edges) HIP_TRY(hipMemsetAsync(g.v.ps, 0, root_edges, st));
        g.max_ne = max_root_ne;
        g.max_nextv = max_root_nv;
        g.n_edges = root_edges;
        if (total_roots) {
            WLStatesArgs w{};
            w.n = total_roots; w.A = A; w.state_off = g.v.eoff; w.ar = g.v.ar; w.ev = g.v.ev;
            w.nv = g.v.nextv; w.out = g.v.hash; w.run_of = g.v.run; w.sid_of = g.v.sid;
            w.runs = d_runs; w.ctr = d_ctr; w.max_ne = max_root_ne; w.max_nv = std::max(1u, max_root_nv);
            w.lds_per_wave = wl_states_lds_per_wave(w.max_ne, w.max_nv, A);
            const uint32_t wpb = waves_per_block(w.lds_per_wave, ctx->max_lds);
            timed(ctx, st, "k_wl_states", root_edges * (4ull * A + 1) + total_roots * 36ull, total_roots,
                  [&] { return launch_wl_states(w, ctx->num_cu * 8, wpb, st); });
        }
        HIP_TRY(hipStreamSynchronize(st));   // host staging vectors end with this scope
        S->gens.push_back(std::move(g));
    }

    // ---- host bookkeeping per run ----
    std::vector<uint64_t> states(R), events(R);
    std::vector<uint8_t> done(R, 0), active(R, 0), truncated(R, 0);
    std::vector<uint32_t> steps_done(R, 0);
    for (uint32_t r = 0; r < R; ++r) states[r] = hruns[r].n_roots;
    uint32_t max_steps = 0;
    for (auto& h : S->runs) max_steps = std::max(max_steps, h.steps);

    uint32_t* d_sbeg = dalloc<uint32_t>(ar, R);
    uint32_t* d_send = dalloc<uint32_t>(ar, R);
    uint64_t* d_tot = dalloc<uint64_t>(ar, R + 1);   // [R] = edges of parents with a match
    std::vector<uint64_t> tot(R + 1);

    for (uint32_t L = 1; L <= max_steps; ++L) {
        Gen& P = S->gens[L - 1];
        bool any = false;
        for (uint32_t r = 0; r < R; ++r) {
            active[r] = (!done[r] && L <= S->runs[r].steps && P.send[r] > P.sbeg[r]) ? 1 : 0;
            any |= active[r] != 0;
        }
        if (!any || P.v.n == 0) break;
        h2d(d_active, active.data(), R, st);
        h2d(d_sbeg, P.sbeg.data(), R, st);
        h2d(d_send, P.send.data(), R, st);

        // count
        uint32_t* cnt = dalloc<uint32_t>(ar, P.v.n + 1);
        HIP_TRY(hipMemsetAsync(cnt + P.v.n, 0, sizeof(uint32_t), st));
        uint64_t* moff = dalloc<uint64_t>(ar, P.v.n + 1);
        MatchArgs ma{};
        ma.P = P.v; ma.rules = d_rules; ma.runs = d_runs; ma.active = d_active; ma.cnt = cnt;
        ma.moff = moff; ma.max_ne = std::max(1u, P.max_ne);
        ma.lds_per_wave = (ma.max_ne * A * 4 + ((ma.max_ne + 7u) & ~7u) + ma.max_ne * 2 + 64 + 15) & ~15u;
        ma.parent_edges = (unsigned long long*)(d_tot + R);
        HIP_TRY(hipMemsetAsync(d_tot + R, 0, sizeof(uint64_t), st));
        const uint32_t wpb_m = waves_per_block(ma.lds_per_wave, ctx->max_lds);
        // algorithmic bytes of a join pass: every parent edge read once, one record per state
        const uint64_t match_bytes = P.n_edges * (4ull * A + 1) + (uint64_t)P.v.n * 24;
        timed(ctx, st, "k_match_count", match_bytes + P.v.n * 4ull, P.v.n,
              [&] { return launch_match(ma, false, ctx->num_cu * 8, wpb_m, st); });
        scan_u32_to_u64(ar, cnt, moff, P.v.n, st);
        HIP_TRY(launch_run_totals(moff, d_sbeg, d_send, R, d_tot, st));
        d2h(tot.data(), d_tot, R + 1, st);
        HIP_TRY(hipStreamSynchronize(st));
        const uint64_t parent_edges = tot[R];

        // deterministic state cap: stop before a generation that would exceed it
        bool masked = false;
        for (uint32_t r = 0; r < R; ++r) {
            if (!active[r]) { tot[r] = 0; continue; }
            const uint32_t cap = S->runs[r].cap;
            if (cap && states[r] + tot[r] > cap) {
                active[r] = 0; done[r] = 1; truncated[r] = 1; tot[r] = 0; masked = true;
            } else if (tot[r] == 0) {
                active[r] = 0; done[r] = 1;
            }
        }
        if (masked) {
            h2d(d_active, active.data(), R, st);
            HIP_TRY(launch_mask_counts(P.v.run, d_active, P.v.n, cnt, st));
            scan_u32_to_u64(ar, cnt, moff, P.v.n, st);
        }
        uint64_t M64 = 0;
        for (uint32_t r = 0; r < R; ++r) M64 += tot[r];
        if (M64 == 0) break;
        if (M64 > 0xFFFFFFF0ull) throw Error{HGX_ERR_UNSUPPORTED, "more than 2^32 events in one generation"};
        const uint32_t M = (uint32_t)M64;

        // state maps of uncapped runs grow ahead of the inserts
        bool runs_dirty = true;
        for (uint32_t r = 0; r < R; ++r) {
            if (!active[r]) continue;
            const uint64_t need = 2 * (states[r] + tot[r]);
            if (need <= (uint64_t)map_cap[r]) continue;
            uint64_t c = map_cap[r];
            while (c < need) c <<= 1;
            if (c > 0x80000000ull) throw Error{HGX_ERR_UNSUPPORTED, "state map above 2^31 slots"};
            uint64_t* keys = dalloc<uint64_t>(ar, c);
            uint32_t* vals = dalloc<uint32_t>(ar, c);
            HIP_TRY(hipMemsetAsync(keys, 0, sizeof(uint64_t) * c, st));
            HIP_TRY(launch_rehash(hruns[r].map_keys, hruns[r].map_vals, map_cap[r], keys, vals,
                                  (uint32_t)(c - 1), st));
            hruns[r].map_keys = keys; hruns[r].map_vals = vals;
            hruns[r].map_mask = (uint32_t)(c - 1);
            map_cap[r] = (uint32_t)c;
        }

        // emit matches
        MatchView mv{};
        mv.n = M; mv.K = K;
        mv.parent = dalloc<uint32_t>(ar, M);
        mv.rule = dalloc<uint32_t>(ar, M);
        mv.q = dalloc<uint16_t>(ar, (size_t)M * K);
        mv.cne = dalloc<uint32_t>(ar, M + 1);
        HIP_TRY(hipMemsetAsync(mv.cne + M, 0, sizeof(uint32_t), st));
        ma.M = mv;
        ma.parent_edges = nullptr;
        timed(ctx, st, "k_match_emit", match_bytes + (uint64_t)M * (12ull + 2ull * K), P.v.n,
              [&] { return launch_match(ma, true, ctx->num_cu * 8, wpb_m, st); });

        // child generation
        Gen C;
        C.v.n = M; C.v.step = L; C.v.A = A;
        C.v.eoff = dalloc<uint64_t>(ar, M + 1);
        scan_u32_to_u64(ar, mv.cne, C.v.eoff, M, st);
        uint64_t EC = 0;
        d2h(&EC, C.v.eoff + M, 1, st);
        // per-run child ranges and event bases
        C.sbeg.resize(R); C.send.resize(R);
        {
            uint32_t o = 0;
            for (uint32_t r = 0; r < R; ++r) {
                C.sbeg[r] = o; o += (uint32_t)tot[r]; C.send[r] = o;
                hruns[r].ev_base = (uint32_t)events[r];
                hruns[r].child_begin = C.sbeg[r];
            }
        }
        if (runs_dirty) h2d(d_runs, hruns.data(), R, st);
        HIP_TRY(hipStreamSynchronize(st));
        C.v.run = dalloc<uint32_t>(ar, M);
        C.v.ne = dalloc<uint32_t>(ar, M);
        C.v.nextv = dalloc<uint32_t>(ar, M);
        C.v.sid = dalloc<uint32_t>(ar, M);
        C.v.hash = dalloc<uint64_t>(ar, M);
        C.v.lanc = dalloc<uint64_t>(ar, (size_t)M * L);
        C.v.ev = dalloc<uint32_t>(ar, EC * A);
        C.v.ar = dalloc<uint8_t>(ar, EC);
        C.v.ps = dalloc<uint8_t>(ar, EC);
        C.ev.in_state = dalloc<uint32_t>(ar, M);
        C.ev.rule = dalloc<uint32_t>(ar, M);
        C.ev.prod = dalloc<uint64_t>(ar, M);
        C.ev.tr = dalloc<uint64_t>(ar, M);
        C.max_ne = P.max_ne + max_rhs_gain;
        C.max_nextv = P.max_nextv + max_new;
        C.n_edges = EC;

        ExpandArgs ea{};
        ea.P = P.v; ea.C = C.v; ea.M = mv; ea.E = C.ev; ea.rules = d_rules; ea.runs = d_runs;
        ea.ctr = d_ctr; ea.max_child_ne = std::max(1u, C.max_ne);
        ea.max_nv = std::max(1u, std::min(C.max_nextv, C.max_ne * A));
        ea.max_nextv = C.max_nextv;
        ea.lds_per_wave = expand_lds_per_wave(ea.max_child_ne, ea.max_nv, ea.max_nextv, A);
        const uint32_t wpb_e = waves_per_block(ea.lds_per_wave, ctx->max_lds);
        // algorithmic bytes: match records and the edges of every parent with a child read once;
        // child edges, child state record (incl. lineage), event record and map slot written once
        const uint64_t expand_bytes = (uint64_t)M * (8ull + 2ull * K) + parent_edges * (4ull * A + 2) +
                                      EC * (4ull * A + 2) + (uint64_t)M * (24ull + 8ull * L) +
                                      (uint64_t)M * 24ull + (uint64_t)M * 12ull;
        timed(ctx, st, "k_expand", expand_bytes, M, [&] { return launch_expand(ea, ctx->num_cu * 8, wpb_e, st); });

        // branchial
        if (any_branch) {
            BranchialArgs ba{};
            ba.P = P.v; ba.M = mv; ba.rules = d_rules; ba.runs = d_runs; ba.cnt = cnt; ba.moff = moff;
            ba.ctr = d_ctr; ba.emit = 0; ba.mask_cap = 512; ba.lds_per_wave = 512 * 8;
            uint32_t* pcnt = nullptr;
            if (any_rel) { pcnt = dalloc<uint32_t>(ar, P.v.n + 1); HIP_TRY(hipMemsetAsync(pcnt, 0, sizeof(uint32_t) * (P.v.n + 1), st)); }
            ba.pcnt = pcnt;
            // runs that do not record branchial are masked out of this launch by their counts
            std::vector<uint8_t> bmask(R);
            bool all_rec = true;
            for (uint32_t r = 0; r < R; ++r) {
                bmask[r] = active[r] && (S->runs[r].flags & HGX_REC_BRANCHIAL);
                all_rec &= (bmask[r] == active[r]);
            }
            uint32_t* bcnt = cnt;
            if (!all_rec) {
                bcnt = dalloc<uint32_t>(ar, P.v.n + 1);
                HIP_TRY(hipMemcpyAsync(bcnt, cnt, sizeof(uint32_t) * (P.v.n + 1), hipMemcpyDeviceToDevice, st));
                uint8_t* d_bm = dalloc<uint8_t>(ar, R);
                h2d(d_bm, bmask.data(), R, st);
                HIP_TRY(launch_mask_counts(P.v.run, d_bm, P.v.n, bcnt, st));
                HIP_TRY(hipStreamSynchronize(st));
            }
            ba.cnt = bcnt;
            timed(ctx, st, "k_branchial", (uint64_t)M * (6ull + 2ull * K) + (uint64_t)P.v.n * 12ull, P.v.n,
                  [&] { return launch_branchial(ba, ctx->num_cu * 8, 4, st); });
            if (any_rel) {
                uint64_t* poff = dalloc<uint64_t>(ar, P.v.n + 1);
                scan_u32_to_u64(ar, pcnt, poff, P.v.n, st);
                uint64_t NB = 0;
                d2h(&NB, poff + P.v.n, 1, st);
                HIP_TRY(hipStreamSynchronize(st));
                if (NB) {
                    C.br_pairs = dalloc<uint32_t>(ar, 2 * NB);
                    C.n_br = NB;
                    ba.emit = 1; ba.poff = poff; ba.pairs = C.br_pairs;
                    HIP_TRY(launch_branchial(ba, ctx->num_cu * 8, 4, st));
                    // per-run offsets: parents of run r are [sbeg, send) in P
                    std::vector<uint64_t> h_poff(P.v.n + 1);
                    d2h(h_poff.data(), poff, P.v.n + 1, st);
                    HIP_TRY(hipStreamSynchronize(st));
                    C.br_run_off.resize(R + 1);
                    for (uint32_t r = 0; r < R; ++r) C.br_run_off[r] = h_poff[P.sbeg[r]];
                    C.br_run_off[R] = NB;
                }
            }
        }

        for (uint32_t r = 0; r < R; ++r) {
            if (!active[r]) continue;
            states[r] += tot[r];
            events[r] += tot[r];
            steps_done[r] = L;
        }
        S->gens.push_back(std::move(C));
    }

    HIP_TRY(hipEventRecord(t1, st));
    std::vector<uint64_t> ctr((size_t)R * CTR_N);
    d2h(ctr.data(), d_ctr, ctr.size(), st);
    HIP_TRY(hipStreamSynchronize(st));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
    prof_resolve(ctx);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    S->device_ms = ms;

    S->counts.resize(R);
    for (uint32_t r = 0; r < R; ++r) {
        hgx_counts& c = S->counts[r];
        std::memset(&c, 0, sizeof(c));
        const uint64_t* x = &ctr[(size_t)r * CTR_N];
        c.raw_states = states[r];
        c.events = events[r];
        c.canonical_states = x[CTR_CANON];
        c.fingerprint = x[CTR_FINGERPRINT];
        c.causal_triples = x[CTR_TRIPLES];
        c.causal_pairs = x[CTR_PAIRS];
        const bool tr = S->runs[r].flags & HGX_OPT_TR;
        c.causal_pairs_tr = tr ? x[CTR_PAIRS_TR] : x[CTR_PAIRS];
        c.causal_triples_tr = tr ? x[CTR_TRIPLES_TR] : x[CTR_TRIPLES];
        c.branchial_pairs = x[CTR_BRANCHIAL];
        c.steps_done = steps_done[r];
        c.truncated = truncated[r];
    }
    return S;
}

// Copy what readback needs to the host (once per batch).
static void fetch_host(BatchStore& S) {
    std::lock_guard<std::mutex> g(S.mu);
    if (S.host_ready) return;
    hipStream_t st = S.st;
    const size_t G = S.gens.size();
    S.h_hash.resize(G); S.h_sid.resize(G); S.h_in.resize(G); S.h_rule.resize(G);
    S.h_prod.resize(G); S.h_tr.resize(G);
    for (size_t i = 0; i < G; ++i) {
        const Gen& g = S.gens[i];
        S.h_hash[i].resize(g.v.n); S.h_sid[i].resize(g.v.n);
        d2h(S.h_hash[i].data(), g.v.hash, g.v.n, st);
        d2h(S.h_sid[i].data(), g.v.sid, g.v.n, st);
        if (i > 0) {
            S.h_in[i].resize(g.v.n); S.h_rule[i].resize(g.v.n);
            S.h_prod[i].resize(g.v.n); S.h_tr[i].resize(g.v.n);
            d2h(S.h_in[i].data(), g.ev.in_state, g.v.n, st);
            d2h(S.h_rule[i].data(), g.ev.rule, g.v.n, st);
            d2h(S.h_prod[i].data(), g.ev.prod, g.v.n, st);
            d2h(S.h_tr[i].data(), g.ev.tr, g.v.n, st);
        }
    }
    HIP_TRY(hipStreamSynchronize(st));
    S.host_ready = true;
}

}  // namespace hgx

using namespace hgx;

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------

static size_t copy_str(const std::string& s, char* buf, size_t len) {
    if (buf && len) {
        const size_t n = std::min(len - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return s.size();
}

extern "C" int hgx_ctx_create(int device, uint64_t budget_bytes, hgx_ctx** out) {
    if (!out) return HGX_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        g_create_error = "no HIP device";
        return HGX_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) { g_create_error = "device index out of range"; return HGX_ERR_INVALID; }
    if (hipSetDevice(device) != hipSuccess) { g_create_error = "hipSetDevice failed"; return HGX_ERR_DEVICE; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { g_create_error = "hipGetDeviceProperties failed"; return HGX_ERR_DEVICE; }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_create_error = std::string("kernels are built for gfx950; device is ") + prop.gcnArchName;
        return HGX_ERR_NO_DEVICE;
    }
    auto* c = new hgx_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        g_create_error = "hipStreamCreate failed";
        return HGX_ERR_DEVICE;
    }
    size_t freeb = 0, totalb = 0;
    (void)hipMemGetInfo(&freeb, &totalb);
    const uint64_t budget = budget_bytes ? budget_bytes : (uint64_t)(freeb * 0.8);
    c->pool = std::make_shared<ChunkPool>(budget, (size_t)256 << 20);
    c->num_cu = (uint32_t)prop.multiProcessorCount;
    c->max_lds = (uint32_t)prop.sharedMemPerBlock;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s, %d CUs, %.1f GiB HBM, LDS/block %u B, budget %.1f GiB",
                  prop.gcnArchName, prop.multiProcessorCount, totalb / 1073741824.0, c->max_lds,
                  budget / 1073741824.0);
    c->info = buf;
    *out = c;
    return HGX_OK;
}

extern "C" void hgx_ctx_destroy(hgx_ctx* ctx) {
    if (!ctx) return;
    serve_destroy(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" size_t hgx_last_error(const hgx_ctx* ctx, char* buf, size_t len) {
    return copy_str(ctx ? ctx->err : g_create_error, buf, len);
}

extern "C" size_t hgx_device_info(const hgx_ctx* ctx, char* buf, size_t len) {
    return copy_str(ctx ? ctx->info : std::string(), buf, len);
}

template <class F>
static int guarded(hgx_ctx* ctx, F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        if (ctx) ctx->err = e.msg;
        return e.code;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return HGX_ERR_DEVICE;
    }
}

extern "C" int hgx_evolve_batch(hgx_ctx* ctx, const hgx_run_desc* runs, uint32_t n, hgx_result** results) {
    if (!ctx || (!runs && n) || (!results && n)) return HGX_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        std::vector<RunHost> hs;
        hs.reserve(n);
        for (uint32_t i = 0; i < n; ++i) hs.push_back(parse_run(&runs[i]));
        auto S = run_batch(ctx, std::move(hs));
        for (uint32_t i = 0; i < n; ++i) {
            auto* r = new hgx_result();
            r->store = S;
            r->run = i;
            r->counts = S->counts[i];
            r->device_ns = (uint64_t)(S->device_ms * 1e6);
            results[i] = r;
        }
        return HGX_OK;
    });
}

extern "C" int hgx_result_counts(const hgx_result* r, hgx_counts* out) {
    if (!r || !out) return HGX_ERR_INVALID;
    *out = r->counts;
    return HGX_OK;
}

extern "C" int hgx_result_timing(const hgx_result* r, uint64_t* device_ns, uint64_t* latency_ns) {
    if (!r) return HGX_ERR_INVALID;
    if (device_ns) *device_ns = r->device_ns;
    if (latency_ns) *latency_ns = r->latency_ns;
    return HGX_OK;
}

extern "C" void hgx_result_release(hgx_result* r) { delete r; }

// Per-run walk over generations: calls f(gen, index in gen) for the run's states in id order.
template <class F>
static void for_run_states(const BatchStore& S, uint32_t run, F&& f) {
    for (size_t g = 0; g < S.gens.size(); ++g) {
        const Gen& G = S.gens[g];
        if (run >= G.sbeg.size()) continue;
        for (uint32_t i = G.sbeg[run]; i < G.send[run]; ++i) f(g, i);
    }
}

// Per-run state/event tables on the host, built from the generation caches.
struct RunTables {
    std::vector<uint64_t> hash;      // by state id
    std::vector<uint32_t> step;      // by state id
    std::vector<uint32_t> creator;   // by state id (event id, INV for roots)
    std::vector<uint32_t> ev_in, ev_rule, ev_step;   // by event id
    std::vector<uint64_t> ev_prod, ev_tr;
};

static RunTables run_tables(BatchStore& S, uint32_t run, uint64_t n_states) {
    fetch_host(S);
    RunTables T;
    T.hash.resize(n_states); T.step.resize(n_states); T.creator.assign(n_states, INV);
    const uint32_t nroots = (uint32_t)S.runs[run].roots.size();
    const uint64_t n_ev = n_states >= nroots ? n_states - nroots : 0;
    T.ev_in.resize(n_ev); T.ev_rule.resize(n_ev); T.ev_step.resize(n_ev);
    T.ev_prod.resize(n_ev); T.ev_tr.resize(n_ev);
    for_run_states(S, run, [&](size_t g, uint32_t i) {
        const uint32_t sid = S.h_sid[g][i];
        if (sid >= n_states) throw Error{HGX_ERR_DEVICE, "state id out of range in readback"};
        T.hash[sid] = S.h_hash[g][i];
        T.step[sid] = (uint32_t)g;
        if (g > 0) {
            const uint32_t e = sid - nroots;
            T.creator[sid] = e;
            T.ev_in[e] = S.h_in[g][i];
            T.ev_rule[e] = S.h_rule[g][i];
            T.ev_step[e] = (uint32_t)g;
            T.ev_prod[e] = S.h_prod[g][i];
            T.ev_tr[e] = S.h_tr[g][i];
        }
    });
    return T;
}

extern "C" int hgx_result_states(const hgx_result* r, uint64_t* hash, uint32_t* step, uint32_t* creator,
                                 uint64_t cap) {
    if (!r) return HGX_ERR_INVALID;
    if (cap < r->counts.raw_states) return HGX_ERR_INVALID;
    try {
        RunTables T = run_tables(*r->store, r->run, r->counts.raw_states);
        for (uint64_t s = 0; s < r->counts.raw_states; ++s) {
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
    if (!r || (!out && r->counts.events)) return HGX_ERR_INVALID;
    if (cap < r->counts.events) return HGX_ERR_INVALID;
    try {
        RunTables T = run_tables(*r->store, r->run, r->counts.raw_states);
        const uint32_t nroots = (uint32_t)r->store->runs[r->run].roots.size();
        for (uint64_t e = 0; e < r->counts.events; ++e)
            out[e] = hgx_event{T.ev_in[e], (uint32_t)(e + nroots), T.ev_rule[e], T.ev_step[e]};
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_causal(const hgx_result* r, hgx_pair* pairs, uint8_t* in_tr, uint64_t cap) {
    if (!r) return HGX_ERR_INVALID;
    const BatchStore& S0 = *r->store;
    if (!(S0.runs[r->run].flags & HGX_REC_CAUSAL)) return HGX_ERR_INVALID;
    if (cap < r->counts.causal_pairs) return HGX_ERR_INVALID;
    try {
        RunTables T = run_tables(*r->store, r->run, r->counts.raw_states);
        uint64_t o = 0;
        for (uint64_t c = 0; c < r->counts.events; ++c) {
            for (uint64_t m = T.ev_prod[c]; m; m &= m - 1) {
                const uint32_t s = (uint32_t)__builtin_ctzll(m);
                // walk the lineage of c's input state up to the generation-s event
                uint32_t x = T.ev_in[c];
                while (T.step[x] > s) x = T.ev_in[T.creator[x]];
                const uint32_t p = T.creator[x];
                if (o >= cap) return HGX_ERR_INVALID;
                if (pairs) pairs[o] = hgx_pair{p, (uint32_t)c};
                if (in_tr) in_tr[o] = (uint8_t)((T.ev_tr[c] >> s) & 1u);
                ++o;
            }
        }
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_branchial(const hgx_result* r, hgx_pair* pairs, uint64_t cap) {
    if (!r) return HGX_ERR_INVALID;
    BatchStore& S = *r->store;
    const uint32_t f = S.runs[r->run].flags;
    if (!(f & HGX_REC_BRANCHIAL) || !(f & HGX_REC_RELATIONS)) return HGX_ERR_INVALID;
    if (cap < r->counts.branchial_pairs) return HGX_ERR_INVALID;
    try {
        uint64_t o = 0;
        for (size_t g = 1; g < S.gens.size(); ++g) {
            const Gen& G = S.gens[g];
            if (!G.n_br || G.br_run_off.empty()) continue;
            const uint64_t b = G.br_run_off[r->run], e = G.br_run_off[r->run + 1];
            if (e <= b) continue;
            std::vector<uint32_t> tmp(2 * (e - b));
            d2h(tmp.data(), G.br_pairs + 2 * b, tmp.size(), S.st);
            HIP_TRY(hipStreamSynchronize(S.st));
            for (uint64_t i = 0; i < e - b; ++i) {
                if (o >= cap) return HGX_ERR_INVALID;
                pairs[o++] = hgx_pair{tmp[2 * i], tmp[2 * i + 1]};
            }
        }
        return HGX_OK;
    } catch (const Error& e) {
        return e.code;
    }
}

extern "C" int hgx_result_state_edges(const hgx_result* r, uint64_t* state_offsets, uint32_t* edge_offsets,
                                      uint32_t* vertices, uint64_t* n_edges_total, uint64_t* n_vertices_total) {
    if (!r) return HGX_ERR_INVALID;
    BatchStore& S = *r->store;
    try {
        uint64_t ne_tot = 0, nv_tot = 0, s_idx = 0;
        const uint32_t A = S.A;
        for (size_t g = 0; g < S.gens.size(); ++g) {
            const Gen& G = S.gens[g];
            if (r->run >= G.sbeg.size()) continue;
            const uint32_t sb = G.sbeg[r->run], se = G.send[r->run];
            if (se <= sb) continue;
            std::vector<uint64_t> eo(se - sb + 1);
            d2h(eo.data(), G.v.eoff + sb, eo.size(), S.st);
            HIP_TRY(hipStreamSynchronize(S.st));
            const uint64_t E0 = eo[0], E1 = eo.back();
            std::vector<uint8_t> ar(E1 - E0);
            std::vector<uint32_t> ev((E1 - E0) * A);
            d2h(ar.data(), G.v.ar + E0, ar.size(), S.st);
            d2h(ev.data(), G.v.ev + E0 * A, ev.size(), S.st);
            HIP_TRY(hipStreamSynchronize(S.st));
            for (uint32_t i = 0; i < se - sb; ++i) {
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

extern "C" int hgx_wl_hash_batch(hgx_ctx* ctx, const uint64_t* state_offsets, uint32_t n_states,
                                 const uint32_t* edge_offsets, const uint32_t* vertices, uint64_t* out_hash) {
    if (!ctx || (n_states && (!state_offsets || !out_hash))) return HGX_ERR_INVALID;
    if (n_states == 0) return HGX_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        const uint64_t E = state_offsets[n_states] - state_offsets[0];
        uint32_t A = 1, max_ne = 1, max_nv = 1;
        for (uint64_t e = state_offsets[0]; e < state_offsets[n_states]; ++e) {
            const uint32_t a = edge_offsets[e + 1] - edge_offsets[e];
            if (a < 1 || a > 16) throw Error{HGX_ERR_INVALID, "edge arity must be 1..16"};
            A = std::max(A, a);
        }
        std::vector<uint32_t> ev(E * A, 0), nv(n_states);
        std::vector<uint8_t> ar(E);
        std::vector<uint64_t> so(n_states + 1);
        for (uint32_t s = 0; s < n_states; ++s) {
            std::unordered_map<uint32_t, uint32_t> lab;
            so[s] = state_offsets[s] - state_offsets[0];
            const uint64_t e0 = state_offsets[s], e1 = state_offsets[s + 1];
            if (e1 < e0) throw Error{HGX_ERR_INVALID, "state_offsets not ascending"};
            max_ne = std::max<uint32_t>(max_ne, (uint32_t)(e1 - e0));
            for (uint64_t e = e0; e < e1; ++e) {
                const uint32_t a = edge_offsets[e + 1] - edge_offsets[e];
                ar[e - state_offsets[0]] = (uint8_t)a;
                for (uint32_t k = 0; k < a; ++k) {
                    auto it = lab.emplace(vertices[edge_offsets[e] + k], (uint32_t)lab.size()).first;
                    ev[(e - state_offsets[0]) * A + k] = it->second;
                }
            }
            nv[s] = (uint32_t)lab.size();
            max_nv = std::max(max_nv, nv[s]);
        }
        so[n_states] = E;
        Arena ar_(ctx->pool.get());
        hipStream_t st = ctx->stream;
        WLStatesArgs w{};
        w.n = n_states; w.A = A;
        uint64_t* d_so = dalloc<uint64_t>(ar_, n_states + 1);
        uint8_t* d_ar = dalloc<uint8_t>(ar_, E);
        uint32_t* d_ev = dalloc<uint32_t>(ar_, E * A);
        uint32_t* d_nv = dalloc<uint32_t>(ar_, n_states);
        uint64_t* d_out = dalloc<uint64_t>(ar_, n_states);
        h2d(d_so, so.data(), n_states + 1, st);
        h2d(d_ar, ar.data(), E, st);
        h2d(d_ev, ev.data(), E * A, st);
        h2d(d_nv, nv.data(), n_states, st);
        w.state_off = d_so; w.ar = d_ar; w.ev = d_ev; w.nv = d_nv; w.out = d_out;
        w.max_ne = max_ne; w.max_nv = max_nv;
        w.lds_per_wave = wl_states_lds_per_wave(max_ne, max_nv, A);
        const uint32_t wpb = waves_per_block(w.lds_per_wave, ctx->max_lds);
        HIP_TRY(launch_wl_states(w, ctx->num_cu * 8, wpb, st));
        d2h(out_hash, d_out, n_states, st);
        HIP_TRY(hipStreamSynchronize(st));
        return HGX_OK;
    });
}

// The served path and the large-state join are implemented in hgx_serve.cpp / hgx_bigjoin.hip.
namespace hgx {
void serve_destroy(hgx_ctx*) {}
}

extern "C" uint64_t hgx_request_footprint(const hgx_run_desc* run) {
    if (!run) return 0;
    // states the request may create (cap, or a generous bound) x bytes per state
    const uint64_t states = run->state_cap ? run->state_cap : 1u << 20;
    const uint64_t ne = run->n_edges + 16ull * run->steps;
    return states * (64 + ne * 10 + 8ull * run->steps) + (uint64_t)ne * 64;
}

extern "C" int hgx_request_submit(hgx_ctx* ctx, const hgx_run_desc*, uint64_t*) {
    if (ctx) ctx->err = "served path not built in this library";
    return HGX_ERR_UNSUPPORTED;
}
extern "C" int hgx_serve_step(hgx_ctx* ctx, uint32_t*) {
    if (ctx) ctx->err = "served path not built in this library";
    return HGX_ERR_UNSUPPORTED;
}
extern "C" int hgx_request_poll(hgx_ctx* ctx, uint64_t, hgx_result**) {
    if (ctx) ctx->err = "served path not built in this library";
    return HGX_ERR_UNSUPPORTED;
}
extern "C" int hgx_request_wait(hgx_ctx* ctx, uint64_t, hgx_result**) {
    if (ctx) ctx->err = "served path not built in this library";
    return HGX_ERR_UNSUPPORTED;
}
extern "C" int hgx_match_count(hgx_ctx* ctx, const uint8_t*, uint32_t, uint32_t, uint32_t, const uint32_t*,
                               const uint32_t*, uint32_t, uint64_t*, uint32_t*, uint64_t, float*) {
    if (ctx) ctx->err = "large-state join not built in this library";
    return HGX_ERR_UNSUPPORTED;
}

extern "C" int hgx_profile_enable(hgx_ctx* ctx, int on) {
    if (!ctx) return HGX_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->prof.on = on != 0;
    return HGX_OK;
}

extern "C" int hgx_profile_read(hgx_ctx* ctx, hgx_kernel_stat* out, uint32_t cap, uint32_t* n) {
    if (!ctx || !n) return HGX_ERR_INVALID;
    std::lock_guard<std::mutex> g(ctx->mu);
    uint32_t k = 0;
    for (auto& kv : ctx->prof.acc) {
        if (out && k < cap) out[k] = kv.second;
        ++k;
    }
    *n = k;
    ctx->prof.acc.clear();
    return HGX_OK;
}
