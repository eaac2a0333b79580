// hgx_runtime.cpp -- host side of the C ABI (include/hgx.h) that is not the generation
// engine: device memory (chunk pool, arenas, slabs), request parsing with the reference's join
// order, the context, kernel profiling, and the stand-alone WL hash entry point.
// The generation engine (batch and served paths, results) is hgx_engine.cpp.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <random>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <unordered_map>

#include "hgx_engine.hpp"
#include "hgx_kernels.hpp"
#include "hgx_runtime.hpp"

namespace hgx {

static std::string g_create_error;

// ------------------------------------------------------------------------------------------
// device memory
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
        in_use_ += c.size;
        return c;
    }
    size_t sz = std::max(min_bytes, chunk_bytes_);
    sz = (sz + (2u << 20) - 1) & ~((size_t)(2u << 20) - 1);
    if (total_ + sz > budget_) {
        // give back idle chunks before refusing (after the device is idle: a chunk returned to
        // the pool may still be read by work queued before it was returned)
        if (!free_.empty()) (void)hipDeviceSynchronize();
        for (auto& c : free_) { (void)hipFree(c.p); total_ -= c.size; }
        free_.clear();
        if (total_ + sz > budget_)
            throw Error{HGX_ERR_CAPACITY, "HBM budget exhausted: need " + std::to_string(sz >> 20) +
                                              " MiB more, budget " + std::to_string(budget_ >> 20) + " MiB"};
    }
    void* p = nullptr;
    if (hipMalloc(&p, sz) != hipSuccess) {
        (void)hipGetLastError();
        throw Error{HGX_ERR_CAPACITY, "the device has no " + std::to_string(sz >> 20) + " MiB left"};
    }
    total_ += sz;
    in_use_ += sz;
    return Chunk{p, sz};
}

void ChunkPool::put(const Chunk& c) {
    std::lock_guard<std::mutex> g(mu_);
    if (holding_) { held_.push_back(c); return; }      // stays in use, owned by the recording
    in_use_ -= c.size;
    free_.push_back(c);
}

void ChunkPool::hold(bool on) {
    std::lock_guard<std::mutex> g(mu_);
    holding_ = on;
}

std::vector<Chunk> ChunkPool::take_held() {
    std::lock_guard<std::mutex> g(mu_);
    std::vector<Chunk> h;
    h.swap(held_);
    return h;
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
        chunks_.push_back(pool_->get(std::max(bytes, min_chunk_)));
        used_ = 0;
    }
    void* p = (uint8_t*)chunks_.back().p + used_;
    used_ += bytes;
    bytes_ += bytes;
    return p;
}


SlabAlloc::~SlabAlloc() {
    for (auto& c : slabs_) pool_->put(c);
    for (auto& kv : big_) pool_->put(kv.second);
}

static uint32_t size_class(size_t bytes, uint32_t min_class) {
    uint32_t c = min_class;
    while (((size_t)1 << c) < bytes) ++c;
    return c;
}

void* SlabAlloc::alloc(size_t bytes) {
    const uint32_t c = size_class(bytes, kMinClass);
    if (c > kMaxClass) {                              // above 16 MiB: a chunk of its own
        Chunk ch = pool_->get(bytes);
        big_[ch.p] = ch;
        return ch.p;
    }
    auto& fl = free_[c];
    if (fl.empty()) {
        const size_t blk = (size_t)1 << c;
        if (slabs_.empty() || slab_used_ + blk > slabs_.back().size) {
            slabs_.push_back(pool_->get((size_t)64 << 20));
            slab_used_ = 0;
        }
        fl.push_back((uint8_t*)slabs_.back().p + slab_used_);
        slab_used_ += blk;
    }
    void* p = fl.back();
    fl.pop_back();
    return p;
}

void SlabAlloc::free(void* p, size_t bytes) {
    if (!p) return;
    const uint32_t c = size_class(bytes, kMinClass);
    if (c > kMaxClass) {
        auto b = big_.find(p);
        if (b != big_.end()) { pool_->put(b->second); big_.erase(b); }
        return;
    }
    free_[c].push_back(p);
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

void parse_rules(const uint8_t* p, uint32_t len, uint32_t n, std::vector<DRule>& out) {
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
        // pivot of each depth: the first position whose variable an earlier depth bound
        std::memset(R.piv, 0xFF, sizeof(R.piv));
        uint32_t bound = 0;
        for (int d = 0; d < R.n_lhs; ++d) {
            const int pe = R.order[d];
            if (d > 0)
                for (int i = 0; i < R.la[pe]; ++i)
                    if (bound & (1u << R.lv[pe][i])) { R.piv[d] = (uint8_t)i; break; }
            bound |= var_mask(R.lv[pe], R.la[pe]);
        }
        std::memset(R.jsrc, 0xFF, sizeof(R.jsrc));
        {
            uint8_t first[32];
            std::memset(first, 0xFF, sizeof(first));
            for (int d = 0; d < R.n_lhs; ++d) {
                const int pe = R.order[d];
                for (int i = 0; i < R.la[pe]; ++i) {
                    const uint8_t v = R.lv[pe][i];
                    if (first[v] != 0xFF) R.jsrc[d][i] = first[v];
                    else first[v] = (uint8_t)((d << 4) | i);
                }
            }
        }
        // anchored orders: position p first, then most shared bound variables (ties: lower
        // position); the choice only affects join cost, never the set of matches
        std::memset(R.aord, 0, sizeof(R.aord));
        std::memset(R.apiv, 0xFF, sizeof(R.apiv));
        for (int p0 = 0; p0 < R.n_lhs; ++p0) {
            bool used[16] = {};
            R.aord[p0][0] = (uint8_t)p0; used[p0] = true;
            uint32_t bnd = var_mask(R.lv[p0], R.la[p0]);
            for (int d = 1; d < R.n_lhs; ++d) {
                int pick = -1, ps = -1;
                for (int e = 0; e < R.n_lhs; ++e) {
                    if (used[e]) continue;
                    const int sh = __builtin_popcount(var_mask(R.lv[e], R.la[e]) & bnd);
                    if (sh > ps) { pick = e; ps = sh; }
                }
                R.aord[p0][d] = (uint8_t)pick; used[pick] = true;
                for (int i = 0; i < R.la[pick]; ++i)
                    if (bnd & (1u << R.lv[pick][i])) { R.apiv[p0][d] = (uint8_t)i; break; }
                bnd |= var_mask(R.lv[pick], R.la[pick]);
            }
        }
        out.push_back(R);
    }
    if (k != len) throw Error{HGX_ERR_INVALID, "rule program has trailing bytes"};
}

RunHost parse_run(const hgx_run_desc* d, bool evolve) {
    if (!d) throw Error{HGX_ERR_INVALID, "null request"};
    if (!d->rules || d->n_rules == 0) throw Error{HGX_ERR_INVALID, "request has no rules"};
    if (d->steps > HGX_MAX_STEPS) throw Error{HGX_ERR_UNSUPPORTED, "steps above HGX_MAX_STEPS"};
    RunHost h;
    parse_rules(d->rules, d->rules_len, d->n_rules, h.rules);
    h.steps = d->steps; h.cap = d->state_cap; h.flags = d->flags;
    if (const hgx_sampling* sp = d->sampling) {
        SampleHost& m = h.smp;
        if (sp->n_rule_weights && !sp->rule_weights) throw Error{HGX_ERR_INVALID, "rule weights missing"};
        m.rate = sp->transition_rate;
        m.weights.assign(sp->rule_weights, sp->rule_weights + sp->n_rule_weights);
        m.seed = sp->seed;
        m.per_step = sp->max_states_per_step;
        m.per_parent = sp->max_successor_states_per_parent;
        m.per_state_rule = sp->matches_per_state_rule;
        float p = (float)sp->exploration_probability;
        if (!(p > 0.0f)) p = 0.0f;
        if (p > 1.0f) p = 1.0f;
        m.thr = p >= 1.0f ? 0xFFFFFFFFu : p <= 0.0f ? 0u : (uint32_t)((double)p * 4294967296.0);
        m.coin_seed = m.seed;
        if (m.coin_seed == 0 && m.thr != 0xFFFFFFFFu) {
            std::random_device rd;
            m.coin_seed = ((uint64_t)rd() << 32) | rd();
            if (m.coin_seed == 0) m.coin_seed = 0xA5A5A5A5A5A5A5A5ull;
        }
        m.on = m.transitions() || m.frontier();
        if (m.on && h.cap)
            throw Error{HGX_ERR_UNSUPPORTED, "a state cap together with sampling (the cap counts states before thinning)"};
    }
    if ((h.flags & HGX_OPT_QUOTIENT_ROOTS) && !(h.flags & HGX_OPT_QUOTIENT))
        throw Error{HGX_ERR_INVALID, "HGX_OPT_QUOTIENT_ROOTS without HGX_OPT_QUOTIENT"};
    if ((h.flags & HGX_OPT_QUOTIENT) && h.cap)
        throw Error{HGX_ERR_UNSUPPORTED, "a state cap together with quotient exploration (the cap counts raw states)"};
    if (d->n_edges && (!d->edge_offsets || !d->vertices)) throw Error{HGX_ERR_INVALID, "null state arrays"};
    const uint32_t n_roots = d->root_offsets ? d->n_roots : 1;
    if (d->root_offsets && (d->n_roots == 0 || d->root_offsets[0] != 0 || d->root_offsets[d->n_roots] != d->n_edges))
        throw Error{HGX_ERR_INVALID, "root_offsets must span the edges"};
    uint32_t amin = 0xFFu, amax = 0, lmax = 0;
    for (uint32_t e = 0; e < d->n_edges; ++e) {
        if (d->edge_offsets[e + 1] < d->edge_offsets[e] || d->edge_offsets[e + 1] - d->edge_offsets[e] < 1 ||
            d->edge_offsets[e + 1] - d->edge_offsets[e] > 16)
            throw Error{HGX_ERR_INVALID, "state edge arity must be 1..16"};
        const uint32_t a = d->edge_offsets[e + 1] - d->edge_offsets[e];
        amin = std::min(amin, a);
        amax = std::max(amax, a);
    }
    const uint32_t base = d->n_edges ? d->edge_offsets[0] : 0;        // offsets may start anywhere
    const uint32_t nvert = d->n_edges ? d->edge_offsets[d->n_edges] - base : 0;
    const uint32_t* vin = d->n_edges ? d->vertices + base : nullptr;
    for (uint32_t i = 0; i < nvert; ++i) lmax = std::max(lmax, vin[i]);
    h.max_arity = std::max(h.max_arity, amax);
    h.edge_vb.resize((size_t)d->n_edges + 1);
    h.edge_vb[0] = 0;
    for (uint32_t e = 0; e < d->n_edges; ++e) h.edge_vb[e + 1] = d->edge_offsets[e + 1] - base;
    h.verts.resize(nvert);
    h.root_eb.resize(n_roots + 1);
    h.root_eb[0] = 0;
    // Renumbering by first appearance within each root (FFI :599-636): a dense table when the
    // labels are small, a hash map otherwise.
    const bool dense = (uint64_t)lmax < 8ull * nvert + 4096;
    std::vector<uint32_t> tab;
    std::unordered_map<uint32_t, uint32_t> lab;
    if (dense) tab.assign((size_t)lmax + 1, INV);
    for (uint32_t r = 0; r < n_roots; ++r) {
        const uint32_t e0 = d->root_offsets ? d->root_offsets[r] : 0;
        const uint32_t e1 = d->root_offsets ? d->root_offsets[r + 1] : d->n_edges;
        if (e1 < e0 || e1 > d->n_edges) throw Error{HGX_ERR_INVALID, "root_offsets not ascending"};
        const uint32_t v0 = h.edge_vb[e0], v1 = h.edge_vb[e1];
        uint32_t nv = 0;
        if (dense) {
            for (uint32_t i = v0; i < v1; ++i) {
                uint32_t& t = tab[vin[i]];
                if (t == INV) t = nv++;
                h.verts[i] = t;
            }
            for (uint32_t i = v0; i < v1; ++i) tab[vin[i]] = INV;
        } else {
            lab.clear();
            for (uint32_t i = v0; i < v1; ++i) h.verts[i] = lab.emplace(vin[i], (uint32_t)lab.size()).first->second;
            nv = (uint32_t)lab.size();
        }
        h.root_eb[r + 1] = e1;
        h.root_nv.push_back(nv);
        h.max_root_ne = std::max(h.max_root_ne, e1 - e0);
        h.max_root_nv = std::max(h.max_root_nv, nv);
    }
    for (auto& R : h.rules) {
        for (int e = 0; e < R.n_rhs; ++e) h.max_arity = std::max<uint32_t>(h.max_arity, R.ra[e]);
        for (int e = 0; e < R.n_lhs; ++e) h.max_arity = std::max<uint32_t>(h.max_arity, R.la[e]);
        h.max_lhs = std::max<uint32_t>(h.max_lhs, R.n_lhs);
        h.max_gain = std::max<uint32_t>(h.max_gain, R.n_rhs > R.n_lhs ? R.n_rhs - R.n_lhs : 0);
        h.max_new = std::max<uint32_t>(h.max_new, R.n_new);
    }
    h.root_edges = d->n_edges;
    h.root_uni = d->n_edges == 0 || (amin == amax && amax >= 2 && amax <= 4);
    for (auto& R : h.rules)
        for (int e = 0; e < R.n_rhs; ++e) { amin = std::min<uint32_t>(amin, R.ra[e]); amax = std::max<uint32_t>(amax, R.ra[e]); }
    h.uni_arity = (amin == amax && amax >= 2 && amax <= 4) ? amax : 0u;
    // edge positions (match list) and local vertex ids (hash workspaces) are 16-bit: refused
    // before any kernel sees a state that could exceed them (one larger state's matches:
    // hgx_match_count)
    if (evolve && (uint64_t)h.max_root_nv + (uint64_t)h.steps * h.max_new > 65535)
        throw Error{HGX_ERR_UNSUPPORTED, "a state of this request can exceed 65535 vertices (16-bit local "
                                         "vertex ids in the hash workspaces)"};
    if (evolve && (uint64_t)h.max_root_ne + (uint64_t)h.steps * h.max_gain > 65535)
        throw Error{HGX_ERR_UNSUPPORTED, "a state of this request can exceed 65535 edges (largest initial state " +
                                             std::to_string(h.max_root_ne) + " edges, " + std::to_string(h.steps) +
                                             " steps); one such state's matches: hgx_match_count"};
    return h;
}


// HBM a request is charged at admission: every state it may create (its cap, or 2^20 when
// uncapped) at its largest edge count, in the generation that holds it and the one built from
// it, plus its state map. With `held` > 0: the same for that many states (an uncapped request is
// re-charged as it grows).
uint64_t footprint_bytes(const RunHost& h, uint64_t held) {
    uint64_t max_ne = h.max_root_ne;
    max_ne += (uint64_t)h.steps * h.max_gain;
    uint64_t states = h.cap ? (uint64_t)h.cap : (uint64_t)1 << 20;
    if (held) states = held;
    const uint64_t per_state = 2 * max_ne * (4ull * h.max_arity + 2) + 96 + 8ull * h.steps;
    uint64_t map = 64;
    while (map < 2 * states) map <<= 1;
    return states * per_state + map * 8 + h.rules.size() * sizeof(DRule) + 4096;
}

struct CastU64 {
    __host__ __device__ uint64_t operator()(uint32_t x) const { return x; }
};

// Exclusive prefix sum of n uint32 values into n+1 uint64 offsets (last = total).
void scan_u32_to_u64(Arena& ar, const uint32_t* in, uint64_t* out, uint32_t n, hipStream_t st) {
    // in must hold n+1 readable values with in[n] == 0
    hipcub::TransformInputIterator<uint64_t, CastU64, const uint32_t*> it(in, CastU64());
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, it, out, n + 1, st));
    void* t = ar.alloc(tmp);
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(t, tmp, it, out, n + 1, st));
}

// Sorts n (key, value) pairs by key bits [0, end_bit), stable.
void sort_pairs_u64(Arena& ar, const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout,
                    uint64_t n, int end_bit, hipStream_t st) {
    if (n >= 0x7FFFFFFFull) throw Error{HGX_ERR_UNSUPPORTED, "sort above 2^31 entries"};
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, vout, (int)n, 0, end_bit, st));
    void* t = ar.alloc(tmp);
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(t, tmp, kin, kout, vin, vout, (int)n, 0, end_bit, st));
}

uint32_t place_wl_states(Arena& ar, WLStatesArgs& w, bool lean, uint32_t max_lds, uint32_t grid_cap,
                         uint32_t forced_team, uint32_t* wpb) {
    w.lean = lean && w.max_ne < 65536 ? 1u : 0u;
    w.lds_per_wave = wl_states_lds_per_wave(w.max_ne, w.max_nv, w.A, w.lean != 0);
    w.scratch = nullptr;
    const uint32_t team = pick_team(w.lds_per_wave, forced_team);
    if (team == 1) {
        *wpb = place_workspace(ar, w.lds_per_wave, w.n, max_lds, grid_cap, &w.scratch);
    } else {
        *wpb = 1;
        if (w.lds_per_wave > max_lds)
            w.scratch = (uint8_t*)ar.alloc((size_t)wl_states_grid(w.n, team, 1, grid_cap) * w.lds_per_wave);
    }
    return team;
}

uint32_t waves_per_block(uint32_t lds_per_wave, uint32_t max_lds) {
    uint32_t w = 4;
    while (w > 1 && (uint64_t)lds_per_wave * w > max_lds) --w;
    if ((uint64_t)lds_per_wave * w > max_lds)
        throw Error{HGX_ERR_UNSUPPORTED, "a state is too large for the LDS-resident path (" +
                                             std::to_string(lds_per_wave) + " bytes per wave)"};
    return w;
}

uint32_t place_workspace(Arena& ar, uint32_t per_wave, uint32_t items, uint32_t max_lds, uint32_t grid_cap,
                         uint8_t** scratch) {
    *scratch = nullptr;
    if (per_wave <= max_lds) return waves_per_block(per_wave, max_lds);
    const uint32_t wpb = 4;
    const uint32_t grid = launch_grid(items, wpb, grid_cap);
    *scratch = (uint8_t*)ar.alloc((size_t)grid * wpb * per_wave);
    return wpb;
}

// After a stream synchronisation: fold finished launches into the per-kernel totals.
void prof_resolve(hgx_ctx* ctx) {
    for (auto& p : ctx->prof.pend) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
        auto add = [&](const std::string& name) {
            hgx_kernel_stat& s = ctx->prof.acc[name];
            std::snprintf(s.name, sizeof(s.name), "%s", name.c_str());
            s.launches += 1;
            s.total_ms += ms;
            s.alg_bytes += p.bytes;
            s.items += p.items;
        };
        add(p.name);
        if (ctx->prof.by_class && p.cls >= 0) add(p.name + "#" + std::to_string(p.cls));
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    ctx->prof.pend.clear();
}


}  // namespace hgx

using namespace hgx;

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
    try {
        noalloc_setup(device);                 // the allocation-free entry point's scratch address
    } catch (const Error& e) {
        g_create_error = e.msg;
        return e.code;
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
    c->pool = std::make_shared<ChunkPool>(budget, (size_t)64 << 20);
    c->num_cu = (uint32_t)prop.multiProcessorCount;
    c->max_lds = (uint32_t)prop.sharedMemPerBlock;
    (void)tune_load(c, default_tune_cache());   // tuned launch shapes of this device, when cached
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
    engine_destroy(ctx->serve);
    if (ctx->flush_buf) (void)hipFree(ctx->flush_buf);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" size_t hgx_last_error(const hgx_ctx* ctx, char* buf, size_t len) {
    return copy_str(ctx ? ctx->err : g_create_error, buf, len);
}

extern "C" size_t hgx_device_info(const hgx_ctx* ctx, char* buf, size_t len) {
    return copy_str(ctx ? ctx->info : std::string(), buf, len);
}

extern "C" int hgx_ctx_set_option(hgx_ctx* ctx, uint32_t option, int64_t value) {
    if (!ctx) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (option == HGX_OPTION_FORWARD && value >= -1 && value <= 1) {
        ctx->forward = (int)value;
        return HGX_OK;
    }
    if (option == HGX_OPTION_DELTA_CAP && value >= 1 && value <= 4096) {
        ctx->delta_cap = (uint32_t)value;
        return HGX_OK;
    }
    if (option == HGX_OPTION_EXPAND_TEAM && (value == 0 || value == 1 || value == 2 || value == 4 || value == 8)) {
        ctx->expand_team = (uint32_t)value;
        return HGX_OK;
    }
    if (option == HGX_OPTION_JOIN_TEAM && (value == 0 || value == 1 || value == 4 || value == 8)) {
        ctx->join_team = (uint32_t)value;
        return HGX_OK;
    }
    if (option == HGX_OPTION_EXPAND_LANES && (value == 0 || value == 16 || value == 32 || value == 64)) {
        ctx->expand_lanes = (uint32_t)value;
        return HGX_OK;
    }
    if (option == HGX_OPTION_JOIN_LANES && (value == 0 || value == 16 || value == 32 || value == 64)) {
        ctx->join_lanes = (uint32_t)value;
        return HGX_OK;
    }
    if (option == HGX_OPTION_COUNTER_FLUSH && (value == 0 || (value >= 16 && value <= 1024))) {
        return guarded(ctx, [&]() -> int {
            HIP_TRY(hipSetDevice(ctx->device));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            if (ctx->flush_buf) HIP_TRY(hipFree(ctx->flush_buf));
            ctx->flush_buf = nullptr;
            ctx->flush_bytes = 0;
            if (value) {
                HIP_TRY(hipMalloc(&ctx->flush_buf, (size_t)value << 20));
                HIP_TRY(launch_fill(ctx->flush_buf, 0, (size_t)value << 20, ctx->stream));
                ctx->flush_bytes = (uint64_t)value << 20;
            }
            return HGX_OK;
        });
    }
    ctx->err = "unknown option or value";
    return HGX_ERR_INVALID;
}


extern "C" int hgx_wl_hash_batch(hgx_ctx* ctx, const uint64_t* state_offsets, uint32_t n_states,
                                 const uint32_t* edge_offsets, const uint32_t* vertices, uint64_t* out_hash) {
    if (!ctx || (n_states && (!state_offsets || !out_hash))) return HGX_ERR_INVALID;
    if (n_states == 0) return HGX_OK;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        const uint64_t E = state_offsets[n_states] - state_offsets[0];
        uint32_t A = 1, max_ne = 1, max_nv = 1, amin = 0xFFu, amax = 0;
        for (uint64_t e = state_offsets[0]; e < state_offsets[n_states]; ++e) {
            const uint32_t a = edge_offsets[e + 1] - edge_offsets[e];
            if (a < 1 || a > 16) throw Error{HGX_ERR_INVALID, "edge arity must be 1..16"};
            A = std::max(A, a);
            amin = std::min(amin, a);
            amax = std::max(amax, a);
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
        uint32_t wpb = 1;
        const uint32_t team = place_wl_states(ar_, w, amin == amax && amax >= 2 && amax <= 4, ctx->max_lds,
                                              ctx->num_cu * 8, ctx->expand_team, &wpb);
        HIP_TRY(launch_wl_states(w, ctx->num_cu * 8, wpb, team, st));
        d2h(out_hash, d_out, n_states, st);
        HIP_TRY(hipStreamSynchronize(st));
        return HGX_OK;
    });
}

extern "C" int hgx_calibration_read(hgx_ctx* ctx, uint64_t bytes) {
    if (!ctx || bytes < 16) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        Arena ar(ctx->pool.get());
        uint8_t* buf = dalloc<uint8_t>(ar, bytes);
        uint64_t* sink = dalloc<uint64_t>(ar, 1);
        uint64_t* ctr = dalloc<uint64_t>(ar, 4096 * 8);
        HIP_TRY(launch_fill(buf, 1, bytes, ctx->stream));
        HIP_TRY(launch_fill(sink, 0, 8, ctx->stream));
        HIP_TRY(launch_fill(ctr, 0, 4096 * 64, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(launch_read_calibration(buf, bytes, sink, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(launch_read_gather(buf, bytes, sink, ctr, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        return HGX_OK;
    });
}

extern "C" int hgx_profile_enable(hgx_ctx* ctx, int on) {
    if (!ctx) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    ctx->prof.on = on != 0;
    return HGX_OK;
}

extern "C" int hgx_profile_read(hgx_ctx* ctx, hgx_kernel_stat* out, uint32_t cap, uint32_t* n) {
    if (!ctx || !n) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    uint32_t k = 0;
    for (auto& kv : ctx->prof.acc) {
        if (out && k < cap) out[k] = kv.second;
        ++k;
    }
    *n = k;
    ctx->prof.acc.clear();
    return HGX_OK;
}
