// hgx_runtime.hpp -- host-side pieces shared by the runtime files: errors, device memory
// (chunk pool, bump arenas, slab allocator), request parsing, scan helper, the context.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/hgx.h"
#include "hgx_device.hpp"

namespace hgx {

struct Error {
    int code;
    std::string msg;
};

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            throw ::hgx::Error{HGX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
    } while (0)

struct Chunk {
    void* p;
    size_t size;
};

// Device memory chunks, allocated on first use and then recycled (best fit), bounded by the
// context's HBM budget. A serving loop in steady state allocates nothing from the runtime.
class ChunkPool {
public:
    ChunkPool(uint64_t budget, size_t chunk_bytes) : budget_(budget), chunk_bytes_(chunk_bytes) {}
    ~ChunkPool();
    Chunk get(size_t min_bytes);
    void put(const Chunk& c);
    uint64_t budget() const { return budget_; }
    uint64_t allocated() const { return total_; }
    uint64_t in_use() const { return in_use_; }
    // While holding, chunks given back are kept aside instead of reused (a pass being recorded into
    // a graph keeps every buffer it touched); take_held() hands them to their new owner, who gives
    // them back with put() when the graph goes.
    void hold(bool on);
    std::vector<Chunk> take_held();

private:
    std::mutex mu_;
    bool holding_ = false;
    std::vector<Chunk> held_;
    std::vector<Chunk> free_;
    uint64_t budget_;
    size_t chunk_bytes_;
    uint64_t total_ = 0, in_use_ = 0;
};

// Bump allocation over pool chunks; everything returns to the pool at once.
class Arena {
public:
    explicit Arena(ChunkPool* pool, size_t min_chunk = 0) : pool_(pool), min_chunk_(min_chunk) {}
    ~Arena();
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    void* alloc(size_t bytes);
    void release();
    uint64_t bytes() const { return bytes_; }

private:
    ChunkPool* pool_;
    size_t min_chunk_;
    std::vector<Chunk> chunks_;
    size_t used_ = 0;
    uint64_t bytes_ = 0;
};

// Power-of-two size classes carved from 64 MiB pool chunks: per-request rules and state maps,
// which come and go in any order on the served path.
class SlabAlloc {
public:
    explicit SlabAlloc(ChunkPool* pool) : pool_(pool) {}
    ~SlabAlloc();
    void* alloc(size_t bytes);
    void free(void* p, size_t bytes);                 // bytes: as passed to alloc

private:
    static constexpr uint32_t kMinClass = 12, kMaxClass = 24;   // 4 KiB .. 16 MiB blocks
    ChunkPool* pool_;
    std::vector<Chunk> slabs_;
    std::vector<void*> free_[kMaxClass + 1];          // class -> free blocks
    std::unordered_map<void*, Chunk> big_;            // blocks above the largest class
    size_t slab_used_ = 0;
};

template <class T>
T* dalloc(Arena& a, size_t n) { return (T*)a.alloc(sizeof(T) * (n ? n : 1)); }

template <class T>
void h2d(T* dst, const T* src, size_t n, hipStream_t st) {
    if (n) HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyHostToDevice, st));
}

template <class T>
void d2h(T* dst, const T* src, size_t n, hipStream_t st) {
    if (n) HIP_TRY(hipMemcpyAsync(dst, src, sizeof(T) * n, hipMemcpyDeviceToHost, st));
}

// Exclusive prefix sum of n uint32 values (in[n] must read 0) into n+1 uint64 offsets.
void scan_u32_to_u64(Arena& ar, const uint32_t* in, uint64_t* out, uint32_t n, hipStream_t st);
void sort_pairs_u64(Arena& ar, const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout,
                    uint64_t n, int end_bit, hipStream_t st);

// One parsed request.
// Device copies of a prepared request's rules and initial states (hgx_batch_prepare).
struct StagedSrc {
    const DRule* rules = nullptr;      // null: the request's inputs are on the host only
    const uint32_t* ne = nullptr;      // [roots] from this request's first root
    const uint32_t* nextv = nullptr;
    const uint64_t* eoff = nullptr;    // [roots + 1] offsets into the batch edge arrays
    const uint8_t* ar = nullptr;       // the batch edge arrays (arity, vertices at stride A)
    const uint32_t* ev = nullptr;
    uint32_t A = 0;
};

// A request's sampling and capping (hgx_sampling), resolved: the exploration coin's threshold
// as the reference encodes it (a float probability scaled to 2^32, gpu/src/evolve.cu:271-285)
// and its seed (a random one when 0 and a coin is flipped, :351-356).
struct SampleHost {
    bool on = false;
    double rate = 1.0;
    std::vector<double> weights;
    uint64_t seed = 0, coin_seed = 0;
    uint32_t thr = 0xFFFFFFFFu;
    uint64_t per_step = 0, per_parent = 0, per_state_rule = 0;
    bool draws() const { return rate < 1.0 || !weights.empty(); }      // rewrite.cu:270
    bool transitions() const { return draws() || per_state_rule != 0; }
    bool frontier() const { return thr != 0xFFFFFFFFu || per_step || per_parent; }
    double rate_for(uint32_t rule) const {                                // sampling_core.hpp:28-34
        const double r = rate * (rule < weights.size() ? weights[rule] : 1.0);
        return r < 0.0 ? 0.0 : (r > 1.0 ? 1.0 : r);
    }
};

// The transition draw: splitmix64 of the key against the seed (mixed by multiplication, so seed
// 0 still draws), compared in [0, 1) through its top 53 bits (sampling_core.hpp:42-57).
inline bool transition_survives(uint64_t key, uint64_t seed, double rate) {
    if (rate >= 1.0) return true;
    if (rate <= 0.0) return false;
    const uint64_t x = splitmix(key ^ (seed * 0x9E3779B97F4A7C15ULL));
    return (double)(x >> 11) * (1.0 / 9007199254740992.0) < rate;
}
// The order a per-(state, rule) cap keeps its k in: a second mix of the key, uncorrelated with
// the draw (sampling_core.hpp:66-71).
inline uint64_t transition_rank(uint64_t key, uint64_t seed) {
    uint64_t x = key ^ (seed * 0x9E3779B97F4A7C15ULL) ^ 0xA5A5A5A5A5A5A5A5ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
// The transition key: the event signature over the input state's exact hash, the rule and the
// consumed edges' canonical ranks in pattern order (EVENT_SIG_TRANSITION, event_core.hpp:55-84).
inline uint64_t transition_key(uint64_t input_hash, uint32_t rule, const uint32_t* ranks, uint32_t n) {
    uint64_t sig = fnv(fnv(FNV_OFFSET, input_hash), (uint64_t)(uint16_t)rule);
    for (uint32_t i = 0; i < n; ++i) sig = fnv(sig, ranks[i]);
    return sig == 0 || sig == FNV_OFFSET ? 1 : sig;
}

struct RunHost {
    StagedSrc staged;
    SampleHost smp;
    std::vector<DRule> rules;
    // Initial states, flat: root k holds edges [root_eb[k], root_eb[k+1]); edge e holds
    // verts[edge_vb[e] .. edge_vb[e+1]), vertices renumbered per root by first appearance.
    std::vector<uint32_t> root_eb{0};
    std::vector<uint32_t> edge_vb{0};
    std::vector<uint32_t> verts;
    std::vector<uint32_t> root_nv;
    uint32_t max_root_ne = 0, max_root_nv = 0;
    uint32_t n_roots() const { return (uint32_t)root_nv.size(); }
    uint32_t root_ne(size_t k) const { return root_eb[k + 1] - root_eb[k]; }
    uint32_t arity(size_t e) const { return edge_vb[e + 1] - edge_vb[e]; }
    uint32_t steps = 0, cap = 0, flags = 0;
    uint32_t max_arity = 1, max_lhs = 1, max_gain = 0, max_new = 0;
    uint32_t uni_arity = 0;      // k when every root edge and every RHS edge has arity k (2..4),
                                 // so every state of the run does; else 0
    bool root_uni = false;       // every root edge has one arity 2..4 (the roots alone)
    uint64_t root_edges = 0;
    // quotient exploration: only the first state of each exact class is expanded
    bool quotient() const { return (flags & HGX_OPT_QUOTIENT) != 0; }
    // the host decides each generation which new states are expanded
    bool host_frontier() const { return smp.frontier() || quotient(); }
};
// evolve = false: the request is one state for the large-state join, not an evolution (no
// limit on its size)
RunHost parse_run(const hgx_run_desc* d, bool evolve = true);
// Looks up, once per device (current), the address of the allocation-free entry point's scratch.
void noalloc_setup(int device);
// The rule program (include/hgx.h) into device rules with the reference's join order.
void parse_rules(const uint8_t* p, uint32_t len, uint32_t n, std::vector<DRule>& out);
uint64_t footprint_bytes(const RunHost& h, uint64_t held = 0);

// Per-state workspace (bytes) up to which one wave handles a state, and up to which four waves
// share one; above, eight (a CU holds 160 KiB of LDS).
constexpr uint32_t kTeamBytes[2] = {24u << 10, 48u << 10};
inline uint32_t pick_team(uint32_t bytes, uint32_t forced) {
    return forced ? forced : bytes <= kTeamBytes[0] ? 1u : bytes <= kTeamBytes[1] ? 4u : 8u;
}
struct WLStatesArgs;
// Workspace of a k_wl_states launch (lean when `lean`): sets lds_per_wave and scratch; returns
// the team size, and the waves per workgroup in *wpb.
uint32_t place_wl_states(Arena& ar, WLStatesArgs& w, bool lean, uint32_t max_lds, uint32_t grid_cap,
                         uint32_t forced_team, uint32_t* wpb);

uint32_t waves_per_block(uint32_t lds_per_wave, uint32_t max_lds);
// Workspace placement for a wave-per-state kernel: LDS when it fits, else a global-memory slice
// per wave from `ar` (4 waves per block). Returns waves per block; *scratch set in the second case.
uint32_t place_workspace(Arena& ar, uint32_t per_wave, uint32_t items, uint32_t max_lds, uint32_t grid_cap,
                         uint8_t** scratch);

struct ProfPending {
    std::string name;
    hipEvent_t a, b;
    uint64_t bytes, items;
    int cls;                  // size class of the launch, or -1
};
struct Profiler {
    bool on = false;
    bool by_class = false;    // also keep "<kernel>#<class>" totals (the autotuner's view)
    std::vector<ProfPending> pend;
    std::map<std::string, hgx_kernel_stat> acc;
};

// Launch shapes per size class, chosen by the autotuner (hgx_autotune) or loaded from its cache;
// the defaults are the hand-fitted rule. None of them changes a result.
constexpr int kTuneExpandClasses = 6;   // children of up to 32, 64, 128, 512, 2048 edges, more
constexpr int kTuneJoinClasses = 6;     // parents of up to 32, 64, 128, 512, 2048 edges, more
struct Tune {
    uint32_t expand_lanes[kTuneExpandClasses] = {16, 32, 64, 64, 64, 64};   // lanes per child
    uint32_t expand_grid[kTuneExpandClasses] = {8, 8, 8, 8, 8, 8};          // workgroups per CU
    uint32_t join_team[kTuneJoinClasses] = {0, 0, 0, 0, 0, 0};              // waves per state, 0 = by size
    uint32_t join_lanes[kTuneJoinClasses] = {32, 32, 64, 64, 64, 64};       // lanes per state (one-wave teams)
    uint32_t join_grid[kTuneJoinClasses] = {16, 16, 16, 16, 16, 16};        // workgroups per CU
    std::string source = "built-in defaults";
};

class Engine;

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
    std::recursive_mutex mu;
    std::unique_ptr<hgx::Engine> serve;          // the served path's engine
    uint64_t next_request_id = 1;
    int forward = -1;                            // HGX_OPTION_FORWARD
    uint32_t delta_cap = 256;                    // HGX_OPTION_DELTA_CAP
    uint32_t expand_team = 0;                    // HGX_OPTION_EXPAND_TEAM
    uint32_t join_team = 0;                      // HGX_OPTION_JOIN_TEAM
    uint32_t expand_lanes = 0;                   // HGX_OPTION_EXPAND_LANES
    uint32_t join_lanes = 0;                     // HGX_OPTION_JOIN_LANES
    uint64_t flush_bytes = 0;                    // HGX_OPTION_COUNTER_FLUSH
    void* flush_buf = nullptr;                   // its buffer (hipMalloc'd, freed with the context)
    hgx::Tune tune;                              // launch shapes per size class
};

namespace hgx {

// Launch with optional event timing; `bytes` are the launch's algorithmic HBM bytes.
// HGX_OPTION_COUNTER_FLUSH: streams the flush buffer through the L2 (counter runs only).
hipError_t launch_flush(const void* buf, uint64_t bytes, hipStream_t st);

template <class F>
void timed(hgx_ctx* ctx, hipStream_t st, const char* name, uint64_t bytes, uint64_t items, F&& launch,
           int cls = -1) {
    if (ctx->flush_bytes) HIP_TRY(launch_flush(ctx->flush_buf, ctx->flush_bytes, st));
    if (!ctx->prof.on) {
        HIP_TRY(launch());
    } else {
        ProfPending p{name, nullptr, nullptr, bytes, items, cls};
        HIP_TRY(hipEventCreate(&p.a));
        HIP_TRY(hipEventCreate(&p.b));
        HIP_TRY(hipEventRecord(p.a, st));
        HIP_TRY(launch());
        HIP_TRY(hipEventRecord(p.b, st));
        ctx->prof.pend.push_back(p);
    }
    if (ctx->flush_bytes) HIP_TRY(launch_flush(ctx->flush_buf, ctx->flush_bytes, st));
}
void prof_resolve(hgx_ctx* ctx);
// Loads the tuned launch shapes of this device (key "<arch>:<CUs>") from a cache written by
// hgx_autotune; false when the file or the key is absent.
bool tune_load(hgx_ctx* ctx, const std::string& path);
// HGX_TUNE_CACHE, else tune_cache.json beside the library.
std::string default_tune_cache();

template <class F>
int guarded(hgx_ctx* ctx, F&& f) {
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

}  // namespace hgx
