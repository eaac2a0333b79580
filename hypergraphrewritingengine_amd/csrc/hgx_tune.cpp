// hgx_tune.cpp -- the autotuner: launch shapes per size class, measured on the device.
//
// For every size class of the expansion (children of up to 32, 64, 128, 512, 2048 edges, more)
// it chooses the lanes per child (16 / 32 / 64: sub-wave teams or one wave) and the workgroups per
// CU; for every size class of the join (parents of up to 32, 64, 128, 512, 2048 edges, more) the
// waves per state, the lanes per state when one wave is the team (16 / 32 / 64), and the
// workgroups per CU. Each candidate is run over a caller's prepared batch with the
// launches event-timed per class, and each class keeps the fastest candidate (classes are separate
// launches, so one sweep serves them all). The choice is kept in the context and written to a JSON
// cache keyed by "<arch>:<CUs>", which hgx_ctx_create loads for its device (from HGX_TUNE_CACHE,
// else tune_cache.json beside the library). None of the shapes changes a result -- the tests run
// every lane count and team size against the oracle. The join order is not tuned: event numbering
// follows the reference's join order (hypergraph/src/pattern.cpp:43-75), which fixes it.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "hgx_runtime.hpp"

namespace hgx {

static std::string device_key(hgx_ctx* ctx) {
    hipDeviceProp_t prop;
    std::string arch = "unknown";
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) {
        arch = prop.gcnArchName;
        const size_t c = arch.find(':');
        if (c != std::string::npos) arch = arch.substr(0, c);
    }
    return arch + ":" + std::to_string(ctx->num_cu);
}

template <size_t N>
static std::string json_array(const uint32_t (&a)[N]) {
    std::string s = "[";
    for (size_t i = 0; i < N; ++i) s += (i ? ", " : "") + std::to_string(a[i]);
    return s + "]";
}

static std::string tune_json(const Tune& t) {
    return "{\"expand_lanes\": " + json_array(t.expand_lanes) + ", \"expand_grid\": " + json_array(t.expand_grid) +
           ", \"join_team\": " + json_array(t.join_team) + ", \"join_lanes\": " + json_array(t.join_lanes) +
           ", \"join_grid\": " + json_array(t.join_grid) +
           ", \"source\": \"" + t.source + "\"}";
}

// The integers of `"key": [a, b, ...]` inside text, or false.
template <size_t N>
static bool read_array(const std::string& text, const char* key, uint32_t (&out)[N]) {
    const std::string k = std::string("\"") + key + "\"";
    size_t p = text.find(k);
    if (p == std::string::npos) return false;
    p = text.find('[', p);
    const size_t e = text.find(']', p);
    if (p == std::string::npos || e == std::string::npos) return false;
    std::stringstream ss(text.substr(p + 1, e - p - 1));
    uint32_t v[N];
    for (size_t i = 0; i < N; ++i) {
        long x = -1;
        char comma;
        if (!(ss >> x) || x < 0) return false;
        v[i] = (uint32_t)x;
        if (i + 1 < N && !(ss >> comma)) return false;
    }
    std::copy(v, v + N, out);
    return true;
}

// The entry of `key` in a cache file: the text between its braces.
static bool cache_entry(const std::string& text, const std::string& key, std::string& entry) {
    const size_t p = text.find("\"" + key + "\"");
    if (p == std::string::npos) return false;
    const size_t b = text.find('{', p), e = text.find('}', b);
    if (b == std::string::npos || e == std::string::npos) return false;
    entry = text.substr(b, e - b + 1);
    return true;
}

static bool valid(const Tune& t) {
    for (uint32_t x : t.expand_lanes) if (x != 16 && x != 32 && x != 64) return false;
    for (uint32_t x : t.expand_grid) if (x < 1 || x > 64) return false;
    for (uint32_t x : t.join_team) if (x != 0 && x != 1 && x != 4 && x != 8) return false;
    for (uint32_t x : t.join_lanes) if (x != 16 && x != 32 && x != 64) return false;
    for (uint32_t x : t.join_grid) if (x < 1 || x > 64) return false;
    return true;
}

bool tune_load(hgx_ctx* ctx, const std::string& path) {
    std::ifstream f(path);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    std::string entry;
    const std::string key = device_key(ctx);
    if (!cache_entry(ss.str(), key, entry)) return false;
    Tune t;
    if (!read_array(entry, "expand_lanes", t.expand_lanes) || !read_array(entry, "expand_grid", t.expand_grid) ||
        !read_array(entry, "join_team", t.join_team) || !read_array(entry, "join_lanes", t.join_lanes) ||
        !read_array(entry, "join_grid", t.join_grid) || !valid(t))
        return false;
    t.source = "tuned (" + key + ", " + path + ")";
    ctx->tune = t;
    return true;
}

// The cache hgx_ctx_create looks in: HGX_TUNE_CACHE, else tune_cache.json beside this library.
std::string default_tune_cache() {
    if (const char* e = std::getenv("HGX_TUNE_CACHE")) return e;
    Dl_info info;
    if (dladdr((void*)&default_tune_cache, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t s = p.rfind('/');
        return (s == std::string::npos ? std::string(".") : p.substr(0, s)) + "/tune_cache.json";
    }
    return "tune_cache.json";
}

}  // namespace hgx

using namespace hgx;

uint32_t hgx_batch_size(const hgx_batch* b);        // requests in a prepared batch (hgx_engine.cpp)

extern "C" int hgx_autotune(hgx_ctx* ctx, const hgx_batch* batch, uint32_t reps, const char* cache_path,
                            char* report, size_t report_len) {
    if (!ctx || !batch) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        const uint32_t n = hgx_batch_size(batch);
        if (!n) throw Error{HGX_ERR_INVALID, "the tuning batch is empty"};
        if (reps < 1) reps = 1;
        // forced options would hide what the candidates change
        const uint32_t saved_lanes = ctx->expand_lanes, saved_jt = ctx->join_team, saved_et = ctx->expand_team,
                       saved_jl = ctx->join_lanes;
        const bool saved_on = ctx->prof.on, saved_cls = ctx->prof.by_class;
        const Tune saved = ctx->tune;
        ctx->expand_lanes = 0; ctx->join_team = 0; ctx->expand_team = 0; ctx->join_lanes = 0;
        std::vector<hgx_result*> res(n, nullptr);
        // one candidate: every class set alike, the best of `reps` passes per class
        auto measure = [&](const Tune& t) {
            ctx->tune = t;
            std::map<std::string, double> best;
            for (uint32_t r = 0; r < reps + 1; ++r) {          // the first pass warms up
                ctx->prof.on = true;
                ctx->prof.by_class = true;
                ctx->prof.acc.clear();
                const int rc = hgx_batch_evolve(ctx, batch, res.data());
                hgx_result_release_many(res.data(), n);
                if (rc != HGX_OK) throw Error{rc, "a tuning pass failed: " + ctx->err};
                if (r == 0) continue;
                for (auto& kv : ctx->prof.acc) {
                    auto it = best.find(kv.first);
                    if (it == best.end() || kv.second.total_ms < it->second) best[kv.first] = kv.second.total_ms;
                }
            }
            return best;
        };
        auto ms_of = [](const std::map<std::string, double>& m, const std::string& k) {
            auto it = m.find(k);
            return it == m.end() ? -1.0 : it->second;
        };
        Tune out = saved;
        std::string log = "{\"expand\": [";
        // expansion: lanes x workgroups per CU
        double eb[kTuneExpandClasses];
        std::fill(eb, eb + kTuneExpandClasses, -1.0);
        bool first = true;
        for (uint32_t lanes : {16u, 32u, 64u})
            for (uint32_t grid : {4u, 8u, 16u}) {
                Tune t = saved;
                std::fill(t.expand_lanes, t.expand_lanes + kTuneExpandClasses, lanes);
                std::fill(t.expand_grid, t.expand_grid + kTuneExpandClasses, grid);
                const auto m = measure(t);
                for (int k = 0; k < kTuneExpandClasses; ++k) {
                    const double ms = ms_of(m, "k_expand#" + std::to_string(k));
                    if (ms < 0) continue;
                    log += std::string(first ? "" : ", ") + "{\"class\": " + std::to_string(k) + ", \"lanes\": " +
                           std::to_string(lanes) + ", \"grid\": " + std::to_string(grid) + ", \"ms\": " +
                           std::to_string(ms) + "}";
                    first = false;
                    if (eb[k] < 0 || ms < eb[k]) { eb[k] = ms; out.expand_lanes[k] = lanes; out.expand_grid[k] = grid; }
                }
            }
        log += "], \"join\": [";
        // join: (waves per state, lanes per state) x workgroups per CU (count and emit passes
        // together); lanes below 64 only with one-wave teams
        double jb[kTuneJoinClasses];
        std::fill(jb, jb + kTuneJoinClasses, -1.0);
        first = true;
        const uint32_t shapes[5][2] = {{1, 16}, {1, 32}, {1, 64}, {0, 64}, {4, 64}};
        for (const auto& sh : shapes)
            for (uint32_t grid : {8u, 16u, 32u}) {
                Tune t = out;
                std::fill(t.join_team, t.join_team + kTuneJoinClasses, sh[0]);
                std::fill(t.join_lanes, t.join_lanes + kTuneJoinClasses, sh[1]);
                std::fill(t.join_grid, t.join_grid + kTuneJoinClasses, grid);
                const auto m = measure(t);
                for (int k = 0; k < kTuneJoinClasses; ++k) {
                    const double c = ms_of(m, "k_match_count#" + std::to_string(k));
                    const double e = ms_of(m, "k_match_emit#" + std::to_string(k));
                    if (c < 0) continue;
                    const double ms = c + (e > 0 ? e : 0.0);
                    log += std::string(first ? "" : ", ") + "{\"class\": " + std::to_string(k) + ", \"team\": " +
                           std::to_string(sh[0]) + ", \"lanes\": " + std::to_string(sh[1]) + ", \"grid\": " +
                           std::to_string(grid) + ", \"ms\": " + std::to_string(ms) + "}";
                    first = false;
                    if (jb[k] < 0 || ms < jb[k]) {
                        jb[k] = ms;
                        out.join_team[k] = sh[0]; out.join_lanes[k] = sh[1]; out.join_grid[k] = grid;
                    }
                }
            }
        log += "]}";
        const std::string key = device_key(ctx);
        out.source = "autotuned (" + key + ", " + std::to_string(n) + "-request batch)";
        ctx->tune = out;
        ctx->expand_lanes = saved_lanes; ctx->join_team = saved_jt; ctx->expand_team = saved_et;
        ctx->join_lanes = saved_jl;
        ctx->prof.on = saved_on; ctx->prof.by_class = saved_cls;
        ctx->prof.acc.clear();
        if (cache_path && *cache_path) {
            // one entry per device key; other keys in an existing cache are kept
            std::map<std::string, std::string> entries;
            {
                std::ifstream f(cache_path);
                if (f) {
                    std::stringstream ss;
                    ss << f.rdbuf();
                    const std::string text = ss.str();
                    size_t p = 0;
                    while ((p = text.find("\"", p)) != std::string::npos) {
                        const size_t q = text.find("\"", p + 1);
                        if (q == std::string::npos) break;
                        const std::string k = text.substr(p + 1, q - p - 1);
                        std::string entry;
                        if (k.find(':') != std::string::npos && text.compare(q + 1, 2, ": ") == 0 &&
                            text[q + 3] == '{' && cache_entry(text.substr(p), k, entry))
                            entries[k] = entry;
                        p = q + 1;
                    }
                }
            }
            entries[key] = tune_json(out);
            std::ofstream f(cache_path);
            if (!f) throw Error{HGX_ERR_INVALID, std::string("cannot write ") + cache_path};
            f << "{\n";
            size_t i = 0;
            for (auto& kv : entries) f << "  \"" << kv.first << "\": " << kv.second << (++i < entries.size() ? ",\n" : "\n");
            f << "}\n";
        }
        if (report && report_len) {
            const std::string r = "{\"chosen\": " + tune_json(out) + ", \"measured\": " + log + "}";
            std::snprintf(report, report_len, "%s", r.c_str());
        }
        return HGX_OK;
    });
}

extern "C" int hgx_tune_load(hgx_ctx* ctx, const char* path) {
    if (!ctx || !path) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    if (!tune_load(ctx, path)) {
        ctx->err = std::string("no tuned shapes for this device in ") + path;
        return HGX_ERR_INVALID;
    }
    return HGX_OK;
}

extern "C" size_t hgx_tune_get(const hgx_ctx* ctx, char* buf, size_t len) {
    if (!ctx) return 0;
    const std::string s = tune_json(ctx->tune);
    if (buf && len) std::snprintf(buf, len, "%s", s.c_str());
    return s.size();
}
