// hgx_bigjoin.hip -- the join over one large state (tens of thousands of edges and up), where
// the state does not fit LDS: the match-join stress workload, one rewriting step of one big
// state (hgx_match_count).
//
// Index build (the role of SignatureIndex / InvertedVertexIndex, hypergraph/src/index.cpp:45-80):
// a CSR keyed by (vertex, position) -> ascending edge ids, by counting sort: histogram with
// atomics, exclusive scan, scatter, then each bucket sorted so the CSR is deterministic.
// Keying by position as well as vertex makes a candidate list exactly the edges that carry the
// bound vertex where the pattern needs it, so the join never re-tests the position.
//
// Join (join_core.hpp:143-208): one lane per seed edge of the first join position (SCAN: the
// arity and repetition test), depth-first completion over the CSR lists (EXPAND), per-seed
// counts prefix-summed so matches are written in (seed, depth-first) order (SINK) -- the order
// of a serial DFS over ascending candidates.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "hgx_device.hpp"
#include "hgx_kernels.hpp"
#include "hgx_runtime.hpp"

namespace hgx {

__global__ void k_index_count(const uint32_t* ev, const uint8_t* ar, uint32_t m, uint32_t A, uint32_t* cnt) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    for (uint32_t p = 0; p < ar[e]; ++p) atomicAdd(&cnt[ev[(size_t)e * A + p] * A + p], 1u);
}

__global__ void k_off_to_u32(const uint64_t* off, uint32_t* cur, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cur[i] = (uint32_t)off[i];
}

__global__ void k_index_scatter(const uint32_t* ev, const uint8_t* ar, uint32_t m, uint32_t A,
                                uint32_t* cursor, uint32_t* idx) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    for (uint32_t p = 0; p < ar[e]; ++p) idx[atomicAdd(&cursor[ev[(size_t)e * A + p] * A + p], 1u)] = e;
}

__global__ void k_index_sort(const uint64_t* off, uint32_t nkeys, uint32_t* idx) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const uint64_t b = off[k], n = off[k + 1] - b;
    for (uint64_t i = 1; i < n; ++i) {
        const uint32_t key = idx[b + i];
        uint64_t j = i;
        while (j > 0 && idx[b + j - 1] > key) { idx[b + j] = idx[b + j - 1]; --j; }
        idx[b + j] = key;
    }
}

struct BigJoinArgs {
    DRule R;
    const uint32_t* ev;
    const uint8_t* ar;
    uint32_t m, A;
    const uint64_t* off;    // [nkeys + 1]
    const uint32_t* idx;
    uint32_t* seed_cnt;     // count pass: [m + 1]
    const uint64_t* seed_off;   // emit pass
    uint32_t* out;          // emit pass: n_lhs edge ids per match, pattern order
    uint64_t cap;
};

__device__ inline bool bj_bind(const uint32_t* ev, uint32_t ea, const uint8_t* pv, uint32_t pa,
                               uint32_t* bind, uint32_t& mask) {   // match_core.hpp:32-49
    if (ea != pa) return false;
    for (uint32_t i = 0; i < ea; ++i) {
        const uint32_t var = pv[i], bit = 1u << var;
        if (mask & bit) { if (bind[var] != ev[i]) return false; }
        else { bind[var] = ev[i]; mask |= bit; }
    }
    return true;
}

template <bool EMIT>
__global__ __launch_bounds__(256) void k_bigjoin(BigJoinArgs a) {
    const DRule& R = a.R;
    const uint32_t n = R.n_lhs;
    uint32_t matched[16], smask[17], cidx[16], chi[16], bind[32];
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t seed = blockIdx.x * blockDim.x + threadIdx.x; seed < a.m; seed += stride) {
        uint32_t found = 0;
        uint64_t o = EMIT ? a.seed_off[seed] : 0;
        uint32_t mask = 0;
        const uint32_t p0 = R.order[0];
        if (bj_bind(a.ev + (size_t)seed * a.A, a.ar[seed], R.lv[p0], R.la[p0], bind, mask)) {
            matched[0] = seed;
            if (n == 1) {
                if (EMIT && o < a.cap) a.out[o * n + p0] = seed;
                found = 1;
            } else {
                auto range = [&](uint32_t d, uint32_t& lo, uint32_t& hi) {
                    const uint32_t pv = R.piv[d];
                    if (pv == 0xFFu) { lo = 0; hi = a.m; return; }
                    const uint64_t key = (uint64_t)bind[R.lv[R.order[d]][pv]] * a.A + pv;
                    lo = (uint32_t)a.off[key];
                    hi = (uint32_t)a.off[key + 1];
                };
                smask[1] = mask;
                range(1, cidx[1], chi[1]);
                uint32_t d = 1;
                while (d >= 1) {
                    const uint32_t p = R.order[d];
                    const bool indexed = R.piv[d] != 0xFFu;
                    bool adv = false;
                    while (cidx[d] < chi[d]) {
                        const uint32_t j = cidx[d]++;
                        const uint32_t e = indexed ? a.idx[j] : j;
                        bool taken = false;
                        for (uint32_t t = 0; t < d; ++t) taken |= (matched[t] == e);
                        if (taken) continue;
                        uint32_t mm = smask[d];
                        if (!bj_bind(a.ev + (size_t)e * a.A, a.ar[e], R.lv[p], R.la[p], bind, mm)) continue;
                        matched[d] = e;
                        if (d + 1 == n) {
                            if (EMIT && o + found < a.cap)
                                for (uint32_t t = 0; t < n; ++t) a.out[(o + found) * n + R.order[t]] = matched[t];
                            ++found;
                            continue;
                        }
                        smask[d + 1] = mm;
                        range(d + 1, cidx[d + 1], chi[d + 1]);
                        ++d;
                        adv = true;
                        break;
                    }
                    if (!adv) --d;
                }
            }
        }
        if (!EMIT) a.seed_cnt[seed] = found;
    }
}

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_match_count(hgx_ctx* ctx, const uint8_t* rules, uint32_t rules_len, uint32_t n_rules,
                               uint32_t rule, const uint32_t* edge_offsets, const uint32_t* vertices,
                               uint32_t n_edges, uint64_t* count, uint32_t* match_edges, uint64_t cap,
                               float* device_ms) {
    if (!ctx || !rules || !count || (n_edges && (!edge_offsets || !vertices)) || rule >= n_rules)
        return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        hgx_run_desc d{};
        d.rules = rules; d.rules_len = rules_len; d.n_rules = n_rules; d.steps = 0;
        d.edge_offsets = edge_offsets; d.vertices = vertices; d.n_edges = n_edges;
        RunHost h = parse_run(&d);                     // validates rules and arities, renumbers
        const DRule R = h.rules[rule];
        const uint32_t m = h.root_ne(0);
        uint32_t A = 1;
        for (uint32_t e = 0; e < m; ++e) A = std::max(A, h.arity(e));
        for (int i = 0; i < R.n_lhs; ++i) A = std::max<uint32_t>(A, R.la[i]);
        const uint32_t nv = std::max(1u, h.root_nv[0]);
        std::vector<uint32_t> ev((size_t)m * A, 0);
        std::vector<uint8_t> ar(m);
        for (uint32_t e = 0; e < m; ++e) {
            ar[e] = (uint8_t)h.arity(e);
            for (uint32_t k = 0; k < h.arity(e); ++k) ev[(size_t)e * A + k] = h.verts[h.edge_vb[e] + k];
        }
        hipStream_t s = ctx->stream;
        Arena arena(ctx->pool.get());
        uint32_t* d_ev = dalloc<uint32_t>(arena, (size_t)m * A);
        uint8_t* d_ar = dalloc<uint8_t>(arena, m);
        const uint64_t nkeys = (uint64_t)nv * A;
        if (nkeys > 0xFFFFFFF0ull) throw Error{HGX_ERR_UNSUPPORTED, "index above 2^32 keys"};
        uint32_t* d_cnt = dalloc<uint32_t>(arena, nkeys + 1);
        uint64_t* d_off = dalloc<uint64_t>(arena, nkeys + 1);
        uint32_t* d_cur = dalloc<uint32_t>(arena, nkeys + 1);
        uint64_t occ = 0;
        for (auto x : ar) occ += x;
        uint32_t* d_idx = dalloc<uint32_t>(arena, occ);
        uint32_t* d_seed = dalloc<uint32_t>(arena, (size_t)m + 1);
        uint64_t* d_soff = dalloc<uint64_t>(arena, (size_t)m + 1);
        h2d(d_ev, ev.data(), ev.size(), s);
        h2d(d_ar, ar.data(), ar.size(), s);
        hipEvent_t t0, t1, t2;
        HIP_TRY(hipEventCreate(&t0)); HIP_TRY(hipEventCreate(&t1)); HIP_TRY(hipEventCreate(&t2));
        HIP_TRY(hipEventRecord(t0, s));
        // ---- index build: histogram, scan, scatter, bucket order ----
        HIP_TRY(hipMemsetAsync(d_cnt, 0, sizeof(uint32_t) * (nkeys + 1), s));
        const uint32_t gb = (m + 255) / 256;
        if (m) hipLaunchKernelGGL(k_index_count, dim3(gb), dim3(256), 0, s, d_ev, d_ar, m, A, d_cnt);
        scan_u32_to_u64(arena, d_cnt, d_off, (uint32_t)nkeys, s);
        // scatter cursors start at the bucket offsets (occurrences < 2^32)
        hipLaunchKernelGGL(k_off_to_u32, dim3((uint32_t)((nkeys + 255) / 256)), dim3(256), 0, s, d_off, d_cur,
                           (uint32_t)nkeys);
        if (m) hipLaunchKernelGGL(k_index_scatter, dim3(gb), dim3(256), 0, s, d_ev, d_ar, m, A, d_cur, d_idx);
        hipLaunchKernelGGL(k_index_sort, dim3((uint32_t)((nkeys + 255) / 256)), dim3(256), 0, s, d_off,
                           (uint32_t)nkeys, d_idx);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(t1, s));
        // ---- join: count, scan, emit ----
        BigJoinArgs ja{};
        ja.R = R; ja.ev = d_ev; ja.ar = d_ar; ja.m = m; ja.A = A; ja.off = d_off; ja.idx = d_idx;
        ja.seed_cnt = d_seed;
        HIP_TRY(hipMemsetAsync(d_seed + m, 0, sizeof(uint32_t), s));
        const uint32_t grid = std::max(1u, std::min(gb, ctx->num_cu * 16));
        if (m) hipLaunchKernelGGL(k_bigjoin<false>, dim3(grid), dim3(256), 0, s, ja);
        scan_u32_to_u64(arena, d_seed, d_soff, m, s);
        uint64_t total = 0;
        d2h(&total, d_soff + m, 1, s);
        uint32_t* d_out = nullptr;
        if (match_edges && cap && total) {
            const uint64_t keep = std::min(cap, total);
            d_out = dalloc<uint32_t>(arena, keep * R.n_lhs);
            ja.seed_off = d_soff; ja.out = d_out; ja.cap = keep;
            hipLaunchKernelGGL(k_bigjoin<true>, dim3(grid), dim3(256), 0, s, ja);
            HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipEventRecord(t2, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (d_out) d2h(match_edges, d_out, std::min(cap, total) * R.n_lhs, s);
        HIP_TRY(hipStreamSynchronize(s));
        float ms_index = 0, ms_join = 0;
        HIP_TRY(hipEventElapsedTime(&ms_index, t0, t1));
        HIP_TRY(hipEventElapsedTime(&ms_join, t1, t2));
        (void)hipEventDestroy(t0); (void)hipEventDestroy(t1); (void)hipEventDestroy(t2);
        if (device_ms) { device_ms[0] = ms_index + ms_join; }
        if (ctx->prof.on) {
            auto& si = ctx->prof.acc["bigjoin_index"];
            std::snprintf(si.name, sizeof(si.name), "bigjoin_index");
            si.launches += 1; si.total_ms += ms_index; si.items += m;
            si.alg_bytes += (uint64_t)m * (4ull * A + 1) + 2ull * 4 * (occ + nkeys + 1);
            auto& sj = ctx->prof.acc["bigjoin_join"];
            std::snprintf(sj.name, sizeof(sj.name), "bigjoin_join");
            sj.launches += 1; sj.total_ms += ms_join; sj.items += m;
            sj.alg_bytes += (uint64_t)m * (4ull * A + 1) + total * 4ull * R.n_lhs;
        }
        *count = total;
        return HGX_OK;
    });
}
