// hgx_unfold.hip -- the raw unfolding beneath an evolution, counted on the device without
// materialising it (hgx_result_unfold).
//
// Quotient exploration expands one state per exact isomorphism class; the multiway system a
// caller is told about is still the raw one, in which every raw state is expanded. The reference
// recovers it by replay (common/include/hgcommon/quotient_replay_core.hpp:112-226): a class's
// expansion is recorded once, in slots of the class, and every raw instance of the class replays
// it. This file does the same, level-synchronously on the device:
//
//   * slots are the edges' canonical ranks (the positions of the exact canonicalizer's canonical
//     form). Any two raw states of a class map onto the canonical form by an isomorphism, so one
//     recorded match replays on every instance; the reference numbers slots by automorphism orbit
//     and edge id instead (slot_core.hpp), an equally valid frame -- the relations and counts
//     are the same under any frame, since a frame change is an isomorphism of the whole unfolding.
//   * an instance is its class, the producing generation of every slot (0xFF: an initial edge),
//     and, for each generation t on its branch, the causal ancestry of the event applied at t
//     (a mask over generations). A raw branch holds one event per generation, so a generation
//     names an event, as in the evolution kernels' lineage masks.
//   * replaying match m on instance i (one work item): the event's producers are the
//     generations P of its consumed slots (distinct pairs = |P|, triples = consumed slots with a
//     producer); a pair (p, event) survives transitive reduction when p is no ancestor of another
//     producer; the child instance takes the survivors' producers, the produced slots this
//     generation, and the ancestry P | ancestry(P) for this generation (qr_apply's causal and
//     descend steps). Sibling applications overlapping in consumed slots (the branchial pairs)
//     depend only on the class, so each class's count is computed once on the host and added
//     per instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/hgx.h"
#include "hgx_engine.hpp"
#include "hgx_runtime.hpp"

namespace hgx {

constexpr uint16_t kUnfoldProduced = 0xFFFFu;     // child slot made by the event itself
constexpr uint8_t kUnfoldInitial = 0xFFu;         // slot holding an initial edge
constexpr uint32_t kUnfoldMaxSlots = 4096;

struct UnfoldArgs {
    uint64_t W;                  // work items: (instance, match) pairs of this level
    uint32_t N;                  // instances of this level
    uint32_t d;                  // this level's generation
    uint32_t S;                  // slot stride
    uint32_t write;              // children are kept (a later generation expands them)
    const uint64_t* wofs;        // [N+1] first work item of each instance
    const uint32_t* cls;         // [N] class of each instance
    const uint8_t* prod;         // [N*S] producing generation of each slot
    const uint64_t* anc;         // [N*d] ancestry of the event of each generation on the branch
    const uint32_t* cls_moff;    // [C+1] recorded matches of each class
    const uint64_t* cls_br;      // [C] overlapping sibling pairs of one instance of the class
    const uint32_t* m_to;        // [M] class of the child
    const uint32_t* m_coff;      // [M+1] consumed slots, pattern order
    const uint16_t* m_cons;
    const uint32_t* m_soff;      // [M+1] one source per child slot
    const uint16_t* m_src;       // parent slot, or kUnfoldProduced
    uint32_t* ccls;              // [W] children (write)
    uint8_t* cprod;              // [W*S]
    uint64_t* canc;              // [W*(d+1)]
    unsigned long long* part;    // [gridDim.x * 5] this launch's per-workgroup sums
};

__global__ void k_unfold_count(uint32_t N, const uint32_t* __restrict__ cls, const uint32_t* __restrict__ cls_moff,
                               uint32_t* __restrict__ nm) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= N; i += gridDim.x * blockDim.x) {
        if (i == N) { nm[i] = 0; continue; }
        const uint32_t c = cls[i];
        nm[i] = cls_moff[c + 1] - cls_moff[c];
    }
}

__device__ inline unsigned long long wave_sum(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__global__ __launch_bounds__(256) void k_unfold_apply(UnfoldArgs a) {
    unsigned long long trip = 0, pairs = 0, ptr = 0, ttr = 0, br = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < a.W; w += stride) {
        // the instance: the last i with wofs[i] <= w (wofs[0] = 0 <= w < wofs[N])
        uint32_t lo = 0, hi = a.N;
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (a.wofs[mid] <= w) lo = mid; else hi = mid;
        }
        const uint32_t i = lo, c = a.cls[i];
        const uint32_t k = (uint32_t)(w - a.wofs[i]);
        const uint32_t m = a.cls_moff[c] + k;
        if (k == 0) br += a.cls_br[c];
        const uint8_t* pp = a.prod + (uint64_t)i * a.S;
        const uint64_t* pa = a.anc + (uint64_t)i * a.d;
        const uint32_t q0 = a.m_coff[m], q1 = a.m_coff[m + 1];
        uint64_t P = 0;
        for (uint32_t q = q0; q < q1; ++q) {
            const uint8_t p = pp[a.m_cons[q]];
            if (p != kUnfoldInitial) { P |= 1ull << p; ++trip; }
        }
        uint64_t up = 0;                                  // ancestors of the producers
        for (uint64_t x = P; x; x &= x - 1) up |= pa[__builtin_ctzll(x)];
        const uint64_t kept = P & ~up;
        for (uint32_t q = q0; q < q1; ++q) {
            const uint8_t p = pp[a.m_cons[q]];
            if (p != kUnfoldInitial && ((kept >> p) & 1ull)) ++ttr;
        }
        pairs += __popcll(P);
        ptr += __popcll(kept);
        if (a.write) {
            a.ccls[w] = a.m_to[m];
            uint8_t* cp = a.cprod + w * a.S;
            const uint32_t s0 = a.m_soff[m], ns = a.m_soff[m + 1] - s0;
            for (uint32_t s = 0; s < ns; ++s) {
                const uint16_t src = a.m_src[s0 + s];
                cp[s] = src == kUnfoldProduced ? (uint8_t)a.d : pp[src];
            }
            uint64_t* ca = a.canc + w * (uint64_t)(a.d + 1);
            for (uint32_t t = 0; t < a.d; ++t) ca[t] = pa[t];
            ca[a.d] = P | up;
        }
    }
    // per-workgroup sums written to the workgroup's own slots: thousands of workgroups adding into
    // the same five words would serialise on them
    __shared__ unsigned long long ws[4][5];
    trip = wave_sum(trip); pairs = wave_sum(pairs); ptr = wave_sum(ptr); ttr = wave_sum(ttr); br = wave_sum(br);
    const uint32_t wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ws[wv][0] = trip; ws[wv][1] = pairs; ws[wv][2] = ptr; ws[wv][3] = ttr; ws[wv][4] = br;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned long long t = 0;
        for (uint32_t k = 0; k < (blockDim.x >> 6); ++k) t += ws[k][threadIdx.x];
        a.part[(uint64_t)blockIdx.x * 5 + threadIdx.x] = t;
    }
}

// The launch's per-workgroup sums into the five totals (one workgroup of 256).
__global__ __launch_bounds__(256) void k_unfold_reduce(const unsigned long long* part, uint32_t nblocks,
                                                       unsigned long long* acc) {
    __shared__ unsigned long long ws[4][5];
    unsigned long long t[5] = {0, 0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < 5; ++k) t[k] += part[(uint64_t)b * 5 + k];
    for (int k = 0; k < 5; ++k) t[k] = wave_sum(t[k]);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 5; ++k) ws[threadIdx.x >> 6][k] = t[k];
    __syncthreads();
    if (threadIdx.x < 5) {
        unsigned long long v = 0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) v += ws[w][threadIdx.x];
        acc[threadIdx.x] += v;
    }
}

namespace {

// The class skeleton on the host: each class's recorded expansion in slots.
struct Skeleton {
    std::vector<uint32_t> cls_slots, cls_moff{0};
    std::vector<uint64_t> cls_br;
    std::vector<uint32_t> m_to, m_coff{0}, m_soff{0};
    std::vector<uint16_t> m_cons, m_src;
    std::vector<uint32_t> roots;     // class of each root instance
    uint32_t S = 1;
};

void check(int rc, const char* what) {
    if (rc != HGX_OK) throw Error{rc, std::string("unfold: reading the result's ") + what + " failed"};
}

// `steps`: the requested generations. The states of generation `steps` are never expanded, and
// a child of that generation is only ever written for an instance of a class first met there --
// which is never replayed below it -- so their classes are not needed: unless `all_states`
// (sampled requests, whose recorded state of a class need not be its first), they are left out
// of the exact canonicalization and point at an empty sentinel class.
Skeleton build_skeleton(const hgx_result* res, const hgx_counts& c, uint32_t steps, bool all_states) {
    const uint64_t ns = c.raw_states, ne = c.events;
    std::vector<uint32_t> creator(ns), step(ns);
    check(hgx_result_states(res, nullptr, step.data(), creator.data(), ns), "states");
    std::vector<hgx_event> ev(ne);
    if (ne) check(hgx_result_events(res, ev.data(), ne), "events");
    uint64_t te = 0, tv = 0;
    if (hgx_result_state_edges(res, nullptr, nullptr, nullptr, &te, &tv) != HGX_OK)
        throw Error{HGX_ERR_INVALID, "unfold: the result keeps no state edges (evolve it with HGX_REC_STATES)"};
    std::vector<uint64_t> soff(ns + 1);
    std::vector<uint32_t> eoff(te + 1), verts(std::max<uint64_t>(tv, 1));
    check(hgx_result_state_edges(res, soff.data(), eoff.data(), verts.data(), nullptr, nullptr), "state edges");
    std::vector<uint32_t> coff(ne + 1);
    check(hgx_result_consumed(res, coff.data(), nullptr, 0), "consumed edges");
    std::vector<uint16_t> cpos(std::max<uint32_t>(coff[ne], 1));
    check(hgx_result_consumed(res, coff.data(), cpos.data(), coff[ne]), "consumed edges");
    for (uint64_t s = 0; s < ns; ++s)
        if (soff[s + 1] - soff[s] > kUnfoldMaxSlots)
            throw Error{HGX_ERR_UNSUPPORTED, "unfold: a state above 4096 edges"};

    // exact hash and canonical rank of every edge of the states that can be replayed (states are
    // numbered generation by generation, so those below generation `steps` are a prefix)
    uint64_t nr = ns;
    if (!all_states)
        while (nr > 0 && step[nr - 1] >= steps) --nr;
    std::vector<uint64_t> ir(nr);
    std::vector<uint32_t> rank(std::max<uint64_t>(soff[nr], 1));
    if (nr) check(hgx_ir_rank_batch(nullptr, soff.data(), (uint32_t)nr, eoff.data(), verts.data(), ir.data(),
                                    rank.data()), "exact hashes");
    std::unordered_map<uint64_t, uint32_t> cls_of;
    std::vector<uint32_t> cls(ns, 0xFFFFFFFFu);
    Skeleton K;
    for (uint64_t s = 0; s < nr; ++s) {
        if (!ir[s] && soff[s + 1] > soff[s])
            throw Error{HGX_ERR_UNSUPPORTED, "unfold: the exact canonicalizer gave up on a state"};
        auto it = cls_of.emplace(ir[s], (uint32_t)K.cls_slots.size());
        if (it.second) K.cls_slots.push_back((uint32_t)(soff[s + 1] - soff[s]));
        cls[s] = it.first->second;
        K.S = std::max<uint32_t>(K.S, (uint32_t)(soff[s + 1] - soff[s]));
        if (creator[s] == 0xFFFFFFFFu) K.roots.push_back(cls[s]);
    }
    const uint32_t C = (uint32_t)K.cls_slots.size();      // the sentinel is class C
    for (uint64_t s = nr; s < ns; ++s) cls[s] = C;

    // each class's expansion, recorded from the first of its states that was expanded (events
    // are numbered by generation, then input state: the first event of a class names that state)
    std::vector<uint32_t> rec(C, 0xFFFFFFFFu);
    std::vector<std::vector<uint32_t>> by_cls(C);
    for (uint64_t e = 0; e < ne; ++e) {
        const uint32_t in = ev[e].input_state, k = cls[in];
        if (k >= C) throw Error{HGX_ERR_DEVICE, "unfold: an event leaves a state of the last generation"};
        if (rec[k] == 0xFFFFFFFFu) rec[k] = in;
        if (rec[k] == in) by_cls[k].push_back((uint32_t)e);
    }
    K.cls_br.assign(C, 0);
    std::vector<uint8_t> taken;
    std::vector<uint32_t> src;
    for (uint32_t k = 0; k < C; ++k) {
        const uint32_t mb = (uint32_t)K.m_to.size();
        for (uint32_t e : by_cls[k]) {
            const hgx_event& x = ev[e];
            const uint64_t r0 = soff[x.input_state], np = soff[x.input_state + 1] - r0;
            const uint64_t x0 = soff[x.output_state], nx = soff[x.output_state + 1] - x0;
            const uint32_t nk = coff[e + 1] - coff[e];
            if (np < nk || nx < np - nk) throw Error{HGX_ERR_DEVICE, "unfold: an event's states are inconsistent"};
            taken.assign(np, 0);
            for (uint32_t q = coff[e]; q < coff[e + 1]; ++q) {
                if (cpos[q] >= np) throw Error{HGX_ERR_DEVICE, "unfold: a consumed position is out of range"};
                taken[cpos[q]] = 1;
                K.m_cons.push_back((uint16_t)rank[r0 + cpos[q]]);
            }
            // the child keeps the parent's unconsumed edges in order, then the produced ones (a
            // child of the last generation has no slots: its class is the sentinel)
            const uint32_t to = cls[x.output_state];
            if (to < C) {
                src.assign(nx, kUnfoldProduced);
                uint64_t j = 0;
                for (uint64_t p = 0; p < np; ++p)
                    if (!taken[p]) src[rank[x0 + j++]] = (uint16_t)rank[r0 + p];
                K.m_src.insert(K.m_src.end(), src.begin(), src.end());
            }
            K.m_to.push_back(to);
            K.m_coff.push_back((uint32_t)K.m_cons.size());
            K.m_soff.push_back((uint32_t)K.m_src.size());
        }
        const uint32_t me = (uint32_t)K.m_to.size();
        K.cls_moff.push_back(me);
        // sibling applications of one instance whose consumed slots overlap (qr_apply's
        // branchial step: each unordered pair once)
        uint64_t b = 0;
        for (uint32_t m1 = mb; m1 < me; ++m1)
            for (uint32_t m2 = m1 + 1; m2 < me; ++m2) {
                bool hit = false;
                for (uint32_t q1 = K.m_coff[m1]; q1 < K.m_coff[m1 + 1] && !hit; ++q1)
                    for (uint32_t q2 = K.m_coff[m2]; q2 < K.m_coff[m2 + 1]; ++q2)
                        if (K.m_cons[q1] == K.m_cons[q2]) { hit = true; break; }
                b += hit;
            }
        K.cls_br[k] = b;
    }
    K.cls_moff.push_back(K.cls_moff.back());              // the sentinel: no slots, no matches
    K.cls_br.push_back(0);
    return K;
}

template <class T>
T* put(Arena& ar, const std::vector<T>& v, hipStream_t st) {
    T* d = dalloc<T>(ar, v.size());
    h2d(d, v.data(), v.size(), st);
    return d;
}

}  // namespace
}  // namespace hgx

using namespace hgx;

extern "C" int hgx_result_unfold(hgx_ctx* ctx, const hgx_result* res, hgx_unfolding* out) {
    if (!ctx || !res || !out) return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        hgx_counts c{};
        check(hgx_result_counts(res, &c), "counts");
        if (hgx_result_status(res, nullptr, 0) != HGX_OK || c.truncated)
            throw Error{HGX_ERR_INVALID, "unfold: the result is partial (a state cap or the HBM budget stopped it)"};
        std::memset(out, 0, sizeof(*out));
        // the requested generations: a quotient run whose new states all fell into known classes
        // stops early, while the raw unfolding beneath it goes on
        const uint32_t steps = res->run->hp->steps;
        const Skeleton K = build_skeleton(res, c, steps, res->run->hp->smp.on || steps == 0);
        out->classes = 0;
        for (size_t k = 0; k + 1 < K.cls_moff.size(); ++k) out->classes += K.cls_moff[k + 1] > K.cls_moff[k];
        out->class_matches = K.m_to.size();
        if (K.roots.empty() || steps == 0) {
            out->instances = K.roots.size();
            return HGX_OK;
        }
        hipStream_t st = ctx->stream;
        Arena tables(ctx->pool.get());
        const uint32_t* d_cls_moff = put(tables, K.cls_moff, st);
        const uint64_t* d_cls_br = put(tables, K.cls_br, st);
        const uint32_t* d_m_to = put(tables, K.m_to, st);
        const uint32_t* d_m_coff = put(tables, K.m_coff, st);
        const uint16_t* d_m_cons = put(tables, K.m_cons, st);
        const uint32_t* d_m_soff = put(tables, K.m_soff, st);
        const uint16_t* d_m_src = put(tables, K.m_src, st);
        unsigned long long* d_acc = dalloc<unsigned long long>(tables, 5);
        unsigned long long* d_part = dalloc<unsigned long long>(tables, 5ull * ctx->num_cu * 8);
        const std::vector<unsigned long long> zero(5, 0);
        h2d(d_acc, zero.data(), 5, st);
        const uint32_t S = K.S;

        // generation 0: one instance per root, every slot an initial edge
        auto lvl = std::make_unique<Arena>(ctx->pool.get());
        uint32_t N = (uint32_t)K.roots.size();
        uint32_t* d_cls = put(*lvl, K.roots, st);
        uint8_t* d_prod = dalloc<uint8_t>(*lvl, (uint64_t)N * S);
        const std::vector<uint8_t> init((uint64_t)N * S, kUnfoldInitial);   // alive until the copies end
        h2d(d_prod, init.data(), init.size(), st);
        uint64_t* d_anc = dalloc<uint64_t>(*lvl, 1);
        const uint32_t grid_cap = ctx->num_cu * 8;
        for (uint32_t d = 0; d < steps && N; ++d) {
            Arena tmp(ctx->pool.get());
            uint32_t* nm = dalloc<uint32_t>(tmp, (uint64_t)N + 1);
            uint64_t* wofs = dalloc<uint64_t>(tmp, (uint64_t)N + 1);
            const uint32_t gc = std::min<uint32_t>(grid_cap, (N + 256) / 256);
            hipLaunchKernelGGL(k_unfold_count, dim3(std::max(1u, gc)), dim3(256), 0, st, N, d_cls, d_cls_moff, nm);
            HIP_TRY(hipGetLastError());
            scan_u32_to_u64(tmp, nm, wofs, N, st);
            uint64_t W = 0;
            d2h(&W, wofs + N, 1, st);
            HIP_TRY(hipStreamSynchronize(st));
            out->instances += N;
            out->raw_events += W;
            if (!W) break;
            const bool write = d + 1 < steps;
            if (write && W >= 0xFFFFFFFFull)
                throw Error{HGX_ERR_UNSUPPORTED, "unfold: a generation of more than 2^32 raw states"};
            auto next = std::make_unique<Arena>(ctx->pool.get());
            UnfoldArgs a{};
            a.W = W; a.N = N; a.d = d; a.S = S; a.write = write ? 1u : 0u;
            a.wofs = wofs; a.cls = d_cls; a.prod = d_prod; a.anc = d_anc;
            a.cls_moff = d_cls_moff; a.cls_br = d_cls_br; a.m_to = d_m_to; a.m_coff = d_m_coff;
            a.m_cons = d_m_cons; a.m_soff = d_m_soff; a.m_src = d_m_src; a.part = d_part;
            if (write) {
                a.ccls = dalloc<uint32_t>(*next, W);
                a.cprod = dalloc<uint8_t>(*next, W * S);
                a.canc = dalloc<uint64_t>(*next, W * (uint64_t)(d + 1));
            }
            const uint64_t blocks = std::min<uint64_t>(grid_cap, (W + 255) / 256);
            // per work item: the instance's slots and ancestry read, the child's written
            const uint64_t bytes = W * (16ull + S + 8ull * d) + (write ? W * (4ull + S + 8ull * (d + 1)) : 0);
            timed(ctx, st, "unfold_apply", bytes, W, [&]() {
                hipLaunchKernelGGL(k_unfold_apply, dim3((uint32_t)blocks), dim3(256), 0, st, a);
                return hipGetLastError();
            });
            hipLaunchKernelGGL(k_unfold_reduce, dim3(1), dim3(256), 0, st, (const unsigned long long*)d_part,
                               (uint32_t)blocks, d_acc);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(st));
            if (!write) break;
            lvl = std::move(next);
            d_cls = a.ccls; d_prod = a.cprod; d_anc = a.canc;
            N = (uint32_t)W;
        }
        unsigned long long acc[5];
        d2h(acc, d_acc, 5, st);
        HIP_TRY(hipStreamSynchronize(st));
        prof_resolve(ctx);
        out->causal_triples = acc[0];
        out->causal_pairs = acc[1];
        out->causal_pairs_tr = acc[2];
        out->causal_triples_tr = acc[3];
        out->branchial_pairs = acc[4];
        return HGX_OK;
    });
}
