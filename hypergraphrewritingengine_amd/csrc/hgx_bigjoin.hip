// hgx_bigjoin.hip -- the join over one large state (tens of thousands of edges and up), where
// the state does not fit LDS: the match-join stress workload, one rewriting step of one big
// state (hgx_match_count, and hgx_match_count_noalloc over a fixed device scratch).
//
// Index (the role of SignatureIndex / InvertedVertexIndex, hypergraph/src/index.cpp:45-80): a CSR
// keyed by (vertex, position) -> ascending edge ids. Offsets from a histogram and a scan; the
// lists from a stable radix sort of (key, edge) pairs laid out in edge order, so each list is in
// ascending edge order with no per-bucket sort. Keying by position as well as vertex makes a
// candidate list exactly the edges carrying the bound vertex where the pattern needs it.
//
// Join (join_core.hpp:143-208), breadth first: the partial matches of depth d (d edges in join
// order) are a frontier held in HBM, in lexicographic order. Depth d+1 takes every (partial,
// candidate) pair -- candidates from the CSR list of the pivot vertex, ascending -- as one item:
// the items are numbered by a scan of the per-partial list lengths and dealt to workgroups in
// tiles of 2,048, each tile finding its partials by a binary search over the scan staged in LDS.
// A count pass sizes the next frontier per tile, a scan of the tile counts places every tile, and
// an emit pass writes the survivors in item order (ballot + mbcnt within a wave, wave totals
// across the workgroup). Item order is (partial, ascending candidate), so the final match list
// is in the order of a serial depth-first join over ascending candidates -- the oracle's order.
// The work per depth is spread over all its candidates, not one thread per seed.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hgx_bigjoin.hpp"
#include "hgx_device.hpp"
#include "hgx_kernels.hpp"
#include "hgx_runtime.hpp"

namespace hgx {

constexpr uint32_t BJ_THREADS = 256;
constexpr uint32_t BJ_IPT = 8;                            // items per thread per tile
constexpr uint32_t BJ_TILE = BJ_THREADS * BJ_IPT;         // items per workgroup tile
constexpr uint32_t BJ_LDS_PART = 514;                     // partials a tile stages in LDS (+1)

// ---- index -------------------------------------------------------------------------------

__global__ void k_index_count(const uint32_t* ev, const uint8_t* ar, uint32_t m, uint32_t A, uint32_t* cnt) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    for (uint32_t p = 0; p < ar[e]; ++p) atomicAdd(&cnt[ev[(size_t)e * A + p] * A + p], 1u);
}

// One (key, edge) pair per vertex slot, in edge order; slots past an edge's arity get the key
// nkeys, which sorts after every real one.
__global__ void k_index_keys(const uint32_t* ev, const uint8_t* ar, uint32_t m, uint32_t A, uint32_t nkeys,
                             uint32_t* key, uint32_t* val) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (uint64_t)m * A) return;
    const uint32_t e = (uint32_t)(s / A), p = (uint32_t)(s % A);
    key[s] = p < ar[e] ? ev[s] * A + p : nkeys;
    val[s] = e;
}

// The key of a vertex tuple (arity, vertices): equal tuples have equal keys; the join verifies
// every vertex of a candidate, so a collision costs a rejected candidate, never a wrong match.
__device__ __forceinline__ uint64_t tuple_key(uint32_t a, const uint32_t* x) {
    uint64_t h = 0x9E3779B97F4A7C15ULL ^ a;
    for (uint32_t i = 0; i < a; ++i) h = mix64(h ^ ((uint64_t)x[i] + 0x632BE59BD9B4E019ULL * (i + 1)));
    return h ? h : 1;
}

// One entry per run of equal keys in the sorted tuple index: (key -> first position, length).
__global__ void k_tuple_table(const uint64_t* tkey, uint32_t m, uint64_t* slot, uint2* run, uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t k = tkey[i];
    if (i && tkey[i - 1] == k) return;                     // not a run start
    uint32_t n = 1;
    while (i + n < m && tkey[i + n] == k) ++n;
    uint32_t h = (uint32_t)mix64(k) & mask;
    while (atomicCAS((unsigned long long*)&slot[h], 0ull, (unsigned long long)k) != 0ull) h = (h + 1) & mask;
    run[h] = make_uint2(i, n);
}

__global__ void k_tuple_keys(const uint32_t* ev, const uint8_t* ar, uint32_t m, uint32_t A, uint64_t* key,
                             uint32_t* val) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    key[e] = tuple_key(ar[e], ev + (size_t)e * A);
    val[e] = e;
}

// Zero fill: a kernel, so the join's stream holds kernels and copies only.
template <class T>
__global__ void k_bj_zero(T* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_u64_to_u32(const uint64_t* in, uint32_t* out, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)in[i];
}

// ---- join --------------------------------------------------------------------------------

// Depth 0 (SCAN): the edges that fit pattern edge order[0] -- its arity, and its repeated
// variables (match_core.hpp:32-49).
__global__ void k_bj_seed(BJView v, uint8_t* flag) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= v.m) return;
    const DRule& R = v.R;
    const uint32_t p = R.order[0], a = R.la[p];
    bool ok = v.ar[e] == a;
    const uint32_t* x = v.ev + (size_t)e * v.A;
    for (uint32_t i = 0; ok && i < a; ++i) {
        const uint32_t s = R.jsrc[0][i];
        if (s != 0xFFu) ok = x[i] == x[s & 15u];
    }
    flag[e] = ok ? 1 : 0;
}

// Candidate list of each partial of depth d: [lo, lo + cnt) of the pivot vertex's CSR list, or
// every edge when pattern edge order[d] shares no bound variable.
__global__ void k_bj_range(BJView v, uint32_t d, const uint32_t* F, uint64_t P, uint32_t* lo, uint32_t* cnt) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    const DRule& R = v.R;
    if ((v.full >> d) & 1u) {
        // a closing edge: the run of edges whose tuple key is the bound tuple's
        const uint32_t a = R.la[R.order[d]];
        uint32_t x[16];
        for (uint32_t i = 0; i < a; ++i) {
            const uint32_t s = R.jsrc[d][i];
            x[i] = v.ev[(size_t)F[q * d + (s >> 4)] * v.A + (s & 15u)];
        }
        const uint64_t k = tuple_key(a, x);
        uint32_t h = (uint32_t)mix64(k) & v.tmask;
        uint32_t l0 = 0, n = 0;
        for (uint32_t probe = 0; probe <= v.tmask; ++probe, h = (h + 1) & v.tmask) {
            const uint64_t s = v.tslot[h];
            if (s == k) { const uint2 r = v.trun[h]; l0 = r.x; n = r.y; break; }
            if (s == 0) break;
        }
        lo[q] = l0;
        cnt[q] = n;
        return;
    }
    const uint32_t pv = R.piv[d];
    if (pv == 0xFFu) { lo[q] = 0; cnt[q] = v.m; return; }
    const uint32_t s = R.jsrc[d][pv];                      // bound at an earlier depth
    const uint32_t vert = v.ev[(size_t)F[q * d + (s >> 4)] * v.A + (s & 15u)];
    const uint32_t key = vert * v.A + pv;
    lo[q] = v.off[key];
    cnt[q] = v.off[key + 1] - v.off[key];
}

// Last q with start[q] <= t, over start[0..n] (start[0] = 0).
__device__ __forceinline__ uint64_t bj_owner(const uint64_t* start, uint64_t n, uint64_t t) {
    uint64_t a = 0, b = n;                                  // start[a] <= t < start[b]
    while (b - a > 1) {
        const uint64_t mid = (a + b) >> 1;
        if (start[mid] <= t) a = mid; else b = mid;
    }
    return a;
}

// Item t of depth d: candidate e of partial q. Fits when it has the arity, is not already in the
// partial (edge-injective) and agrees with every binding the partial and its own earlier slots
// fixed (vertex-non-injective: distinct variables may share a vertex).
__device__ __forceinline__ bool bj_fits(const BJView& v, uint32_t d, const uint32_t* part, uint32_t e) {
    const DRule& R = v.R;
    const uint32_t p = R.order[d], a = R.la[p];
    if (v.ar[e] != a) return false;
    for (uint32_t t = 0; t < d; ++t)
        if (part[t] == e) return false;
    const uint32_t* x = v.ev + (size_t)e * v.A;
    for (uint32_t i = 0; i < a; ++i) {
        const uint32_t s = R.jsrc[d][i];
        if (s == 0xFFu) continue;
        const uint32_t sd = s >> 4, si = s & 15u;
        const uint32_t want = sd < d ? v.ev[(size_t)part[sd] * v.A + si] : x[si];
        if (x[i] != want) return false;
    }
    return true;
}

// bj_fits with the partial's edges and the vertices its bound slots need staged in LDS. The pivot
// slot needs no test: the candidate came from the pivot vertex's list at that position.
__device__ __forceinline__ bool bj_fits_staged(const BJView& v, uint32_t d, uint32_t pv, const uint32_t* part,
                                               const uint32_t* want, uint32_t e) {
    const DRule& R = v.R;
    const uint32_t p = R.order[d], a = R.la[p];
    if (v.ar[e] != a) return false;
    for (uint32_t t = 0; t < d; ++t)
        if (part[t] == e) return false;
    const uint32_t* x = v.ev + (size_t)e * v.A;
    for (uint32_t i = 0; i < a; ++i) {
        const uint32_t s = R.jsrc[d][i];
        if (s == 0xFFu || i == pv) continue;
        const uint32_t want_i = (s >> 4) < d ? want[i] : x[s & 15u];
        if (x[i] != want_i) return false;
    }
    return true;
}

struct BJLevel {
    uint32_t d;               // depth being added (1 .. n-1)
    const uint32_t* F;        // [P * d] partials, join order
    uint64_t P;
    const uint32_t* lo;       // [P]
    const uint64_t* start;    // [P + 1] scan of the list lengths; start[P] = items (device)
    uint64_t ntiles;          // tiles provisioned: an upper bound of the items' tiles
    uint64_t* tile_cnt;       // count pass: survivors per tile
    const uint64_t* tile_off; // emit pass: scan of tile_cnt
    uint32_t* Fn;             // emit, inner depth: [S * (d+1)]
    uint32_t* out;            // emit, last depth: matches, n_lhs edges each in pattern order
    const uint64_t* found;    //   (device: matches written before this call)
    uint64_t cap;             //   (room for cap matches in all)
    uint64_t* slot_q;         // last depth: [ntiles * BJ_SLOT] survivors' partials
    uint32_t* slot_e;         //   and candidates, in item order
};

// Passes over the items of a depth, a tile of BJ_TILE items per workgroup iteration. The number of
// items is read on the device (the scan's total), so no host round trip sizes the launch; tiles
// past it count zero. The tile's partials -- their item offsets, candidate-list starts and (up to
// BJ_STAGE_D edges) their edges -- are staged in LDS, so an item reads only its candidate from
// global memory.
//   BJ_COUNT   survivors per tile (an inner depth, sizing the next frontier)
//   BJ_EMIT    survivors written in item order: the next frontier (inner depth), or, at the last
//              depth, the matches of the tiles whose survivors overflowed their slots
//   BJ_SLOTS   the last depth in one pass: survivors per tile, and the first BJ_SLOT of each tile
//              kept in item order as (partial, candidate), placed into the match list after the
//              scan by k_bj_gather (matches are rare among the items, so a second pass over them
//              is the exception)
enum { BJ_COUNT = 0, BJ_EMIT = 1, BJ_SLOTS = 2 };
constexpr uint32_t BJ_SLOT = 16;
constexpr uint32_t BJ_STAGE_D = 3;
constexpr uint32_t BJ_STAGE_A = 4;

template <int MODE, bool LAST>
__global__ __launch_bounds__(BJ_THREADS) void k_bj_items(BJView v, BJLevel L) {
    __shared__ uint64_t sst[BJ_LDS_PART];
    __shared__ uint32_t slo[BJ_LDS_PART];
    __shared__ uint32_t spart[BJ_LDS_PART * BJ_STAGE_D];
    __shared__ uint32_t swant[BJ_LDS_PART * BJ_STAGE_A];   // per partial: the vertex each slot needs
    __shared__ uint32_t wtot[BJ_THREADS / WAVE];
    __shared__ uint64_t sq[2];
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / WAVE;
    const uint32_t d = L.d;
    const bool full = (v.full >> d) & 1u;                 // candidates from the tuple index
    const uint32_t pv = full ? 0xFFFFFFFEu : v.R.piv[d];  // (a closing edge checks every slot)
    const uint32_t pa = v.R.la[v.R.order[d]];
    const uint64_t T = L.start[L.P];
    const uint64_t nreal = (T + BJ_TILE - 1) / BJ_TILE;
    const uint64_t base = (MODE == BJ_EMIT && LAST) ? *L.found : 0;
    for (uint64_t tile = blockIdx.x; tile < L.ntiles; tile += gridDim.x) {
        // the same for the whole workgroup: past the items, or (last depth, emit) a tile whose
        // survivors all fit its slots
        if (tile >= nreal || (MODE == BJ_EMIT && LAST && L.tile_cnt[tile] <= BJ_SLOT)) {
            if (MODE != BJ_EMIT && tid == 0 && tile >= nreal) L.tile_cnt[tile] = 0;
            continue;
        }
        const uint64_t t0 = tile * BJ_TILE;
        const uint64_t t1 = t0 + BJ_TILE < T ? t0 + BJ_TILE : T;
        if (tid == 0) {
            sq[0] = bj_owner(L.start, L.P, t0);
            sq[1] = bj_owner(L.start, L.P, t1 - 1);
        }
        __syncthreads();
        const uint64_t q0 = sq[0], nq = sq[1] - sq[0] + 1;  // partials this tile touches
        const bool staged = nq + 1 <= BJ_LDS_PART;
        const bool sparts = staged && d <= BJ_STAGE_D && pa <= BJ_STAGE_A;
        if (staged) {
            for (uint64_t i = tid; i <= nq; i += BJ_THREADS) sst[i] = L.start[q0 + i];
            for (uint64_t i = tid; i < nq; i += BJ_THREADS) slo[i] = L.lo[q0 + i];
            if (sparts) {
                for (uint64_t i = tid; i < nq * d; i += BJ_THREADS) spart[i] = L.F[q0 * d + i];
                // the vertex every slot bound at an earlier depth must carry, once per partial
                for (uint64_t i = tid; i < nq * pa; i += BJ_THREADS) {
                    const uint32_t k = (uint32_t)(i % pa), sk = v.R.jsrc[d][k];
                    uint32_t want = 0;
                    if (sk != 0xFFu && (sk >> 4) < d)
                        want = v.ev[(size_t)L.F[(q0 + i / pa) * d + (sk >> 4)] * v.A + (sk & 15u)];
                    swant[i] = want;
                }
            }
        }
        __syncthreads();
        uint64_t rank = MODE == BJ_EMIT ? L.tile_off[tile] : 0;   // running place of the tile's survivors
        uint32_t found = 0;
        for (uint32_t k = 0; k < BJ_IPT; ++k) {
            const uint64_t t = t0 + (uint64_t)k * BJ_THREADS + tid;
            bool ok = false;
            uint64_t q = 0;
            uint32_t e = 0;
            const uint32_t* part = nullptr;
            if (t < t1) {
                uint64_t j;
                if (staged) {
                    const uint64_t r = bj_owner(sst, nq, t);
                    q = q0 + r;
                    j = t - sst[r];
                    e = full ? v.tidx[slo[r] + j] : pv == 0xFFu ? (uint32_t)j : v.idx[slo[r] + j];
                    part = sparts ? spart + r * d : L.F + q * d;
                    if (sparts) ok = bj_fits_staged(v, d, pv, part, swant + r * pa, e);
                    else ok = bj_fits(v, d, part, e);
                } else {
                    q = bj_owner(L.start, L.P, t);
                    j = t - L.start[q];
                    e = full ? v.tidx[L.lo[q] + j] : pv == 0xFFu ? (uint32_t)j : v.idx[L.lo[q] + j];
                    part = L.F + q * d;
                    ok = bj_fits(v, d, part, e);
                }
            }
            if (MODE == BJ_COUNT) { found += ok ? 1u : 0u; continue; }
            // place among this round's survivors: within the wave, then across the waves
            const uint64_t bal = __ballot(ok);
            if (lane == 0) wtot[w] = (uint32_t)__popcll(bal);
            __syncthreads();
            uint32_t before = 0, round = 0;
            for (uint32_t x = 0; x < BJ_THREADS / WAVE; ++x) {
                before += x < w ? wtot[x] : 0u;
                round += wtot[x];
            }
            __syncthreads();
            if (ok) {
                const uint64_t o = rank + before + lanes_below(bal);
                if (MODE == BJ_SLOTS) {
                    if (o < BJ_SLOT) {
                        L.slot_q[tile * BJ_SLOT + o] = q;
                        L.slot_e[tile * BJ_SLOT + o] = e;
                    }
                } else if (LAST) {
                    const uint32_t n = v.R.n_lhs;
                    const uint64_t g = base + o;
                    if (g < L.cap) {
                        for (uint32_t x = 0; x < d; ++x) L.out[g * n + v.R.order[x]] = part[x];
                        L.out[g * n + v.R.order[d]] = e;
                    }
                } else {
                    for (uint32_t x = 0; x < d; ++x) L.Fn[o * (d + 1) + x] = part[x];
                    L.Fn[o * (d + 1) + d] = e;
                }
            }
            rank += round;
        }
        if (MODE == BJ_COUNT) {
            const uint32_t ws = wave_sum_u32(found);
            if (lane == 0) wtot[w] = ws;
            __syncthreads();
            if (tid == 0) {
                uint64_t sm = 0;
                for (uint32_t x = 0; x < BJ_THREADS / WAVE; ++x) sm += wtot[x];
                L.tile_cnt[tile] = sm;
            }
        } else if (MODE == BJ_SLOTS) {
            if (tid == 0) L.tile_cnt[tile] = rank;
        }
        __syncthreads();
    }
}

// The slotted survivors of the last depth into the match list: slot i of a tile whose survivors all
// fit its slots goes to place found + tile_off[tile] + i.
__global__ void k_bj_gather(BJView v, BJLevel L) {
    const uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t tile = x / BJ_SLOT, i = x % BJ_SLOT;
    if (tile >= L.ntiles) return;
    const uint64_t c = L.tile_cnt[tile];
    if (c > BJ_SLOT || i >= c) return;
    const uint64_t g = *L.found + L.tile_off[tile] + i;
    if (g >= L.cap) return;
    const uint32_t n = v.R.n_lhs, d = L.d;
    const uint64_t q = L.slot_q[tile * BJ_SLOT + i];
    const uint32_t* part = L.F + q * d;
    for (uint32_t t = 0; t < d; ++t) L.out[g * n + v.R.order[t]] = part[t];
    L.out[g * n + v.R.order[d]] = L.slot_e[tile * BJ_SLOT + i];
}

// Device-side bookkeeping of a depth: acc[0] matches found, acc[1] algorithmic bytes (the
// candidate reads of every item, the partials' lists, the frontier or the matches written).
__global__ void k_bj_tally(uint64_t* acc, const uint64_t* T, const uint64_t* S, uint64_t P, uint32_t d,
                           uint32_t a, uint32_t n, uint32_t last, uint64_t cap) {
    const uint64_t s = *S, t = *T;
    uint64_t bytes = t * 4ull * (1ull + a) + P * (4ull * d + 12ull);
    if (last) {
        const uint64_t before = acc[0];
        const uint64_t room = cap > before ? cap - before : 0;
        bytes += (s < room ? s : room) * 4ull * n;
        acc[0] = before + s;
    } else {
        bytes += s * 4ull * (d + 1);
    }
    acc[1] += bytes;
}

}  // namespace hgx

namespace hgx {

struct BJCast64 {
    __host__ __device__ uint64_t operator()(uint32_t x) const { return x; }
};

template <class T>
static T* bget(BJMem& m, size_t n) { return (T*)m.get(sizeof(T) * (n ? n : 1)); }

template <class T>
static T read1(const T* d, hipStream_t st) {
    T x{};
    HIP_TRY(hipMemcpyAsync(&x, d, sizeof(T), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return x;
}

// exclusive sums over n + 1 entries (the last input zero): out[n] is the total
static void bj_scan(BJMem& m, const uint32_t* in, uint64_t* out, uint64_t n, hipStream_t st) {
    if (n + 1 > 0x7FFFFFFFull) throw Error{HGX_ERR_UNSUPPORTED, "join frontier above 2^31 partials"};
    hipcub::TransformInputIterator<uint64_t, BJCast64, const uint32_t*> it(in, BJCast64());
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, it, out, (int)(n + 1), st));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(m.get(tmp), tmp, it, out, (int)(n + 1), st));
}
static void bj_scan(BJMem& m, const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t st) {
    if (n + 1 > 0x7FFFFFFFull) throw Error{HGX_ERR_UNSUPPORTED, "join above 2^31 tiles"};
    size_t tmp = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, (int)(n + 1), st));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(m.get(tmp), tmp, in, out, (int)(n + 1), st));
}

static uint32_t blocks_for(uint64_t n, uint32_t threads) { return (uint32_t)std::max<uint64_t>(1, (n + threads - 1) / threads); }

// CSR (vertex, position) -> ascending edges of a state held on the device as ev [m * A], ar [m],
// vertex ids below nv.
static void bj_index(BJMem& mem, BJView& v, uint32_t nv, hipStream_t st, uint64_t* alg_bytes, uint32_t* maxlist) {
    const uint32_t m = v.m, A = v.A;
    const uint64_t nkeys = (uint64_t)nv * A, slots = (uint64_t)m * A;
    if (nkeys >= 0xFFFFFFF0ull || slots >= 0x7FFFFFF0ull)
        throw Error{HGX_ERR_UNSUPPORTED, "index above 2^31 vertex slots"};
    uint32_t* cnt = bget<uint32_t>(mem, nkeys + 1);
    uint64_t* off64 = bget<uint64_t>(mem, nkeys + 1);
    uint32_t* off = bget<uint32_t>(mem, nkeys + 1);
    hipLaunchKernelGGL(k_bj_zero<uint32_t>, dim3(blocks_for(nkeys + 1, 256)), dim3(256), 0, st, cnt, nkeys + 1);
    if (m) hipLaunchKernelGGL(k_index_count, dim3(blocks_for(m, 256)), dim3(256), 0, st, v.ev, v.ar, m, A, cnt);
    bj_scan(mem, cnt, off64, nkeys, st);
    {   // the longest list, which bounds the candidates of every partial match
        size_t tmp = 0;
        HIP_TRY(hipcub::DeviceReduce::Max(nullptr, tmp, cnt, maxlist, (int)(nkeys + 1), st));
        HIP_TRY(hipcub::DeviceReduce::Max(mem.get(tmp), tmp, cnt, maxlist, (int)(nkeys + 1), st));
    }
    hipLaunchKernelGGL(k_u64_to_u32, dim3(blocks_for(nkeys + 1, 256)), dim3(256), 0, st, off64, off,
                       (uint32_t)(nkeys + 1));
    uint32_t* kin = bget<uint32_t>(mem, slots);
    uint32_t* vin = bget<uint32_t>(mem, slots);
    uint32_t* kout = bget<uint32_t>(mem, slots);
    uint32_t* idx = bget<uint32_t>(mem, slots);
    if (slots) {
        hipLaunchKernelGGL(k_index_keys, dim3(blocks_for(slots, 256)), dim3(256), 0, st, v.ev, v.ar, m, A,
                           (uint32_t)nkeys, kin, vin);
        int end_bit = 1;
        while (end_bit < 32 && (nkeys >> end_bit) != 0) ++end_bit;      // keys <= nkeys
        size_t tmp = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, kin, kout, vin, idx, (int)slots, 0, end_bit, st));
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(mem.get(tmp), tmp, kin, kout, vin, idx, (int)slots, 0, end_bit, st));
    }
    HIP_TRY(hipGetLastError());
    v.off = off;
    v.idx = idx;
    // read the state once, write and read the pairs through the sort, write the offsets
    *alg_bytes += (uint64_t)m * (4ull * A + 1) + slots * 8ull * 3 + (nkeys + 1) * 16ull;
    // closing edges: the depths whose pattern edge is bound in every slot by earlier depths
    const DRule& R = v.R;
    v.full = 0;
    for (uint32_t d = 1; d < R.n_lhs; ++d) {
        const uint32_t a = R.la[R.order[d]];
        bool all = true;
        for (uint32_t i = 0; i < a; ++i) {
            const uint32_t src = R.jsrc[d][i];
            all &= src != 0xFFu && (src >> 4) < d;
        }
        if (all) v.full |= 1u << d;
    }
    v.tkey = nullptr;
    v.tidx = nullptr;
    if (v.full && m) {
        uint64_t* tk_in = bget<uint64_t>(mem, m);
        uint32_t* tv_in = bget<uint32_t>(mem, m);
        uint64_t* tk = bget<uint64_t>(mem, m);
        uint32_t* tv = bget<uint32_t>(mem, m);
        hipLaunchKernelGGL(k_tuple_keys, dim3(blocks_for(m, 256)), dim3(256), 0, st, v.ev, v.ar, m, A, tk_in, tv_in);
        size_t tmp = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, tk_in, tk, tv_in, tv, (int)m, 0, 64, st));
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(mem.get(tmp), tmp, tk_in, tk, tv_in, tv, (int)m, 0, 64, st));
        HIP_TRY(hipGetLastError());
        v.tkey = tk;
        v.tidx = tv;
        uint32_t cap = 1;
        while (cap < 2 * m) cap <<= 1;                      // load factor <= 1/2
        v.tslot = bget<uint64_t>(mem, cap);
        v.trun = bget<uint2>(mem, cap);
        v.tmask = cap - 1;
        hipLaunchKernelGGL(k_bj_zero<uint64_t>, dim3(blocks_for(cap, 256)), dim3(256), 0, st, v.tslot, (uint64_t)cap);
        hipLaunchKernelGGL(k_tuple_table, dim3(blocks_for(m, 256)), dim3(256), 0, st, tk, m, v.tslot, v.trun, v.tmask);
        HIP_TRY(hipGetLastError());
        *alg_bytes += (uint64_t)m * (4ull * A + 1) + (uint64_t)m * 12ull * 3 + (uint64_t)cap * 16ull;
    } else {
        v.full = 0;
    }
}

// Extends the partials F (P of them, d edges each) by pattern edge order[d], and on to the last
// depth; splits F in two, in order, when the next frontier would outgrow the memory at hand. The
// host learns a depth's survivors only where it must size the next frontier; the last depth
// writes its matches at the device-side running count and leaves the host nothing to wait for.
static void bj_depth(BJRun& J, uint32_t d, const uint32_t* F, uint64_t P) {
    BJMem& mem = *J.mem;
    const uint32_t n = J.v.R.n_lhs, a = J.v.R.la[J.v.R.order[d]];
    const bool last = d + 1 == n;
    const size_t mk = mem.mark();
    uint32_t* lo = bget<uint32_t>(mem, P);
    uint32_t* cnt = bget<uint32_t>(mem, P + 1);
    uint64_t* start = bget<uint64_t>(mem, P + 1);
    hipLaunchKernelGGL(k_bj_zero<uint32_t>, dim3(1), dim3(64), 0, J.st, cnt + P, (uint64_t)1);
    hipLaunchKernelGGL(k_bj_range, dim3(blocks_for(P, 256)), dim3(256), 0, J.st, J.v, d, F, P, lo, cnt);
    bj_scan(mem, cnt, start, P, J.st);
    // tiles to provision: the longest candidate list bounds every partial's; an exact count (one
    // round trip) only when that bound is loose beyond use
    const bool full = (J.v.full >> d) & 1u;
    const uint64_t per = J.v.R.piv[d] == 0xFFu ? (uint64_t)J.v.m : std::max<uint64_t>(1, J.maxlist);
    uint64_t ntiles = (P * per + BJ_TILE - 1) / BJ_TILE;
    // a closing edge has about one candidate per partial: the exact count, not the list bound
    if (full || P * per / std::max<uint64_t>(1, P) != per || ntiles > ((uint64_t)1 << 20)) {
        const uint64_t T = read1(start + P, J.st);
        ntiles = (T + BJ_TILE - 1) / BJ_TILE;
    }
    if (ntiles == 0) { mem.rewind(mk); return; }
    uint64_t* tcnt = bget<uint64_t>(mem, ntiles + 1);
    uint64_t* toff = bget<uint64_t>(mem, ntiles + 1);
    hipLaunchKernelGGL(k_bj_zero<uint64_t>, dim3(1), dim3(64), 0, J.st, tcnt + ntiles, (uint64_t)1);
    BJLevel L{};
    L.d = d; L.F = F; L.P = P; L.lo = lo; L.start = start; L.ntiles = ntiles; L.tile_cnt = tcnt; L.tile_off = toff;
    L.found = J.acc; L.cap = J.cap;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(ntiles, J.grid);
    if (last) {
        if (J.out) {
            L.slot_q = bget<uint64_t>(mem, ntiles * BJ_SLOT);
            L.slot_e = bget<uint32_t>(mem, ntiles * BJ_SLOT);
            hipLaunchKernelGGL((k_bj_items<BJ_SLOTS, true>), dim3(grid), dim3(BJ_THREADS), 0, J.st, J.v, L);
        } else {
            hipLaunchKernelGGL((k_bj_items<BJ_COUNT, true>), dim3(grid), dim3(BJ_THREADS), 0, J.st, J.v, L);
        }
    } else {
        hipLaunchKernelGGL((k_bj_items<BJ_COUNT, false>), dim3(grid), dim3(BJ_THREADS), 0, J.st, J.v, L);
    }
    HIP_TRY(hipGetLastError());
    bj_scan(mem, tcnt, toff, ntiles, J.st);
    if (last) {
        if (J.out) {
            L.out = J.out;
            hipLaunchKernelGGL(k_bj_gather, dim3(blocks_for(ntiles * BJ_SLOT, 256)), dim3(256), 0, J.st, J.v, L);
            // tiles whose survivors overflowed their slots: a second pass over those tiles only
            hipLaunchKernelGGL((k_bj_items<BJ_EMIT, true>), dim3(grid), dim3(BJ_THREADS), 0, J.st, J.v, L);
            HIP_TRY(hipGetLastError());
        }
        hipLaunchKernelGGL(k_bj_tally, dim3(1), dim3(1), 0, J.st, J.acc, start + P, toff + ntiles, P, d, a, n,
                           1u, J.out ? J.cap : 0ull);
        mem.rewind(mk);
        return;
    }
    const uint64_t S = read1(toff + ntiles, J.st);
    if (S == 0) { mem.rewind(mk); return; }
    // the next frontier, the next depth's lists and scan, and its tiles (counts, offsets, slots)
    const uint64_t need = S * (4ull * (d + 1) + 16ull) +
                          (S * std::max<uint64_t>(1, J.maxlist) / BJ_TILE + 2) * (16ull + 12ull * BJ_SLOT);
    if (need > mem.room()) {
        if (P == 1) throw Error{HGX_ERR_CAPACITY, "one partial match of the join has more extensions than memory holds"};
        mem.rewind(mk);
        bj_depth(J, d, F, P / 2);
        bj_depth(J, d, F + (P / 2) * d, P - P / 2);
        return;
    }
    uint32_t* Fn = bget<uint32_t>(mem, S * (d + 1));
    L.Fn = Fn;
    hipLaunchKernelGGL((k_bj_items<BJ_EMIT, false>), dim3(grid), dim3(BJ_THREADS), 0, J.st, J.v, L);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_bj_tally, dim3(1), dim3(1), 0, J.st, J.acc, start + P, toff + ntiles, P, d, a, n, 0u, 0ull);
    bj_depth(J, d + 1, Fn, S);
    mem.rewind(mk);
}

// All matches of rule R in the state (ev, ar) on the device: the index, the seeds, the depths.
void bj_match(BJRun& J, uint32_t nv, float* ms_index, float* ms_join, bool timed) {
    BJMem& mem = *J.mem;
    if (!mem.arena) {
        // a fixed region: refuse a state whose index alone cannot fit before launching anything
        const uint64_t nkeys = (uint64_t)nv * J.v.A, slots = (uint64_t)J.v.m * J.v.A;
        const uint64_t need = (nkeys + 1) * 16ull + slots * 16ull + (uint64_t)J.v.m * 93ull + (1ull << 20);
        if (need > mem.cap - mem.used)
            throw Error{HGX_ERR_CAPACITY, "a state of " + std::to_string(J.v.m) + " edges needs " +
                                              std::to_string(need >> 20) + " MiB for its index, more than the " +
                                              std::to_string(mem.cap >> 20) + " MiB fixed scratch"};
    }
    hipEvent_t t0 = nullptr, t1 = nullptr, t2 = nullptr;
    if (timed) {
        HIP_TRY(hipEventCreate(&t0)); HIP_TRY(hipEventCreate(&t1)); HIP_TRY(hipEventCreate(&t2));
        HIP_TRY(hipEventRecord(t0, J.st));
    }
    uint32_t* scal = bget<uint32_t>(mem, 2);          // seeds, longest candidate list
    uint32_t* nsel = scal;
    J.acc = bget<uint64_t>(mem, 2);
    hipLaunchKernelGGL(k_bj_zero<uint64_t>, dim3(1), dim3(64), 0, J.st, J.acc, (uint64_t)2);
    bj_index(mem, J.v, nv, J.st, &J.index_bytes, scal + 1);
    if (timed) HIP_TRY(hipEventRecord(t1, J.st));
    const uint32_t m = J.v.m;
    uint8_t* flag = bget<uint8_t>(mem, m);
    uint32_t* seeds = bget<uint32_t>(mem, m);
    if (m) {
        hipLaunchKernelGGL(k_bj_seed, dim3(blocks_for(m, 256)), dim3(256), 0, J.st, J.v, flag);
        hipcub::CountingInputIterator<uint32_t> ids(0);
        size_t tmp = 0;
        HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, tmp, ids, flag, seeds, nsel, (int)m, J.st));
        HIP_TRY(hipcub::DeviceSelect::Flagged(mem.get(tmp), tmp, ids, flag, seeds, nsel, (int)m, J.st));
    } else {
        hipLaunchKernelGGL(k_bj_zero<uint32_t>, dim3(1), dim3(64), 0, J.st, nsel, (uint64_t)1);
    }
    uint32_t hs[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(hs, scal, sizeof(hs), hipMemcpyDeviceToHost, J.st));
    HIP_TRY(hipStreamSynchronize(J.st));
    const uint64_t P0 = hs[0];
    J.maxlist = hs[1];
    J.alg_bytes += (uint64_t)m * (4ull * J.v.A + 1) + P0 * 4ull;
    if (J.v.R.n_lhs == 1) {
        if (J.out && P0) {
            const uint64_t k = std::min(P0, J.cap);
            HIP_TRY(hipMemcpyAsync(J.out, seeds, sizeof(uint32_t) * k, hipMemcpyDeviceToDevice, J.st));
        }
        J.found = P0;
    } else if (P0) {
        bj_depth(J, 1, seeds, P0);
    }
    if (timed) HIP_TRY(hipEventRecord(t2, J.st));
    uint64_t acc[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(acc, J.acc, sizeof(acc), hipMemcpyDeviceToHost, J.st));
    HIP_TRY(hipStreamSynchronize(J.st));
    if (J.v.R.n_lhs > 1) J.found = acc[0];
    J.alg_bytes += acc[1];
    if (timed) {
        HIP_TRY(hipEventElapsedTime(ms_index, t0, t1));
        HIP_TRY(hipEventElapsedTime(ms_join, t1, t2));
        (void)hipEventDestroy(t0); (void)hipEventDestroy(t1); (void)hipEventDestroy(t2);
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
        RunHost h = parse_run(&d, false);              // validates rules and arities, renumbers
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
        BJMem mem;
        mem.arena = &arena;
        uint32_t* d_ev = bget<uint32_t>(mem, (size_t)m * A);
        uint8_t* d_ar = bget<uint8_t>(mem, m);
        h2d(d_ev, ev.data(), ev.size(), s);
        h2d(d_ar, ar.data(), ar.size(), s);
        BJRun J{};
        J.st = s; J.mem = &mem; J.grid = ctx->num_cu * 8;
        J.v.R = R; J.v.ev = d_ev; J.v.ar = d_ar; J.v.m = m; J.v.A = A;
        uint32_t* d_out = nullptr;
        float ms_index = 0, ms_join = 0;
        if (match_edges && cap && cap * R.n_lhs * 4ull > (256ull << 20)) {
            // a large caller buffer: count first, then hold only the matches there are
            BJRun C = J;
            Arena counting(ctx->pool.get());
            BJMem cm;
            cm.arena = &counting;
            C.mem = &cm;
            bj_match(C, nv, &ms_index, &ms_join, false);
            cap = std::min<uint64_t>(cap, std::max<uint64_t>(1, C.found));
        }
        if (match_edges && cap) {
            d_out = bget<uint32_t>(mem, cap * R.n_lhs);
            J.out = d_out; J.cap = cap;
        }
        bj_match(J, nv, &ms_index, &ms_join, true);
        const uint64_t total = J.found;
        if (d_out && total) d2h(match_edges, d_out, std::min(cap, total) * R.n_lhs, s);
        HIP_TRY(hipStreamSynchronize(s));
        if (device_ms) device_ms[0] = ms_index + ms_join;
        if (ctx->prof.on) {
            auto& si = ctx->prof.acc["bigjoin_index"];
            std::snprintf(si.name, sizeof(si.name), "bigjoin_index");
            si.launches += 1; si.total_ms += ms_index; si.items += m; si.alg_bytes += J.index_bytes;
            auto& sj = ctx->prof.acc["bigjoin_join"];
            std::snprintf(sj.name, sizeof(sj.name), "bigjoin_join");
            sj.launches += 1; sj.total_ms += ms_join; sj.items += m; sj.alg_bytes += J.alg_bytes;
        }
        *count = total;
        return HGX_OK;
    });
}

