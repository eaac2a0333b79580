// hgx_noalloc.hip -- hgx_match_count_noalloc: the large-state join (hgx_match_count) for callers
// that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). The caller passes the state and the output buffer in device
// memory; everything else the join needs -- the CSR index, the frontiers of partial matches, the
// scans' temporary storage -- comes from `scratch`, a fixed array in device memory of
// HGX_NOALLOC_SCRATCH_MIB MiB per device, held from the library's first context on that device
// until the process ends. No hipMalloc, hipFree or allocating library routine is called: hipCUB's
// scans, sort and selection get their temporary storage from the scratch as well.
//
// The scratch is shared by every call on a device and keeps the last call's contents; what a call
// needs zeroed it zeroes itself. Calls through one context are serialised by its lock; calls that
// overlap on one device through different contexts race on the scratch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "hgx_bigjoin.hpp"
#include "hgx_runtime.hpp"

namespace hgx {

constexpr size_t kNoAllocScratch = (size_t)HGX_NOALLOC_SCRATCH_MIB << 20;

// One copy per device. In host code the name stands for a host-side symbol, not a device
// address: every use goes through the address hipGetSymbolAddress returned for the device.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratch];

static std::mutex g_scratch_mu;
static std::vector<uint8_t*> g_scratch_ptr;          // per device, null until looked up

// The device address of `scratch` on `device`, which must be the current device.
static uint8_t* scratch_on(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (device < 0) throw Error{HGX_ERR_INVALID, "negative device index"};
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            throw Error{HGX_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
        }
        g_scratch_ptr[device] = (uint8_t*)p;
    }
    return g_scratch_ptr[device];
}

// Called by hgx_ctx_create with the context's device current.
void noalloc_setup(int device) { (void)scratch_on(device); }

}  // namespace hgx

using namespace hgx;

extern "C" int hgx_match_count_noalloc(hgx_ctx* ctx, const uint8_t* rules, uint32_t rules_len, uint32_t n_rules,
                                       uint32_t rule, const uint32_t* vertices, const uint8_t* arity,
                                       uint32_t n_edges, uint32_t stride, uint32_t n_vertices, uint64_t* count,
                                       uint32_t* match_edges, uint64_t cap) {
    if (!ctx || !rules || !count || (n_edges && (!vertices || !arity)) || rule >= n_rules || stride < 1 ||
        stride > HGX_MAX_ARITY || (cap && !match_edges))
        return HGX_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    return guarded(ctx, [&]() -> int {
        HIP_TRY(hipSetDevice(ctx->device));
        std::vector<DRule> rs;
        parse_rules(rules, rules_len, n_rules, rs);
        const DRule& R = rs[rule];
        for (int i = 0; i < R.n_lhs; ++i)
            if (R.la[i] > stride) throw Error{HGX_ERR_INVALID, "a pattern edge is wider than the vertex stride"};
        if ((uint64_t)std::max(1u, n_vertices) * stride >= 0xFFFFFFF0ull)
            throw Error{HGX_ERR_INVALID, "n_vertices * stride above 2^32"};
        BJMem mem;
        mem.base = scratch_on(ctx->device);
        mem.cap = kNoAllocScratch;
        BJRun J{};
        J.st = ctx->stream; J.mem = &mem; J.grid = ctx->num_cu * 8;
        J.v.R = R; J.v.ev = vertices; J.v.ar = arity; J.v.m = n_edges; J.v.A = stride;
        J.out = cap ? match_edges : nullptr; J.cap = cap;
        bj_match(J, std::max(1u, n_vertices), nullptr, nullptr, false);
        *count = J.found;
        return HGX_OK;
    });
}
