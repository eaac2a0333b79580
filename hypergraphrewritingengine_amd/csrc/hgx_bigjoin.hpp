// hgx_bigjoin.hpp -- the large-state join's host interface, shared by hgx_match_count and the
// allocation-free entry point: the state and rule as the kernels see them, the device memory
// the join draws from, and the join itself.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "hgx_device.hpp"
#include "hgx_runtime.hpp"

namespace hgx {

struct BJView {
    DRule R;
    const uint32_t* ev;     // [m * A]
    const uint8_t* ar;      // [m]
    uint32_t m, A;
    const uint32_t* off;    // [nkeys + 1]
    const uint32_t* idx;    // [occurrences]
    // Closing edges: a depth whose pattern edge has every vertex bound by earlier depths takes
    // its candidates from the edges with exactly that vertex tuple -- the edges sorted by a
    // 64-bit key of (arity, vertices), stably, so equal keys stay in edge order. Bit d of
    // `full` marks such a depth (0: no tuple index).
    const uint64_t* tkey;   // [m] sorted keys
    const uint32_t* tidx;   // [m] edges in key order
    uint32_t full;
    // Each distinct key's run [lo, lo + cnt) of tidx, in an open-addressing table (key 0 empty;
    // a zero key is stored as 1, which only widens its run's verification)
    uint64_t* tslot;        // [tmask + 1] keys
    uint2* trun;            // [tmask + 1] (lo, cnt)
    uint32_t tmask;
};

// Device memory of one join: an Arena, or a fixed region (the allocation-free entry point)
// handed out by a bump pointer that each depth rewinds when its subtree is done.
struct BJMem {
    Arena* arena = nullptr;
    uint8_t* base = nullptr;
    size_t cap = 0, used = 0;
    void* get(size_t bytes) {
        if (arena) return arena->alloc(bytes ? bytes : 1);
        const size_t o = (used + 255) & ~(size_t)255;
        if (o + bytes > cap)
            throw Error{HGX_ERR_CAPACITY, "the join needs more than the " + std::to_string(cap >> 20) +
                                              " MiB of its fixed scratch"};
        used = o + bytes;
        return base + o;
    }
    size_t mark() const { return used; }
    void rewind(size_t m) { if (!arena) used = m; }
    // bytes a next frontier may take before the depth splits its partials in two
    size_t room() const { return arena ? ((size_t)1 << 30) : (cap > used + (8u << 20) ? cap - used - (8u << 20) : 0); }
};

struct BJRun {
    hipStream_t st;
    BJView v;
    BJMem* mem;
    uint32_t grid;            // workgroup cap of the item passes
    uint32_t* out;            // matches (device), or null to count
    uint64_t cap = 0;
    uint64_t found = 0;
    uint32_t maxlist = 0;     // the longest candidate list of the index
    uint64_t* acc = nullptr;  // device: matches found, join bytes (written by the depths)
    uint64_t index_bytes = 0; // algorithmic bytes of the index build
    uint64_t alg_bytes = 0;   // of the join: adjacency reads of every item, frontiers, matches
};

// All matches of J.v.R in the state J.v (on the device, vertex ids below nv): J.found counts them
// and, with J.out, the first J.cap are written (n_lhs edge ids each, pattern order), in the order
// of a serial depth-first join over ascending candidates. With `timed`, the index and join times.
void bj_match(BJRun& J, uint32_t nv, float* ms_index, float* ms_join, bool timed);

}  // namespace hgx
