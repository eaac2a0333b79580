// hgx_runtime.hpp -- host-side device-memory management shared by the batch and served paths.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

struct hgx_ctx;

namespace hgx {

struct Chunk {
    void* p;
    size_t size;
};

// Large device chunks, allocated on first use and then recycled: the steady state of a serving
// loop allocates nothing from the HIP runtime. Bounded by the context's HBM budget.
class ChunkPool {
public:
    ChunkPool(uint64_t budget, size_t chunk_bytes) : budget_(budget), chunk_bytes_(chunk_bytes) {}
    ~ChunkPool();
    Chunk get(size_t min_bytes);
    void put(const Chunk& c);
    uint64_t budget() const { return budget_; }
    uint64_t allocated() const { return total_; }

private:
    std::mutex mu_;
    std::vector<Chunk> free_;
    uint64_t budget_;
    size_t chunk_bytes_;
    uint64_t total_ = 0;
};

// Bump allocation over pool chunks; everything returns to the pool at once.
class Arena {
public:
    explicit Arena(ChunkPool* pool) : pool_(pool) {}
    ~Arena();
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    void* alloc(size_t bytes);
    void release();
    uint64_t bytes() const { return bytes_; }

private:
    ChunkPool* pool_;
    std::vector<Chunk> chunks_;
    size_t used_ = 0;
    uint64_t bytes_ = 0;
};

struct ServeState;
void serve_destroy(hgx_ctx* ctx);

}  // namespace hgx
