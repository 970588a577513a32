// dheap.hpp - device heap: one bump pool feeding two lock-free size-class heaps.
//
// Replaces the reference's DevicePool/DeviceHeap (cuda/allocator.cuh:66-362,
// cuda/mm.cu:14-42): a bump pool with an "output" heap and a "scratch" heap,
// whose counters back Context.pool_usage() and Context.heap_stats().
//
// Design for AMD wave64 instead of per-arena spin locks:
//   * size classes are powers of two from 256 B (class c holds blocks of
//     256 << c bytes, header included); a block never changes class, so there
//     is no coalescing and no boundary tags;
//   * each class keeps a Treiber stack whose head packs (ABA tag << 32 | block
//     offset / 256) in one 64-bit word and is updated with a single agent-scope
//     CAS; nothing ever spins on a lock held by another lane (a lane spinning
//     on a lock held by a lane of its own wave would never finish on hardware
//     without independent thread scheduling);
//   * an empty class refills by bumping the pool with one atomic add.
// One lane per wave calls heap_alloc/heap_free. Blocks are 256-byte aligned;
// user data starts 16 bytes in (the header) and is 16-byte aligned.
#pragma once
#include "wave.hpp"
#include "kerr.h"

namespace cuacd {

constexpr int DH_NCLASS = 40;            // 256 B .. 128 TiB classes (only small ones used)
constexpr uint32_t DH_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t DH_MAGIC_LIVE = 0xC0DEA11Cu;
constexpr uint32_t DH_MAGIC_FREE = 0xF4EEB10Cu;
constexpr int DH_HDR = 16;

struct DHeap {
    unsigned long long heads[DH_NCLASS];  // tagged Treiber-stack heads
    unsigned long long outstanding;       // live user bytes
    unsigned long long allocs;            // cumulative alloc calls
    unsigned long long frees;             // cumulative free calls
};

struct DPool {
    char* base;                    // device pointer to the pool memory
    unsigned long long capacity;   // bytes
    unsigned long long bump;       // bytes handed out by the bump pointer (monotonic)
    DHeap heap;                    // output heap (meshes that outlive a kernel)
    DHeap scratch;                 // scratch heap (workspaces freed inside one call)
};

struct DBlockHdr {
    uint32_t magic;
    uint32_t next;      // free-list link (offset / 256) while the block is free
    unsigned long long meta;  // user bytes (bits 0-47) | heap id << 48 | class << 56
};
__device__ __forceinline__ unsigned long long dh_meta(int c, int h, unsigned long long bytes) {
    return (bytes & 0xFFFFFFFFFFFFull) | ((unsigned long long)h << 48) | ((unsigned long long)c << 56);
}

__device__ __forceinline__ int dh_class_of(unsigned long long total) {
    // smallest c with (256 << c) >= total
    if (total <= 256ull) return 0;
    return 64 - __builtin_clzll(total - 1) - 8;
}

// Allocate `bytes` user bytes from heap `h` (0 = output heap, 1 = scratch).
// Returns nullptr and sets KERR_HEAP_OOM in *err on exhaustion.
__device__ inline void* heap_alloc(DPool* pool, int h, unsigned long long bytes, int* err) {
    DHeap* hp = h ? &pool->scratch : &pool->heap;
    const int c = dh_class_of(bytes + DH_HDR);
    if (c >= DH_NCLASS) { err_or(err, KERR_HEAP_OOM); return nullptr; }
    char* base = pool->base;
    char* blk = nullptr;
    unsigned long long old = ld_acquire(&hp->heads[c]);
    while ((uint32_t)old != DH_EMPTY) {
        uint32_t off = (uint32_t)old;
        DBlockHdr* b = (DBlockHdr*)(base + ((unsigned long long)off << 8));
        uint32_t nxt = __hip_atomic_load(&b->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long long neu = (((old >> 32) + 1ull) << 32) | nxt;
        if (cas_acq_rel(&hp->heads[c], old, neu)) { blk = (char*)b; break; }
    }
    if (!blk) {
        unsigned long long sz = 256ull << c;
        unsigned long long o = fetch_add_relaxed(&pool->bump, sz);
        if (o + sz > pool->capacity) { err_or(err, KERR_HEAP_OOM); return nullptr; }
        blk = base + o;
    }
    DBlockHdr* hd = (DBlockHdr*)blk;
    hd->meta = dh_meta(c, h, bytes);
    hd->magic = DH_MAGIC_LIVE;
    hd->next = DH_EMPTY;
    fetch_add_relaxed(&hp->outstanding, bytes);
    fetch_add_relaxed(&hp->allocs, 1ull);
    return blk + DH_HDR;
}

// Return a block obtained from heap_alloc (to the heap it came from). Null is a
// no-op; a block whose header does not carry the live magic sets KERR_HEAP_CORRUPT.
__device__ inline void heap_free(DPool* pool, void* p, int* err) {
    if (!p) return;
    char* blk = (char*)p - DH_HDR;
    DBlockHdr* hd = (DBlockHdr*)blk;
    const unsigned long long meta = hd->meta;
    const int c = (int)(meta >> 56);
    const int hid = (int)((meta >> 48) & 0xFF);
    const unsigned long long bytes = meta & 0xFFFFFFFFFFFFull;
    if (hd->magic != DH_MAGIC_LIVE || c >= DH_NCLASS || hid > 1) {
        err_or(err, KERR_HEAP_CORRUPT);
        return;
    }
    DHeap* hp = hid ? &pool->scratch : &pool->heap;
    hd->magic = DH_MAGIC_FREE;
    const uint32_t off = (uint32_t)((unsigned long long)(blk - pool->base) >> 8);
    unsigned long long old = ld_acquire(&hp->heads[c]);
    for (;;) {
        __hip_atomic_store(&hd->next, (uint32_t)old, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long long neu = (((old >> 32) + 1ull) << 32) | off;
        if (cas_acq_rel(&hp->heads[c], old, neu)) break;
    }
    fetch_add_relaxed(&hp->outstanding, (unsigned long long)0 - bytes);
    fetch_add_relaxed(&hp->frees, 1ull);
}

}  // namespace cuacd
