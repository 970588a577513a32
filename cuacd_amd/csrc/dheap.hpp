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
//   * each class keeps DH_STRIPES Treiber stacks; a head packs (ABA tag << 32 |
//     block offset / 256) in one 64-bit word, sits alone in a 128-byte line
//     and is updated with a single agent-scope CAS. A workgroup pushes to and
//     pops from the stripe of its block index (then tries the others before
//     bumping), so thousands of concurrent workgroups spread their CAS
//     traffic over 32 lines instead of retrying on one; nothing ever spins on
//     a lock held by another lane (a lane spinning on a lock held by a lane of
//     its own wave would never finish without independent thread scheduling);
//   * an empty class refills by bumping the pool with one atomic add;
//   * heap_free_batch returns many blocks at once: a workgroup links its
//     blocks of one (heap, class) into a chain in LDS and splices the chain
//     with one CAS.
// One lane per workgroup calls heap_alloc/heap_free. Blocks are 256-byte
// aligned; user data starts 16 bytes in (the header) and is 16-byte aligned.
#pragma once
#include "wave.hpp"
#include "kerr.h"

namespace cuacd {

constexpr int DH_NCLASS = 40;            // 256 B .. 128 TiB classes (only small ones used)
constexpr int DH_STRIPES = 32;
constexpr uint32_t DH_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t DH_MAGIC_LIVE = 0xC0DEA11Cu;
constexpr uint32_t DH_MAGIC_FREE = 0xF4EEB10Cu;
constexpr int DH_HDR = 16;

struct alignas(128) DHead {
    unsigned long long v;  // tagged Treiber-stack head
};
struct alignas(128) DCount {
    unsigned long long outstanding;  // live user bytes
    unsigned long long allocs;       // cumulative alloc calls
    unsigned long long frees;        // cumulative free calls
};

struct DHeap {
    DHead heads[DH_NCLASS][DH_STRIPES];
    DCount cnt[DH_STRIPES];  // summed by the host
};

struct DPool {
    char* base;                    // device pointer to the pool memory
    unsigned long long capacity;   // bytes
    alignas(128) unsigned long long bump;  // bytes handed out by the bump pointer (monotonic)
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

__device__ __forceinline__ int dh_stripe() { return (int)(blockIdx.x & (DH_STRIPES - 1)); }

// Heads and links are read with relaxed agent-scope atomics: they live in L2,
// the point of coherence, and the link's address depends on the head just
// read. Only a successful pop acquires (one L1 invalidation per block taken,
// not one per stripe looked at).
__device__ __forceinline__ unsigned long long dh_head(DHeap* hp, int c, int s) {
    return __hip_atomic_load(&hp->heads[c][s].v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pop one block of class c from stripe s starting from head value `old`;
// nullptr if the stripe is (or becomes) empty.
__device__ inline char* dh_pop_from(DPool* pool, DHeap* hp, int c, int s, unsigned long long old) {
    char* base = pool->base;
    while ((uint32_t)old != DH_EMPTY) {
        uint32_t off = (uint32_t)old;
        DBlockHdr* b = (DBlockHdr*)(base + ((unsigned long long)off << 8));
        uint32_t nxt = __hip_atomic_load(&b->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long long neu = (((old >> 32) + 1ull) << 32) | nxt;
        if (__hip_atomic_compare_exchange_strong(&hp->heads[c][s].v, &old, neu, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return (char*)b;
    }
    return nullptr;
}

// Push the chain first..last (linked through next, offsets / 256) onto stripe s.
// The CAS releases (the link and the blocks' last contents are visible to the
// next owner's acquiring pop); it acquires nothing, since the pusher only
// stores the old head into the link. On CDNA an agent-scope release writes
// back the XCD's L2 and an acquire invalidates it, so a push pays the first
// and not the second.
__device__ inline void dh_push_chain(DHeap* hp, int c, int s, DBlockHdr* last, uint32_t first_off) {
    unsigned long long old = dh_head(hp, c, s);
    for (;;) {
        __hip_atomic_store(&last->next, (uint32_t)old, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long long neu = (((old >> 32) + 1ull) << 32) | first_off;
        if (__hip_atomic_compare_exchange_strong(&hp->heads[c][s].v, &old, neu, __ATOMIC_RELEASE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            break;
    }
}

// Allocate `bytes` user bytes from heap `h` (0 = output heap, 1 = scratch).
// Returns nullptr and sets KERR_HEAP_OOM in *err on exhaustion.
__device__ inline void* heap_alloc(DPool* pool, int h, unsigned long long bytes, int* err) {
    DHeap* hp = h ? &pool->scratch : &pool->heap;
    const int c = dh_class_of(bytes + DH_HDR);
    if (c >= DH_NCLASS) { err_or(err, KERR_HEAP_OOM); return nullptr; }
    const int s0 = dh_stripe();
    char* blk = dh_pop_from(pool, hp, c, s0, dh_head(hp, c, s0));
    if (!blk) {
        // look at every other stripe at once (independent loads, a register
        // bit mask - no private array), then try the non-empty ones in order
        uint32_t live = 0;
#pragma unroll
        for (int k = 1; k < DH_STRIPES; ++k)
            if ((uint32_t)dh_head(hp, c, (s0 + k) & (DH_STRIPES - 1)) != DH_EMPTY) live |= 1u << k;
        while (!blk && live) {
            const int k = __builtin_ctz(live);
            live &= live - 1;
            const int st = (s0 + k) & (DH_STRIPES - 1);
            blk = dh_pop_from(pool, hp, c, st, dh_head(hp, c, st));
        }
    }
    if (!blk) {
        unsigned long long sz = 256ull << c;
        unsigned long long o = fetch_add_relaxed(&pool->bump, sz);
        if (o + sz > pool->capacity) { err_or(err, KERR_HEAP_OOM); return nullptr; }
        blk = pool->base + o;
    }
    DBlockHdr* hd = (DBlockHdr*)blk;
    hd->meta = dh_meta(c, h, bytes);
    hd->magic = DH_MAGIC_LIVE;
    hd->next = DH_EMPTY;
    fetch_add_relaxed(&hp->cnt[s0].outstanding, bytes);
    fetch_add_relaxed(&hp->cnt[s0].allocs, 1ull);
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
    const int s = dh_stripe();
    dh_push_chain(hp, c, s, hd, off);
    fetch_add_relaxed(&hp->cnt[s].outstanding, (unsigned long long)0 - bytes);
    fetch_add_relaxed(&hp->cnt[s].frees, 1ull);
}

// Free N blocks (nullptr entries are skipped) from one thread with a single
// release: every link is stored against the head read just before, then one
// agent-scope release fence, then one relaxed CAS per block. The fence
// synchronizes with the acquiring pop that reads a head value written by any
// of these CASes (or by a later CAS on the same head), so the links and the
// blocks' contents are visible to the next owner as with a releasing CAS. A
// CAS that finds the head moved (another push or pop, or an earlier block of
// this call with the same class) pushes that block again with dh_push_chain.
// An agent-scope release is a write-back of the XCD's whole L2 on CDNA (and
// the wave waits for it): one per batch instead of one per block.
template <int N>
__device__ inline void heap_free_n(DPool* pool, void* const (&ps)[N], int* err) {
    DBlockHdr* hd[N];
    unsigned long long old[N];
    const int s = dh_stripe();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        hd[i] = nullptr;
        old[i] = 0;
        if (!ps[i]) continue;
        DBlockHdr* h = (DBlockHdr*)((char*)ps[i] - DH_HDR);
        const unsigned long long meta = h->meta;
        const int c = (int)(meta >> 56);
        const int hid = (int)((meta >> 48) & 0xFF);
        if (h->magic != DH_MAGIC_LIVE || c >= DH_NCLASS || hid > 1) {
            err_or(err, KERR_HEAP_CORRUPT);
            continue;
        }
        DHeap* hp = hid ? &pool->scratch : &pool->heap;
        h->magic = DH_MAGIC_FREE;
        old[i] = dh_head(hp, c, s);
        __hip_atomic_store(&h->next, (uint32_t)old[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        hd[i] = h;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if (!hd[i]) continue;
        const unsigned long long meta = hd[i]->meta;
        const int c = (int)(meta >> 56);
        DHeap* hp = ((meta >> 48) & 0xFF) ? &pool->scratch : &pool->heap;
        const uint32_t off = (uint32_t)((unsigned long long)((char*)hd[i] - pool->base) >> 8);
        const unsigned long long neu = (((old[i] >> 32) + 1ull) << 32) | off;
        unsigned long long exp = old[i];
        if (!__hip_atomic_compare_exchange_strong(&hp->heads[c][s].v, &exp, neu, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT))
            dh_push_chain(hp, c, s, hd[i], off);
        fetch_add_relaxed(&hp->cnt[s].outstanding, (unsigned long long)0 - (meta & 0xFFFFFFFFFFFFull));
        fetch_add_relaxed(&hp->cnt[s].frees, 1ull);
    }
}

// Free up to blockDim.x blocks, one per thread (p == nullptr: nothing), with
// one CAS per (heap, class) present in the workgroup. Every thread of the
// workgroup must call it. `lds` needs DH_BATCH_LDS 32-bit words.
constexpr int DH_KEYS = 2 * DH_NCLASS;
constexpr int DH_BATCH_LDS = 6 * DH_KEYS;
__device__ inline void heap_free_batch(DPool* pool, void* p, int* err, uint32_t* lds) {
    uint32_t* first = lds;                 // chain head per key (block offset / 256)
    uint32_t* last = lds + DH_KEYS;        // chain tail per key
    uint32_t* n = lds + 2 * DH_KEYS;       // blocks per key
    unsigned long long* by = (unsigned long long*)(lds + 4 * DH_KEYS);  // 8-byte aligned: bytes per key
    for (int k = threadIdx.x; k < DH_KEYS; k += blockDim.x) {
        first[k] = DH_EMPTY;
        last[k] = DH_EMPTY;
        n[k] = 0;
        by[k] = 0;
    }
    __syncthreads();
    if (p) {
        char* blk = (char*)p - DH_HDR;
        DBlockHdr* hd = (DBlockHdr*)blk;
        const unsigned long long meta = hd->meta;
        const int c = (int)(meta >> 56);
        const int hid = (int)((meta >> 48) & 0xFF);
        if (hd->magic != DH_MAGIC_LIVE || c >= DH_NCLASS || hid > 1) {
            err_or(err, KERR_HEAP_CORRUPT);
        } else {
            const int key = hid * DH_NCLASS + c;
            hd->magic = DH_MAGIC_FREE;
            const uint32_t off = (uint32_t)((unsigned long long)(blk - pool->base) >> 8);
            const uint32_t prev = atomicExch(&first[key], off);
            __hip_atomic_store(&hd->next, prev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (prev == DH_EMPTY) last[key] = off;
            atomicAdd(&n[key], 1u);
            atomicAdd(&by[key], meta & 0xFFFFFFFFFFFFull);
        }
    }
    __threadfence();  // the links are visible device-wide before a chain is published
    __syncthreads();
    const int s = dh_stripe();
    for (int k = threadIdx.x; k < DH_KEYS; k += blockDim.x) {
        if (first[k] == DH_EMPTY) continue;
        DHeap* hp = k >= DH_NCLASS ? &pool->scratch : &pool->heap;
        DBlockHdr* tail = (DBlockHdr*)(pool->base + ((unsigned long long)last[k] << 8));
        dh_push_chain(hp, k % DH_NCLASS, s, tail, first[k]);
        fetch_add_relaxed(&hp->cnt[s].outstanding, (unsigned long long)0 - by[k]);
        fetch_add_relaxed(&hp->cnt[s].frees, (unsigned long long)n[k]);
    }
}

}  // namespace cuacd
