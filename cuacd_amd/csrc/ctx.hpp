// ctx.hpp - host-side context shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include "cuacd_amd.h"
#include "dheap.hpp"

#include <vector>

namespace cuacd {
// Per-kernel timing (HIP events around individual launches on the context's
// stream) for the benchmark's roofline figures. Enabled by the lookahead
// search while it runs; ids name kernel classes.
enum KtId {
    KT_HULL_W1, KT_HULL_W2, KT_HULL_W4, KT_HULL_W8, KT_HULL_W16,  // one-workgroup LDS hulls by size class
    KT_HULL_B4, KT_HULL_B8, KT_HULL_B16, KT_HULL_B32,              // multi-wave hulls, workspace in HBM
    KT_KDOP,
    KT_HULL_R8, KT_HULL_R16, KT_HULL_RB4,  // full-capacity retries of the half-capacity first passes
    KT_HULL_RB8,                           // HBM retries of the 2049-4096-point LDS first pass
    KT_COUNT
};
struct KTimers {
    struct Rec {
        int kid;
        hipEvent_t a, b;
        double bytes;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> free_ev;
    double stats[KT_COUNT][3] = {};  // ms, launches, algorithmic bytes
};
}  // namespace cuacd

struct gpu_mlp;

struct gpu_ctx {
    int device;
    int n_cu;
    hipStream_t stream;
    char* d_pool_mem;        // pool backing memory
    cuacd::DPool* d_pool;    // pool/heap bookkeeping on the device
    size_t pool_bytes;
    int* d_err;              // scratch kernel error word
    hipEvent_t ev_a, ev_b;   // timer events
    int inject_err;          // KERR bits to OR into the next error readback
    // grow-only staging buffers for the host-pointer batch entry points
    void* stage[8];
    size_t stage_cap[8];
    char last_error[512];
    cuacd::KTimers* kt;      // per-kernel timers, null when off
    // side streams (default priority) for independent launches of one call:
    // they fork from `stream` after ev_fork and join back through ev_join
    hipStream_t aux[3];
    hipEvent_t ev_fork;
    hipEvent_t ev_join[3];
    // learned proposal for the lookahead (null: plain search)
    struct gpu_mlp* proposal;
    int proposal_topk;
    void* record;  // level-0 training record of the search (lookahead.hip), null when off
    long long chunk_subtasks;  // chunk hulls of oversized sets, counted since gpu_init
    hipEvent_t ev_tier[2];     // first passes of the eight-wave LDS hull tiers (launch order)
    long long host_syncs;      // host waits on the context's streams, counted since gpu_init
    // grow-only pinned staging of the search's batched readbacks; kept across
    // calls (pinning tens of MB costs milliseconds), freed with the context
    char* la_pinned;
    size_t la_pinned_cap;
};

// A host wait for a stream of the context, counted in host_syncs.
inline hipError_t ctx_stream_sync(gpu_ctx_t ctx, hipStream_t s) {
    ++ctx->host_syncs;
    return hipStreamSynchronize(s);
}

namespace cuacd {

void set_error(gpu_ctx_t ctx, const char* fmt, ...);

// Grow-only device staging buffer `slot` of at least `bytes`.
void* stage_buf(gpu_ctx_t ctx, int slot, size_t bytes);

// Read the kernel error word written by the last launches (synchronises) and
// clear it. Applies host-injected bits. Returns the word.
int read_and_clear_err(gpu_ctx_t ctx);

// Per-kernel timers: kt_begin records the start event (a = nullptr when off),
// kt_end the stop event with the launch's algorithmic bytes, both on stream s
// (the context's stream when null); kt_resolve
// (after the stream has passed them) accumulates the elapsed times.
hipEvent_t kt_begin(gpu_ctx_t ctx, hipStream_t s = nullptr);
void kt_end(gpu_ctx_t ctx, int kid, hipEvent_t a, double bytes, hipStream_t s = nullptr);
int kt_resolve(gpu_ctx_t ctx);

}  // namespace cuacd

#define HIP_TRY(call)                                                                   \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess) {                                                         \
            cuacd::set_error(ctx, "%s failed at %s:%d: %s", #call, __FILE__, __LINE__,  \
                             hipGetErrorString(_e));                                    \
            return (int)_e;                                                             \
        }                                                                               \
    } while (0)

#define REQUIRE(cond, ...)                      \
    do {                                        \
        if (!(cond)) {                          \
            cuacd::set_error(ctx, __VA_ARGS__); \
            return -1;                          \
        }                                       \
    } while (0)
