// ctx.hpp - host-side context shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include "cuacd_amd.h"
#include "dheap.hpp"

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
};

namespace cuacd {

void set_error(gpu_ctx_t ctx, const char* fmt, ...);

// Grow-only device staging buffer `slot` of at least `bytes`.
void* stage_buf(gpu_ctx_t ctx, int slot, size_t bytes);

// Read the kernel error word written by the last launches (synchronises) and
// clear it. Applies host-injected bits. Returns the word.
int read_and_clear_err(gpu_ctx_t ctx);

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
