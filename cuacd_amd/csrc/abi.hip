// abi.hip - context, device pool and plumbing of the C ABI.
//
// Mirrors the reference's context lifecycle (csrc/heap.c:29-294): init picks
// the device, reserves the pool and initialises the heaps; pool_usage and
// heap_stats read the device bookkeeping back. There is no module loading:
// the kernels are linked into this library.
#include <stdio.h>
#include <stdlib.h>
#include "ctx.hpp"
#include "kerr.h"

using namespace cuacd;

static char g_init_error[512] = "";

void cuacd::set_error(gpu_ctx_t ctx, const char* fmt, ...) {
    char* dst = ctx ? ctx->last_error : g_init_error;
    size_t cap = ctx ? sizeof(ctx->last_error) : sizeof(g_init_error);
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(dst, cap, fmt, ap);
    va_end(ap);
}

void* cuacd::stage_buf(gpu_ctx_t ctx, int slot, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (ctx->stage_cap[slot] >= bytes) return ctx->stage[slot];
    if (ctx->stage[slot]) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->stage[slot]);
        ctx->stage[slot] = nullptr;
        ctx->stage_cap[slot] = 0;
    }
    size_t cap = bytes + bytes / 4;
    if (hipMalloc(&ctx->stage[slot], cap) != hipSuccess) {
        ctx->stage[slot] = nullptr;
        return nullptr;
    }
    ctx->stage_cap[slot] = cap;
    return ctx->stage[slot];
}

hipEvent_t cuacd::kt_begin(gpu_ctx_t ctx, hipStream_t s) {
    KTimers* k = ctx->kt;
    if (!k) return nullptr;
    hipEvent_t e = nullptr;
    if (!k->free_ev.empty()) {
        e = k->free_ev.back();
        k->free_ev.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        return nullptr;
    }
    (void)hipEventRecord(e, s ? s : ctx->stream);
    return e;
}

void cuacd::kt_end(gpu_ctx_t ctx, int kid, hipEvent_t a, double bytes, hipStream_t s) {
    KTimers* k = ctx->kt;
    if (!k || !a) return;
    hipEvent_t b = nullptr;
    if (!k->free_ev.empty()) {
        b = k->free_ev.back();
        k->free_ev.pop_back();
    } else if (hipEventCreate(&b) != hipSuccess) {
        k->free_ev.push_back(a);
        return;
    }
    (void)hipEventRecord(b, s ? s : ctx->stream);
    k->pending.push_back({kid, a, b, bytes});
}

int cuacd::kt_resolve(gpu_ctx_t ctx) {
    KTimers* k = ctx->kt;
    if (!k) return 0;
    for (auto& r : k->pending) {
        float ms = 0.f;
        HIP_TRY(hipEventSynchronize(r.b));
        HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
        k->stats[r.kid][0] += ms;
        k->stats[r.kid][1] += 1;
        k->stats[r.kid][2] += r.bytes;
        k->free_ev.push_back(r.a);
        k->free_ev.push_back(r.b);
    }
    k->pending.clear();
    return 0;
}

int cuacd::read_and_clear_err(gpu_ctx_t ctx) {
    int h = 0;
    if (hipMemcpyAsync(&h, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
        return -1;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1;
    if (h) (void)hipMemsetAsync(ctx->d_err, 0, sizeof(int), ctx->stream);
    if (ctx->inject_err) {
        h |= ctx->inject_err;
        ctx->inject_err = 0;
    }
    return h;
}

__global__ void dpool_init_kernel(DPool* p, char* base, unsigned long long cap) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        p->base = base;
        p->capacity = cap;
        p->bump = 0;
    }
    for (int i = threadIdx.x; i < DH_NCLASS * DH_STRIPES; i += blockDim.x) {
        p->heap.heads[i / DH_STRIPES][i % DH_STRIPES].v = (unsigned long long)DH_EMPTY;
        p->scratch.heads[i / DH_STRIPES][i % DH_STRIPES].v = (unsigned long long)DH_EMPTY;
    }
    for (int i = threadIdx.x; i < DH_STRIPES; i += blockDim.x) {
        p->heap.cnt[i] = DCount{0, 0, 0};
        p->scratch.cnt[i] = DCount{0, 0, 0};
    }
}

extern "C" {

void gpu_set_fatbin(const void* data, size_t len) {
    (void)data;
    (void)len;
}

const char* gpu_last_error(gpu_ctx_t ctx) { return ctx ? ctx->last_error : g_init_error; }

void gpu_destroy(gpu_ctx_t ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < 8; ++i)
        if (ctx->stage[i]) (void)hipFree(ctx->stage[i]);
    if (ctx->d_pool_mem) (void)hipFree(ctx->d_pool_mem);
    if (ctx->d_pool) (void)hipFree(ctx->d_pool);
    if (ctx->d_err) (void)hipFree(ctx->d_err);
    if (ctx->ev_a) (void)hipEventDestroy(ctx->ev_a);
    if (ctx->ev_b) (void)hipEventDestroy(ctx->ev_b);
    for (int i = 0; i < 3; ++i) {
        if (ctx->aux[i]) (void)hipStreamSynchronize(ctx->aux[i]);
        if (ctx->aux[i]) (void)hipStreamDestroy(ctx->aux[i]);
        if (ctx->ev_join[i]) (void)hipEventDestroy(ctx->ev_join[i]);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (int i = 0; i < 2; ++i)
        if (ctx->ev_tier[i]) (void)hipEventDestroy(ctx->ev_tier[i]);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->la_pinned) (void)hipHostFree(ctx->la_pinned);
    free(ctx);
}

int gpu_init(gpu_ctx_t* out, int device, size_t pool_bytes) {
    *out = nullptr;
    gpu_ctx_t ctx = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        set_error(nullptr, "no HIP device available");
        return -1;
    }
    if (device < 0) {
        if (hipGetDevice(&device) != hipSuccess) device = 0;
    }
    if (device >= ndev) {
        set_error(nullptr, "device %d out of range (%d devices)", device, ndev);
        return -1;
    }
    ctx = (gpu_ctx_t)calloc(1, sizeof(struct gpu_ctx));
    if (!ctx) {
        set_error(nullptr, "out of host memory");
        return -1;
    }
    ctx->device = device;
    *out = ctx;
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    ctx->n_cu = prop.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    for (int i = 0; i < 3; ++i) {
        HIP_TRY(hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ctx->ev_join[i], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    for (int i = 0; i < 2; ++i) HIP_TRY(hipEventCreateWithFlags(&ctx->ev_tier[i], hipEventDisableTiming));
    HIP_TRY(hipEventCreate(&ctx->ev_a));
    HIP_TRY(hipEventCreate(&ctx->ev_b));
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    if (pool_bytes == 0) {
        pool_bytes = (size_t)((double)free_b * 0.8);
        const size_t cap = (size_t)64 << 30;
        if (pool_bytes > cap) pool_bytes = cap;
        if (pool_bytes < ((size_t)64 << 20)) pool_bytes = (size_t)64 << 20;
    }
    pool_bytes = (pool_bytes + 255) & ~(size_t)255;
    ctx->pool_bytes = pool_bytes;
    HIP_TRY(hipMalloc(&ctx->d_pool_mem, pool_bytes));
    HIP_TRY(hipMalloc(&ctx->d_pool, sizeof(DPool)));
    HIP_TRY(hipMalloc(&ctx->d_err, sizeof(int)));
    HIP_TRY(hipMemsetAsync(ctx->d_err, 0, sizeof(int), ctx->stream));
    dpool_init_kernel<<<1, 64, 0, ctx->stream>>>(ctx->d_pool, ctx->d_pool_mem,
                                                 (unsigned long long)pool_bytes);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int gpu_heap_compact(gpu_ctx_t ctx) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

size_t gpu_pool_usage(gpu_ctx_t ctx) {
    unsigned long long bump = 0;
    if (hipMemcpyAsync(&bump, &ctx->d_pool->bump, sizeof(bump), hipMemcpyDeviceToHost, ctx->stream) !=
        hipSuccess)
        return 0;
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return 0;
    return (size_t)(bump < ctx->pool_bytes ? bump : ctx->pool_bytes);
}

int gpu_heap_stats(gpu_ctx_t ctx, unsigned long long out[6]) {
    DCount c[2][DH_STRIPES];
    HIP_TRY(hipMemcpyAsync(c[0], ctx->d_pool->heap.cnt, sizeof(c[0]), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(c[1], ctx->d_pool->scratch.cnt, sizeof(c[1]), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int h = 0; h < 2; ++h) {
        unsigned long long o = 0, a = 0, f = 0;
        for (int s = 0; s < DH_STRIPES; ++s) { o += c[h][s].outstanding; a += c[h][s].allocs; f += c[h][s].frees; }
        out[3 * h] = o;
        out[3 * h + 1] = a;
        out[3 * h + 2] = f;
    }
    return 0;
}

void gpu_result_free(struct gpu_result* r) {
    if (!r) return;
    if (r->parts) {
        for (int i = 0; i < r->nparts; ++i) {
            free(r->parts[i].verts);
            free(r->parts[i].tris);
            free(r->parts[i].hull_verts);
            free(r->parts[i].hull_tris);
        }
        free(r->parts);
    }
    r->parts = nullptr;
    r->nparts = 0;
}

int gpu_preprocess(gpu_ctx_t ctx, const float*, int, const int*, int, int, float** out_verts, int* out_nv,
                   int** out_tris, int* out_nt) {
    if (out_verts) *out_verts = nullptr;
    if (out_tris) *out_tris = nullptr;
    if (out_nv) *out_nv = 0;
    if (out_nt) *out_nt = 0;
    set_error(ctx, "preprocess (PaMO remesh) is not part of this library");
    return -1;
}

int gpu_dev_alloc(gpu_ctx_t ctx, size_t bytes, void** out_ptr) {
    *out_ptr = nullptr;
    HIP_TRY(hipMalloc(out_ptr, bytes ? bytes : 16));
    return 0;
}
int gpu_dev_free(gpu_ctx_t ctx, void* ptr) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipFree(ptr));
    return 0;
}
int gpu_memcpy_h2d(gpu_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}
int gpu_memcpy_d2h(gpu_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}
int gpu_memset_dev(gpu_ctx_t ctx, void* dst, int value, size_t bytes) {
    HIP_TRY(hipMemsetAsync(dst, value, bytes, ctx->stream));
    return 0;
}
int gpu_sync(gpu_ctx_t ctx) {
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}
int gpu_timer_start(gpu_ctx_t ctx) {
    HIP_TRY(hipEventRecord(ctx->ev_a, ctx->stream));
    return 0;
}
int gpu_timer_stop(gpu_ctx_t ctx, float* ms) {
    HIP_TRY(hipEventRecord(ctx->ev_b, ctx->stream));
    HIP_TRY(hipEventSynchronize(ctx->ev_b));
    HIP_TRY(hipEventElapsedTime(ms, ctx->ev_a, ctx->ev_b));
    return 0;
}
// Marketing name of the board: the runtime's device name, else the name the
// amdgpu driver publishes for the PCI device (boards with a FRU record), else
// the PCI device id.
static void board_name(const hipDeviceProp_t& prop, int device, char* out, size_t len) {
    if (prop.name[0]) {
        snprintf(out, len, "%s", prop.name);
        return;
    }
    char nm[256] = {0};
    if (hipDeviceGetName(nm, (int)sizeof(nm), device) == hipSuccess && nm[0]) {
        snprintf(out, len, "%s", nm);
        return;
    }
    char path[128];
    snprintf(path, sizeof(path), "/sys/bus/pci/devices/%04x:%02x:%02x.0/product_name", prop.pciDomainID,
             prop.pciBusID, prop.pciDeviceID);
    if (FILE* f = fopen(path, "r")) {
        size_t k = fread(nm, 1, sizeof(nm) - 1, f);
        fclose(f);
        while (k > 0 && (nm[k - 1] == '\n' || nm[k - 1] == ' ')) nm[--k] = 0;
        if (k > 0) {
            snprintf(out, len, "%s", nm);
            return;
        }
    }
    snprintf(out, len, "AMD GPU at PCI %04x:%02x:%02x", prop.pciDomainID, prop.pciBusID, prop.pciDeviceID);
}

int gpu_device_info(gpu_ctx_t ctx, char* name, int name_len, int* n_cu, size_t* hbm_bytes) {
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, ctx->device));
    if (name && name_len > 0) {
        char board[200];
        board_name(prop, ctx->device, board, sizeof(board));
        snprintf(name, (size_t)name_len, "%s (%s)", board, prop.gcnArchName);
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return 0;
}
void gpu_debug_inject_error(gpu_ctx_t ctx, int kerr_bits) {
    if (ctx) ctx->inject_err |= kerr_bits;
}

}  // extern "C"
