// hausdorff_kernels.hip - batched directed Hausdorff tasks.
//
// A task = (source mesh, target mesh, sampler): build the target's implicit
// BVH in the task's HBM workspace, lay out the source's samples by area, and
// give each lane samples to answer; the task's value is the maximum squared
// nearest distance. One 256-thread workgroup per task.
//
// Entry points: gpu_test_hausdorff (reference csrc/test.c:434-486,
// cuda/test_hausdorff.cu:8-39) and the batch API with plans (no reference
// counterpart) used for concavity scoring of many mesh/hull pairs.
#include <stdlib.h>
#include <algorithm>
#include <vector>
#include "ctx.hpp"
#include "hausdorff.hpp"

using namespace cuacd;

// A plan's launches (preparation + query) are captured once as a HIP graph;
// each run replays it. The kernels clear their own outputs (no memset nodes).
struct hd_plan {
    HdTask* d_tasks;
    char* d_ws;
    int n_tasks;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace {

// Block-wide exclusive scan of a per-element int value (computed by f(i)),
// written to out[0..n); returns the total. Contiguous chunk per thread.
template <class F>
__device__ int block_excl_scan_int(int n, int* out, HdLDS* s, F f) {
    const int tid = threadIdx.x, wv = tid >> 6;
    const int per = (n + HD_T - 1) / HD_T;
    const int b = tid * per, e = min(n, b + per);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += f(i);
    const int incl = wave_incl_scan(sum);
    if (lane_id() == 63) s->ired[wv] = incl;
    __syncthreads();
    int base = 0, tot = 0;
    for (int q = 0; q < HD_T / 64; ++q) {
        if (q < wv) base += s->ired[q];
        tot += s->ired[q];
    }
    int run = base + incl - sum;
    for (int i = b; i < e; ++i) {
        out[i] = run;
        run += f(i);
    }
    __syncthreads();
    return tot;
}

// Inclusive prefix of per-element float areas in FP64 (stored as float).
template <class F>
__device__ float block_incl_scan_area(int n, float* out, HdLDS* s, F f) {
    const int tid = threadIdx.x, wv = tid >> 6;
    const int per = (n + HD_T - 1) / HD_T;
    const int b = tid * per, e = min(n, b + per);
    double sum = 0.0;
    for (int i = b; i < e; ++i) sum += (double)f(i);
    // wave inclusive scan in double
    double incl = sum;
    const int l = lane_id();
    for (int o = 1; o < 64; o <<= 1) {
        double w = __shfl_up(incl, o);
        if (l >= o) incl += w;
    }
    if (l == 63) s->dred[wv] = incl;
    __syncthreads();
    double base = 0.0, tot = 0.0;
    for (int q = 0; q < HD_T / 64; ++q) {
        if (q < wv) base += s->dred[q];
        tot += s->dred[q];
    }
    double run = base + incl - sum;
    for (int i = b; i < e; ++i) {
        run += (double)f(i);
        out[i] = (float)run;
    }
    __syncthreads();
    return (float)tot;
}

// Sample point from two hashes (evaluated as written, no FMA contraction, so
// that a serial C evaluation reproduces the point bit for bit).
__device__ __forceinline__ V3 bary(V3 a, V3 b, V3 c, uint32_t ha, uint32_t hb) {
#pragma clang fp contract(off)
    const float ra = (float)(ha & 0xFFFFFFu) / (float)0xFFFFFFu;
    const float rb = (float)(hb & 0xFFFFFFu) / (float)0xFFFFFFu;
    const float sqa = sqrtf(ra);
    const float wa = 1.0f - sqa, wb = sqa * (1.0f - rb), wc = sqa * rb;
    return {wa * a.x + wb * b.x + wc * c.x, wa * a.y + wb * b.y + wc * c.y, wa * a.z + wb * b.z + wc * c.z};
}

// Preparation, one workgroup per task: the target's BVH, the source's
// sample layout (mode 0: per-triangle counts and their exclusive prefix;
// mode 1: the cumulative area), the sample count; clears the result.
__global__ __launch_bounds__(HD_T) void hd_prep_kernel(HdTask* tasks, int n_tasks, const int* dcount,
                                                       char* __restrict__ ws_base) {
    __shared__ HdLDS s;
    // tasks[0, n_tasks), or the first min(*dcount, n_tasks) when the count is on the device
    if ((int)blockIdx.x >= (dcount ? min(*dcount, n_tasks) : n_tasks)) return;
    HdTask& tk = tasks[blockIdx.x];
    const int tid = threadIdx.x, wv = tid >> 6;
    if (tk.snt <= 0 || tk.tnt <= 0) {
        if (tid == 0) {
            *tk.out_d2 = 0.0f;
            tk.ns = 0;
            if (tk.out_ns) *tk.out_ns = 0;
        }
        return;
    }
    const MeshView src{tk.sv, tk.st, tk.snv, tk.snt};
    const MeshView tgt{tk.tv, tk.tt, tk.tnv, tk.tnt};
    char* ws = ws_base + tk.ws_off;
    (void)hd_build(tgt, ws, &s);
    int nbox = 0;
    hd_bvh_levels(tgt.nt, nullptr, nullptr, &nbox);
    char* ws2 = ws + (unsigned long long)((tgt.nt + 3) / 4 * 4) * 48ull + (unsigned long long)nbox * 32ull;
    auto area_of = [&](int k) {
        V3 a, b, c;
        src.tri(k, a, b, c);
        return tri_area(a, b, c);
    };
    int ns = 0;
    float area = 0.0f;
    if (tk.mode == 0) {
        // reference sampler: total area (FP64 sum of the fp32 triangle areas,
        // rounded once, so the counts do not depend on the summation order),
        // per-triangle counts, exclusive prefix
        double ta = 0.0;
        for (int k = tid; k < src.nt; k += HD_T) ta += (double)area_of(k);
        ta = wave_sum_d(ta);
        if (lane_id() == 0) s.dred[wv] = ta;
        __syncthreads();
        double tsum = 0.0;
        for (int q = 0; q < HD_T / 64; ++q) tsum += s.dred[q];
        __syncthreads();
        area = (float)tsum;
        int res = (int)((float)HD_RESOLUTION * area);
        if (res < HD_MIN_SAMPLES) res = HD_MIN_SAMPLES;
        const int nt = src.nt;
        const int step = nt > res ? nt / res : 2;
        const float dens = area > 1e-20f ? (float)res / area : 0.0f;
        auto count_of = [&](int k) {
            const int ac = area > 1e-20f ? (int)(dens * area_of(k)) : 0;
            return ac > 0 ? ac : ((k % step == 0) ? 1 : 0);
        };
        ns = block_excl_scan_int(nt, (int*)ws2, &s, count_of);
    } else {
        area = block_incl_scan_area(src.nt, (float*)ws2, &s, area_of);
        ns = tk.K;
    }
    if (tid == 0) {
        *tk.out_d2 = 0.0f;
        tk.ns = ns;
        tk.area = area;
        if (tk.out_ns) *tk.out_ns = ns;
    }
}

// Queries: `chunks` workgroups per task, each answering a contiguous share
// of the task's samples; the maximum is merged with an atomic max on the
// result's bits (non-negative floats order as unsigned integers).
__global__ __launch_bounds__(HD_T) void hd_query_kernel(const HdTask* __restrict__ tasks, int n_tasks,
                                                        const int* dcount, int chunks, char* __restrict__ ws_base) {
    __shared__ int qoff[HD_MAX_LEVELS], qcnt[HD_MAX_LEVELS];
    __shared__ float qred[HD_T / 64];
    // XCD-aware order: the dispatcher hands consecutive workgroups to the 8
    // XCDs in turn, so slot L (task L / chunks, chunk L % chunks) is renumbered
    // to give each XCD a contiguous range of slots - the chunks of a task (and
    // neighbouring tasks) share one XCD's L2 instead of fetching the task's
    // BVH and triangles into several. A bijection of [0, nb), nb the slots of
    // the tasks present (all of the grid unless the count is on the device).
    const unsigned nb = dcount ? (unsigned)min(*dcount, n_tasks) * (unsigned)chunks : gridDim.x, b = blockIdx.x;
    if (b >= nb) return;
    const unsigned q8 = nb / 8u, r8 = nb % 8u, x8 = b % 8u, j8 = b / 8u;
    const unsigned slot = (x8 < r8 ? x8 * (q8 + 1u) : r8 * (q8 + 1u) + (x8 - r8) * q8) + j8;
    const int ti = (int)(slot / (unsigned)chunks), ch = (int)(slot % (unsigned)chunks);
    if (ti >= n_tasks) return;
    const HdTask tk = tasks[ti];
    const int tid = threadIdx.x, wv = tid >> 6;
    const int ns = tk.ns;
    if (ns <= 0 || tk.snt <= 0 || tk.tnt <= 0) return;
    const int per = (ns + chunks - 1) / chunks;
    const int i0 = ch * per, i1 = min(ns, i0 + per);
    if (i0 >= i1) return;
    const MeshView src{tk.sv, tk.st, tk.snv, tk.snt};
    char* ws = ws_base + tk.ws_off;
    HdBvh bvh;
    {
        int lo[HD_MAX_LEVELS], lc[HD_MAX_LEVELS], nbox = 0;
        bvh.top = hd_bvh_levels(tk.tnt, lo, lc, &nbox);
        if (tid == 0)
            for (int L = 0; L <= bvh.top; ++L) { qoff[L] = lo[L]; qcnt[L] = lc[L]; }
        bvh.nt = tk.tnt;
        bvh.off = (asptr<AS_LDS, int>)qoff;
        bvh.cnt = (asptr<AS_LDS, int>)qcnt;
        const int ntp = (tk.tnt + 3) / 4 * 4;
        bvh.tri = (f4v*)ws;
        bvh.box = (f4v*)(ws + (unsigned long long)ntp * 48ull);
        __syncthreads();
    }
    int nbox = 0;
    hd_bvh_levels(tk.tnt, nullptr, nullptr, &nbox);
    char* ws2 = ws + (unsigned long long)((tk.tnt + 3) / 4 * 4) * 48ull + (unsigned long long)nbox * 32ull;
    float local = 0.0f;
    if (tk.mode == 0) {
        const int* pre = (const int*)ws2;
        const int nt = src.nt;
        for (int i = i0 + tid; i < i1; i += HD_T) {
            int lo = 0, hi = nt - 1;  // last t with pre[t] <= i
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pre[mid] <= i) lo = mid;
                else hi = mid - 1;
            }
            const int t = lo, k = i - pre[t];
            V3 a, b, c;
            src.tri(t, a, b, c);
            const uint32_t ha = wang_hash((uint32_t)t * 65537u + (uint32_t)k);
            const uint32_t hb = wang_hash(ha + 0x9e3779b9u);
            const V3 p = bary(a, b, c, ha, hb);
            local = fmaxf(local, hd_nearest_d2(bvh, p));
        }
    } else {
        const float* cum = (const float*)ws2;
        const float scale = tk.area / (float)tk.K;
        for (int i = i0 + tid; i < i1; i += HD_T) {
            const uint32_t h1 = wang_hash((uint32_t)i * 0x9E3779B1u + 0x632BE5ABu);
            const uint32_t h2 = wang_hash(h1 + 0x9e3779b9u);
            const uint32_t h3 = wang_hash(h2 + 0x9e3779b9u);
            const float u = ((float)i + (float)(h1 & 0xFFFFFFu) * (1.0f / 16777216.0f)) * scale;
            int lo = 0, hi = src.nt - 1;  // first t with cum[t] > u
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cum[mid] > u) hi = mid;
                else lo = mid + 1;
            }
            V3 a, b, c;
            src.tri(lo, a, b, c);
            const V3 p = bary(a, b, c, h2, h3);
            if (tk.out_samples) {
                tk.out_samples[3 * i] = p.x;
                tk.out_samples[3 * i + 1] = p.y;
                tk.out_samples[3 * i + 2] = p.z;
            }
            local = fmaxf(local, hd_nearest_d2(bvh, p));
        }
    }
    local = wave_max_f(local);
    if (lane_id() == 0) qred[wv] = local;
    __syncthreads();
    if (tid == 0) {
        float m = 0.0f;
        for (int q = 0; q < HD_T / 64; ++q) m = fmaxf(m, qred[q]);
        atomicMax((unsigned int*)tk.out_d2, __float_as_uint(m));
    }
}

int launch_tasks(gpu_ctx_t ctx, const HdTask* d_tasks, int n, char* d_ws) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(hd_prep_kernel, dim3(n), dim3(HD_T), 0, ctx->stream, (HdTask*)d_tasks, n, (const int*)nullptr,
                       d_ws);
    HIP_TRY(hipGetLastError());
    // about 4096 query workgroups per launch, at least one per task
    int chunks = (4096 + n - 1) / n;
    if (chunks > 64) chunks = 64;
    if (chunks < 1) chunks = 1;
    hipLaunchKernelGGL(hd_query_kernel, dim3((unsigned)(n * chunks)), dim3(HD_T), 0, ctx->stream, d_tasks, n,
                       (const int*)nullptr, chunks,
                       d_ws);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

// Internal entry for other kernels' hosts (lookahead): run a device task table.
int hd_launch_tasks(gpu_ctx_t ctx, const HdTask* d_tasks, int n, char* d_ws) {
    return launch_tasks(ctx, d_tasks, n, d_ws);
}

// The same with the task count on the device: at most cap tasks, *dcount of
// them present; `chunks` query workgroups per task.
int hd_launch_tasks_dev(gpu_ctx_t ctx, const HdTask* d_tasks, int cap, const int* dcount, int chunks, char* d_ws) {
    if (cap <= 0) return 0;
    chunks = std::max(1, std::min(64, chunks));
    hipLaunchKernelGGL(hd_prep_kernel, dim3(cap), dim3(HD_T), 0, ctx->stream, (HdTask*)d_tasks, cap, dcount, d_ws);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(hd_query_kernel, dim3((unsigned)(cap * chunks)), dim3(HD_T), 0, ctx->stream, d_tasks, cap,
                       dcount, chunks, d_ws);
    HIP_TRY(hipGetLastError());
    return 0;
}

extern "C" {

int gpu_test_hausdorff(gpu_ctx_t ctx, const float* hv, int hnv, const int* ht, int hnt, const float* mv, int mnv,
                       const int* mt, int mnt, float* out) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(hnv >= 0 && hnt >= 0 && mnv >= 0 && mnt >= 0, "invalid sizes");
    const size_t bhv = 12ull * hnv, bht = 12ull * hnt, bmv = 12ull * mnv, bmt = 12ull * mnt;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const unsigned long long w0 = hd_ws_bytes(mnt, hnt), w1 = hd_ws_bytes(hnt, mnt);
    const size_t total = al(bhv) + al(bht) + al(bmv) + al(bmt) + al(2 * sizeof(HdTask)) + 256 + al(w0) + al(w1);
    char* d = (char*)stage_buf(ctx, 0, total);
    REQUIRE(d, "staging allocation failed");
    float* d_hv = (float*)d;
    int* d_ht = (int*)(d + al(bhv));
    float* d_mv = (float*)(d + al(bhv) + al(bht));
    int* d_mt = (int*)(d + al(bhv) + al(bht) + al(bmv));
    HdTask* d_tasks = (HdTask*)(d + al(bhv) + al(bht) + al(bmv) + al(bmt));
    float* d_out = (float*)((char*)d_tasks + al(2 * sizeof(HdTask)));
    char* d_ws = (char*)d_out + 256;
    HIP_TRY(hipMemcpyAsync(d_hv, hv, bhv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_ht, ht, bht, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_mv, mv, bmv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_mt, mt, bmt, hipMemcpyHostToDevice, ctx->stream));
    HdTask t[2];
    memset(t, 0, sizeof(t));
    // hull samples -> mesh, mesh samples -> hull (reference sampler both ways)
    t[0] = HdTask{d_hv, d_ht, d_mv, d_mt, hnv, hnt, mnv, mnt, 0, 0, 0, d_out, nullptr, nullptr};
    t[1] = HdTask{d_mv, d_mt, d_hv, d_ht, mnv, mnt, hnv, hnt, al(w0), 0, 0, d_out + 1, nullptr, nullptr};
    HIP_TRY(hipMemcpyAsync(d_tasks, t, sizeof(t), hipMemcpyHostToDevice, ctx->stream));
    int rc = launch_tasks(ctx, d_tasks, 2, d_ws);
    if (rc) return rc;
    float h[2];
    HIP_TRY(hipMemcpyAsync(h, d_out, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
    *out = sqrtf(fmaxf(h[0], h[1]));
    return 0;
}

// Plan: device task table + workspaces for n_pairs pairs (A_i, B_i) whose data
// already lives on the device. Two tasks per pair: A samples -> B, B samples -> A.
static int plan_create(gpu_ctx_t ctx, const float* d_av, const int* d_at, const float* d_bv, const int* d_bt,
                       const int* h_avoff, const int* h_atoff, const int* h_bvoff, const int* h_btoff, int n_pairs,
                       int mode, int n_samples, float* d_out_ab, float* d_out_ba, float* d_samples_a,
                       float* d_samples_b, bool capture, hd_plan_t* out_plan) {
    *out_plan = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(n_pairs >= 0 && (mode == 0 || mode == 1), "invalid hausdorff plan arguments");
    REQUIRE(mode == 0 || n_samples > 0, "stratified sampler needs n_samples > 0");
    std::vector<HdTask> t(2 * (size_t)n_pairs);
    unsigned long long ws = 0;
    for (int i = 0; i < n_pairs; ++i) {
        const int anv = h_avoff[i + 1] - h_avoff[i], ant = h_atoff[i + 1] - h_atoff[i];
        const int bnv = h_bvoff[i + 1] - h_bvoff[i], bnt = h_btoff[i + 1] - h_btoff[i];
        const float* av = d_av + 3ll * h_avoff[i];
        const int* at = d_at + 3ll * h_atoff[i];
        const float* bv = d_bv + 3ll * h_bvoff[i];
        const int* bt = d_bt + 3ll * h_btoff[i];
        HdTask& x = t[2 * i];
        x = HdTask{av, at, bv, bt, anv, ant, bnv, bnt, ws, mode, n_samples, d_out_ab + i, nullptr,
                   d_samples_a ? d_samples_a + 3ll * n_samples * i : nullptr};
        ws += (hd_ws_bytes(bnt, ant) + 255ull) & ~255ull;
        HdTask& y = t[2 * i + 1];
        y = HdTask{bv, bt, av, at, bnv, bnt, anv, ant, ws, mode, n_samples, d_out_ba + i, nullptr,
                   d_samples_b ? d_samples_b + 3ll * n_samples * i : nullptr};
        ws += (hd_ws_bytes(ant, bnt) + 255ull) & ~255ull;
    }
    hd_plan_t p = (hd_plan_t)calloc(1, sizeof(struct hd_plan));
    REQUIRE(p, "out of host memory");
    p->n_tasks = 2 * n_pairs;
    if (hipMalloc(&p->d_tasks, sizeof(HdTask) * t.size() + 16) != hipSuccess ||
        hipMalloc(&p->d_ws, ws + 256) != hipSuccess) {
        if (p->d_tasks) (void)hipFree(p->d_tasks);
        free(p);
        set_error(ctx, "hausdorff plan allocation failed (%llu bytes)", ws);
        return -1;
    }
    HIP_TRY(hipMemcpyAsync(p->d_tasks, t.data(), sizeof(HdTask) * t.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
    if (capture && p->n_tasks > 0) {
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        const int rc = launch_tasks(ctx, p->d_tasks, p->n_tasks, p->d_ws);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        hipError_t ei = hipErrorUnknown;
        if (rc == 0 && e == hipSuccess) ei = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
        if (rc != 0 || e != hipSuccess || ei != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            (void)hipFree(p->d_tasks);
            (void)hipFree(p->d_ws);
            free(p);
            set_error(ctx, "hausdorff plan graph capture failed (%s)",
                      hipGetErrorString(e != hipSuccess ? e : ei));
            return rc ? rc : -1;
        }
        p->graph = g;
    }
    *out_plan = p;
    return 0;
}

int gpu_hausdorff_plan_create(gpu_ctx_t ctx, const float* d_av, const int* d_at, const float* d_bv,
                              const int* d_bt, const int* h_avoff, const int* h_atoff, const int* h_bvoff,
                              const int* h_btoff, int n_pairs, int mode, int n_samples, float* d_out_ab,
                              float* d_out_ba, float* d_samples_a, float* d_samples_b, hd_plan_t* out_plan) {
    return plan_create(ctx, d_av, d_at, d_bv, d_bt, h_avoff, h_atoff, h_bvoff, h_btoff, n_pairs, mode, n_samples,
                       d_out_ab, d_out_ba, d_samples_a, d_samples_b, true, out_plan);
}

int gpu_hausdorff_plan_run(gpu_ctx_t ctx, hd_plan_t plan) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(plan, "null plan");
    if (plan->n_tasks == 0) return 0;
    HIP_TRY(hipGraphLaunch(plan->exec, ctx->stream));
    return 0;
}

void gpu_hausdorff_plan_destroy(gpu_ctx_t ctx, hd_plan_t plan) {
    if (!plan) return;
    if (ctx) (void)ctx_stream_sync(ctx, ctx->stream);
    if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph) (void)hipGraphDestroy(plan->graph);
    (void)hipFree(plan->d_tasks);
    (void)hipFree(plan->d_ws);
    free(plan);
}

int gpu_hausdorff_batch(gpu_ctx_t ctx, const float* av, const int* avoff, const int* at, const int* atoff,
                        const float* bv, const int* bvoff, const int* bt, const int* btoff, int n_pairs, int mode,
                        int n_samples, float* out_ab, float* out_ba, float* out_samples_a, float* out_samples_b) {
    if (n_pairs <= 0) return 0;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t nav = (size_t)avoff[n_pairs], nat = (size_t)atoff[n_pairs];
    const size_t nbv = (size_t)bvoff[n_pairs], nbt = (size_t)btoff[n_pairs];
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t smp = (mode == 1 && (out_samples_a || out_samples_b)) ? 12ull * n_samples * n_pairs : 0;
    const size_t total = al(12 * nav) + al(12 * nat) + al(12 * nbv) + al(12 * nbt) + 2 * al(4ull * n_pairs) + 2 * al(smp);
    char* d = (char*)stage_buf(ctx, 1, total);
    REQUIRE(d, "staging allocation failed");
    size_t o = 0;
    float* d_av = (float*)(d + o); o += al(12 * nav);
    int* d_at = (int*)(d + o); o += al(12 * nat);
    float* d_bv = (float*)(d + o); o += al(12 * nbv);
    int* d_bt = (int*)(d + o); o += al(12 * nbt);
    float* d_ab = (float*)(d + o); o += al(4ull * n_pairs);
    float* d_ba = (float*)(d + o); o += al(4ull * n_pairs);
    float* d_sa = smp ? (float*)(d + o) : nullptr; o += al(smp);
    float* d_sb = smp ? (float*)(d + o) : nullptr;
    HIP_TRY(hipMemcpyAsync(d_av, av, 12 * nav, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_at, at, 12 * nat, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_bv, bv, 12 * nbv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_bt, bt, 12 * nbt, hipMemcpyHostToDevice, ctx->stream));
    hd_plan_t plan = nullptr;
    // one run: launched directly, no graph
    int rc = plan_create(ctx, d_av, d_at, d_bv, d_bt, avoff, atoff, bvoff, btoff, n_pairs, mode, n_samples, d_ab,
                         d_ba, d_sa, d_sb, false, &plan);
    if (rc) return rc;
    rc = launch_tasks(ctx, plan->d_tasks, plan->n_tasks, plan->d_ws);
    if (rc == 0) {
        HIP_TRY(hipMemcpyAsync(out_ab, d_ab, 4ull * n_pairs, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(out_ba, d_ba, 4ull * n_pairs, hipMemcpyDeviceToHost, ctx->stream));
        if (out_samples_a && d_sa)
            HIP_TRY(hipMemcpyAsync(out_samples_a, d_sa, smp, hipMemcpyDeviceToHost, ctx->stream));
        if (out_samples_b && d_sb)
            HIP_TRY(hipMemcpyAsync(out_samples_b, d_sb, smp, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        for (int i = 0; i < n_pairs; ++i) {
            out_ab[i] = sqrtf(out_ab[i]);
            out_ba[i] = sqrtf(out_ba[i]);
        }
    }
    gpu_hausdorff_plan_destroy(ctx, plan);
    return rc;
}

}  // extern "C"
