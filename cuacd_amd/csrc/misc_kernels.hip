// misc_kernels.hip - mesh volumes and the record-sort entry point.
//
// gpu_test_mesh_volume / gpu_batch_mesh_volume (reference csrc/test.c:160-241,
// cuda/test_mesh_volume.cu:13-63, cuda/mesh_volume.cuh:20-57): |sum of signed
// tetrahedra anchored at vertex 0| / 6, one wave per mesh. Accumulated in
// FP64 (the reference uses fp32) and returned as float.
//
// gpu_test_warp_sort (reference csrc/test.c, cuda/test_warp_sort.cu:14-33):
// sorts (x, y, z, index) int records by (y, x, z, index): runs of 4096
// records sorted by one workgroup each with an LDS bitonic network, then
// merged pairwise in HBM until each array is one run (any array length).
#include <vector>
#include "ctx.hpp"
#include "sort.hpp"

using namespace cuacd;

namespace {

__global__ __launch_bounds__(64) void mesh_volume_kernel(const float* __restrict__ verts,
                                                         const int* __restrict__ tris,
                                                         const int* __restrict__ voff,
                                                         const int* __restrict__ toff, int n,
                                                         float* __restrict__ out) {
    const int m = blockIdx.x;
    if (m >= n) return;
    const float* v = verts + 3ll * (voff ? voff[m] : 0);
    const int t0 = toff[m], nt = toff[m + 1] - t0;
    const int* t = tris + 3ll * t0;
    if (nt <= 0) {
        if (threadIdx.x == 0) out[m] = 0.0f;
        return;
    }
    const double ax = v[0], ay = v[1], az = v[2];
    double acc = 0.0;
    for (int i = threadIdx.x; i < nt; i += 64) {
        const int a = t[3 * i], b = t[3 * i + 1], c = t[3 * i + 2];
        const double x0 = v[3 * a] - ax, y0 = v[3 * a + 1] - ay, z0 = v[3 * a + 2] - az;
        const double x1 = v[3 * b] - ax, y1 = v[3 * b + 1] - ay, z1 = v[3 * b + 2] - az;
        const double x2 = v[3 * c] - ax, y2 = v[3 * c + 1] - ay, z2 = v[3 * c + 2] - az;
        acc += x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
    }
    acc = wave_sum_d(acc);
    if (threadIdx.x == 0) out[m] = (float)fabs(acc / 6.0);
}

struct Rec {
    int x, y, z, idx;
};
struct RecLess {
    __device__ bool operator()(const Rec& p, const Rec& q) const {
        if (p.y != q.y) return p.y < q.y;
        if (p.x != q.x) return p.x < q.x;
        if (p.z != q.z) return p.z < q.z;
        return p.idx < q.idx;
    }
};

// Records sorted by one workgroup in LDS (64 KB of the CU's 160 KB).
constexpr int SORT_RUN = 4096;

// Pass 1: each run of SORT_RUN records of each array (runs start at the
// array's first record) sorted in LDS. Grid: (runs of the longest array,
// arrays).
__global__ __launch_bounds__(256) void rec_sort_kernel(int* __restrict__ recs, const int* __restrict__ off,
                                                       int n_arrays) {
    __shared__ Rec buf[SORT_RUN];
    const int a = blockIdx.y;
    if (a >= n_arrays) return;
    const int s = off[a], len = off[a + 1] - s;
    const int r0 = blockIdx.x * SORT_RUN;
    if (r0 >= len) return;
    const int n = min(SORT_RUN, len - r0);
    Rec* g = (Rec*)(recs + 4ll * (s + r0));
    const int P = pow2_ceil(n < 2 ? 2 : n);
    for (int i = threadIdx.x; i < P; i += blockDim.x)
        buf[i] = i < n ? g[i] : Rec{0x7fffffff, 0x7fffffff, 0x7fffffff, 0x7fffffff};
    __syncthreads();
    block_bitonic_sort(buf, P, RecLess());
    for (int i = threadIdx.x; i < n; i += blockDim.x) g[i] = buf[i];
}

// Later passes: merge neighbouring sorted runs of `run` records into runs of
// 2 run, src -> dst. Every record finds its place directly: its index in its
// own run plus the number of records of the partner run that precede it
// (strictly smaller for the left run, smaller or equal for the right one, so
// equal records keep their run order). Grid: (blocks over the longest array,
// arrays).
__global__ __launch_bounds__(256) void rec_merge_kernel(const Rec* __restrict__ src, Rec* __restrict__ dst,
                                                        const int* __restrict__ off, int n_arrays, int run) {
    const int a = blockIdx.y;
    if (a >= n_arrays) return;
    const int s = off[a], len = off[a + 1] - s;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const Rec x = src[s + i];
    const int r = i / run, pr = r ^ 1;
    const int lo = min(pr * run, len), hi = min(lo + run, len);
    const int base = min(r, pr) * run;
    if (lo >= hi) {
        dst[s + i] = x;
        return;
    }
    const RecLess less;
    int l = lo, h = hi;  // first partner record that must follow x
    while (l < h) {
        const int m = (l + h) >> 1;
        const bool before = (r & 1) ? !less(x, src[s + m]) : less(src[s + m], x);
        if (before) l = m + 1;
        else h = m;
    }
    dst[s + base + (i - r * run) + (l - lo)] = x;
}

}  // namespace

extern "C" {

int gpu_batch_mesh_volume(gpu_ctx_t ctx, const float* verts, int total_verts, const int* tris, int total_tris,
                          const int* tri_offsets, const int* vert_offsets, int n_meshes, float* out_volumes) {
    if (n_meshes <= 0) return 0;
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(total_verts >= 0 && total_tris >= 0, "invalid sizes");
    float* d_v = (float*)stage_buf(ctx, 0, 12ull * total_verts);
    int* d_t = (int*)stage_buf(ctx, 1, 12ull * total_tris);
    int* d_to = (int*)stage_buf(ctx, 2, 4ull * (n_meshes + 1));
    int* d_vo = vert_offsets ? (int*)stage_buf(ctx, 3, 4ull * (n_meshes + 1)) : nullptr;
    float* d_o = (float*)stage_buf(ctx, 4, 4ull * n_meshes);
    REQUIRE(d_v && d_t && d_to && d_o && (!vert_offsets || d_vo), "staging allocation failed");
    HIP_TRY(hipMemcpyAsync(d_v, verts, 12ull * total_verts, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_t, tris, 12ull * total_tris, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_to, tri_offsets, 4ull * (n_meshes + 1), hipMemcpyHostToDevice, ctx->stream));
    if (d_vo)
        HIP_TRY(hipMemcpyAsync(d_vo, vert_offsets, 4ull * (n_meshes + 1), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(mesh_volume_kernel, dim3(n_meshes), dim3(64), 0, ctx->stream, d_v, d_t, d_vo, d_to,
                       n_meshes, d_o);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out_volumes, d_o, 4ull * n_meshes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int gpu_test_mesh_volume(gpu_ctx_t ctx, const float* verts, int n_verts, const int* tris, int n_tris,
                         float* out_volume) {
    const int toff[2] = {0, n_tris};
    return gpu_batch_mesh_volume(ctx, verts, n_verts, tris, n_tris, toff, nullptr, 1, out_volume);
}

int gpu_test_warp_sort(gpu_ctx_t ctx, int* points, int total_pts, const int* offsets, int n_arrays) {
    if (n_arrays <= 0) return 0;
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(total_pts >= 0 && offsets[0] == 0 && offsets[n_arrays] == total_pts, "test_warp_sort: bad offsets");
    REQUIRE(n_arrays <= 65535, "test_warp_sort: at most 65535 arrays per call");
    int longest = 0;
    for (int a = 0; a < n_arrays; ++a) {
        const int n = offsets[a + 1] - offsets[a];
        REQUIRE(n >= 0, "test_warp_sort: offsets must not decrease");
        longest = n > longest ? n : longest;
    }
    if (total_pts == 0) return 0;
    int* d_p = (int*)stage_buf(ctx, 0, 16ull * total_pts);
    int* d_o = (int*)stage_buf(ctx, 1, 4ull * (n_arrays + 1));
    int* d_q = longest > SORT_RUN ? (int*)stage_buf(ctx, 2, 16ull * total_pts) : nullptr;
    REQUIRE(d_p && d_o && (longest <= SORT_RUN || d_q), "staging allocation failed");
    HIP_TRY(hipMemcpyAsync(d_p, points, 16ull * total_pts, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_o, offsets, 4ull * (n_arrays + 1), hipMemcpyHostToDevice, ctx->stream));
    const unsigned runs = (unsigned)((longest + SORT_RUN - 1) / SORT_RUN);
    hipLaunchKernelGGL(rec_sort_kernel, dim3(runs, (unsigned)n_arrays), dim3(256), 0, ctx->stream, d_p, d_o,
                       n_arrays);
    HIP_TRY(hipGetLastError());
    Rec* src = (Rec*)d_p;
    Rec* dst = (Rec*)d_q;
    for (long long run = SORT_RUN; run < longest; run *= 2) {
        hipLaunchKernelGGL(rec_merge_kernel, dim3((unsigned)((longest + 255) / 256), (unsigned)n_arrays), dim3(256),
                           0, ctx->stream, src, dst, d_o, n_arrays, (int)run);
        HIP_TRY(hipGetLastError());
        Rec* t = src;
        src = dst;
        dst = t;
    }
    HIP_TRY(hipMemcpyAsync(points, src, 16ull * total_pts, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

}  // extern "C"
