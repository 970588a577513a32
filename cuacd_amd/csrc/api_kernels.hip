// api_kernels.hip - the remaining Context entry points: mesh audit
// (check_mesh), the k-DOP prefiltered hull batch and the connected-components
// debug entry. The last two drive the lookahead's task kernels.
#include <algorithm>
#include <cstring>
#include <vector>
#include "lookahead.hpp"

#pragma clang fp contract(off)

using namespace cuacd;

namespace {

// Mesh audit (reference cuda/mesh_audit.cuh:90-158): per undirected edge, the
// number of incident triangles and how many traverse it lo -> hi. The edge
// table is an open-addressing hash of 64-bit keys claimed with CAS, so a slot
// is its own record (no publish step); any number of workgroups insert.
constexpr int AU_T = 256;

__global__ __launch_bounds__(AU_T) void audit_insert_kernel(const float* V, const int* T, int nv, int nt,
                                                            unsigned long long* keys, unsigned* total, unsigned* fwd,
                                                            unsigned mask, unsigned* cnt) {
    const int t = blockIdx.x * AU_T + threadIdx.x;
    if (t >= nt) return;
    const int a = T[3 * t], b = T[3 * t + 1], c = T[3 * t + 2];
    if ((unsigned)a >= (unsigned)nv || (unsigned)b >= (unsigned)nv || (unsigned)c >= (unsigned)nv) {
        atomicAdd(&cnt[1], 1u);
        return;
    }
    if (a == b || b == c || a == c) {
        atomicAdd(&cnt[0], 1u);
    } else {
        const float ax = V[3 * a], ay = V[3 * a + 1], az = V[3 * a + 2];
        const float abx = V[3 * b] - ax, aby = V[3 * b + 1] - ay, abz = V[3 * b + 2] - az;
        const float acx = V[3 * c] - ax, acy = V[3 * c + 1] - ay, acz = V[3 * c + 2] - az;
        const float nx = aby * acz - abz * acy, ny = abz * acx - abx * acz, nz = abx * acy - aby * acx;
        if (nx * nx + ny * ny + nz * nz <= 0.0f) atomicAdd(&cnt[0], 1u);
    }
    const int e[3][2] = {{a, b}, {b, c}, {c, a}};
    for (int k = 0; k < 3; ++k) {
        const unsigned x = (unsigned)e[k][0], y = (unsigned)e[k][1];
        const unsigned lo = x < y ? x : y, hi = x < y ? y : x;
        const unsigned long long key = ((unsigned long long)lo << 32) | hi;
        unsigned h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
        for (unsigned probe = 0; probe <= mask; ++probe) {
            const unsigned long long old = atomicCAS(&keys[h], ~0ull, key);
            if (old == ~0ull || old == key) {
                atomicAdd(&total[h], 1u);
                if (x < y) atomicAdd(&fwd[h], 1u);
                break;
            }
            h = (h + 1) & mask;
        }
    }
}

__global__ __launch_bounds__(AU_T) void audit_classify_kernel(const unsigned long long* keys, const unsigned* total,
                                                              const unsigned* fwd, unsigned n, unsigned* cnt) {
    const unsigned i = blockIdx.x * AU_T + threadIdx.x;
    if (i >= n || keys[i] == ~0ull) return;
    const unsigned tot = total[i], f = fwd[i], rev = tot - f;
    if (tot == 1) atomicAdd(&cnt[3], 1u);
    else if (tot > 2) atomicAdd(&cnt[4], 1u);
    else if (tot == 2 && (f == 2 || rev == 2)) atomicAdd(&cnt[5], 1u);
}

}  // namespace

extern "C" {

int gpu_mesh_audit(gpu_ctx_t ctx, const float* verts, int nv, const int* tris, int nt, unsigned int* out_flags) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(nv >= 0 && nt >= 0 && out_flags, "mesh_audit: invalid arguments");
    unsigned cap = 1024;
    while ((unsigned long long)cap < 6ull * (unsigned long long)nt + 1ull) cap <<= 1;
    char* d = nullptr;
    const size_t bv = (12ull * nv + 255) & ~255ull, bt = (12ull * nt + 255) & ~255ull;
    const size_t total = bv + bt + 8ull * cap + 8ull * cap + 64;
    HIP_TRY(hipMallocAsync((void**)&d, total, ctx->stream));
    float* d_v = (float*)d;
    int* d_t = (int*)(d + bv);
    unsigned long long* keys = (unsigned long long*)(d + bv + bt);
    unsigned* tot = (unsigned*)(keys + cap);
    unsigned* fw = tot + cap;
    unsigned* cnt = fw + cap;
    if (nv) HIP_TRY(hipMemcpyAsync(d_v, verts, 12ull * nv, hipMemcpyHostToDevice, ctx->stream));
    if (nt) HIP_TRY(hipMemcpyAsync(d_t, tris, 12ull * nt, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemsetAsync(keys, 0xFF, 8ull * cap, ctx->stream));
    HIP_TRY(hipMemsetAsync(tot, 0, 8ull * cap + 64, ctx->stream));
    if (nt > 0) {
        hipLaunchKernelGGL(audit_insert_kernel, dim3((nt + AU_T - 1) / AU_T), dim3(AU_T), 0, ctx->stream, d_v, d_t,
                           nv, nt, keys, tot, fw, cap - 1, cnt);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(audit_classify_kernel, dim3((cap + AU_T - 1) / AU_T), dim3(AU_T), 0, ctx->stream, keys,
                           tot, fw, cap, cnt);
        HIP_TRY(hipGetLastError());
    }
    unsigned h[8];
    HIP_TRY(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipFreeAsync(d, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    unsigned f = 0;
    if (h[0]) f |= 1u << 0;  // degenerate triangle
    if (h[1]) f |= 1u << 1;  // index out of range
    if (h[3]) f |= 1u << 2;  // open edge
    if (h[4]) f |= 1u << 3;  // non-manifold edge
    if (h[5]) f |= 1u << 4;  // edge traversed twice the same way
    *out_flags = f;
    return 0;
}

// k-DOP prefiltered hull of every cloud (reference csrc/test.c:364-428,
// cuda/test_kdop_hull.cu:7-63): exact hull meshes into flat caller buffers.
int gpu_kdop_hull(gpu_ctx_t ctx, const float* pts, int total_pts, const int* offsets, int n_hulls,
                  int max_hull_verts, int max_hull_tris, float* out_verts, int* out_tris, int* out_nv, int* out_nt,
                  float* out_volumes, int* out_errors) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(total_pts >= 0 && n_hulls >= 0, "kdop_hull: invalid arguments");
    if (n_hulls == 0) return 0;
    for (int i = 0; i < n_hulls; ++i)
        REQUIRE(offsets[i] >= 0 && offsets[i] <= offsets[i + 1] && offsets[i + 1] <= total_pts,
                "kdop_hull: offsets out of range");
    float* d_pts = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_pts, 12ull * total_pts + 16, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_pts, pts, 12ull * total_pts, hipMemcpyHostToDevice, ctx->stream));
    std::vector<LaHullTask> h(n_hulls);
    for (int i = 0; i < n_hulls; ++i) {
        h[i] = LaHullTask{};
        h[i].pts = d_pts + 3ll * offsets[i];
        h[i].n = offsets[i + 1] - offsets[i];
        h[i].emit = 1;
        h[i].prune_gap = 1e30;
    }
    LaHullTask* d_t = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_t, sizeof(LaHullTask) * n_hulls, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_t, h.data(), sizeof(LaHullTask) * n_hulls, hipMemcpyHostToDevice, ctx->stream));
    int rc = la_launch_hull(ctx, d_t, h.data(), n_hulls);
    if (rc == 0) {
        HIP_TRY(hipMemcpyAsync(h.data(), d_t, sizeof(LaHullTask) * n_hulls, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<void*> chunks;
        for (int i = 0; i < n_hulls; ++i) {
            const LaHullTask& t = h[i];
            out_nv[i] = t.hnv;
            out_nt[i] = t.hnt;
            out_volumes[i] = (float)t.vol;
            out_errors[i] = t.err;
            if (t.hv) {
                chunks.push_back(t.hv);
                if (t.hnv <= max_hull_verts && t.hnt <= max_hull_tris) {
                    HIP_TRY(hipMemcpyAsync(out_verts + 3ll * max_hull_verts * i, t.hv, 12ull * t.hnv,
                                           hipMemcpyDeviceToHost, ctx->stream));
                    HIP_TRY(hipMemcpyAsync(out_tris + 3ll * max_hull_tris * i, t.ht, 12ull * t.hnt,
                                           hipMemcpyDeviceToHost, ctx->stream));
                }
            }
        }
        if (!chunks.empty()) {
            void** d_c = nullptr;
            HIP_TRY(hipMallocAsync((void**)&d_c, sizeof(void*) * chunks.size(), ctx->stream));
            HIP_TRY(hipMemcpyAsync(d_c, chunks.data(), sizeof(void*) * chunks.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            rc = la_launch_free(ctx, d_c, (int)chunks.size());
            HIP_TRY(hipFreeAsync(d_c, ctx->stream));
        }
    }
    HIP_TRY(hipFreeAsync(d_t, ctx->stream));
    HIP_TRY(hipFreeAsync(d_pts, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (rc) return rc;
    const int e = read_and_clear_err(ctx);
    if (e) {
        char buf[256];
        kerr_decode(e, buf, sizeof(buf));
        set_error(ctx, "kdop_hull: %s", buf);
        return e;
    }
    return 0;
}

// Connected components of one mesh (reference csrc/postprocess.h:20-27,
// cuda/test_postprocess.cu): pieces numbered by their lowest vertex, inner
// shells dropped when mixed with outer ones; one piece returns the input.
int gpu_test_postprocess_dc(gpu_ctx_t ctx, const float* verts, int nv, const int* tris, int nt, int max_components,
                            int max_verts_per, int max_tris_per, float* out_verts, int* out_tris, int* out_nv,
                            int* out_nt, int* out_n_components) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(nv >= 0 && nt >= 0 && max_components > 0, "postprocess_dc: invalid arguments");
    for (long long k = 0; k < 3ll * nt; ++k)
        REQUIRE(tris[k] >= 0 && tris[k] < nv, "postprocess_dc: triangle index out of range");
    char* d = nullptr;
    const size_t bv = (12ull * nv + 255) & ~255ull, bt = (12ull * nt + 255) & ~255ull;
    const size_t bo = sizeof(LaCompOut) * LA_MAX_COMP;
    HIP_TRY(hipMallocAsync((void**)&d, bv + bt + bo + sizeof(LaCompTask) + 256, ctx->stream));
    float* d_v = (float*)d;
    int* d_t = (int*)(d + bv);
    LaCompOut* d_o = (LaCompOut*)(d + bv + bt);
    LaCompTask* d_task = (LaCompTask*)(d + bv + bt + bo);
    if (nv) HIP_TRY(hipMemcpyAsync(d_v, verts, 12ull * nv, hipMemcpyHostToDevice, ctx->stream));
    if (nt) HIP_TRY(hipMemcpyAsync(d_t, tris, 12ull * nt, hipMemcpyHostToDevice, ctx->stream));
    LaCompTask t{d_v, d_t, nv, nt, 1, d_o};
    HIP_TRY(hipMemcpyAsync(d_task, &t, sizeof(t), hipMemcpyHostToDevice, ctx->stream));
    int rc = la_launch_components(ctx, d_task, 1);
    if (rc == 0) {
        HIP_TRY(hipMemcpyAsync(&t, d_task, sizeof(t), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (t.n_comp == 0) {
            *out_n_components = (nv > 0 && nt > 0) ? 1 : 0;
            out_nv[0] = nv;
            out_nt[0] = nt;
            if (nv <= max_verts_per && nt <= max_tris_per) {
                memcpy(out_verts, verts, 12ull * nv);
                memcpy(out_tris, tris, 12ull * nt);
            }
        } else {
            std::vector<LaCompOut> o(t.n_comp);
            HIP_TRY(hipMemcpyAsync(o.data(), d_o, sizeof(LaCompOut) * t.n_comp, hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            const int nc = std::min(t.n_comp, max_components);
            *out_n_components = nc;
            std::vector<void*> chunks;
            for (int c = 0; c < t.n_comp; ++c) {
                if (o[c].v) chunks.push_back(o[c].v);
                if (c >= nc) continue;
                out_nv[c] = o[c].nv;
                out_nt[c] = o[c].nt;
                if (o[c].v && o[c].nv <= max_verts_per && o[c].nt <= max_tris_per) {
                    HIP_TRY(hipMemcpyAsync(out_verts + 3ll * max_verts_per * c, o[c].v, 12ull * o[c].nv,
                                           hipMemcpyDeviceToHost, ctx->stream));
                    HIP_TRY(hipMemcpyAsync(out_tris + 3ll * max_tris_per * c, o[c].t, 12ull * o[c].nt,
                                           hipMemcpyDeviceToHost, ctx->stream));
                }
            }
            void** d_c = nullptr;
            HIP_TRY(hipMallocAsync((void**)&d_c, sizeof(void*) * chunks.size() + 8, ctx->stream));
            HIP_TRY(hipMemcpyAsync(d_c, chunks.data(), sizeof(void*) * chunks.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            rc = la_launch_free(ctx, d_c, (int)chunks.size());
            HIP_TRY(hipFreeAsync(d_c, ctx->stream));
        }
    }
    HIP_TRY(hipFreeAsync(d, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (rc) return rc;
    const int e = read_and_clear_err(ctx);
    if (e) {
        char buf[256];
        kerr_decode(e, buf, sizeof(buf));
        set_error(ctx, "postprocess_dc: %s", buf);
        return e;
    }
    return 0;
}

}  // extern "C"
