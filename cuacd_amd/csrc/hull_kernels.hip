// hull_kernels.hip - batch exact hulls.
//
// Backs gpu_hull_dandc / batch_hull_volume / batch_hull_dandc_mesh
// (reference: csrc/test.c:95-160, cuda/test_hull_dandc.cu:24-77).
//
// Three tiers, all exact on the same quantization grid:
//   * clouds of <= 1024 points: one wave64 workgroup, points in registers,
//     workspace in LDS (hull_wave_kernel);
//   * clouds of <= 16384 points: eight waves, points in registers, workspace
//     in global memory (hull_block_kernel);
//   * larger clouds: divide and conquer - chunks of 1024 points are hulled
//     independently on the cloud's grid, the union of their hull vertices
//     (kept in ascending input order, so lowest-index duplicate resolution is
//     preserved) replaces the cloud, until the block kernel can finish it.
#include <stdlib.h>
#include <algorithm>
#include <vector>
#include "ctx.hpp"
#include "hull.hpp"

using namespace cuacd;

namespace {

constexpr int WAVE_MAX = 1024;    // points per cloud for the one-wave tier
constexpr int BLOCK_NW = 8;       // waves of the block tier
constexpr int BLOCK_MAX = 16384;  // points per cloud for the block tier
constexpr int CHUNK = 1024;

struct HullItemArgs {
    const float* pts;        // all clouds, xyz
    const int* offsets;      // [n_clouds+1]
    const int* sel;          // optional subset (indices relative to the cloud start)
    const int* sel_off;      // [n_items+1] ranges into sel
    const int* item_cloud;   // optional item -> cloud map (identity if null)
    const QFrame* frames;    // optional frame per cloud (else from the item's points)
    int n_items;
    int VC;                  // vertex capacity of the workspace
    int NP;                  // point capacity of the workspace
    int min_n, max_n;        // items outside [min_n, max_n] are left to another tier
    char* gws;               // global workspaces (block tier), gws_stride bytes each
    unsigned long long gws_stride;
    // mesh output (mode 0), indexed by cloud
    int max_hv, max_ht;
    float* out_verts;
    int* out_tris;
    int* out_nv;
    int* out_nt;
    int* out_err;
    double* out_vol;
    // survivor output (mode 1): hull vertex indices, same ranges as sel
    int* out_sel;
    int* out_cnt;
};

template <int PPL, int AS, int NW>
__device__ __forceinline__ void hull_item(const HullItemArgs& a, int mode, int item, const HullWS<AS> w,
                                          HullBlockLDS* sb) {
    constexpr int T = 64 * NW;
    const int tid = threadIdx.x;
    const int cloud = a.item_cloud ? a.item_cloud[item] : item;
    const float* cpts = a.pts + 3ll * a.offsets[cloud];
    int n;
    const int* sel = nullptr;
    if (a.sel) {
        sel = a.sel + a.sel_off[item];
        n = a.sel_off[item + 1] - a.sel_off[item];
    } else {
        n = a.offsets[cloud + 1] - a.offsets[cloud];
    }
    if (n < a.min_n || n > a.max_n) return;  // another tier owns this item
    auto orig = [&](int i) { return sel ? sel[i] : i; };
    if (tid == 0) {
        if (mode == 0) {
            a.out_nv[cloud] = 0;
            a.out_nt[cloud] = 0;
            a.out_err[cloud] = 0;
            if (a.out_vol) a.out_vol[cloud] = 0.0;
        } else {
            a.out_cnt[item] = 0;
        }
    }
    if (n < 4 || n > T * PPL) {
        if (n > T * PPL) {
            if (tid == 0) {
                if (mode == 0) a.out_err[cloud] = KERR_HULL_CAPACITY;
                else a.out_cnt[item] = -1;
            }
        } else if (mode == 1) {
            // fewer than 4 points: no hull (reference hull_dandc_warp_mesh n<4); keep them all
            for (int i = tid; i < n; i += T) a.out_sel[a.sel_off[item] + i] = orig(i);
            if (tid == 0) a.out_cnt[item] = n;
        }
        return;
    }

    float xs[PPL], ys[PPL], zs[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        xs[k] = ys[k] = zs[k] = 0.0f;
        const int i = k * T + tid;
        if (i < n) {
            const float* p = cpts + 3ll * orig(i);
            xs[k] = p[0];
            ys[k] = p[1];
            zs[k] = p[2];
        }
    }
    QFrame fr;
    if (a.frames) {
        fr = a.frames[cloud];
    } else {
        float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            if (k * T + tid < n) {
                mn[0] = fminf(mn[0], xs[k]); mx[0] = fmaxf(mx[0], xs[k]);
                mn[1] = fminf(mn[1], ys[k]); mx[1] = fmaxf(mx[1], ys[k]);
                mn[2] = fminf(mn[2], zs[k]); mx[2] = fmaxf(mx[2], zs[k]);
            }
        }
        for (int j = 0; j < 3; ++j) {
            mn[j] = wave_min_f(mn[j]);
            mx[j] = wave_max_f(mx[j]);
        }
        if (NW > 1) {
            const int wv = tid >> 6;
            if (lane_id() == 0)
                for (int j = 0; j < 3; ++j) {
                    sb->bmn[wv][j] = mn[j];
                    sb->bmx[wv][j] = mx[j];
                }
            __syncthreads();
            for (int j = 0; j < 3; ++j) {
                for (int q = 0; q < NW; ++q) {
                    mn[j] = fminf(mn[j], sb->bmn[q][j]);
                    mx[j] = fmaxf(mx[j], sb->bmx[q][j]);
                }
            }
            __syncthreads();
        }
        fr = qframe_make(mn, mx);
    }
    int px[PPL], py[PPL], pz[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        px[k] = py[k] = pz[k] = 0;
        if (k * T + tid < n) {
            const float v[3] = {xs[k], ys[k], zs[k]};
            px[k] = qcoord(fr, 0, sel3(v, fr.ax[0]));
            py[k] = qcoord(fr, 1, sel3(v, fr.ax[1]));
            pz[k] = qcoord(fr, 2, sel3(v, fr.ax[2]));
        }
    }

    HullResult r = block_quickhull<PPL, AS, NW>(w, sb, n, px, py, pz);
    if (r.status == HULL_OVERFLOW) {
        if (tid == 0) {
            if (mode == 0) a.out_err[cloud] = KERR_HULL_CAPACITY;
            else a.out_cnt[item] = -1;
        }
        return;
    }
    if (r.status == HULL_DEGENERATE) {
        // flat or collinear input: no solid hull. Survivor mode keeps all points.
        if (mode == 1) {
            for (int i = tid; i < n; i += T) a.out_sel[a.sel_off[item] + i] = orig(i);
            if (tid == 0) a.out_cnt[item] = n;
        }
        return;
    }
    if (mode == 1) {
        // survivors: every hull vertex (a superset of the corners is enough)
        const int nv = hull_rank_vertices<AS, NW>(w, sb, r, n);
        int* dst = a.out_sel + a.sel_off[item];
        for (int v = tid; v < r.nvert; v += T)
            if (w.vref[v]) dst[w.vout[v]] = orig(w.vpid[v]);
        if (tid == 0) a.out_cnt[item] = nv;
        return;
    }
    auto pos = [&](int p, float* o3) {
        const float* q = cpts + 3ll * orig(p);
        o3[0] = q[0];
        o3[1] = q[1];
        o3[2] = q[2];
    };
    const HullCanon hc = hull_canon<AS, NW>(w, sb, r, n, pos);
    if (hc.nt < 0) {
        if (tid == 0) a.out_err[cloud] = KERR_HULL_CAPACITY;
        return;
    }
    const int nv = hc.nv, nt = hc.nt;
    if (tid == 0) {
        a.out_nv[cloud] = nv;
        a.out_nt[cloud] = nt;
    }
    if (nv <= a.max_hv) {
        float* dv = a.out_verts + 3ll * a.max_hv * cloud;
        for (int v = tid; v < r.nvert; v += T)
            if (w.vref[v]) {
                const float* p = cpts + 3ll * orig(w.vpid[v]);
                const int o = w.vout[v];
                dv[3 * o + 0] = p[0];
                dv[3 * o + 1] = p[1];
                dv[3 * o + 2] = p[2];
            }
    }
    if (nt <= a.max_ht) hull_canon_write<AS, NW>(w, r, a.out_tris + 3ll * a.max_ht * cloud);
    if (a.out_vol && tid == 0) a.out_vol[cloud] = hc.vol;
}

template <int PPL>
__global__ __launch_bounds__(64) void hull_wave_kernel(HullItemArgs a, int mode) {
    extern __shared__ __align__(16) char lds[];
    __shared__ HullBlockLDS sb;
    __shared__ HullBits<64 * PPL> hb;
    const int item = blockIdx.x;
    if (item >= a.n_items) return;
    // compile-time capacities: every sub-array address folds into ds offsets
    const HullWS<AS_LDS> w = hull_ws_attach(hull_ws_carve<AS_LDS>(lds, 64 * PPL, 64 * PPL), &sb, &hb);
    hull_item<PPL, AS_LDS, 1>(a, mode, item, w, &sb);
}

template <int PPL>
__global__ __launch_bounds__(64 * BLOCK_NW) void hull_block_kernel(HullItemArgs a, int mode) {
    __shared__ HullBlockLDS sb;
    __shared__ HullBits<HULL_NP_MAX> hb;
    const int item = blockIdx.x;
    if (item >= a.n_items) return;
    const HullWS<AS_GLOBAL> w = hull_ws_attach(
        hull_ws_carve<AS_GLOBAL>(a.gws + (unsigned long long)item * a.gws_stride, a.VC, a.NP), &sb, &hb);
    hull_item<PPL, AS_GLOBAL, BLOCK_NW>(a, mode, item, w, &sb);
}

// Per-cloud AABB -> quantization frame (for the chunked path).
__global__ __launch_bounds__(256) void cloud_frame_kernel(const float* pts, const int* offsets,
                                                          const int* clouds, int n, QFrame* frames) {
    __shared__ float smn[3][4], smx[3][4];
    if ((int)blockIdx.x >= n) return;
    const int c = clouds[blockIdx.x];
    const int s = offsets[c], e = offsets[c + 1];
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = s + threadIdx.x; i < e; i += 256) {
        for (int j = 0; j < 3; ++j) {
            const float v = pts[3ll * i + j];
            mn[j] = fminf(mn[j], v);
            mx[j] = fmaxf(mx[j], v);
        }
    }
    const int wv = threadIdx.x >> 6;
    for (int j = 0; j < 3; ++j) {
        mn[j] = wave_min_f(mn[j]);
        mx[j] = wave_max_f(mx[j]);
        if (lane_id() == 0) { smn[j][wv] = mn[j]; smx[j][wv] = mx[j]; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float a[3], b[3];
        for (int j = 0; j < 3; ++j) {
            a[j] = fminf(fminf(smn[j][0], smn[j][1]), fminf(smn[j][2], smn[j][3]));
            b[j] = fmaxf(fmaxf(smx[j][0], smx[j][1]), fmaxf(smx[j][2], smx[j][3]));
        }
        frames[c] = qframe_make(a, b);
    }
}

using KernelFn = void (*)(HullItemArgs, int);

int launch_wave(gpu_ctx_t ctx, HullItemArgs a, int maxn, int mode) {
    if (a.n_items <= 0) return 0;
    KernelFn k;
    int ppl;
    if (maxn <= 64) { ppl = 1; k = hull_wave_kernel<1>; }
    else if (maxn <= 128) { ppl = 2; k = hull_wave_kernel<2>; }
    else if (maxn <= 256) { ppl = 4; k = hull_wave_kernel<4>; }
    else if (maxn <= 512) { ppl = 8; k = hull_wave_kernel<8>; }
    else { ppl = 16; k = hull_wave_kernel<16>; }
    a.VC = 64 * ppl;
    a.NP = 64 * ppl;
    const unsigned lds = hull_ws_bytes<AS_LDS>(a.VC, a.NP);
    if (lds > 65536u)
        HIP_TRY(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k, dim3(a.n_items), dim3(64), lds, ctx->stream, a, mode);
    HIP_TRY(hipGetLastError());
    return 0;
}

int launch_block(gpu_ctx_t ctx, HullItemArgs a, int maxn, int mode) {
    if (a.n_items <= 0) return 0;
    KernelFn k;
    int ppl;
    constexpr int T = 64 * BLOCK_NW;
    if (maxn <= 4 * T) { ppl = 4; k = hull_block_kernel<4>; }
    else if (maxn <= 8 * T) { ppl = 8; k = hull_block_kernel<8>; }
    else if (maxn <= 16 * T) { ppl = 16; k = hull_block_kernel<16>; }
    else { ppl = 32; k = hull_block_kernel<32>; }
    int np = maxn;
    np = (np + 63) / 64 * 64;
    a.NP = np;
    a.VC = np < 32704 ? np : 32704;
    a.gws_stride = (hull_ws_bytes<AS_GLOBAL>(a.VC, a.NP) + 255ull) & ~255ull;
    a.gws = (char*)stage_buf(ctx, 6, a.gws_stride * (unsigned long long)a.n_items);
    REQUIRE(a.gws, "hull workspace allocation failed (%llu bytes)", a.gws_stride * (unsigned long long)a.n_items);
    hipLaunchKernelGGL(k, dim3(a.n_items), dim3(T), 0, ctx->stream, a, mode);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ---- the allocation-free entry point --------------------------------------
// gpu_hull_dandc_dev_noalloc's block tier: the work of hull_block_kernel for
// items item0 + b (workgroup b), each in workspace b of the caller's region.
template <int PPL>
__global__ __launch_bounds__(64 * BLOCK_NW) void hull_block_pass_kernel(HullItemArgs a, int mode, int item0) {
    __shared__ HullBlockLDS sb;
    __shared__ HullBits<HULL_NP_MAX> hb;
    const int item = item0 + (int)blockIdx.x;
    if (item >= a.n_items) return;
    const HullWS<AS_GLOBAL> w = hull_ws_attach(
        hull_ws_carve<AS_GLOBAL>(a.gws + (unsigned long long)blockIdx.x * a.gws_stride, a.VC, a.NP), &sb, &hb);
    hull_item<PPL, AS_GLOBAL, BLOCK_NW>(a, mode, item, w, &sb);
}

// Block tier over a fixed workspace region of ws_cap bytes: passes of as
// many items as fit (same kernel choice and workspace layout as launch_block).
int launch_block_passes(gpu_ctx_t ctx, HullItemArgs a, int maxn, char* ws, unsigned long long ws_cap) {
    if (a.n_items <= 0) return 0;
    using PassFn = void (*)(HullItemArgs, int, int);
    PassFn k;
    constexpr int T = 64 * BLOCK_NW;
    if (maxn <= 4 * T) k = hull_block_pass_kernel<4>;
    else if (maxn <= 8 * T) k = hull_block_pass_kernel<8>;
    else if (maxn <= 16 * T) k = hull_block_pass_kernel<16>;
    else k = hull_block_pass_kernel<32>;
    const int np = (maxn + 63) / 64 * 64;
    a.NP = np;
    a.VC = np < 32704 ? np : 32704;
    a.gws_stride = (hull_ws_bytes<AS_GLOBAL>(a.VC, a.NP) + 255ull) & ~255ull;
    const unsigned long long per = ws_cap / a.gws_stride;
    REQUIRE(per >= 1, "a %d-point cloud needs %llu workspace bytes, above the %llu-byte scratch", maxn,
            a.gws_stride, ws_cap);
    a.gws = ws;
    for (long long i0 = 0; i0 < a.n_items; i0 += (long long)per) {
        const long long cnt = std::min<long long>((long long)per, a.n_items - i0);
        hipLaunchKernelGGL(k, dim3((unsigned)cnt), dim3(T), 0, ctx->stream, a, 0, (int)i0);
        HIP_TRY(hipGetLastError());
    }
    return 0;
}

// Its workspace: a file-scope device array, one copy per GPU (each device that
// loads the library holds its own until the process ends). On the host the
// array's name is not a device address: host code uses the address looked up
// for the device (scratch_ptr), never the name.
constexpr unsigned long long HULL_SCRATCH_BYTES = 64ull << 20;
__device__ __align__(256) char scratch[HULL_SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
char* scratch_ptr[SCRATCH_MAX_DEVICES];  // per device, set at the entry's first call there

}  // namespace

// Device-pointer entry: all buffers on the device; stream-ordered. Clouds
// above BLOCK_MAX points need h_offsets (host sizing of the chunk rounds).
static int hull_batch_device(gpu_ctx_t ctx, const float* d_pts, const int* d_offsets, const int* h_offsets,
                             int n_hulls, int max_pts, int max_hv, int max_ht, float* d_verts, int* d_tris,
                             int* d_nv, int* d_nt, int* d_err, double* d_vol) {
    HullItemArgs a{};
    a.pts = d_pts;
    a.offsets = d_offsets;
    a.n_items = n_hulls;
    a.max_hv = max_hv;
    a.max_ht = max_ht;
    a.out_verts = d_verts;
    a.out_tris = d_tris;
    a.out_nv = d_nv;
    a.out_nt = d_nt;
    a.out_err = d_err;
    a.out_vol = d_vol;
    a.min_n = 0;
    a.max_n = WAVE_MAX;
    int rc = launch_wave(ctx, a, max_pts < WAVE_MAX ? max_pts : WAVE_MAX, 0);
    if (rc || max_pts <= WAVE_MAX) return rc;
    a.min_n = WAVE_MAX + 1;
    a.max_n = BLOCK_MAX;
    rc = launch_block(ctx, a, max_pts < BLOCK_MAX ? max_pts : BLOCK_MAX, 0);
    if (rc || max_pts <= BLOCK_MAX) return rc;

    REQUIRE(h_offsets != nullptr, "clouds above %d points need the host entry point", BLOCK_MAX);
    // ---- divide and conquer for the largest clouds --------------------------
    std::vector<int> big;
    for (int c = 0; c < n_hulls; ++c)
        if (h_offsets[c + 1] - h_offsets[c] > BLOCK_MAX) big.push_back(c);
    const int nb = (int)big.size();
    int* d_big = (int*)stage_buf(ctx, 4, sizeof(int) * nb);
    QFrame* d_frames = (QFrame*)stage_buf(ctx, 5, sizeof(QFrame) * (size_t)n_hulls);
    REQUIRE(d_big && d_frames, "staging allocation failed");
    HIP_TRY(hipMemcpyAsync(d_big, big.data(), sizeof(int) * nb, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(cloud_frame_kernel, dim3(nb), dim3(256), 0, ctx->stream, d_pts, d_offsets, d_big, nb,
                       d_frames);
    HIP_TRY(hipGetLastError());

    std::vector<std::vector<int>> cand(nb);
    for (int b = 0; b < nb; ++b) {
        int n = h_offsets[big[b] + 1] - h_offsets[big[b]];
        cand[b].resize(n);
        for (int i = 0; i < n; ++i) cand[b][i] = i;
    }
    for (int round = 0; round < 32; ++round) {
        std::vector<int> item_cloud, sel, sel_off{0}, owner;
        for (int b = 0; b < nb; ++b) {
            const int n = (int)cand[b].size();
            if (n <= BLOCK_MAX) continue;
            for (int s = 0; s < n; s += CHUNK) {
                const int e = s + CHUNK < n ? s + CHUNK : n;
                item_cloud.push_back(big[b]);
                owner.push_back(b);
                sel.insert(sel.end(), cand[b].begin() + s, cand[b].begin() + e);
                sel_off.push_back((int)sel.size());
            }
        }
        if (item_cloud.empty()) break;
        const int ni = (int)item_cloud.size();
        int* d_ic = (int*)stage_buf(ctx, 0, sizeof(int) * ni);
        int* d_sel = (int*)stage_buf(ctx, 1, sizeof(int) * sel.size());
        int* d_off = (int*)stage_buf(ctx, 2, sizeof(int) * (ni + 1));
        int* d_out = (int*)stage_buf(ctx, 3, sizeof(int) * (sel.size() + ni));
        REQUIRE(d_ic && d_sel && d_off && d_out, "staging allocation failed");
        HIP_TRY(hipMemcpyAsync(d_ic, item_cloud.data(), sizeof(int) * ni, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(d_sel, sel.data(), sizeof(int) * sel.size(), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(d_off, sel_off.data(), sizeof(int) * (ni + 1), hipMemcpyHostToDevice, ctx->stream));
        HullItemArgs c{};
        c.pts = d_pts;
        c.offsets = d_offsets;
        c.sel = d_sel;
        c.sel_off = d_off;
        c.item_cloud = d_ic;
        c.frames = d_frames;
        c.n_items = ni;
        c.min_n = 0;
        c.max_n = CHUNK;
        c.out_sel = d_out;
        c.out_cnt = d_out + sel.size();
        rc = launch_wave(ctx, c, CHUNK, 1);
        if (rc) return rc;
        std::vector<int> out(sel.size() + ni);
        HIP_TRY(hipMemcpyAsync(out.data(), d_out, sizeof(int) * out.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        std::vector<size_t> before(nb);
        for (int b = 0; b < nb; ++b) {
            before[b] = cand[b].size();
            if ((int)cand[b].size() > BLOCK_MAX) cand[b].clear();
        }
        for (int i = 0; i < ni; ++i) {
            const int cnt = out[sel.size() + i];
            REQUIRE(cnt >= 0, "hull chunk exceeded capacity");
            std::vector<int>& dst = cand[owner[i]];
            dst.insert(dst.end(), out.begin() + sel_off[i], out.begin() + sel_off[i] + cnt);
        }
        for (int b = 0; b < nb; ++b)
            REQUIRE(cand[b].size() <= BLOCK_MAX || cand[b].size() < before[b],
                    "hull of a %zu-point cloud has more than %d vertices", before[b], BLOCK_MAX);
    }
    std::vector<int> sel, sel_off{0};
    int maxn = 4;
    for (int b = 0; b < nb; ++b) {
        REQUIRE((int)cand[b].size() <= BLOCK_MAX, "hull divide-and-conquer did not converge");
        sel.insert(sel.end(), cand[b].begin(), cand[b].end());
        sel_off.push_back((int)sel.size());
        maxn = (int)cand[b].size() > maxn ? (int)cand[b].size() : maxn;
    }
    int* d_sel = (int*)stage_buf(ctx, 1, sizeof(int) * (sel.size() + 1));
    int* d_off = (int*)stage_buf(ctx, 2, sizeof(int) * (nb + 1));
    REQUIRE(d_sel && d_off, "staging allocation failed");
    HIP_TRY(hipMemcpyAsync(d_sel, sel.data(), sizeof(int) * sel.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_off, sel_off.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, ctx->stream));
    HullItemArgs f = a;
    f.sel = d_sel;
    f.sel_off = d_off;
    f.item_cloud = d_big;
    f.frames = d_frames;
    f.n_items = nb;
    f.min_n = 0;
    f.max_n = BLOCK_MAX;
    return launch_block(ctx, f, maxn, 0);
}

extern "C" {

int gpu_hull_dandc_dev_noalloc(gpu_ctx_t ctx, const float* d_pts, const int* d_offsets, int n_hulls, int max_pts,
                               int max_hv, int max_ht, float* d_verts, int* d_tris, int* d_nv, int* d_nt,
                               int* d_err, double* d_vol) {
    REQUIRE(ctx->device >= 0 && ctx->device < SCRATCH_MAX_DEVICES, "gpu_hull_dandc_dev_noalloc: device %d",
            ctx->device);
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(n_hulls >= 0 && max_pts >= 0 && max_hv >= 0 && max_ht >= 0,
            "gpu_hull_dandc_dev_noalloc: invalid arguments");
    REQUIRE(max_pts <= BLOCK_MAX, "gpu_hull_dandc_dev_noalloc: clouds above %d points are not supported",
            BLOCK_MAX);
    if (n_hulls == 0) return 0;
    // first call on this device (now current): its copy of the array
    if (!scratch_ptr[ctx->device]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        REQUIRE(p != nullptr, "gpu_hull_dandc_dev_noalloc: no device address for the scratch array");
        scratch_ptr[ctx->device] = (char*)p;
    }
    HullItemArgs a{};
    a.pts = d_pts;
    a.offsets = d_offsets;
    a.n_items = n_hulls;
    a.max_hv = max_hv;
    a.max_ht = max_ht;
    a.out_verts = d_verts;
    a.out_tris = d_tris;
    a.out_nv = d_nv;
    a.out_nt = d_nt;
    a.out_err = d_err;
    a.out_vol = d_vol;
    a.min_n = 0;
    a.max_n = WAVE_MAX;
    int rc = launch_wave(ctx, a, max_pts < WAVE_MAX ? max_pts : WAVE_MAX, 0);
    if (rc || max_pts <= WAVE_MAX) return rc;
    a.min_n = WAVE_MAX + 1;
    a.max_n = BLOCK_MAX;
    return launch_block_passes(ctx, a, max_pts, scratch_ptr[ctx->device], HULL_SCRATCH_BYTES);
}

int gpu_hull_dandc_dev(gpu_ctx_t ctx, const float* d_pts, const int* d_offsets, int n_hulls, int max_pts,
                       int max_hv, int max_ht, float* d_verts, int* d_tris, int* d_nv, int* d_nt, int* d_err,
                       double* d_vol) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(max_pts <= BLOCK_MAX, "gpu_hull_dandc_dev: clouds above %d points need the host entry point",
            BLOCK_MAX);
    return hull_batch_device(ctx, d_pts, d_offsets, nullptr, n_hulls, max_pts, max_hv, max_ht, d_verts, d_tris,
                             d_nv, d_nt, d_err, d_vol);
}

int gpu_hull_dandc_vol(gpu_ctx_t ctx, const float* pts, int total_pts, const int* offsets, int n_hulls,
                       int max_pts, int max_hv, int max_ht, float* out_verts, int* out_tris, int* out_nv,
                       int* out_nt, int* out_errors, double* out_volumes) {
    if (n_hulls <= 0) return 0;
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(total_pts >= 0 && max_hv >= 0 && max_ht >= 0, "invalid sizes");
    // true maximum cloud size from the offsets (max_pts is advisory)
    int mx = 0;
    for (int c = 0; c < n_hulls; ++c) {
        const int n = offsets[c + 1] - offsets[c];
        REQUIRE(n >= 0 && offsets[c + 1] <= total_pts, "offsets out of range");
        mx = n > mx ? n : mx;
    }
    (void)max_pts;
    const size_t b_pts = sizeof(float) * 3ull * (size_t)total_pts;
    const size_t b_off = sizeof(int) * (size_t)(n_hulls + 1);
    const size_t b_v = sizeof(float) * 3ull * (size_t)max_hv * n_hulls;
    const size_t b_t = sizeof(int) * 3ull * (size_t)max_ht * n_hulls;
    const size_t b_i = sizeof(int) * (size_t)n_hulls;
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t in_bytes = al(b_pts) + al(b_off);
    const size_t out_bytes = al(b_v) + al(b_t) + 3 * al(b_i) + al(sizeof(double) * n_hulls);
    char* d_inbuf = nullptr;
    char* d_outbuf = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_inbuf, in_bytes, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&d_outbuf, out_bytes, ctx->stream));
    float* d_pts = (float*)d_inbuf;
    int* d_off = (int*)(d_inbuf + al(b_pts));
    float* d_v = (float*)d_outbuf;
    int* d_t = (int*)(d_outbuf + al(b_v));
    int* d_nv = (int*)(d_outbuf + al(b_v) + al(b_t));
    int* d_nt = d_nv + al(b_i) / sizeof(int);
    int* d_e = d_nt + al(b_i) / sizeof(int);
    double* d_vol = (double*)((char*)d_e + al(b_i));
    HIP_TRY(hipMemcpyAsync(d_pts, pts, b_pts, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_off, offsets, b_off, hipMemcpyHostToDevice, ctx->stream));
    int rc = hull_batch_device(ctx, d_pts, d_off, offsets, n_hulls, mx, max_hv, max_ht, d_v, d_t, d_nv, d_nt, d_e,
                               out_volumes ? d_vol : nullptr);
    if (rc == 0) {
        if (b_v) HIP_TRY(hipMemcpyAsync(out_verts, d_v, b_v, hipMemcpyDeviceToHost, ctx->stream));
        if (b_t) HIP_TRY(hipMemcpyAsync(out_tris, d_t, b_t, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(out_nv, d_nv, b_i, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(out_nt, d_nt, b_i, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(out_errors, d_e, b_i, hipMemcpyDeviceToHost, ctx->stream));
        if (out_volumes)
            HIP_TRY(hipMemcpyAsync(out_volumes, d_vol, sizeof(double) * n_hulls, hipMemcpyDeviceToHost,
                                   ctx->stream));
    }
    HIP_TRY(hipFreeAsync(d_inbuf, ctx->stream));
    HIP_TRY(hipFreeAsync(d_outbuf, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (rc == 0 && ctx->inject_err) {
        out_errors[0] |= ctx->inject_err;
        ctx->inject_err = 0;
    }
    return rc;
}

int gpu_hull_dandc(gpu_ctx_t ctx, const float* pts, int total_pts, const int* offsets, int n_hulls, int max_pts,
                   int max_hv, int max_ht, float* out_verts, int* out_tris, int* out_nv, int* out_nt,
                   int* out_errors) {
    return gpu_hull_dandc_vol(ctx, pts, total_pts, offsets, n_hulls, max_pts, max_hv, max_ht, out_verts, out_tris,
                              out_nv, out_nt, out_errors, nullptr);
}

}  // extern "C"
