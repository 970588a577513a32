// plane_cut_kernels.hip - the plane-cut entry point and a batched cut kernel.
//
// gpu_test_plane_cut (reference csrc/test.c:249-356, cuda/test_plane_cut.cu:11-90):
// cut one mesh, return both closed halves in caller buffers.
// gpu_plane_cut_batch (no reference counterpart): one workgroup per (mesh,
// plane) pair, returning per-cut volumes and sizes; the lookahead search and
// the candidate scorer use the device function directly.
#include <algorithm>
#include <vector>
#include "ctx.hpp"
#include "lookahead.hpp"
#include "plane_cut.hpp"

using namespace cuacd;

struct PcCutOut {
    float* pv;
    int* pt;
    float* nv;
    int* nt;
    int pnv, pnt, nnv, nnt;
    double pvol, nvol;
    int status;
    int pad;
};

namespace {

__global__ __launch_bounds__(PC_T) void pc_single_kernel(const float* V, const int* T, int nv, int nt, float pa,
                                                         float pb, float pc, float pd, DPool* pool, int* err,
                                                         PcCutOut* out, int lds_nv) {
    __shared__ PcLDS s;
    extern __shared__ __align__(16) unsigned char pc_dyn[];
    PcResult r = nv <= lds_nv ? plane_cut_block_lds(V, T, nv, nt, pa, pb, pc, pd, pool, err, &s, pc_dyn)
                              : plane_cut_block(V, T, nv, nt, pa, pb, pc, pd, pool, err, &s);
    if (threadIdx.x == 0) {
        out->pv = r.pos.v; out->pt = r.pos.t; out->pnv = r.pos.nv; out->pnt = r.pos.nt; out->pvol = r.pos.vol;
        out->nv = r.neg.v; out->nt = r.neg.t; out->nnv = r.neg.nv; out->nnt = r.neg.nt; out->nvol = r.neg.vol;
        out->status = r.status;
    }
}

__global__ void pc_free_kernel(DPool* pool, PcCutOut* out, int n, int* err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    heap_free(pool, out[i].pv, err);
    heap_free(pool, out[i].nv, err);
}

// Batched cuts: pair i cuts mesh mesh_of[i] with plane planes[i].
struct PcBatchArgs {
    const float* verts;
    const int* vert_off;
    const int* tris;
    const int* tri_off;
    const int* mesh_of;
    const float4* planes;
    int n;
    DPool* pool;
    int* err;
    PcCutOut* out;
    int lds_nv;  // meshes with up to this many vertices take the LDS-resident cut
};

__global__ __launch_bounds__(PC_T) void pc_batch_kernel(PcBatchArgs a) {
    __shared__ PcLDS s;
    extern __shared__ __align__(16) unsigned char pc_dyn[];
    const int i = blockIdx.x;
    if (i >= a.n) return;
    const int m = a.mesh_of[i];
    const float* V = a.verts + 3ll * a.vert_off[m];
    const int* T = a.tris + 3ll * a.tri_off[m];
    const int nv = a.vert_off[m + 1] - a.vert_off[m], nt = a.tri_off[m + 1] - a.tri_off[m];
    const float4 p = a.planes[i];
    PcResult r = nv <= a.lds_nv ? plane_cut_block_lds(V, T, nv, nt, p.x, p.y, p.z, p.w, a.pool, a.err, &s, pc_dyn)
                                : plane_cut_block(V, T, nv, nt, p.x, p.y, p.z, p.w, a.pool, a.err, &s);
    if (threadIdx.x == 0) {
        PcCutOut& o = a.out[i];
        o.pv = r.pos.v; o.pt = r.pos.t; o.pnv = r.pos.nv; o.pnt = r.pos.nt; o.pvol = r.pos.vol;
        o.nv = r.neg.v; o.nt = r.neg.t; o.nnv = r.neg.nv; o.nnt = r.neg.nt; o.nvol = r.neg.vol;
        o.status = r.status;
    }
}

}  // namespace

extern "C" {

int gpu_test_plane_cut(gpu_ctx_t ctx, const float* verts, int nv, const int* tris, int nt, float pa, float pb,
                       float pc, float pd, float* opv, int opv_cap, int* opt, int opt_cap, float* onv, int onv_cap,
                       int* ont, int ont_cap, int* n_pv, int* n_pt, int* n_nv, int* n_nt) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(nv >= 0 && nt >= 0, "invalid sizes");
    const size_t bv = 12ull * nv, bt = 12ull * nt;
    char* d = (char*)stage_buf(ctx, 0, bv + bt + 1024);
    REQUIRE(d, "staging allocation failed");
    float* d_v = (float*)d;
    int* d_t = (int*)(d + ((bv + 255) & ~255ull));
    PcCutOut* d_out = (PcCutOut*)(d + ((bv + 255) & ~255ull) + ((bt + 255) & ~255ull));
    HIP_TRY(hipMemcpyAsync(d_v, verts, bv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_t, tris, bt, hipMemcpyHostToDevice, ctx->stream));
    int lds_nv;
    size_t lds;
    if (int rc = pc_lds_launch(ctx, (const void*)pc_single_kernel, nv, &lds_nv, &lds)) return rc;
    hipLaunchKernelGGL(pc_single_kernel, dim3(1), dim3(PC_T), lds, ctx->stream, d_v, d_t, nv, nt, pa, pb, pc, pd,
                       ctx->d_pool, ctx->d_err, d_out, lds_nv);
    HIP_TRY(hipGetLastError());
    PcCutOut h;
    HIP_TRY(hipMemcpyAsync(&h, d_out, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int e = read_and_clear_err(ctx);
    if (e) {
        char buf[256];
        kerr_decode(e, buf, sizeof(buf));
        set_error(ctx, "plane_cut: %s", buf);
        return e;
    }
    *n_pv = h.pnv;
    *n_pt = h.pnt;
    *n_nv = h.nnv;
    *n_nt = h.nnt;
    if (h.pv && h.pnv <= opv_cap && h.pnt <= opt_cap) {
        HIP_TRY(hipMemcpyAsync(opv, h.pv, 12ull * h.pnv, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(opt, h.pt, 12ull * h.pnt, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (h.nv && h.nnv <= onv_cap && h.nnt <= ont_cap) {
        HIP_TRY(hipMemcpyAsync(onv, h.nv, 12ull * h.nnv, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ont, h.nt, 12ull * h.nnt, hipMemcpyDeviceToHost, ctx->stream));
    }
    hipLaunchKernelGGL(pc_free_kernel, dim3(1), dim3(64), 0, ctx->stream, ctx->d_pool, d_out, 1, ctx->d_err);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

// Batch of cuts of device-resident meshes: cut i = (mesh mesh_of[i], plane
// planes[i] = (a, b, c, d)). Writes per-cut volumes (FP64) and sizes; the
// halves themselves are freed before returning. Host arrays in, host out.
int gpu_plane_cut_batch(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris, const int* tri_off,
                        int n_meshes, const int* mesh_of, const float* planes, int n_cuts, double* out_vol_pos,
                        double* out_vol_neg, int* out_sizes, int* out_status) {
    if (n_cuts <= 0) return 0;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t tv = (size_t)vert_off[n_meshes], tt = (size_t)tri_off[n_meshes];
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t total = al(12 * tv) + al(12 * tt) + 2 * al(4ull * (n_meshes + 1)) + al(4ull * n_cuts) +
                         al(16ull * n_cuts) + al(sizeof(PcCutOut) * n_cuts);
    char* d = (char*)stage_buf(ctx, 0, total);
    REQUIRE(d, "staging allocation failed");
    size_t o = 0;
    float* d_v = (float*)(d + o); o += al(12 * tv);
    int* d_t = (int*)(d + o); o += al(12 * tt);
    int* d_vo = (int*)(d + o); o += al(4ull * (n_meshes + 1));
    int* d_to = (int*)(d + o); o += al(4ull * (n_meshes + 1));
    int* d_mo = (int*)(d + o); o += al(4ull * n_cuts);
    float4* d_pl = (float4*)(d + o); o += al(16ull * n_cuts);
    PcCutOut* d_out = (PcCutOut*)(d + o);
    HIP_TRY(hipMemcpyAsync(d_v, verts, 12 * tv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_t, tris, 12 * tt, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_vo, vert_off, 4ull * (n_meshes + 1), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_to, tri_off, 4ull * (n_meshes + 1), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_mo, mesh_of, 4ull * n_cuts, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_pl, planes, 16ull * n_cuts, hipMemcpyHostToDevice, ctx->stream));
    int max_nv = 0;
    for (int i = 0; i < n_cuts; ++i) {
        REQUIRE(mesh_of[i] >= 0 && mesh_of[i] < n_meshes, "plane_cut_batch: cut %d names mesh %d of %d", i,
                mesh_of[i], n_meshes);
        max_nv = std::max(max_nv, vert_off[mesh_of[i] + 1] - vert_off[mesh_of[i]]);
    }
    int lds_nv;
    size_t lds;
    if (int rc = pc_lds_launch(ctx, (const void*)pc_batch_kernel, max_nv, &lds_nv, &lds)) return rc;
    PcBatchArgs a{d_v, d_vo, d_t, d_to, d_mo, d_pl, n_cuts, ctx->d_pool, ctx->d_err, d_out, lds_nv};
    hipLaunchKernelGGL(pc_batch_kernel, dim3(n_cuts), dim3(PC_T), lds, ctx->stream, a);
    HIP_TRY(hipGetLastError());
    std::vector<PcCutOut> h(n_cuts);
    HIP_TRY(hipMemcpyAsync(h.data(), d_out, sizeof(PcCutOut) * n_cuts, hipMemcpyDeviceToHost, ctx->stream));
    hipLaunchKernelGGL(pc_free_kernel, dim3((n_cuts + 63) / 64), dim3(64), 0, ctx->stream, ctx->d_pool, d_out, n_cuts,
                       ctx->d_err);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < n_cuts; ++i) {
        out_vol_pos[i] = h[i].pvol;
        out_vol_neg[i] = h[i].nvol;
        out_sizes[4 * i] = h[i].pnv;
        out_sizes[4 * i + 1] = h[i].pnt;
        out_sizes[4 * i + 2] = h[i].nnv;
        out_sizes[4 * i + 3] = h[i].nnt;
        out_status[i] = h[i].status;
    }
    int e = read_and_clear_err(ctx);
    if (e) {
        char buf[256];
        kerr_decode(e, buf, sizeof(buf));
        set_error(ctx, "plane_cut_batch: %s", buf);
        return e;
    }
    return 0;
}

}  // extern "C"
