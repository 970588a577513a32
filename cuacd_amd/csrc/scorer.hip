// scorer.hip - batched exact scoring of candidate cuts.
//
// No reference counterpart as an entry point: it is the per-candidate work of
// the lookahead's level 0 (cut the part, hull both closed halves, cost of each
// half; reference cuda/la_expand.cu:65-424) with the stop cost of
// cuda/la_common.cuh:21-34 - max(0.3 rv, H), where the Hausdorff distance H of
// a half is measured only when its volume cost 0.3 rv is under tau (the lazy
// rule of cuda/la_refine.cu:19-45). A cut's score is the larger cost of its two
// halves. Every (mesh, plane) pair of a batch goes through each stage in one
// launch sequence: all cuts, then all hulls (size tiers on side streams), then
// all Hausdorff pairs; the host only reads back sizes between stages.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "lookahead.hpp"

#pragma clang fp contract(off)

using namespace cuacd;

#define LA_SC_TRY(x)         \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)

namespace {

template <class T> struct Tmp {
    T* p = nullptr;
    gpu_ctx_t ctx = nullptr;
    int alloc(gpu_ctx_t c, size_t n) {
        ctx = c;
        return hipMallocAsync((void**)&p, n * sizeof(T) + 16, c->stream) == hipSuccess ? 0 : -1;
    }
    ~Tmp() {
        if (p) (void)hipFreeAsync(p, ctx->stream);
    }
};

// 0.3 * cbrt(3/(4 pi) * max(hv - mv, 0)), FP64 then float (cuda/la_common.cuh:30-34)
float sc_cost_rv(double hv, double mv) {
    const double g = hv - mv > 0.0 ? hv - mv : 0.0;
    return (float)(0.3 * std::cbrt(3.0 / (4.0 * M_PI) * g));
}

int sc_check(gpu_ctx_t ctx, const char* where) {
    HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
    const int e = read_and_clear_err(ctx);
    if (e) {
        char buf[256];
        kerr_decode(e, buf, sizeof(buf));
        set_error(ctx, "score_cuts: %s (%s)", buf, where);
        return e;
    }
    return 0;
}

// Sizes an eager run saw (for a plan of the same batch).
struct ScoreSizes {
    int hd_tasks = 0;                  // Hausdorff tasks (two per measured half)
    unsigned long long hd_ws = 0;      // their workspace bytes
};

// Per cut: the score from both halves' costs; the detail row when asked
// (shared by the eager run and the plan). mv/hv/crv/hd per half 2i + side,
// has_hull[q] whether the half had a hull task.
void sc_finish(int n_cuts, const std::vector<LaCutTask>& ct, const std::vector<char>& has_hull,
               const std::vector<double>& mv, const std::vector<double>& hv, const std::vector<float>& crv,
               const std::vector<float>& hd, float* out_score, double* out_detail) {
    for (int i = 0; i < n_cuts; ++i) {
        const size_t p = 2 * (size_t)i, n = p + 1;
        const bool ok = has_hull[p] && has_hull[n];
        float sc = 1e30f;
        if (ok) {
            const float cp = std::max(crv[p], hd[p] > 0.0f ? hd[p] : 0.0f);
            const float cn = std::max(crv[n], hd[n] > 0.0f ? hd[n] : 0.0f);
            sc = std::max(cp, cn);
        }
        out_score[i] = sc;
        if (out_detail) {
            double* o = out_detail + 10 * (size_t)i;
            o[0] = ct[i].status;
            o[1] = mv[p]; o[2] = mv[n];
            o[3] = hv[p]; o[4] = hv[n];
            o[5] = crv[p]; o[6] = crv[n];
            o[7] = hd[p]; o[8] = hd[n];
            o[9] = ok ? 1.0 : 0.0;
        }
    }
}

int score_run(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris, const int* tri_off,
              int n_meshes, const int* mesh_of, const float* planes, int n_cuts, float tau, float* out_score,
              double* out_detail, double* out_stats, ScoreSizes* sizes = nullptr) {
    REQUIRE(n_meshes >= 0 && n_cuts >= 0, "score_cuts: negative sizes");
    if (n_cuts == 0) return 0;
    REQUIRE(n_meshes > 0, "score_cuts: cuts without meshes");
    for (int m = 0; m < n_meshes; ++m)
        REQUIRE(vert_off[m + 1] >= vert_off[m] && tri_off[m + 1] >= tri_off[m], "score_cuts: bad offsets");
    for (int i = 0; i < n_cuts; ++i) REQUIRE(mesh_of[i] >= 0 && mesh_of[i] < n_meshes, "score_cuts: bad mesh index");
    const size_t tv = (size_t)vert_off[n_meshes], tt = (size_t)tri_off[n_meshes];
    // per-kernel timers of the hull launches (for the stage statistics)
    struct KtGuard {
        gpu_ctx_t c;
        KTimers k;
        explicit KtGuard(gpu_ctx_t c_) : c(c_) { c->kt = &k; }
        ~KtGuard() {
            for (auto& r : k.pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
            for (hipEvent_t e : k.free_ev) (void)hipEventDestroy(e);
            c->kt = nullptr;
        }
    } ktg(ctx);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    double ph_ms[3] = {0, 0, 0};
    auto phase_end = [&](int k) -> int {
        HIP_TRY(hipEventRecord(e1, ctx->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        ph_ms[k] += ms;
        return 0;
    };
    // meshes (triangle indices made absolute into the packed vertex array)
    Tmp<float> dv;
    Tmp<int> dt;
    REQUIRE(dv.alloc(ctx, 3 * tv) == 0 && dt.alloc(ctx, 3 * tt) == 0, "score_cuts: mesh allocation failed");
    HIP_TRY(hipMemcpyAsync(dv.p, verts, 12 * tv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dt.p, tris, 12 * tt, hipMemcpyHostToDevice, ctx->stream));

    // ---- stage 1: every cut (both closed halves into heap chunks) ----------
    std::vector<LaCutTask> ct(n_cuts);
    for (int i = 0; i < n_cuts; ++i) {
        const int m = mesh_of[i];
        LaCutTask& c = ct[i];
        memset(&c, 0, sizeof(c));
        c.v = dv.p + 3 * (size_t)vert_off[m];
        c.t = dt.p + 3 * (size_t)tri_off[m];
        c.nv = vert_off[m + 1] - vert_off[m];
        c.nt = tri_off[m + 1] - tri_off[m];
        for (int j = 0; j < 4; ++j) c.plane[j] = planes[4 * i + j];
    }
    Tmp<LaCutTask> dct;
    REQUIRE(dct.alloc(ctx, ct.size()) == 0, "score_cuts: task allocation failed");
    HIP_TRY(hipMemcpyAsync(dct.p, ct.data(), sizeof(LaCutTask) * ct.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipEventRecord(e0, ctx->stream));
    int max_nv = 0;
    for (const LaCutTask& c : ct) max_nv = std::max(max_nv, c.nv);
    LA_SC_TRY(la_launch_cut(ctx, dct.p, n_cuts, max_nv));
    LA_SC_TRY(phase_end(0));
    HIP_TRY(hipMemcpyAsync(ct.data(), dct.p, sizeof(LaCutTask) * ct.size(), hipMemcpyDeviceToHost, ctx->stream));
    LA_SC_TRY(sc_check(ctx, "cuts"));
    std::vector<void*> chunks;  // heap chunks to free at the end
    for (const LaCutTask& c : ct) {
        if (c.pv) chunks.push_back(c.pv);
        if (c.nvp) chunks.push_back(c.nvp);
    }
    int rc = 0;
    auto free_chunks = [&]() -> int {
        if (chunks.empty()) return 0;
        Tmp<void*> d;
        REQUIRE(d.alloc(ctx, chunks.size()) == 0, "score_cuts: free list allocation failed");
        HIP_TRY(hipMemcpyAsync(d.p, chunks.data(), sizeof(void*) * chunks.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        LA_SC_TRY(la_launch_free(ctx, d.p, (int)chunks.size()));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        chunks.clear();
        return 0;
    };

    // ---- stage 2: exact hulls of both halves (meshes kept for Hausdorff) ----
    std::vector<LaHullTask> ht;
    std::vector<int> ht_cut;  // 2 * cut + side
    for (int i = 0; i < n_cuts; ++i) {
        const LaCutTask& c = ct[i];
        if (c.status != 0 || !c.pv || !c.nvp) continue;
        for (int s = 0; s < 2; ++s) {
            LaHullTask h{};
            h.pts = s == 0 ? c.pv : c.nvp;
            h.n = s == 0 ? c.pnv : c.nnv;
            h.emit = 1;
            h.prune_gap = 1e30;
            ht.push_back(h);
            ht_cut.push_back(2 * i + s);
        }
    }
    if (!ht.empty()) {
        Tmp<LaHullTask> dht;
        if (dht.alloc(ctx, ht.size()) != 0) {
            set_error(ctx, "score_cuts: task allocation failed");
            rc = -1;
        }
        if (rc == 0 && hipMemcpyAsync(dht.p, ht.data(), sizeof(LaHullTask) * ht.size(), hipMemcpyHostToDevice,
                                      ctx->stream) != hipSuccess)
            rc = -1;
        if (rc == 0 && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = -1;
        if (rc == 0) rc = la_launch_hull(ctx, dht.p, ht.data(), (int)ht.size());
        if (rc == 0) rc = phase_end(1);
        if (rc == 0) rc = kt_resolve(ctx);
        if (rc == 0 && hipMemcpyAsync(ht.data(), dht.p, sizeof(LaHullTask) * ht.size(), hipMemcpyDeviceToHost,
                                      ctx->stream) != hipSuccess)
            rc = -1;
        if (rc == 0) rc = sc_check(ctx, "hulls");
        for (const LaHullTask& h : ht)
            if (h.hv) chunks.push_back(h.hv);
    }
    // per half: mesh volume (clamped by the hull volume), hull volume, rv cost
    std::vector<double> mv(2 * (size_t)n_cuts, 0.0), hv(2 * (size_t)n_cuts, 0.0);
    std::vector<float> crv(2 * (size_t)n_cuts, 0.0f), hd(2 * (size_t)n_cuts, -1.0f);
    std::vector<const LaHullTask*> hof(2 * (size_t)n_cuts, nullptr);
    if (rc == 0)
        for (size_t k = 0; k < ht.size(); ++k) hof[ht_cut[k]] = &ht[k];
    for (int i = 0; i < n_cuts && rc == 0; ++i) {
        const LaCutTask& c = ct[i];
        for (int s = 0; s < 2; ++s) {
            const size_t q = 2 * (size_t)i + s;
            mv[q] = s == 0 ? c.pvol : c.nvol;
            if (!hof[q]) continue;
            hv[q] = hof[q]->vol;
            if (mv[q] > hv[q]) mv[q] = hv[q];  // cuda/la_lifecycle.cu:292-296
            crv[q] = sc_cost_rv(hv[q], mv[q]);
        }
    }

    // ---- stage 3: Hausdorff of the halves already under tau by volume ------
    if (rc == 0) {
        std::vector<HdTask> h;
        std::vector<size_t> who;
        unsigned long long ws = 0;
        for (size_t q = 0; q < hof.size(); ++q) {
            const LaHullTask* x = hof[q];
            if (!x || !x->hv || !(crv[q] < tau)) continue;
            const LaCutTask& c = ct[q / 2];
            const float* pv = (q & 1) ? c.nvp : c.pv;
            const int* pt = (q & 1) ? c.ntp : c.pt;
            const int pnv = (q & 1) ? c.nnv : c.pnv;
            const int pnt = (q & 1) ? c.nnt : c.pnt;
            HdTask a{x->hv, x->ht, pv, pt, x->hnv, x->hnt, pnv, pnt, ws, 0, 0, nullptr, nullptr, nullptr};
            ws += (hd_ws_bytes(pnt, x->hnt) + 255ull) & ~255ull;
            HdTask b{pv, pt, x->hv, x->ht, pnv, pnt, x->hnv, x->hnt, ws, 0, 0, nullptr, nullptr, nullptr};
            ws += (hd_ws_bytes(x->hnt, pnt) + 255ull) & ~255ull;
            h.push_back(a);
            h.push_back(b);
            who.push_back(q);
        }
        if (!h.empty()) {
            Tmp<float> d2;
            Tmp<HdTask> d;
            Tmp<char> dws;
            if (d2.alloc(ctx, h.size()) != 0 || d.alloc(ctx, h.size()) != 0 || dws.alloc(ctx, ws + 256) != 0) {
                set_error(ctx, "score_cuts: Hausdorff buffers allocation failed");
                rc = -1;
            }
            if (rc == 0) {
                for (size_t k = 0; k < h.size(); ++k) h[k].out_d2 = d2.p + k;
                if (hipMemcpyAsync(d.p, h.data(), sizeof(HdTask) * h.size(), hipMemcpyHostToDevice, ctx->stream) !=
                        hipSuccess ||
                    hipEventRecord(e0, ctx->stream) != hipSuccess)
                    rc = -1;
            }
            if (rc == 0) rc = hd_launch_tasks(ctx, d.p, (int)h.size(), dws.p);
            if (rc == 0) rc = phase_end(2);
            std::vector<float> r(h.size());
            if (rc == 0 &&
                hipMemcpyAsync(r.data(), d2.p, 4 * h.size(), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
                rc = -1;
            if (rc == 0) rc = sc_check(ctx, "hausdorff");
            if (rc == 0)
                for (size_t k = 0; k < who.size(); ++k) hd[who[k]] = std::sqrt(std::max(r[2 * k], r[2 * k + 1]));
        }
        if (sizes) {
            sizes->hd_tasks = (int)h.size();
            sizes->hd_ws = ws;
        }
    }
    const int frc = free_chunks();
    if (rc == 0) rc = frc;
    if (rc) return rc;

    // ---- scores --------------------------------------------------------------
    std::vector<char> has_hull(hof.size());
    for (size_t q = 0; q < hof.size(); ++q) has_hull[q] = hof[q] != nullptr;
    sc_finish(n_cuts, ct, has_hull, mv, hv, crv, hd, out_score, out_detail);
    if (out_stats) {
        double cut_bytes = 0, hull_bytes = 0;
        for (const LaCutTask& c : ct) cut_bytes += 12.0 * (c.nv + c.nt + c.pnv + c.pnt + c.nnv + c.nnt) + 16.0;
        for (const LaHullTask& h : ht) hull_bytes += 12.0 * (h.n + h.hnv + h.hnt) + 8.0;
        out_stats[0] = ph_ms[0];
        out_stats[1] = ph_ms[1];
        out_stats[2] = ph_ms[2];
        out_stats[3] = cut_bytes;
        out_stats[4] = hull_bytes;
        out_stats[5] = (double)ht.size();
        for (int kc = 0; kc < KT_COUNT; ++kc)
            for (int j = 0; j < 3; ++j) out_stats[6 + 3 * kc + j] = ctx->kt ? ctx->kt->stats[kc][j] : 0.0;
    }
    return 0;
}

}  // namespace

extern "C" int gpu_score_cuts(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris,
                              const int* tri_off, int n_meshes, const int* mesh_of, const float* planes, int n_cuts,
                              float tau, float* out_score, double* out_detail, double* out_stats) {
    if (!ctx) return -1;
    HIP_TRY(hipSetDevice(ctx->device));
    (void)read_and_clear_err(ctx);
    return score_run(ctx, verts, vert_off, tris, tri_off, n_meshes, mesh_of, planes, n_cuts, tau, out_score,
                     out_detail, out_stats);
}

// ---------------------------------------------------------------------------
// Score plan: the same scoring of a fixed batch, with every stage sized on the
// device (cut -> hull tasks -> k-DOP prefilter -> tier lists -> tiered hulls ->
// Hausdorff task table and workspace -> Hausdorff -> frees), captured once as
// a HIP graph and replayed. A replay has no host synchronisation inside; the
// host reads the per-half results back once and forms the scores exactly as
// gpu_score_cuts does (same FP64 volume costs, same decision which halves'
// Hausdorff distances count).
// ---------------------------------------------------------------------------
namespace {

// The hull task of each half: the cut's vertices of that side when the cut
// produced both halves, else an empty task (skipped by the hull call).
__global__ void sc_hull_tasks_kernel(const LaCutTask* ct, int n_cuts, LaHullTask* ht) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * n_cuts; q += gridDim.x * blockDim.x) {
        const LaCutTask& c = ct[q >> 1];
        const bool ok = c.status == 0 && c.pv && c.nvp;
        LaHullTask h{};
        h.pts = (q & 1) ? c.nvp : c.pv;
        h.n = ok ? ((q & 1) ? c.nnv : c.pnv) : 0;
        h.emit = 1;
        h.prune_gap = 1e30;
        ht[q] = h;
    }
}

// Hausdorff tasks (half -> hull and hull -> half) of every half with a hull
// whose volume gap is under g_max - a bound a little above the gap at which
// 0.3 rv reaches tau, so the host's exact rule picks from a superset.
// Workspace offsets come from an atomic bump of ws_top; a task table or
// workspace overflow is flagged in flags (the run is then redone eagerly).
__global__ void sc_hd_tasks_kernel(const LaCutTask* ct, const LaHullTask* ht, int n_cuts, double g_max,
                                   HdTask* hd, float* d2, int* half_of, int* counts, unsigned long long* ws_top,
                                   unsigned long long ws_cap, int hd_cap) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * n_cuts; q += gridDim.x * blockDim.x) {
        const LaHullTask& x = ht[q];
        if (x.n <= 0 || !x.hv) continue;
        const LaCutTask& c = ct[q >> 1];
        const bool neg = q & 1;
        double mv = neg ? c.nvol : c.pvol;
        if (mv > x.vol) mv = x.vol;
        const double g = x.vol - mv > 0.0 ? x.vol - mv : 0.0;
        if (!(g < g_max)) continue;
        const float* pv = neg ? c.nvp : c.pv;
        const int* pt = neg ? c.ntp : c.pt;
        const int pnv = neg ? c.nnv : c.pnv, pnt = neg ? c.nnt : c.pnt;
        const unsigned long long w0 = (hd_ws_bytes(pnt, x.hnt) + 255ull) & ~255ull;
        const unsigned long long w1 = (hd_ws_bytes(x.hnt, pnt) + 255ull) & ~255ull;
        const int k = atomicAdd(&counts[0], 2);
        if (k + 2 > hd_cap) {
            atomicOr(&counts[1], 1);
            continue;
        }
        const unsigned long long off = atomicAdd(ws_top, w0 + w1);
        HdTask a{x.hv, x.ht, pv, pt, x.hnv, x.hnt, pnv, pnt, off, 0, 0, d2 + k, nullptr, nullptr};
        HdTask b{pv, pt, x.hv, x.ht, pnv, pnt, x.hnv, x.hnt, off + w0, 0, 0, d2 + k + 1, nullptr, nullptr};
        if (off + w0 + w1 > ws_cap) {
            atomicOr(&counts[1], 2);
            a.snt = a.tnt = b.snt = b.tnt = 0;  // empty tasks: their results read 0
        }
        hd[k] = a;
        hd[k + 1] = b;
        half_of[k >> 1] = q;
    }
}

// Tasks [count, cap) of the Hausdorff table left from an earlier replay would
// be skipped by the count; the first count are rewritten each replay.
__global__ void sc_zero_kernel(int* counts, unsigned long long* ws_top) {
    if (threadIdx.x < 4) counts[threadIdx.x] = 0;
    if (threadIdx.x == 0) *ws_top = 0;
}

// Frees the halves and the hull meshes of a replay.
__global__ void sc_free_kernel(const LaCutTask* ct, const LaHullTask* ht, int n_cuts, DPool* pool, int* err) {
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < 2 * n_cuts; q += gridDim.x * blockDim.x) {
        const LaCutTask& c = ct[q >> 1];
        heap_free(pool, (q & 1) ? c.nvp : c.pv, err);
        if (ht[q].n > 0) heap_free(pool, ht[q].hv, err);
    }
}

}  // namespace

struct score_plan {
    gpu_ctx_t ctx = nullptr;
    int n_cuts = 0, max_nv = 0, hd_cap = 0, hd_chunks = 1;
    float tau = 0.0f;
    double g_max = 0.0;
    float* dv = nullptr;
    int* dt = nullptr;
    LaCutTask* dct = nullptr;
    LaHullTask* dht = nullptr;
    char* hwork = nullptr;
    char* harena = nullptr;
    size_t harena_bytes = 0;
    LaHullDevChunks chunks;  // chunk table for the oversized halves (cap 0: none seen)
    HdTask* dhd = nullptr;
    float* dd2 = nullptr;
    int* dhalf = nullptr;
    int* dcounts = nullptr;  // [0] Hausdorff tasks, [1] overflow flags; ws_top after them
    char* dws = nullptr;
    unsigned long long ws_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<LaCutTask> ct;
    std::vector<LaHullTask> ht;
};

namespace {

void plan_free(score_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    void* bufs[] = {p->dv,  p->dt,  p->dct,    p->dht,     p->hwork, p->harena,         p->dhd,
                    p->dd2, p->dhalf, p->dcounts, p->dws, p->chunks.sub, p->chunks.work, p->chunks.arena};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

// The launches of one replay, on the context's stream (the hull tiers fork
// to the side streams and join back).
int plan_launch(score_plan* p) {
    gpu_ctx_t ctx = p->ctx;
    const int nc = p->n_cuts, nh = 2 * nc;
    LA_SC_TRY(la_launch_cut(ctx, p->dct, nc, p->max_nv));
    const unsigned g = (unsigned)std::max(1, std::min(1024, (nh + 255) / 256));
    hipLaunchKernelGGL(sc_hull_tasks_kernel, dim3(g), dim3(256), 0, ctx->stream, (const LaCutTask*)p->dct, nc, p->dht);
    HIP_TRY(hipGetLastError());
    LA_SC_TRY(la_launch_hull_dev(ctx, p->dht, nh, p->hwork, p->harena, p->harena_bytes,
                                 p->chunks.cap > 0 ? &p->chunks : nullptr));
    unsigned long long* ws_top = (unsigned long long*)(p->dcounts + 4);
    hipLaunchKernelGGL(sc_zero_kernel, dim3(1), dim3(64), 0, ctx->stream, p->dcounts, ws_top);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(sc_hd_tasks_kernel, dim3(g), dim3(256), 0, ctx->stream, (const LaCutTask*)p->dct,
                       (const LaHullTask*)p->dht, nc, p->g_max, p->dhd, p->dd2, p->dhalf, p->dcounts, ws_top,
                       p->ws_cap, p->hd_cap);
    HIP_TRY(hipGetLastError());
    LA_SC_TRY(hd_launch_tasks_dev(ctx, p->dhd, p->hd_cap, p->dcounts, p->hd_chunks, p->dws));
    hipLaunchKernelGGL(sc_free_kernel, dim3(g), dim3(256), 0, ctx->stream, (const LaCutTask*)p->dct,
                       (const LaHullTask*)p->dht, nc, ctx->d_pool, ctx->d_err);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Reads a finished replay back and forms the scores. Returns 1 (and leaves
// the outputs alone) when a device table overflowed.
int plan_finish(score_plan* p, float* out_score, double* out_detail) {
    gpu_ctx_t ctx = p->ctx;
    const int nc = p->n_cuts, nh = 2 * nc;
    int counts[4];
    HIP_TRY(hipMemcpyAsync(p->ct.data(), p->dct, sizeof(LaCutTask) * nc, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(p->ht.data(), p->dht, sizeof(LaHullTask) * nh, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(counts, p->dcounts, sizeof(counts), hipMemcpyDeviceToHost, ctx->stream));
    int over = 0;  // halves left unhulled by the device-sized call
    HIP_TRY(hipMemcpyAsync(&over, p->hwork + 4 * LA_HDEV_FAIL, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
    const int e = read_and_clear_err(ctx);
    if (e) {
        char buf[256];
        kerr_decode(e, buf, sizeof(buf));
        set_error(ctx, "score_plan: %s", buf);
        return e;
    }
    if (counts[1] || over) return 1;
    const int nhd = std::min(counts[0], p->hd_cap);
    std::vector<float> d2(nhd);
    std::vector<int> half(nhd / 2);
    if (nhd > 0) {
        HIP_TRY(hipMemcpyAsync(d2.data(), p->dd2, 4ull * nhd, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(half.data(), p->dhalf, 4ull * (nhd / 2), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
    }
    std::vector<double> mv(nh, 0.0), hv(nh, 0.0);
    std::vector<float> crv(nh, 0.0f), hd(nh, -1.0f);
    std::vector<char> has_hull(nh, 0);
    for (int q = 0; q < nh; ++q) {
        const LaCutTask& c = p->ct[q / 2];
        mv[q] = (q & 1) ? c.nvol : c.pvol;
        if (p->ht[q].n <= 0) continue;
        has_hull[q] = 1;
        hv[q] = p->ht[q].vol;
        if (mv[q] > hv[q]) mv[q] = hv[q];  // cuda/la_lifecycle.cu:292-296
        crv[q] = sc_cost_rv(hv[q], mv[q]);
    }
    // the Hausdorff distance counts for the halves the eager rule picks
    for (int k = 0; k < nhd / 2; ++k) {
        const int q = half[k];
        if (q >= 0 && q < nh && has_hull[q] && crv[q] < p->tau)
            hd[q] = std::sqrt(std::max(d2[2 * k], d2[2 * k + 1]));
    }
    sc_finish(nc, p->ct, has_hull, mv, hv, crv, hd, out_score, out_detail);
    return 0;
}

}  // namespace

extern "C" int gpu_score_plan_create(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris,
                                     const int* tri_off, int n_meshes, const int* mesh_of, const float* planes,
                                     int n_cuts, float tau, score_plan_t* out_plan) {
    if (!ctx || !out_plan) return -1;
    *out_plan = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    (void)read_and_clear_err(ctx);
    REQUIRE(n_cuts > 0, "score_plan: no cuts");
    // one eager run: validates the batch, sizes the Hausdorff stage
    std::vector<float> s0(n_cuts);
    ScoreSizes sz;
    const long long chunks0 = ctx->chunk_subtasks;
    if (int rc = score_run(ctx, verts, vert_off, tris, tri_off, n_meshes, mesh_of, planes, n_cuts, tau, s0.data(),
                           nullptr, nullptr, &sz))
        return rc;
    const long long n_chunks = ctx->chunk_subtasks - chunks0;
    score_plan* p = new score_plan();
    p->ctx = ctx;
    p->n_cuts = n_cuts;
    p->tau = tau;
    // the gap at which 0.3 rv = tau, widened by 1e-6 (rounding of cbrt and of
    // the cast to float cannot move a cost by that much)
    const double r = (double)tau / 0.3;
    p->g_max = r * r * r * (4.0 * M_PI / 3.0) * (1.0 + 1e-6);
    const size_t tv = (size_t)vert_off[n_meshes], tt = (size_t)tri_off[n_meshes];
    const int nh = 2 * n_cuts;
    p->hd_cap = 2 * nh;
    p->hd_chunks = std::max(1, std::min(64, (4096 + std::max(sz.hd_tasks, 1) - 1) / std::max(sz.hd_tasks, 1)));
    p->ws_cap = sz.hd_ws + sz.hd_ws / 4 + (1ull << 20);
    p->harena_bytes = la_hull_dev_arena_bytes(ctx, nh);
    bool ok = hipMalloc(&p->dv, 12 * tv + 16) == hipSuccess && hipMalloc(&p->dt, 12 * tt + 16) == hipSuccess &&
              hipMalloc(&p->dct, sizeof(LaCutTask) * n_cuts) == hipSuccess &&
              hipMalloc(&p->dht, sizeof(LaHullTask) * nh) == hipSuccess &&
              hipMalloc(&p->hwork, la_hull_dev_work_bytes(nh)) == hipSuccess &&
              hipMalloc(&p->harena, p->harena_bytes) == hipSuccess &&
              hipMalloc(&p->dhd, sizeof(HdTask) * p->hd_cap) == hipSuccess &&
              hipMalloc(&p->dd2, 4ull * p->hd_cap) == hipSuccess &&
              hipMalloc(&p->dhalf, 4ull * (p->hd_cap / 2) + 16) == hipSuccess &&
              hipMalloc(&p->dcounts, 64) == hipSuccess && hipMalloc(&p->dws, p->ws_cap + 256) == hipSuccess;
    if (ok && n_chunks > 0) {
        // the eager run hulled n_chunks chunks of oversized halves
        LaHullDevChunks& c = p->chunks;
        c.cap = (int)std::min<long long>(n_chunks + n_chunks / 4 + 16, 1 << 30);
        c.arena_bytes = la_hull_dev_arena_bytes(ctx, c.cap);
        ok = hipMalloc(&c.sub, sizeof(LaHullTask) * c.cap) == hipSuccess &&
             hipMalloc(&c.work, la_hull_dev_work_bytes(c.cap)) == hipSuccess &&
             hipMalloc(&c.arena, c.arena_bytes) == hipSuccess;
    }
    if (!ok) {
        plan_free(p);
        set_error(ctx, "score_plan: device allocation failed");
        return -1;
    }
    p->ct.resize(n_cuts);
    p->ht.resize(nh);
    for (int i = 0; i < n_cuts; ++i) {
        const int m = mesh_of[i];
        LaCutTask& c = p->ct[i];
        memset(&c, 0, sizeof(c));
        c.v = p->dv + 3 * (size_t)vert_off[m];
        c.t = p->dt + 3 * (size_t)tri_off[m];
        c.nv = vert_off[m + 1] - vert_off[m];
        c.nt = tri_off[m + 1] - tri_off[m];
        for (int j = 0; j < 4; ++j) c.plane[j] = planes[4 * i + j];
        p->max_nv = std::max(p->max_nv, c.nv);
    }
    int rc = 0;
    if (hipMemcpyAsync(p->dv, verts, 12 * tv, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(p->dt, tris, 12 * tt, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
        hipMemcpyAsync(p->dct, p->ct.data(), sizeof(LaCutTask) * n_cuts, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess ||
        ctx_stream_sync(ctx, ctx->stream) != hipSuccess)
        rc = -1;
    // one uncaptured pass of the device-sized launches (kernel attributes,
    // constant tables), checked against the eager scores, then the capture
    KTimers* kt = ctx->kt;
    ctx->kt = nullptr;
    if (rc == 0) rc = plan_launch(p);
    std::vector<float> s1(n_cuts);
    if (rc == 0) rc = plan_finish(p, s1.data(), nullptr);
    if (rc == 0 && memcmp(s0.data(), s1.data(), 4ull * n_cuts) != 0) {
        set_error(ctx, "score_plan: the device-sized pass disagrees with the eager scores");
        rc = -1;
    } else if (rc == 1) {
        set_error(ctx, "score_plan: a set above the one-workgroup hull size or a full table (use gpu_score_cuts)");
        rc = -1;
    }
    if (rc == 0) {
        if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = -1;
        } else {
            const int lrc = plan_launch(p);
            const hipError_t e = hipStreamEndCapture(ctx->stream, &p->graph);
            if (lrc != 0 || e != hipSuccess) {
                rc = lrc ? lrc : -1;
                set_error(ctx, "score_plan: graph capture failed (%s)", hipGetErrorString(e));
            } else if (hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess) {
                rc = -1;
                set_error(ctx, "score_plan: graph instantiation failed");
            }
        }
    }
    ctx->kt = kt;
    if (rc) {
        plan_free(p);
        return rc;
    }
    *out_plan = p;
    return 0;
}

extern "C" int gpu_score_plan_run(gpu_ctx_t ctx, score_plan_t plan, float* out_score, double* out_detail) {
    if (!ctx || !plan) return -1;
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(plan->ctx == ctx, "score_plan: plan of another context");
    HIP_TRY(hipGraphLaunch(plan->exec, ctx->stream));
    const int rc = plan_finish(plan, out_score, out_detail);
    REQUIRE(rc != 1, "score_plan: a device table overflowed on replay");
    return rc;
}

extern "C" void gpu_score_plan_destroy(gpu_ctx_t ctx, score_plan_t plan) {
    if (!plan) return;
    if (ctx) (void)ctx_stream_sync(ctx, ctx->stream);
    plan_free(plan);
}
