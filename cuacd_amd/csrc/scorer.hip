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
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int e = read_and_clear_err(ctx);
    if (e) {
        char buf[256];
        kerr_decode(e, buf, sizeof(buf));
        set_error(ctx, "score_cuts: %s (%s)", buf, where);
        return e;
    }
    return 0;
}

int score_run(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris, const int* tri_off,
              int n_meshes, const int* mesh_of, const float* planes, int n_cuts, float tau, float* out_score,
              double* out_detail, double* out_stats) {
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
    LA_SC_TRY(la_launch_cut(ctx, dct.p, n_cuts));
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
        HIP_TRY(hipStreamSynchronize(ctx->stream));
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
    }
    const int frc = free_chunks();
    if (rc == 0) rc = frc;
    if (rc) return rc;

    // ---- scores --------------------------------------------------------------
    for (int i = 0; i < n_cuts; ++i) {
        const size_t p = 2 * (size_t)i, n = p + 1;
        const bool ok = hof[p] && hof[n];
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
