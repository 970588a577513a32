// lookahead.hip - lookahead_decompose (reference csrc/lookahead.c:144-1124)
// and its batched form over many meshes.
//
// The search is the reference's: each iteration splits the parts into
// connected components, hulls new parts, measures the Hausdorff distance of
// the parts whose volume cost is already below the threshold, takes up to
// max_n_cutting parts whose full cost max(0.3 rv, H) is at or above the
// threshold, and for each one expands a lookahead tree: level 0 tries `width`
// axis cuts plus up to 4 planes per sampled concave edge, deeper levels cut
// the worst part of each item with `width2` axis cuts, and optional quick
// levels cut it once more at its box midpoint. A leaf's path cost is
// the mean of the worst-part rv costs of its levels; the initial cut of the
// best leaf is applied. Branch-and-bound pruning at the last level, extra
// leaves for parts too small to cut, and the greedy level-0 fallback follow
// the reference (cuda/la_expand.cu, cuda/la_refine.cu, cuda/la_lifecycle.cu).
//
// What is this library's own: every phase is one launch over all meshes of a
// batch (a mesh drops out when it converges); the last full level, unless
// quick levels follow, scores a cut without building it (exact half-space
// volumes + the hull point sets of both sides) since its halves are never
// used again; volumes are FP64 and the
// costs are rounded to float, the reference's cost precision; new parts are
// appended in cutting order, so runs are deterministic.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "lookahead.hpp"
#include "mlp.hpp"

#pragma clang fp contract(off)

using namespace cuacd;

namespace {

constexpr int LA_MAX_PARTS = 8;
constexpr int LA_MAX_LEVELS = 4;
constexpr int LA_MAX_DECOMP = 16384;
constexpr float LA_PI_F = 3.14159265358979f;

struct Params {
    int max_iters, width, width2, depth, quick_depth, max_n_cutting, verbose;
    float threshold;
    int dc_final, dc_per_iter, n_concave, concave_iters, merge;
    float concave_eps, concave_threshold;
};

struct Part {
    float* v = nullptr;   // mesh chunk (device)
    int* t = nullptr;
    int nv = 0, nt = 0;
    float* hv = nullptr;  // hull chunk (device) or null
    int* ht = nullptr;
    int hnv = 0, hnt = 0;
    double mvol = 0.0, hvol = 0.0;
    float haus = 0.0f;
    int cc = 0;
    bool own = false;     // mesh chunk belongs to the search (heap-allocated)
    bool own_hull = false;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

// rv-only cost (reference cuda/la_common.cuh:30-34), FP64 then rounded to float
float cost_rv(double hv, double mv) {
    const double g = hv - mv > 0.0 ? hv - mv : 0.0;
    return (float)(0.3 * std::cbrt(3.0 / (4.0 * M_PI) * g));
}
float cost_full(const Part& p) { return std::max(cost_rv(p.hvol, p.mvol), p.haus); }

struct AxisCut {
    int axis, slice, cpa;
};
// reference cuda/la_axis_cut.cuh:10-19
AxisCut axis_cut(int width, int k) {
    const int count = width / 3, rem = width % 3;
    const int xc = count + (rem > 0), yc = count + (rem > 1);
    if (k < xc) return {0, k, xc};
    if (k < xc + yc) return {1, k - xc, yc};
    return {2, k - xc - yc, count};
}

// One part inside a search item: a reference to a mesh plus its volumes.
struct IPart {
    float* v = nullptr;
    int* t = nullptr;
    int nv = 0, nt = 0;
    double mvol = 0.0, hvol = 0.0;
    float lo[3], hi[3];
    int cc = 0;
    int store = -1;       // index of the owning half in the iteration's store, -1 = decomp part
    // level-0 halves: their hull mesh (device heap chunk), adopted with the half when its cut is applied
    float* hv = nullptr;
    int* ht = nullptr;
    int hnv = 0, hnt = 0;
};
struct Item {
    int mesh = 0, src = 0, cut0 = 0;
    int np = 0;
    IPart parts[LA_MAX_PARTS];
    float lc[LA_MAX_LEVELS] = {0, 0, 0, 0};
    int nl = 0;
    bool pruned = false;
};

float item_path_cost(const Item& it) {
    float s = 0.0f;
    for (int l = 0; l < it.nl; ++l) s += it.lc[l];
    return s / (float)it.nl;
}

// stable insertion sort by rv cost (reference cuda/la_refine.cu:376-387)
void sort_item_parts(Item& it) {
    for (int i = 1; i < it.np; ++i) {
        IPart key = it.parts[i];
        const float kc = cost_rv(key.hvol, key.mvol);
        int j = i - 1;
        while (j >= 0 && cost_rv(it.parts[j].hvol, it.parts[j].mvol) > kc) {
            it.parts[j + 1] = it.parts[j];
            --j;
        }
        it.parts[j + 1] = key;
    }
}

// Level-0 record of the search for the proposal network's training data:
// per cutting part its 256 surface samples (box frame) and, per level-0
// candidate, the plane in the same frame and the candidate's path cost (the
// value the search minimises; 1e30 when the cut yielded no leaf).
struct LaRecord {
    std::vector<float> pts, planes, target;
    std::vector<int> part_of;
    int n_parts = 0;
};

// plane (a, b, c, d) in the box frame of [lo, hi]: unit normal, offset / scale
void frame_plane(const float lo[3], const float hi[3], const float pl[4], float out[4]) {
    float sc = 0.0f, ctr[3];
    for (int j = 0; j < 3; ++j) {
        ctr[j] = 0.5f * (lo[j] + hi[j]);
        sc = std::max(sc, 0.5f * (hi[j] - lo[j]));
    }
    if (!(sc > 0.0f)) sc = 1.0f;
    const float len = std::sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
    const float il = len > 0.0f ? 1.0f / len : 0.0f;
    float dd = pl[3] * il;
    for (int j = 0; j < 3; ++j) {
        out[j] = pl[j] * il;
        dd += pl[j] * il * ctr[j];
    }
    out[3] = dd / sc;
}

// smallest power of two >= x (x >= 1): the size of the device-heap block
// that holds x bytes
static unsigned long long pow2_ceil_ull(unsigned long long x) {
    unsigned long long p = 256;
    while (p < x) p <<= 1;
    return p;
}

struct MeshState {
    std::vector<Part> parts;
    int cc_counter = 0;
    int iters = 0;
    bool done = false;
    float* in_v = nullptr;  // caller's mesh uploaded (not heap)
    int* in_t = nullptr;
    float* in_hv = nullptr;
    int* in_ht = nullptr;
};

template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    gpu_ctx_t ctx = nullptr;
    int alloc(gpu_ctx_t c, size_t count) {
        ctx = c;
        n = count;
        if (!count) return 0;
        return hipMallocAsync((void**)&p, sizeof(T) * count + 16, c->stream) == hipSuccess ? 0 : -1;
    }
    ~DevBuf() {
        if (p) (void)hipFreeAsync(p, ctx->stream);
    }
};

#define LA_TRY(x)                                \
    do {                                         \
        int _rc = (x);                           \
        if (_rc) return _rc;                     \
    } while (0)

// Per-phase accounting (GPU time from events around each phase's launches,
// launch count, algorithmic bytes: what the phase must read and write).
enum LaPhase { PH_COMP, PH_PHULL, PH_HD, PH_EDGE, PH_CUT, PH_SCORE, PH_IHULL, PH_MERGE, PH_PROP, PH_COUNT };
// Host wall time by section of the search (launches, waits and host work):
// input set-up, per-iteration refresh + selection, concave planes, level 0,
// deeper levels, evaluate + apply, post-loop passes and merge, result readback.
enum LaSect { S_SETUP, S_PRE, S_CONC, S_LEVEL0, S_LEVELN, S_APPLY, S_POST, S_OUT, S_COUNT };

struct Engine {
    gpu_ctx_t ctx;
    Params P;
    std::vector<MeshState> M;
    std::vector<void*> to_free;  // device heap chunks to free at the next flush
    double stats[PH_COUNT][3] = {};
    double sect[S_COUNT] = {};
    double sect_gpu[S_COUNT] = {}, sect_sync[S_COUNT] = {}, gpu_pending = 0.0;
    long long sync_mark = 0;
    std::chrono::steady_clock::time_point t_mark = std::chrono::steady_clock::now();
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int hull_phase = PH_IHULL;
    long long proposal_kept = 0, proposal_total = 0, proposal_fallbacks = 0;
    // Device -> host for many small device ranges at once: one segmented
    // device copy into a packed staging buffer, one transfer into pinned
    // memory (the context's grow-only pinned staging, kept across calls).
    // Returns the packed host copy (range k at offs[k], 16-byte aligned),
    // valid until the next call.
    const char* gather_to_host(const std::vector<std::pair<const void*, size_t>>& src, std::vector<size_t>& offs,
                               int* rc) {
        *rc = 0;
        offs.resize(src.size());
        size_t total = 0;
        for (size_t k = 0; k < src.size(); ++k) {
            offs[k] = total;
            total += (src[k].second + 15) / 16 * 16;
        }
        char*& pinned = ctx->la_pinned;
        if (total == 0) return pinned;
        if (total > ctx->la_pinned_cap) {
            if (pinned) (void)hipHostFree(pinned);
            pinned = nullptr;
            ctx->la_pinned_cap = 0;
            const size_t cap = std::max(total + total / 2, (size_t)(1 << 20));
            if (hipHostMalloc((void**)&pinned, cap, hipHostMallocDefault) != hipSuccess) {
                pinned = nullptr;
                set_error(ctx, "lookahead: pinned staging allocation failed");
                *rc = -1;
                return nullptr;
            }
            ctx->la_pinned_cap = cap;
        }
        DevBuf<char> stage;
        DevBuf<LaSeg> dsegs;
        std::vector<LaSeg> segs;
        for (size_t k = 0; k < src.size(); ++k)
            if (src[k].second) segs.push_back({src[k].first, nullptr, src[k].second});
        if (stage.alloc(ctx, total) != 0 || dsegs.alloc(ctx, segs.size() + 1) != 0) {
            set_error(ctx, "lookahead: readback staging allocation failed");
            *rc = -1;
            return nullptr;
        }
        for (size_t k = 0, j = 0; k < src.size(); ++k)
            if (src[k].second) segs[j++].dst = stage.p + offs[k];
        if (!segs.empty()) {
            if (hipMemcpyAsync(dsegs.p, segs.data(), sizeof(LaSeg) * segs.size(), hipMemcpyHostToDevice,
                               ctx->stream) != hipSuccess ||
                la_launch_segcopy(ctx, dsegs.p, (int)segs.size()) != 0 ||
                hipMemcpyAsync(pinned, stage.p, total, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                ctx_stream_sync(ctx, ctx->stream) != hipSuccess) {
                set_error(ctx, "lookahead: batched readback failed");
                *rc = -1;
                return nullptr;
            }
        }
        return pinned;
    }

    // a section's wall time, the GPU-phase time inside it and its host waits
    void lap(int s) {
        const auto t = std::chrono::steady_clock::now();
        sect[s] += std::chrono::duration<double, std::milli>(t - t_mark).count();
        t_mark = t;
        sect_gpu[s] += gpu_pending;
        gpu_pending = 0.0;
        sect_sync[s] += (double)(ctx->host_syncs - sync_mark);
        sync_mark = ctx->host_syncs;
    }
    int tbeg() {
        if (!ev0) {
            HIP_TRY(hipEventCreate(&ev0));
            HIP_TRY(hipEventCreate(&ev1));
        }
        HIP_TRY(hipEventRecord(ev0, ctx->stream));
        return 0;
    }
    int tend(int ph, double bytes, int launches = 1) {
        HIP_TRY(hipEventRecord(ev1, ctx->stream));
        ++ctx->host_syncs;
        HIP_TRY(hipEventSynchronize(ev1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
        LA_TRY(kt_resolve(ctx));
        stats[ph][0] += ms;
        stats[ph][1] += launches;
        stats[ph][2] += bytes;
        gpu_pending += ms;
        return 0;
    }
    ~Engine() {
        print_hull_stats();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }

    int check_err(const char* where) {
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        int e = read_and_clear_err(ctx);
        if (e) {
            char buf[256];
            kerr_decode(e, buf, sizeof(buf));
            set_error(ctx, "lookahead_decompose: %s (%s)", buf, where);
            return e;
        }
        return 0;
    }

    int flush_frees() {
        if (to_free.empty()) return 0;
        DevBuf<void*> d;
        REQUIRE(d.alloc(ctx, to_free.size()) == 0, "lookahead: free list allocation failed");
        HIP_TRY(hipMemcpyAsync(d.p, to_free.data(), sizeof(void*) * to_free.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        LA_TRY(la_launch_free(ctx, d.p, (int)to_free.size()));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        to_free.clear();
        return 0;
    }

    void drop_part(Part& p) {
        if (p.own && p.v) to_free.push_back(p.v);
        if (p.own_hull && p.hv) to_free.push_back(p.hv);
        p.v = nullptr;
        p.hv = nullptr;
    }

    // ---- phase: connected components of every part (cuda/la_lifecycle.cu:310-407)
    int components(const std::vector<int>& meshes) {
        std::vector<LaCompTask> h;
        std::vector<std::pair<int, int>> who;
        for (int m : meshes)
            for (int i = 0; i < (int)M[m].parts.size(); ++i) {
                Part& p = M[m].parts[i];
                if (p.nv <= 0 || p.nt <= 0) continue;
                LaCompTask t{};
                t.v = p.v; t.t = p.t; t.nv = p.nv; t.nt = p.nt;
                h.push_back(t);
                who.push_back({m, i});
            }
        if (h.empty()) return 0;
        DevBuf<LaCompOut> out;
        REQUIRE(out.alloc(ctx, h.size() * LA_MAX_COMP) == 0, "lookahead: component table allocation failed");
        for (size_t k = 0; k < h.size(); ++k) h[k].out = out.p + k * LA_MAX_COMP;
        DevBuf<LaCompTask> d;
        REQUIRE(d.alloc(ctx, h.size()) == 0, "lookahead: task allocation failed");
        HIP_TRY(hipMemcpyAsync(d.p, h.data(), sizeof(LaCompTask) * h.size(), hipMemcpyHostToDevice, ctx->stream));
        LA_TRY(tbeg());
        LA_TRY(la_launch_components(ctx, d.p, (int)h.size()));
        {
            double by = 0;
            for (auto& x : h) by += 12.0 * (x.nv + x.nt);
            LA_TRY(tend(PH_COMP, by));
        }
        HIP_TRY(hipMemcpyAsync(h.data(), d.p, sizeof(LaCompTask) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        LA_TRY(check_err("components"));
        // split parts (ascending mesh, part order: new pieces append in that order)
        std::vector<std::pair<const void*, size_t>> osrc(h.size());
        for (size_t k = 0; k < h.size(); ++k) osrc[k] = {h[k].out, sizeof(LaCompOut) * (size_t)h[k].n_comp};
        std::vector<size_t> ooff;
        int grc = 0;
        const char* oh = gather_to_host(osrc, ooff, &grc);
        if (grc) return grc;
        std::vector<LaCompOut> o(LA_MAX_COMP);
        for (size_t k = 0; k < h.size(); ++k) {
            const int m = who[k].first, i = who[k].second;
            MeshState& S = M[m];
            const int nc = h[k].n_comp;
            if (nc == 0 && S.parts[i].cc != 0) continue;
            if (nc == 0) {
                S.parts[i].cc = ++S.cc_counter;  // first pass over this lineage: fresh id
                continue;
            }
            memcpy(o.data(), oh + ooff[k], sizeof(LaCompOut) * nc);
            if ((int)S.parts.size() + nc - 1 > LA_MAX_DECOMP) {
                // this part's pieces and those of the parts after it belong to no
                // part yet: free them before failing
                for (size_t k2 = k; k2 < h.size(); ++k2)
                    for (int c = 0; c < h[k2].n_comp; ++c) {
                        LaCompOut x;
                        memcpy(&x, oh + ooff[k2] + sizeof(LaCompOut) * c, sizeof(LaCompOut));
                        to_free.push_back(x.v);
                    }
                (void)flush_frees();
                REQUIRE(false, "lookahead: more than %d parts", LA_MAX_DECOMP);
            }
            Part old = S.parts[i];
            int cc_base = 0;
            const bool fresh = old.cc == 0;
            if (fresh) { cc_base = S.cc_counter; S.cc_counter += nc; }
            drop_part(old);
            for (int c = 0; c < nc; ++c) {
                Part q;
                q.v = o[c].v; q.t = o[c].t; q.nv = o[c].nv; q.nt = o[c].nt;
                q.mvol = std::fabs(o[c].svol);
                q.hvol = 0.0;
                q.haus = 0.0f;
                q.own = true;
                q.cc = fresh ? cc_base + c + 1 : old.cc;
                for (int j = 0; j < 3; ++j) { q.lo[j] = o[c].lo[j]; q.hi[j] = o[c].hi[j]; }
                if (c == 0) S.parts[i] = q;
                else S.parts.push_back(q);
            }
        }
        return flush_frees();
    }

    // ---- phase: hulls of the parts without one (cuda/la_lifecycle.cu:268-300).
    // A part adopted from a level-0 cut already carries that half's hull
    // volume (the same point set, the same canonical hull); its mesh is built
    // only when the Hausdorff stage needs it (rv under the threshold) or, with
    // all_meshes, for the results.
    int hull_parts(const std::vector<int>& meshes, bool all_meshes = false) {
        std::vector<LaHullTask> h;
        std::vector<std::pair<int, int>> who;
        for (int m : meshes)
            for (int i = 0; i < (int)M[m].parts.size(); ++i) {
                Part& p = M[m].parts[i];
                if (p.nv == 0) continue;
                const bool need_vol = !(p.hvol > 0.0);
                const bool need_mesh = !p.hv && (all_meshes || cost_rv(p.hvol, p.mvol) < P.threshold);
                if (!need_vol && !need_mesh) continue;
                LaHullTask t{};
                t.pts = p.v; t.n = p.nv; t.emit = 1; t.prune_gap = 1e30;
                h.push_back(t);
                who.push_back({m, i});
            }
        if (h.empty()) return 0;
        DevBuf<LaHullTask> d;
        REQUIRE(d.alloc(ctx, h.size()) == 0, "lookahead: task allocation failed");
        HIP_TRY(hipMemcpyAsync(d.p, h.data(), sizeof(LaHullTask) * h.size(), hipMemcpyHostToDevice, ctx->stream));
        LA_TRY(tbeg());
        LA_TRY(la_launch_hull(ctx, d.p, h.data(), (int)h.size()));
        HIP_TRY(hipMemcpyAsync(h.data(), d.p, sizeof(LaHullTask) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        LA_TRY(tend(PH_PHULL, 0.0));
        for (auto& x : h) stats[PH_PHULL][2] += 12.0 * (x.n + x.hnv + x.hnt) + 8.0;
        LA_TRY(check_err("part hulls"));
        for (size_t k = 0; k < h.size(); ++k) {
            Part& p = M[who[k].first].parts[who[k].second];
            if (p.own_hull && p.hv) to_free.push_back(p.hv);
            p.hv = h[k].hv; p.ht = h[k].ht; p.hnv = h[k].hnv; p.hnt = h[k].hnt;
            p.own_hull = p.hv != nullptr;
            p.hvol = h[k].vol;
            if (p.mvol > p.hvol) p.mvol = p.hvol;  // cuda/la_lifecycle.cu:292-296
        }
        return flush_frees();
    }

    // ---- phase: Hausdorff of parts already under the threshold by volume (cuda/la_refine.cu:19-45)
    int hausdorff_parts(const std::vector<int>& meshes) {
        std::vector<HdTask> h;
        std::vector<std::pair<int, int>> who;
        unsigned long long ws = 0;
        for (int m : meshes)
            for (int i = 0; i < (int)M[m].parts.size(); ++i) {
                Part& p = M[m].parts[i];
                if (!p.hv || p.haus > 0.0f || cost_rv(p.hvol, p.mvol) >= P.threshold) continue;
                HdTask a{p.hv, p.ht, p.v, p.t, p.hnv, p.hnt, p.nv, p.nt, ws, 0, 0, nullptr, nullptr, nullptr};
                ws += (hd_ws_bytes(p.nt, p.hnt) + 255ull) & ~255ull;
                HdTask b{p.v, p.t, p.hv, p.ht, p.nv, p.nt, p.hnv, p.hnt, ws, 0, 0, nullptr, nullptr, nullptr};
                ws += (hd_ws_bytes(p.hnt, p.nt) + 255ull) & ~255ull;
                h.push_back(a);
                h.push_back(b);
                who.push_back({m, i});
            }
        if (h.empty()) return 0;
        DevBuf<float> d2;
        DevBuf<HdTask> d;
        DevBuf<char> dws;
        REQUIRE(d2.alloc(ctx, h.size()) == 0 && d.alloc(ctx, h.size()) == 0 && dws.alloc(ctx, ws + 256) == 0,
                "lookahead: Hausdorff buffers allocation failed");
        for (size_t k = 0; k < h.size(); ++k) h[k].out_d2 = d2.p + k;
        HIP_TRY(hipMemcpyAsync(d.p, h.data(), sizeof(HdTask) * h.size(), hipMemcpyHostToDevice, ctx->stream));
        LA_TRY(tbeg());
        LA_TRY(hd_launch_tasks(ctx, d.p, (int)h.size(), dws.p));
        {
            double by = 0;
            for (auto& x : h) by += 12.0 * (x.snv + x.snt + x.tnv + x.tnt) + 4.0;
            LA_TRY(tend(PH_HD, by));
        }
        std::vector<float> r(h.size());
        HIP_TRY(hipMemcpyAsync(r.data(), d2.p, 4 * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        LA_TRY(check_err("hausdorff"));
        for (size_t k = 0; k < who.size(); ++k)
            M[who[k].first].parts[who[k].second].haus = std::sqrt(std::max(r[2 * k], r[2 * k + 1]));
        return 0;
    }

    // ---- concave-edge planes for the cutting parts (cuda/la_concave.cu:48-363)
    int concave_planes(const std::vector<std::pair<int, int>>& cutting, std::vector<std::vector<float>>& planes) {
        planes.assign(cutting.size(), {});
        std::vector<LaEdgeTask> h(cutting.size());
        size_t cap_total = 0;
        std::vector<size_t> off(cutting.size());
        for (size_t c = 0; c < cutting.size(); ++c) {
            const Part& p = M[cutting[c].first].parts[cutting[c].second];
            off[c] = cap_total;
            const int cap = (3 * p.nt) / 2 + 16;
            h[c] = LaEdgeTask{p.v, p.t, p.nv, p.nt, nullptr, cap, 0};
            cap_total += cap;
        }
        DevBuf<LaEdgeRec> recs;
        DevBuf<LaEdgeTask> d;
        REQUIRE(recs.alloc(ctx, cap_total) == 0 && d.alloc(ctx, h.size()) == 0,
                "lookahead: concave-edge buffers allocation failed");
        for (size_t c = 0; c < h.size(); ++c) h[c].out = recs.p + off[c];
        // dihedral > thr  <=>  (d > 0 and acos(ca) > thr - pi) or (d <= 0 and acos(ca) < pi - thr)
        const double thr = P.concave_threshold;
        const double x1 = thr - M_PI, x2 = M_PI - thr;
        int pmode, nmode;
        float cpos = 0.f, cneg = 0.f;
        if (x1 < 0.0) pmode = 1; else if (x1 >= M_PI) pmode = 0; else { pmode = 2; cpos = (float)std::cos(x1); }
        if (x2 <= 0.0) nmode = 0; else if (x2 > M_PI) nmode = 1; else { nmode = 2; cneg = (float)std::cos(x2); }
        HIP_TRY(hipMemcpyAsync(d.p, h.data(), sizeof(LaEdgeTask) * h.size(), hipMemcpyHostToDevice, ctx->stream));
        LA_TRY(tbeg());
        LA_TRY(la_launch_edges(ctx, d.p, (int)h.size(), cpos, cneg, pmode, nmode));
        {
            double by = 0;
            for (auto& x : h) by += 12.0 * (x.nv + x.nt);
            LA_TRY(tend(PH_EDGE, by));
        }
        HIP_TRY(hipMemcpyAsync(h.data(), d.p, sizeof(LaEdgeTask) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        LA_TRY(check_err("concave edges"));
        std::vector<LaEdgeRec> r;
        const int K = P.n_concave;
        std::vector<std::pair<const void*, size_t>> rsrc(h.size());
        for (size_t c = 0; c < h.size(); ++c)
            rsrc[c] = {h[c].out, sizeof(LaEdgeRec) * (size_t)std::min(h[c].count, h[c].cap)};
        std::vector<size_t> roff;
        int grc = 0;
        const char* rh = gather_to_host(rsrc, roff, &grc);
        if (grc) return grc;
        for (size_t c = 0; c < h.size(); ++c) {
            const int cnt = std::min(h[c].count, h[c].cap);
            r.resize(cnt);
            if (cnt) memcpy(r.data(), rh + roff[c], sizeof(LaEdgeRec) * cnt);
            // edges in (lo, hi) order, reservoir sampling with the reference's LCG
            std::sort(r.begin(), r.end(), [](const LaEdgeRec& a, const LaEdgeRec& b) {
                return a.lo != b.lo ? a.lo < b.lo : (a.hi != b.hi ? a.hi < b.hi : a.f0 < b.f0);
            });
            std::vector<int> res;
            unsigned int rng = 1u;
            int seen = 0;
            for (int e = 0; e < cnt; ++e) {
                ++seen;
                if ((int)res.size() < K) {
                    res.push_back(e);
                } else {
                    rng = rng * 1664525u + 1013904223u;
                    const unsigned int thr_r = (unsigned int)((float)K / (float)seen * 4294967295.0f);
                    if (rng < thr_r) {
                        rng = rng * 1664525u + 1013904223u;
                        res[rng % (unsigned)K] = e;
                    }
                }
            }
            std::vector<float>& pl = planes[c];
            const float eps = P.concave_eps;
            for (int e : res) {
                const LaEdgeRec& x = r[e];
                const float mx = x.mid[0], my = x.mid[1], mz = x.mid[2];
                const float* n0 = x.n0;
                const float* n1 = x.n1;
                pl.insert(pl.end(), {n1[0], n1[1], n1[2], -(n1[0] * mx + n1[1] * my + n1[2] * mz + eps)});
                pl.insert(pl.end(), {n0[0], n0[1], n0[2], -(n0[0] * mx + n0[1] * my + n0[2] * mz + eps)});
                float bx = n0[0] + n1[0], by = n0[1] + n1[1], bz = n0[2] + n1[2];
                const float bl = std::sqrt(bx * bx + by * by + bz * bz);
                if (bl >= 1e-4f) {
                    bx /= bl; by /= bl; bz /= bl;
                    pl.insert(pl.end(), {bx, by, bz, -(bx * mx + by * my + bz * mz + eps)});
                    pl.insert(pl.end(), {bx, by, bz, -(bx * mx + by * my + bz * mz - eps)});
                }
            }
        }
        return 0;
    }

    // Hull volumes for item halves. prune: per task gap (>=1e29 none).
    // Full cuts with caps (halves, volumes, boxes) of a task list, read back.
    int cut_batch(std::vector<LaCutTask>& h) {
        if (h.empty()) return 0;
        DevBuf<LaCutTask> dt;
        REQUIRE(dt.alloc(ctx, h.size()) == 0, "lookahead: task allocation failed");
        HIP_TRY(hipMemcpyAsync(dt.p, h.data(), sizeof(LaCutTask) * h.size(), hipMemcpyHostToDevice, ctx->stream));
        LA_TRY(tbeg());
        int max_nv = 0;
        for (const LaCutTask& c : h) max_nv = std::max(max_nv, c.nv);
        LA_TRY(la_launch_cut(ctx, dt.p, (int)h.size(), max_nv));
        LA_TRY(tend(PH_CUT, 0.0));
        HIP_TRY(hipMemcpyAsync(h.data(), dt.p, sizeof(LaCutTask) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        // mesh read, both halves written, volumes
        for (auto& x : h) stats[PH_CUT][2] += 12.0 * (x.nv + x.nt + x.pnv + x.pnt + x.nnv + x.nnt) + 16.0;
        return check_err("cuts");
    }

    // ---- quick levels (csrc/lookahead.c:656-727, cuda/la_refine.cu:52-351): one
    // child per item, its worst part cut at the box midpoint along the axis
    // whose halves have the most balanced volumes (max / sum; the first axis
    // on ties). An item whose worst part is too small along every axis becomes
    // an extra leaf, one with no valid cut is dropped, and a mesh whose level
    // produced no child keeps its items as they are. The last quick level
    // prunes with the branch-and-bound bound over the items entering it.
    // Quick halves are never applied: all of them go to the iteration's free
    // list, which is flushed after the apply step.
    int quick_levels(const Params& P, float med, int total_levels, std::vector<Item>& cur, std::vector<int>& cur_c,
                     std::vector<Item>& extra) {
        std::vector<Item> done;
        std::vector<int> done_c;
        const float lim = 2.f * med + 1e-6f;
        for (int q = 0; q < P.quick_depth && !cur.empty(); ++q) {
            const bool lastq = q == P.quick_depth - 1;
            // bound per mesh (cuda/la_expand.cu:434-457), last quick level only
            std::vector<float> best_ub(M.size(), 1e30f);
            if (lastq)
                for (const Item& it : cur) {
                    if (it.nl <= 0) continue;
                    float ps = 0.0f;
                    for (int l = 0; l < it.nl; ++l) ps += it.lc[l];
                    const float ub = ps + it.lc[it.nl - 1] * (float)(total_levels - it.nl);
                    if (ub < best_ub[it.mesh]) best_ub[it.mesh] = ub;
                }
            // up to three midpoint cuts per item, in axis order
            std::vector<LaCutTask> h;
            std::vector<int> h_item;
            for (size_t ii = 0; ii < cur.size(); ++ii) {
                const Item& it = cur[ii];
                const IPart& wp = it.parts[it.np - 1];
                REQUIRE(wp.nv > 0 && wp.nt > 0 && wp.nv <= 1000000 && wp.nt <= 1000000,
                        "lookahead: invalid part mesh");
                bool any = false;
                for (int a = 0; a < 3; ++a) any |= wp.hi[a] - wp.lo[a] > lim;
                if (!any) {
                    Item x = it;
                    for (int l = x.nl; l < total_levels && l < LA_MAX_LEVELS; ++l) x.lc[l] = 0.0f;
                    x.nl = std::min(total_levels, LA_MAX_LEVELS);
                    extra.push_back(x);
                    continue;
                }
                for (int a = 0; a < 3; ++a) {
                    if (wp.hi[a] - wp.lo[a] <= lim) continue;
                    LaCutTask t{};
                    t.v = wp.v; t.t = wp.t; t.nv = wp.nv; t.nt = wp.nt;
                    t.plane[a] = 1.f;
                    t.plane[3] = -((wp.lo[a] + wp.hi[a]) * 0.5f);
                    h.push_back(t);
                    h_item.push_back((int)ii);
                }
            }
            LA_TRY(cut_batch(h));
            std::vector<int> pick(cur.size(), -1);
            std::vector<float> pick_cost(cur.size(), 1e30f);
            for (size_t u = 0; u < h.size(); ++u) {
                const LaCutTask& t = h[u];
                if (t.status != 0 || !t.pv || !t.nvp) continue;
                to_free.push_back(t.pv);
                to_free.push_back(t.nvp);
                const float pv = (float)t.pvol, nv = (float)t.nvol;
                const float cost = std::max(pv, nv) / (pv + nv + 1e-10f);
                if (cost < pick_cost[h_item[u]]) {
                    pick_cost[h_item[u]] = cost;
                    pick[h_item[u]] = (int)u;
                }
            }
            // children and the hulls of their two new parts
            std::vector<Item> nxt;
            std::vector<int> nxt_c;
            std::vector<LaHullTask> ht;
            for (size_t ii = 0; ii < cur.size(); ++ii) {
                if (pick[ii] < 0) continue;
                const Item& par = cur[ii];
                REQUIRE(par.np + 1 <= LA_MAX_PARTS, "lookahead: item part capacity exceeded");
                const LaCutTask& t = h[pick[ii]];
                Item ch = par;
                IPart pp, nn;
                pp.v = t.pv; pp.t = t.pt; pp.nv = t.pnv; pp.nt = t.pnt; pp.mvol = t.pvol;
                nn.v = t.nvp; nn.t = t.ntp; nn.nv = t.nnv; nn.nt = t.nnt; nn.mvol = t.nvol;
                pp.cc = nn.cc = par.parts[par.np - 1].cc;
                for (int j = 0; j < 3; ++j) {
                    pp.lo[j] = t.plo[j]; pp.hi[j] = t.phi[j];
                    nn.lo[j] = t.nlo[j]; nn.hi[j] = t.nhi[j];
                }
                ch.parts[par.np - 1] = pp;
                ch.parts[par.np] = nn;
                ch.np = par.np + 1;
                // pruning threshold (cuda/la_expand.cu:380-399)
                double gap = 1e30;
                const float ub = best_ub[par.mesh];
                if (ub < 1e29f && par.nl > 0 && par.nl < total_levels) {
                    float ps = 0.0f;
                    for (int l = 0; l < par.nl; ++l) ps += par.lc[l];
                    const float rv_thr = ub - ps;
                    if (rv_thr <= 0.0f) gap = -1.0;
                    else {
                        const float r = rv_thr / 0.3f;
                        gap = (double)(r * r * r * (4.0f * LA_PI_F / 3.0f));
                    }
                }
                LaHullTask a{};
                a.pts = t.pv; a.n = t.pnv; a.prune_gap = gap; a.mesh_vol = t.pvol;
                LaHullTask b{};
                b.pts = t.nvp; b.n = t.nnv; b.prune_gap = gap; b.mesh_vol = t.nvol;
                ht.push_back(a);
                ht.push_back(b);
                nxt.push_back(ch);
                nxt_c.push_back(cur_c[ii]);
            }
            LA_TRY(hull_volumes(ht));
            for (size_t i = 0; i < nxt.size(); ++i) {
                Item& ch = nxt[i];
                ch.parts[ch.np - 2].hvol = ht[2 * i].vol;
                ch.parts[ch.np - 1].hvol = ht[2 * i + 1].vol;
                // a pruned hull's volume is a lower bound: the item is out of the
                // evaluation (cuda/la_expand.cu:408-420)
                if (ht[2 * i].pruned || ht[2 * i + 1].pruned) ch.pruned = true;
                sort_item_parts(ch);
                if (ch.nl < LA_MAX_LEVELS) {
                    ch.lc[ch.nl] = cost_rv(ch.parts[ch.np - 1].hvol, ch.parts[ch.np - 1].mvol);
                    ch.nl++;
                }
            }
            std::vector<char> has_child(M.size(), 0);
            for (const Item& ch : nxt) has_child[ch.mesh] = 1;
            for (size_t i = 0; i < cur.size(); ++i)
                if (!has_child[cur[i].mesh]) {
                    done.push_back(cur[i]);
                    done_c.push_back(cur_c[i]);
                }
            cur.swap(nxt);
            cur_c.swap(nxt_c);
            lap(S_LEVELN);
        }
        cur.insert(cur.end(), done.begin(), done.end());
        cur_c.insert(cur_c.end(), done_c.begin(), done_c.end());
        return 0;
    }

    int hull_volumes(std::vector<LaHullTask>& h) {
        if (h.empty()) return 0;
        DevBuf<LaHullTask> d;
        REQUIRE(d.alloc(ctx, h.size()) == 0, "lookahead: task allocation failed");
        HIP_TRY(hipMemcpyAsync(d.p, h.data(), sizeof(LaHullTask) * h.size(), hipMemcpyHostToDevice, ctx->stream));
        LA_TRY(tbeg());
        LA_TRY(la_launch_hull(ctx, d.p, h.data(), (int)h.size()));
        HIP_TRY(hipMemcpyAsync(h.data(), d.p, sizeof(LaHullTask) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
        LA_TRY(tend(hull_phase, 0.0));
        for (auto& x : h) stats[hull_phase][2] += 12.0 * (x.n + x.hnv + x.hnt) + 8.0;
        if (hull_stats_on())
            for (auto& x : h) {
                if (x.pruned) { ++hstat[5][0]; continue; }
                const int m = x.use_frame && x.nf > 0 ? x.nf : x.n;  // points the exact hull saw
                const int c = m <= 256 ? 0 : m <= 512 ? 1 : m <= 1024 ? 2 : m <= 2048 ? 3 : 4;
                const int b = m > 0 ? std::min(7, (8 * x.ncorner) / std::max(m, 1)) : 0;
                ++hstat[c][b];
                hin[c] += x.n;
            }
        return check_err("item hulls");
    }

    // Optional statistics of the item hulls (CUACD_HULL_STATS=1): per class
    // of points seen by the exact hull, the corners-per-point ratio in eighths.
    static bool hull_stats_on() {
        static int v = -1;
        if (v < 0) v = getenv("CUACD_HULL_STATS") ? 1 : 0;
        return v == 1;
    }
    long long hstat[6][8] = {};
    double hin[5] = {};
    long long bnb_scored = 0, bnb_skipped = 0;
    // Two-round leaf scoring with the decision-preserving bound
    // (CUACD_LEAF_BNB=0 scores every leaf).
    // items of a cutting part scored in the first leaf round, out of its n
    // items: max(3, ceil(0.3 n)) (measured best for both 16 and ~94 items per
    // part), or CUACD_LEAF_K1 (>= 1) when set. Only the split of the work
    // between the rounds changes, never a decision.
    static int leaf_k1(int n) {
        static int v = -1;
        if (v < 0) {
            const char* e = getenv("CUACD_LEAF_K1");
            v = e ? std::max(1, atoi(e)) : 0;
        }
        return v > 0 ? v : std::max(3, (3 * n + 9) / 10);
    }
    static bool leaf_bnb_on() {
        static int v = -1;
        if (v < 0) {
            const char* e = getenv("CUACD_LEAF_BNB");
            v = e ? (atoi(e) != 0) : 1;
        }
        return v == 1;
    }
    void print_hull_stats() const {
        if (!hull_stats_on()) return;
        const char* nm[5] = {"<=256", "<=512", "<=1024", "<=2048", ">2048"};
        for (int c = 0; c < 5; ++c) {
            long long tot = 0;
            for (int b = 0; b < 8; ++b) tot += hstat[c][b];
            fprintf(stderr, "[hull-stats] %-7s %9lld hulls (mean input %.0f) corner/point eighths:", nm[c], tot,
                    tot ? hin[c] / (double)tot : 0.0);
            for (int b = 0; b < 8; ++b) fprintf(stderr, " %lld", hstat[c][b]);
            fprintf(stderr, "\n");
        }
        fprintf(stderr, "[hull-stats] pruned %lld; leaves scored %lld, skipped by the bound %lld\n", hstat[5][0],
                bnb_scored, bnb_skipped);
    }

    std::vector<void*> ext_bufs;  // hipMallocAsync buffers owned by merged parts

    // ---- merge pass (reference cuda/postprocess_merge.cu, csrc/lookahead.c:988-1067)
    // One round: every pair of parts of a mesh is scored by the rv cost of its
    // merged hull (after volume and box rejects), then by max(rv, Hausdorff of
    // the concatenated mesh - minus triangles on a shared plane - against the
    // merged hull); pairs under the threshold are matched greedily by cost and
    // merged. Pair scoring runs on the GPU for all meshes at once.
    int merge(const std::vector<int>& meshes) {
        const float tau = P.threshold;
        struct MPair {
            int m, p1, p2;
            float cost;
            int hull_task;  // index into the hull tasks, -1 if rejected before
            float rv;
        };
        std::vector<MPair> pairs;
        std::vector<LaGatherTask> pg;       // hull-point concatenations
        size_t pts_total = 0;
        for (int m : meshes) {
            MeshState& S = M[m];
            const int np = (int)S.parts.size();
            for (int p1 = 1; p1 < np; ++p1)
                for (int p2 = 0; p2 < p1; ++p2) {
                    const Part& A = S.parts[p1];
                    const Part& B = S.parts[p2];
                    MPair q{m, p1, p2, 1e20f, -1, 0.f};
                    if (A.nv == 0 || B.nv == 0 || !A.v || !B.v || !A.hv || !B.hv) { pairs.push_back(q); continue; }
                    if (A.cc != 0 && B.cc != 0 && A.cc != B.cc) { pairs.push_back(q); continue; }
                    const double msum = A.mvol + B.mvol;
                    if (cost_rv(A.hvol + B.hvol, msum) >= tau) { q.cost = 1e10f; pairs.push_back(q); continue; }
                    float g2 = 0.f;
                    {
                        const float gx = std::max(0.0f, std::max(A.lo[0] - B.hi[0], B.lo[0] - A.hi[0]));
                        const float gy = std::max(0.0f, std::max(A.lo[1] - B.hi[1], B.lo[1] - A.hi[1]));
                        const float gz = std::max(0.0f, std::max(A.lo[2] - B.hi[2], B.lo[2] - A.hi[2]));
                        g2 = gx * gx + gy * gy + gz * gz;
                    }
                    const float half = tau * 0.5f;
                    if (g2 >= half * half) { q.cost = 1e9f; pairs.push_back(q); continue; }
                    q.hull_task = (int)pg.size();
                    LaGatherTask gt{};
                    gt.v1 = A.hv; gt.n1 = A.hnv; gt.v2 = B.hv; gt.n2 = B.hnv;
                    gt.dv = (float*)(uintptr_t)pts_total;  // offset, patched below
                    pts_total += (size_t)(A.hnv + B.hnv);
                    pg.push_back(gt);
                    pairs.push_back(q);
                }
        }
        std::vector<LaHullTask> ht(pg.size());
        DevBuf<float> cat;
        if (!pg.empty()) {
            REQUIRE(cat.alloc(ctx, 3 * pts_total + 3) == 0, "lookahead: merge buffers allocation failed");
            for (size_t i = 0; i < pg.size(); ++i) {
                pg[i].dv = cat.p + 3 * (size_t)(uintptr_t)pg[i].dv;
                ht[i] = LaHullTask{};
                ht[i].pts = pg[i].dv;
                ht[i].n = pg[i].n1 + pg[i].n2;
                ht[i].emit = 1;
                ht[i].prune_gap = 1e30;
            }
            DevBuf<LaGatherTask> dg;
            REQUIRE(dg.alloc(ctx, pg.size()) == 0, "lookahead: merge buffers allocation failed");
            HIP_TRY(hipMemcpyAsync(dg.p, pg.data(), sizeof(LaGatherTask) * pg.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            LA_TRY(la_launch_gather(ctx, dg.p, (int)pg.size()));
            hull_phase = PH_MERGE;
            LA_TRY(hull_volumes(ht));
            hull_phase = PH_IHULL;
        }
        // rv cost of the merged hull; the survivors go on to the Hausdorff stage
        std::vector<int> hd_pairs;
        for (size_t i = 0; i < pairs.size(); ++i) {
            MPair& q = pairs[i];
            if (q.hull_task < 0) continue;
            const LaHullTask& h = ht[q.hull_task];
            const Part& A = M[q.m].parts[q.p1];
            const Part& B = M[q.m].parts[q.p2];
            if (!h.hv) { q.cost = 1e20f; continue; }
            q.rv = cost_rv(h.vol, A.mvol + B.mvol);
            q.cost = q.rv;
            if (q.rv >= tau) { to_free.push_back(h.hv); ht[q.hull_task].hv = nullptr; continue; }
            hd_pairs.push_back((int)i);
        }
        if (!hd_pairs.empty()) {
            // host copies of the parts' hulls for the shared-plane search (one readback)
            std::vector<std::pair<const void*, size_t>> hsrc;
            for (size_t k = 0; k < hd_pairs.size(); ++k) {
                const MPair& q = pairs[hd_pairs[k]];
                const Part& A = M[q.m].parts[q.p1];
                const Part& B = M[q.m].parts[q.p2];
                hsrc.push_back({A.hv, 12ull * A.hnv});
                hsrc.push_back({A.ht, 12ull * A.hnt});
                hsrc.push_back({B.hv, 12ull * B.hnv});
            }
            std::vector<size_t> hoff;
            int grc = 0;
            const char* hh = gather_to_host(hsrc, hoff, &grc);
            if (grc) return grc;
            std::vector<LaGatherTask> mg(hd_pairs.size());
            size_t vtot = 0, ttot = 0;
            for (size_t k = 0; k < hd_pairs.size(); ++k) {
                const MPair& q = pairs[hd_pairs[k]];
                const Part& A = M[q.m].parts[q.p1];
                const Part& B = M[q.m].parts[q.p2];
                LaGatherTask g{};
                g.v1 = A.v; g.n1 = A.nv; g.v2 = B.v; g.n2 = B.nv;
                g.t1 = A.t; g.m1 = A.nt; g.t2 = B.t; g.m2 = B.nt;
                float pl[4];
                g.has_plane = shared_plane((const float*)(hh + hoff[3 * k]), A.hnv, (const int*)(hh + hoff[3 * k + 1]),
                                           A.hnt, (const float*)(hh + hoff[3 * k + 2]), B.hnv, pl);
                for (int j = 0; j < 4; ++j) g.plane[j] = pl[j];
                g.dv = (float*)(uintptr_t)vtot;
                g.dt = (int*)(uintptr_t)ttot;
                vtot += (size_t)(A.nv + B.nv);
                ttot += (size_t)(A.nt + B.nt);
                mg[k] = g;
            }
            DevBuf<float> mv;
            DevBuf<int> mt;
            DevBuf<LaGatherTask> dg;
            REQUIRE(mv.alloc(ctx, 3 * vtot + 3) == 0 && mt.alloc(ctx, 3 * ttot + 3) == 0 &&
                        dg.alloc(ctx, mg.size()) == 0,
                    "lookahead: merge buffers allocation failed");
            for (auto& g : mg) {
                g.dv = mv.p + 3 * (size_t)(uintptr_t)g.dv;
                g.dt = mt.p + 3 * (size_t)(uintptr_t)g.dt;
            }
            HIP_TRY(hipMemcpyAsync(dg.p, mg.data(), sizeof(LaGatherTask) * mg.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
            LA_TRY(la_launch_gather(ctx, dg.p, (int)mg.size()));
            HIP_TRY(hipMemcpyAsync(mg.data(), dg.p, sizeof(LaGatherTask) * mg.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
            LA_TRY(check_err("merge gather"));
            std::vector<HdTask> hd;
            unsigned long long ws = 0;
            for (size_t k = 0; k < hd_pairs.size(); ++k) {
                const LaHullTask& h = ht[pairs[hd_pairs[k]].hull_task];
                const LaGatherTask& g = mg[k];
                const int cnv = g.n1 + g.n2, cnt = g.nt_out;
                HdTask a{g.dv, g.dt, h.hv, h.ht, cnv, cnt, h.hnv, h.hnt, ws, 0, 0, nullptr, nullptr, nullptr};
                ws += (hd_ws_bytes(h.hnt, cnt) + 255ull) & ~255ull;
                HdTask b{h.hv, h.ht, g.dv, g.dt, h.hnv, h.hnt, cnv, cnt, ws, 0, 0, nullptr, nullptr, nullptr};
                ws += (hd_ws_bytes(cnt, h.hnt) + 255ull) & ~255ull;
                hd.push_back(a);
                hd.push_back(b);
            }
            DevBuf<float> d2;
            DevBuf<HdTask> dh;
            DevBuf<char> dws;
            REQUIRE(d2.alloc(ctx, hd.size()) == 0 && dh.alloc(ctx, hd.size()) == 0 && dws.alloc(ctx, ws + 256) == 0,
                    "lookahead: merge Hausdorff buffers allocation failed");
            for (size_t k = 0; k < hd.size(); ++k) hd[k].out_d2 = d2.p + k;
            HIP_TRY(hipMemcpyAsync(dh.p, hd.data(), sizeof(HdTask) * hd.size(), hipMemcpyHostToDevice, ctx->stream));
            LA_TRY(tbeg());
            LA_TRY(hd_launch_tasks(ctx, dh.p, (int)hd.size(), dws.p));
            {
                double by = 0;
                for (auto& x : hd) by += 12.0 * (x.snv + x.snt + x.tnv + x.tnt) + 4.0;
                LA_TRY(tend(PH_MERGE, by));
            }
            std::vector<float> r(hd.size());
            HIP_TRY(hipMemcpyAsync(r.data(), d2.p, 4 * hd.size(), hipMemcpyDeviceToHost, ctx->stream));
            LA_TRY(check_err("merge hausdorff"));
            for (size_t k = 0; k < hd_pairs.size(); ++k) {
                MPair& q = pairs[hd_pairs[k]];
                // an empty concatenation (every triangle on the shared plane) scores 0
                const float h = mg[k].nt_out > 0 ? std::sqrt(std::max(r[2 * k], r[2 * k + 1])) : 0.0f;
                q.cost = std::max(q.rv, h);
                if (q.cost >= tau) {
                    to_free.push_back(ht[q.hull_task].hv);
                    ht[q.hull_task].hv = nullptr;
                }
            }
        }
        // greedy matching per mesh: repeatedly the cheapest pair under the
        // threshold, then drop its two parts. Ties go the way the reference's
        // one-block matching resolves them (cuda/postprocess_merge.cu:404-480):
        // 256 threads each keep the first minimum of their stride of the pair
        // list, and a tree reduction keeps the left operand unless the right one
        // is strictly smaller.
        size_t b = 0;
        while (b < pairs.size()) {
            const int m = pairs[b].m;
            size_t e = b;
            while (e < pairs.size() && pairs[e].m == m) ++e;
            MeshState& S = M[m];
            const int np = (int)S.parts.size();
            std::vector<char> used(np, 0);
            std::vector<std::pair<int, int>> match;
            constexpr int MT = 256;
            float vals[MT];
            long long idxs[MT];
            for (;;) {
                for (int t = 0; t < MT; ++t) {
                    vals[t] = tau;
                    idxs[t] = -1;
                    for (size_t i = b + t; i < e; i += MT) {
                        const MPair& q = pairs[i];
                        if (used[q.p1] || used[q.p2]) continue;
                        if (q.cost < vals[t]) { vals[t] = q.cost; idxs[t] = (long long)i; }
                    }
                }
                for (int off = MT / 2; off > 0; off >>= 1)
                    for (int t = 0; t < off; ++t)
                        if (vals[t + off] < vals[t]) { vals[t] = vals[t + off]; idxs[t] = idxs[t + off]; }
                if (idxs[0] < 0 || vals[0] >= tau) break;
                const size_t bi = (size_t)idxs[0];
                used[pairs[bi].p1] = used[pairs[bi].p2] = 1;
                match.push_back({(int)bi, 0});
            }
            // apply: merged mesh = A's then B's vertices and triangles; hull = merged hull
            std::vector<LaGatherTask> ag;
            for (auto& mm : match) {
                const MPair& q = pairs[mm.first];
                const Part& A = S.parts[q.p1];
                const Part& B = S.parts[q.p2];
                LaGatherTask g{};
                g.v1 = A.v; g.n1 = A.nv; g.v2 = B.v; g.n2 = B.nv;
                g.t1 = A.t; g.m1 = A.nt; g.t2 = B.t; g.m2 = B.nt;
                void* buf = nullptr;
                HIP_TRY(hipMallocAsync(&buf, 12ull * (A.nv + B.nv) + 16 + 12ull * (A.nt + B.nt) + 16, ctx->stream));
                ext_bufs.push_back(buf);
                g.dv = (float*)buf;
                g.dt = (int*)((char*)buf + ((12ull * (A.nv + B.nv) + 15) & ~15ull));
                ag.push_back(g);
            }
            if (!ag.empty()) {
                DevBuf<LaGatherTask> dg;
                REQUIRE(dg.alloc(ctx, ag.size()) == 0, "lookahead: merge buffers allocation failed");
                HIP_TRY(hipMemcpyAsync(dg.p, ag.data(), sizeof(LaGatherTask) * ag.size(), hipMemcpyHostToDevice,
                                       ctx->stream));
                LA_TRY(la_launch_gather(ctx, dg.p, (int)ag.size()));
                LA_TRY(check_err("merge apply"));
            }
            std::vector<char> dead(np, 0);
            for (size_t k = 0; k < match.size(); ++k) {
                MPair& q = pairs[match[k].first];
                Part& A = S.parts[q.p1];
                Part& B = S.parts[q.p2];
                LaHullTask& h = ht[q.hull_task];
                Part mp;
                mp.v = ag[k].dv; mp.t = ag[k].dt; mp.nv = A.nv + B.nv; mp.nt = A.nt + B.nt;
                mp.hv = h.hv; mp.ht = h.ht; mp.hnv = h.hnv; mp.hnt = h.hnt; mp.own_hull = true;
                mp.mvol = A.mvol + B.mvol;
                mp.hvol = h.vol;
                mp.haus = 0.0f;
                mp.cc = A.cc ? A.cc : B.cc;
                mp.own = false;  // an ext buffer, freed with the engine
                for (int j = 0; j < 3; ++j) { mp.lo[j] = std::min(A.lo[j], B.lo[j]); mp.hi[j] = std::max(A.hi[j], B.hi[j]); }
                h.hv = nullptr;  // adopted
                drop_part(A);
                drop_part(B);
                A = mp;
                dead[q.p2] = 1;
            }
            std::vector<Part> keep;
            for (int i = 0; i < np; ++i)
                if (!dead[i]) keep.push_back(S.parts[i]);
            S.parts.swap(keep);
            b = e;
        }
        // merged hulls of pairs not chosen
        for (auto& h : ht)
            if (h.hv) to_free.push_back(h.hv);
        return flush_frees();
    }

    // Shared plane (reference cuda/merge.cuh:111-165): the first face of A's
    // hull with A's hull on its inner side and B's hull within 1e-3 of it.
    // av/at: A's hull (host copy), bv: B's hull vertices.
    bool shared_plane(const float* av, int a_hnv, const int* at, int a_hnt, const float* bv, int b_hnv, float* out) {
        const float tol = 1e-3f;
        for (int t = 0; t < a_hnt; ++t) {
            const float* a = &av[3 * at[3 * t]];
            const float* bq = &av[3 * at[3 * t + 1]];
            const float* c = &av[3 * at[3 * t + 2]];
            const float e1x = bq[0] - a[0], e1y = bq[1] - a[1], e1z = bq[2] - a[2];
            const float e2x = c[0] - a[0], e2y = c[1] - a[1], e2z = c[2] - a[2];
            float nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
            const float len = std::sqrt(nx * nx + ny * ny + nz * nz);
            if (len < 1e-20f) continue;
            const float inv = 1.0f / len;
            nx *= inv; ny *= inv; nz *= inv;
            const float d = nx * a[0] + ny * a[1] + nz * a[2];
            // B on the plane first: it fails at once unless B is flat, while A
            // (whose own hull face this is) always passes its test
            bool ok = true;
            for (int j = 0; j < b_hnv && ok; ++j)
                if (std::fabs(nx * bv[3 * j] + ny * bv[3 * j + 1] + nz * bv[3 * j + 2] - d) > tol) ok = false;
            for (int j = 0; j < a_hnv && ok; ++j)
                if (nx * av[3 * j] + ny * av[3 * j + 1] + nz * av[3 * j + 2] - d > tol) ok = false;
            if (ok) {
                out[0] = nx; out[1] = ny; out[2] = nz; out[3] = d;
                return true;
            }
        }
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return false;
    }

    // append the iteration's level-0 record: samples of each cutting part with
    // at least two candidates, its candidates' frame planes and path costs
    int record_level0(const std::vector<std::pair<int, int>>& cutting, const std::vector<std::vector<int>>& rk,
                      const std::vector<std::vector<float>>& rpl, const std::vector<std::vector<float>>& best) {
        LaRecord* R = (LaRecord*)ctx->record;
        std::vector<int> sel;
        for (size_t c = 0; c < cutting.size(); ++c)
            if (rk[c].size() >= 2) sel.push_back((int)c);
        if (sel.empty()) return 0;
        std::vector<MlpSamplePart> sp(sel.size());
        for (size_t i = 0; i < sel.size(); ++i) {
            const Part& p = M[cutting[sel[i]].first].parts[cutting[sel[i]].second];
            sp[i] = MlpSamplePart{p.v, p.t, p.nv, p.nt, {p.lo[0], p.lo[1], p.lo[2]}, {p.hi[0], p.hi[1], p.hi[2]}};
        }
        DevBuf<float> d;
        REQUIRE(d.alloc(ctx, sel.size() * MLP_SAMPLES * 3) == 0, "lookahead: record buffer allocation failed");
        LA_TRY(mlp_sample_dev(ctx, sp.data(), (int)sel.size(), d.p));
        const size_t base = R->pts.size();
        R->pts.resize(base + sel.size() * MLP_SAMPLES * 3);
        HIP_TRY(hipMemcpyAsync(R->pts.data() + base, d.p, sizeof(float) * sel.size() * MLP_SAMPLES * 3,
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        for (size_t i = 0; i < sel.size(); ++i) {
            const int c = sel[i];
            const Part& p = M[cutting[c].first].parts[cutting[c].second];
            for (size_t j = 0; j < rk[c].size(); ++j) {
                float fp[4];
                frame_plane(p.lo, p.hi, &rpl[c][4 * j], fp);
                R->planes.insert(R->planes.end(), fp, fp + 4);
                const int k = rk[c][j];
                R->target.push_back(k < (int)best[c].size() ? best[c][k] : 1e30f);
                R->part_of.push_back(R->n_parts);
            }
            R->n_parts++;
        }
        return 0;
    }

    // ---- observability (reference csrc/lookahead.c:284-436, 745-892) --------
    // CUACD_LEAK_BISECT: after each phase, the live block counts (allocations
    // minus frees) of both heaps and their change since the last snapshot.
    static bool bisect_on() {
        static int v = -1;
        if (v < 0) {
            const char* e = getenv("CUACD_LEAK_BISECT");
            v = (e && e[0] && e[0] != '0') ? 1 : 0;
        }
        return v == 1;
    }
    long long bisect_prev[2] = {0, 0};
    int bisect_snap(int iter, const char* label) {
        if (!bisect_on()) return 0;
        unsigned long long st[6];
        LA_TRY(gpu_heap_stats(ctx, st));
        const long long h = (long long)(st[1] - st[2]), sc = (long long)(st[4] - st[5]);
        fprintf(stderr,
                "[bisect iter=%d] %-24s heap_diff=%lld (%+lld) [ha=%llu hf=%llu]  scratch_diff=%lld (%+lld) "
                "[sa=%llu sf=%llu]\n",
                iter, label, h, h - bisect_prev[0], st[1], st[2], sc, sc - bisect_prev[1], st[4], st[5]);
        bisect_prev[0] = h;
        bisect_prev[1] = sc;
        return 0;
    }

    // verbose: per-part table of every active mesh after the iteration's sort
    // ('*' = at or above the threshold); CUACD_DUMP_PARTS_DIR=<dir>: every such
    // part written to <dir>/iter%03d_part%03d.bin (mesh%03d_ prefixed when the
    // batch has several meshes) as int32 {nv, nt}, float32 {mesh_vol,
    // hull_vol}, float32 verts [nv*3], int32 tris [nt*3].
    int diag_parts(const std::vector<int>& act, int iter) {
        const char* dir = getenv("CUACD_DUMP_PARTS_DIR");
        if (dir && !dir[0]) dir = nullptr;
        if (!P.verbose && !dir) return 0;
        for (int m : act) {
            const MeshState& S = M[m];
            if (P.verbose) {
                if (M.size() > 1) fprintf(stderr, "[la] mesh %d iter %d\n", m, iter);
                fprintf(stderr, "[la]   part   nv   nt  mesh_vol  hull_vol    rv_cost  hausdorff  full_cost\n");
                for (size_t i = 0; i < S.parts.size(); ++i) {
                    const Part& q = S.parts[i];
                    const float rc = cost_rv(q.hvol, q.mvol), fc = cost_full(q);
                    fprintf(stderr, "[la]   %3zu  %4d %4d  %9.6f %9.6f  %9.6f %10.6f %10.6f%s\n", i, q.nv, q.nt,
                            (float)q.mvol, (float)q.hvol, rc, q.haus, fc, fc >= P.threshold ? " *" : "");
                }
            }
        }
        if (!dir) return 0;
        std::vector<std::pair<const void*, size_t>> src;
        std::vector<std::pair<int, int>> who;
        for (int m : act)
            for (size_t i = 0; i < M[m].parts.size(); ++i) {
                const Part& q = M[m].parts[i];
                if (q.nv <= 0 || q.nt <= 0 || !q.v || !q.t) continue;
                src.push_back({q.v, 12ull * q.nv});
                src.push_back({q.t, 12ull * q.nt});
                who.push_back({m, (int)i});
            }
        std::vector<size_t> offs;
        int rc = 0;
        const char* hb = gather_to_host(src, offs, &rc);
        if (rc) return rc;
        for (size_t k = 0; k < who.size(); ++k) {
            const Part& q = M[who[k].first].parts[who[k].second];
            char path[1200];
            if (M.size() > 1)
                snprintf(path, sizeof(path), "%s/mesh%03d_iter%03d_part%03d.bin", dir, who[k].first, iter,
                         who[k].second);
            else
                snprintf(path, sizeof(path), "%s/iter%03d_part%03d.bin", dir, iter, who[k].second);
            FILE* f = fopen(path, "wb");
            if (!f) continue;
            const int hdr[2] = {q.nv, q.nt};
            const float vols[2] = {(float)q.mvol, (float)q.hvol};
            fwrite(hdr, sizeof(int), 2, f);
            fwrite(vols, sizeof(float), 2, f);
            fwrite(hb + offs[2 * k], 1, 12ull * q.nv, f);
            fwrite(hb + offs[2 * k + 1], 1, 12ull * q.nt, f);
            fclose(f);
        }
        return 0;
    }

    // ---- one search iteration over the meshes in `act`; returns meshes still active
    int iterate(std::vector<int>& act, int iter) {
        const auto iter_t0 = std::chrono::steady_clock::now();
        LA_TRY(bisect_snap(iter, "iter_start"));
        if (P.dc_per_iter) LA_TRY(components(act));
        LA_TRY(bisect_snap(iter, "components"));
        LA_TRY(hull_parts(act));
        LA_TRY(bisect_snap(iter, "part_hulls"));
        LA_TRY(hausdorff_parts(act));
        LA_TRY(bisect_snap(iter, "hausdorff_parts"));
        // sort by full cost, pick the cutting parts (cuda/la_lifecycle.cu:90-144)
        std::vector<std::pair<int, int>> cutting;   // (mesh, part index)
        std::vector<int> cut_begin(M.size(), 0), cut_count(M.size(), 0);
        std::vector<int> next;
        for (int m : act) {
            MeshState& S = M[m];
            std::stable_sort(S.parts.begin(), S.parts.end(),
                             [](const Part& a, const Part& b) { return cost_full(a) < cost_full(b); });
            const int np = (int)S.parts.size();
            int first = np;
            for (int i = 0; i < np; ++i)
                if (cost_full(S.parts[i]) >= P.threshold) { first = i; break; }
            int start = first;
            if (np - first > P.max_n_cutting) start = np - P.max_n_cutting;
            cut_begin[m] = (int)cutting.size();
            for (int i = start; i < np && cut_count[m] < P.max_n_cutting; ++i)
                if (cost_full(S.parts[i]) >= P.threshold) {
                    cutting.push_back({m, i});
                    ++cut_count[m];
                }
            if (cut_count[m] == 0) S.done = true;
            else next.push_back(m);
        }
        LA_TRY(diag_parts(act, iter));
        if (P.verbose)
            for (int m : act) {
                if (M.size() > 1) fprintf(stderr, "[la] mesh %d", m);
                fprintf(stderr, "[la] iter %d: n_cutting=%d", iter, cut_count[m]);
                if (P.verbose >= 2) {
                    fprintf(stderr, " cutting_indices=[");
                    for (int c = 0; c < cut_count[m]; ++c)
                        fprintf(stderr, "%s%d", c ? " " : "", cutting[cut_begin[m] + c].second);
                    fprintf(stderr, "]");
                }
                fprintf(stderr, "\n");
            }
        act = next;
        lap(S_PRE);
        if (cutting.empty()) return 0;
        for (int m : act) M[m].iters = iter + 1;
        float med = P.threshold * 0.25f;
        if (med > 0.015f) med = 0.015f;
        // concave-edge planes
        std::vector<std::vector<float>> eplanes(cutting.size());
        std::vector<int> n_edge(M.size(), 0);
        if (P.n_concave > 0 && iter < P.concave_iters) {
            LA_TRY(concave_planes(cutting, eplanes));
            for (size_t c = 0; c < cutting.size(); ++c)
                n_edge[cutting[c].first] = std::max(n_edge[cutting[c].first], (int)eplanes[c].size() / 4);
        }
        lap(S_CONC);
        LA_TRY(bisect_snap(iter, "concave_edges"));
        const int total_levels = P.depth + P.quick_depth;
        // level-0 candidates per cutting part, for the training record
        std::vector<std::vector<int>> rec_k(cutting.size());
        std::vector<std::vector<float>> rec_pl(cutting.size());
        // the iteration's store of halves built by full cuts (and the hull
        // meshes of the level-0 ones)
        struct Half {
            float* v;
            bool keep;
            float* hv;
        };
        std::vector<Half> store;
        // an error return leaves the halves not yet adopted to the free list
        // (the caller flushes it); the normal path releases them itself
        struct StoreGuard {
            std::vector<Half>& st;
            std::vector<void*>& tf;
            bool released = false;
            ~StoreGuard() {
                if (released) return;
                for (const Half& hf : st)
                    if (!hf.keep) {
                        tf.push_back(hf.v);
                        if (hf.hv) tf.push_back(hf.hv);
                    }
            }
        } store_guard{store, to_free};
        std::vector<Item> extra;       // leaves for parts too small to cut
        // level-0 items per cutting part and initial cut, -1 = no item
        std::vector<std::vector<int>> l0(cutting.size());
        std::vector<Item> cur;         // items entering the next level
        // seeds
        std::vector<Item> seeds(cutting.size());
        for (size_t c = 0; c < cutting.size(); ++c) {
            const Part& p = M[cutting[c].first].parts[cutting[c].second];
            Item& it = seeds[c];
            it.mesh = cutting[c].first;
            it.src = (int)c - cut_begin[it.mesh];
            it.cut0 = 0;
            it.np = 1;
            IPart& q = it.parts[0];
            q.v = p.v; q.t = p.t; q.nv = p.nv; q.nt = p.nt; q.mvol = p.mvol; q.hvol = p.hvol; q.cc = p.cc;
            for (int j = 0; j < 3; ++j) { q.lo[j] = p.lo[j]; q.hi[j] = p.hi[j]; }
            REQUIRE(p.nv > 0 && p.nt > 0 && p.nv <= 1000000 && p.nt <= 1000000, "lookahead: invalid part mesh");
            l0[c].assign(P.width + n_edge[it.mesh], -1);
        }
        std::vector<Item> level0_items;  // all valid level-0 items (index = l0 entries)
        // items of a mesh whose next level produced nothing stay its leaves
        // (reference: the level loop breaks with next_n == 0, csrc/lookahead.c:579-581)
        std::vector<Item> leaves;
        std::vector<int> leaves_c;
        // items of a mesh whose level produced no child: the mesh's levels end
        // there, and these items go on to the quick levels (or are its leaves)
        std::vector<Item> stay;
        std::vector<int> stay_c;
        struct Leaf {
            int c, cut0;
            float path;
            bool pruned;
        };
        std::vector<Leaf> leaf_recs;
        cur = seeds;
        std::vector<int> cur_c(cutting.size());  // cutting slot of each item
        for (size_t c = 0; c < cutting.size(); ++c) cur_c[c] = (int)c;
        for (int d = 0; d < P.depth; ++d) {
            const bool last = d == P.depth - 1;
            // the last full level is scored without building its children
            // unless quick levels follow (they cut the children's parts)
            const bool full = !last || P.depth == 1 || P.quick_depth > 0;
            // candidates
            struct Cand {
                int item, k;
                float pl[4];
            };
            std::vector<Cand> cands;
            for (size_t ii = 0; ii < cur.size(); ++ii) {
                const Item& it = cur[ii];
                const int w = d == 0 ? P.width : P.width2;
                const int ne = d == 0 ? n_edge[it.mesh] : 0;
                const int tw = w + ne;
                const IPart& wp = it.parts[it.np - 1];
                REQUIRE(it.np + 1 <= LA_MAX_PARTS, "lookahead: item part capacity exceeded");
                for (int k = 0; k < tw; ++k) {
                    Cand cd;
                    cd.item = (int)ii;
                    cd.k = k;
                    if (k >= w) {
                        const std::vector<float>& ep = eplanes[cur_c[ii]];
                        const int e = k - w;
                        if (4 * e + 3 < (int)ep.size()) {
                            for (int j = 0; j < 4; ++j) cd.pl[j] = ep[4 * e + j];
                        } else {
                            cd.pl[0] = cd.pl[1] = cd.pl[2] = cd.pl[3] = 0.f;
                        }
                        if (cd.pl[0] == 0.f && cd.pl[1] == 0.f && cd.pl[2] == 0.f) continue;
                    }
                    const float eps = 1e-6f;
                    const bool small = (wp.hi[0] - wp.lo[0] <= 2.f * med + eps) && (wp.hi[1] - wp.lo[1] <= 2.f * med + eps) &&
                                       (wp.hi[2] - wp.lo[2] <= 2.f * med + eps);
                    if (small) {
                        if (k == 0) {
                            Item x = it;
                            for (int l = x.nl; l < total_levels && l < LA_MAX_LEVELS; ++l) x.lc[l] = 0.0f;
                            x.nl = std::min(total_levels, LA_MAX_LEVELS);
                            extra.push_back(x);
                        }
                        continue;
                    }
                    if (k < w) {
                        const AxisCut a = axis_cut(w, k);
                        const float lov = wp.lo[a.axis] + med, hiv = wp.hi[a.axis] - med;
                        if (lov >= hiv) continue;
                        const float frac = (a.slice + 1.f) / (a.cpa + 1.f);
                        const float pos = lov + frac * (hiv - lov);
                        cd.pl[0] = a.axis == 0 ? 1.f : 0.f;
                        cd.pl[1] = a.axis == 1 ? 1.f : 0.f;
                        cd.pl[2] = a.axis == 2 ? 1.f : 0.f;
                        cd.pl[3] = -pos;
                    }
                    cands.push_back(cd);
                }
            }
            // learned proposal: per cutting part, only the top_k candidates by
            // predicted score are evaluated exactly; the others are the
            // fallback for a part whose kept candidates yield no cut
            std::vector<Cand> rest;
            if (d == 0 && ctx->record)
                for (const Cand& cd : cands) {
                    rec_k[cur_c[cd.item]].push_back(cd.k);
                    rec_pl[cur_c[cd.item]].insert(rec_pl[cur_c[cd.item]].end(), cd.pl, cd.pl + 4);
                }
            if (d == 0 && ctx->proposal && ctx->proposal_topk > 0 && !cands.empty()) {
                const int P = (int)cur.size(), C = (int)cands.size();
                std::vector<MlpSamplePart> sp(P);
                for (int i = 0; i < P; ++i) {
                    const IPart& wp = cur[i].parts[cur[i].np - 1];
                    sp[i] = MlpSamplePart{wp.v, wp.t, wp.nv, wp.nt, {wp.lo[0], wp.lo[1], wp.lo[2]},
                                          {wp.hi[0], wp.hi[1], wp.hi[2]}};
                }
                // planes in their part's frame: unit normal, offset
                std::vector<float> pl(4 * (size_t)C);
                std::vector<int> po(C);
                for (int q = 0; q < C; ++q) {
                    const IPart& wp = cur[cands[q].item].parts[cur[cands[q].item].np - 1];
                    frame_plane(wp.lo, wp.hi, cands[q].pl, &pl[4 * q]);
                    po[q] = cands[q].item;
                }
                DevBuf<float> dpts, dpl, dsc;
                DevBuf<int> dpo;
                REQUIRE(dpts.alloc(ctx, (size_t)P * MLP_SAMPLES * 3) == 0 && dpl.alloc(ctx, 4 * (size_t)C) == 0 &&
                            dsc.alloc(ctx, C) == 0 && dpo.alloc(ctx, C) == 0,
                        "lookahead: proposal buffers allocation failed");
                HIP_TRY(hipMemcpyAsync(dpl.p, pl.data(), 16 * (size_t)C, hipMemcpyHostToDevice, ctx->stream));
                HIP_TRY(hipMemcpyAsync(dpo.p, po.data(), 4 * (size_t)C, hipMemcpyHostToDevice, ctx->stream));
                LA_TRY(tbeg());
                LA_TRY(mlp_sample_dev(ctx, sp.data(), P, dpts.p));
                LA_TRY(mlp_forward_dev(ctx, ctx->proposal, dpts.p, P, dpl.p, dpo.p, C, dsc.p));
                LA_TRY(tend(PH_PROP, 12.0 * P * MLP_SAMPLES + 20.0 * C, 2));
                std::vector<float> score(C);
                HIP_TRY(hipMemcpyAsync(score.data(), dsc.p, 4 * (size_t)C, hipMemcpyDeviceToHost, ctx->stream));
                LA_TRY(check_err("proposal"));
                std::vector<std::vector<int>> by_item(P);
                for (int q = 0; q < C; ++q) by_item[cands[q].item].push_back(q);
                std::vector<Cand> kept;
                for (int i = 0; i < P; ++i) {
                    std::vector<int>& v = by_item[i];
                    std::stable_sort(v.begin(), v.end(), [&](int x, int y) { return score[x] < score[y]; });
                    for (int j = 0; j < (int)v.size(); ++j)
                        (j < ctx->proposal_topk ? kept : rest).push_back(cands[v[j]]);
                }
                proposal_kept += (long long)kept.size();
                proposal_total += (long long)C;
                cands.swap(kept);
            }
            // build the children: halves + volumes
            std::vector<Item> nxt;
            std::vector<int> nxt_c;
            std::vector<LaHullTask> ht;
            std::vector<std::pair<int, int>> ht_who;  // (child, which half)
            std::vector<void*> pts_free;
            // last-level children are scored, not built: parent, cut, volumes
            struct SRec {
                int item, k;
                double pvol, nvol, phv, nhv;
                bool pruned;
            };
            std::vector<SRec> srec;
            auto build_children = [&](const std::vector<Cand>& cs) -> int {
                    std::vector<LaCutTask> h(cs.size());
                    for (size_t q = 0; q < cs.size(); ++q) {
                        const IPart& wp = cur[cs[q].item].parts[cur[cs[q].item].np - 1];
                        LaCutTask t{};
                        t.v = wp.v; t.t = wp.t; t.nv = wp.nv; t.nt = wp.nt;
                        for (int j = 0; j < 4; ++j) t.plane[j] = cs[q].pl[j];
                        h[q] = t;
                    }
                    LA_TRY(cut_batch(h));
                    // level-0 halves keep their hull mesh (the two of the applied
                    // cut become parts that need it) while the meshes' worst-case
                    // blocks (n vertices, 2n triangles, 12 B each, plus the header,
                    // rounded up to the power-of-two block) fit in a quarter of the
                    // pool's unused space; the emission is optional (emit 2): a
                    // mesh that finds no room is built by the part-hull phase
                    int emit0 = 0;
                    if (d == 0) {
                        double worst = 0.0;
                        for (const LaCutTask& t : h)
                            if (t.status == 0 && t.pv && t.nvp)
                                worst += (double)pow2_ceil_ull(36ull * t.pnv + 288) +
                                         (double)pow2_ceil_ull(36ull * t.nnv + 288);
                        const double room = (double)ctx->pool_bytes - (double)gpu_pool_usage(ctx);
                        emit0 = worst <= 0.25 * room ? 2 : 0;
                    }
                    for (size_t q = 0; q < cs.size(); ++q) {
                        const LaCutTask& t = h[q];
                        if (t.status != 0 || !t.pv || !t.nvp) continue;
                        const Item& par = cur[cs[q].item];
                        Item ch = par;
                        IPart pp, nn;
                        pp.v = t.pv; pp.t = t.pt; pp.nv = t.pnv; pp.nt = t.pnt; pp.mvol = t.pvol;
                        nn.v = t.nvp; nn.t = t.ntp; nn.nv = t.nnv; nn.nt = t.nnt; nn.mvol = t.nvol;
                        pp.cc = nn.cc = par.parts[par.np - 1].cc;
                        for (int j = 0; j < 3; ++j) {
                            pp.lo[j] = t.plo[j]; pp.hi[j] = t.phi[j];
                            nn.lo[j] = t.nlo[j]; nn.hi[j] = t.nhi[j];
                        }
                        pp.store = (int)store.size();
                        store.push_back({t.pv, false, nullptr});
                        nn.store = (int)store.size();
                        store.push_back({t.nvp, false, nullptr});
                        ch.parts[par.np - 1] = pp;
                        ch.parts[par.np] = nn;
                        ch.np = par.np + 1;
                        if (par.nl == 0) ch.cut0 = cs[q].k;
                        const int ci = (int)nxt.size();
                        nxt.push_back(ch);
                        nxt_c.push_back(cur_c[cs[q].item]);
                        LaHullTask a{};
                        a.pts = t.pv; a.n = t.pnv; a.prune_gap = 1e30; a.emit = emit0;
                        LaHullTask b{};
                        b.pts = t.nvp; b.n = t.nnv; b.prune_gap = 1e30; b.emit = emit0;
                        ht.push_back(a);
                        ht_who.push_back({ci, 0});
                        ht.push_back(b);
                        ht_who.push_back({ci, 1});
                    }
                return 0;
            };
            if (full) {
                LA_TRY(build_children(cands));
                // exact fallback of the learned proposal: a cutting part whose
                // kept candidates produced no cut gets the others
                if (d == 0 && !rest.empty()) {
                    std::vector<char> got(cutting.size(), 0);
                    for (int c : nxt_c) got[c] = 1;
                    std::vector<Cand> fb;
                    for (const Cand& cd : rest)
                        if (!got[cur_c[cd.item]]) fb.push_back(cd);
                    if (!fb.empty()) {
                        proposal_fallbacks += 1;
                        LA_TRY(build_children(fb));
                    }
                }
            } else {
                // branch-and-bound bound per mesh (cuda/la_expand.cu:434-457): min over the
                // items entering this level of sum(lc) + worst * (levels left)
                std::vector<float> best_ub(M.size(), 1e30f);
                for (const Item& it : cur) {
                    if (it.nl <= 0) continue;
                    float ps = 0.0f;
                    for (int l = 0; l < it.nl; ++l) ps += it.lc[l];
                    const float ub = ps + it.lc[it.nl - 1] * (float)(total_levels - it.nl);
                    if (ub < best_ub[it.mesh]) best_ub[it.mesh] = ub;
                }
                // Lower bound of every leaf of an item: the leaf's path cost with
                // the two new halves left out of its worst part (same float
                // operations, max only grows, so the bound never exceeds the cost).
                std::vector<float> item_lb(cur.size());
                for (size_t i = 0; i < cur.size(); ++i) {
                    const Item& par = cur[i];
                    float wlb = 0.0f;
                    for (int j = 0; j < par.np - 1; ++j) wlb = std::max(wlb, cost_rv(par.parts[j].hvol, par.parts[j].mvol));
                    float sum = 0.0f;
                    for (int l = 0; l < par.nl; ++l) sum += par.lc[l];
                    sum += wlb;
                    item_lb[i] = sum / (float)(par.nl + 1);
                }
                std::vector<char> mesh_has(M.size(), 0);
                // score a set of candidates: volumes and point sets, their hulls, the leaf records
                auto run_round = [&](const std::vector<int>& qs) -> int {
                    if (qs.empty()) return 0;
                    std::vector<LaScoreTask> h(qs.size());
                    for (size_t u = 0; u < qs.size(); ++u) {
                        const Cand& cd = cands[qs[u]];
                        const IPart& wp = cur[cd.item].parts[cur[cd.item].np - 1];
                        LaScoreTask t{};
                        t.v = wp.v; t.t = wp.t; t.nv = wp.nv; t.nt = wp.nt;
                        for (int j = 0; j < 4; ++j) t.plane[j] = cd.pl[j];
                        h[u] = t;
                    }
                    DevBuf<LaScoreTask> dt;
                    REQUIRE(dt.alloc(ctx, h.size()) == 0, "lookahead: task allocation failed");
                    HIP_TRY(hipMemcpyAsync(dt.p, h.data(), sizeof(LaScoreTask) * h.size(), hipMemcpyHostToDevice,
                                           ctx->stream));
                    LA_TRY(tbeg());
                    int max_nv = 0;
                    for (const LaScoreTask& t : h) max_nv = std::max(max_nv, t.nv);
                    LA_TRY(la_launch_score(ctx, dt.p, (int)h.size(), max_nv));
                    LA_TRY(tend(PH_SCORE, 0.0));
                    HIP_TRY(hipMemcpyAsync(h.data(), dt.p, sizeof(LaScoreTask) * h.size(), hipMemcpyDeviceToHost,
                                           ctx->stream));
                    HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
                    {
                        // mesh read, both point sets written, volumes and counts
                        double by = 0;
                        for (auto& x : h) by += 12.0 * (x.nv + x.nt + x.np_pts + x.nn_pts) + 16.0;
                        stats[PH_SCORE][2] += by;
                    }
                    LA_TRY(check_err("leaf scoring"));
                    std::vector<LaHullTask> rh;
                    std::vector<std::pair<int, int>> rwho;
                    const size_t s0 = srec.size();
                    for (size_t u = 0; u < qs.size(); ++u) {
                        const LaScoreTask& t = h[u];
                        const Cand& cd = cands[qs[u]];
                        if (t.ppts) pts_free.push_back(t.ppts);  // the negative set shares its chunk
                        if (t.npt == 0 || t.nnt == 0 || !t.ppts || !t.npts) continue;
                        const Item& par = cur[cd.item];
                        // pruning threshold (cuda/la_expand.cu:380-399)
                        double gap = 1e30;
                        const float ub = best_ub[par.mesh];
                        const int nl = par.nl;
                        if (ub < 1e29f && nl > 0 && nl < total_levels) {
                            float ps = 0.0f;
                            for (int l = 0; l < nl; ++l) ps += par.lc[l];
                            const float rv_thr = ub - ps;
                            if (rv_thr <= 0.0f) gap = -1.0;
                            else {
                                const float r = rv_thr / 0.3f;
                                gap = (double)(r * r * r * (4.0f * LA_PI_F / 3.0f));
                            }
                        }
                        const int ci = (int)srec.size();
                        srec.push_back({cd.item, cd.k, t.pvol, t.nvol, 0.0, 0.0, false});
                        LaHullTask a{};
                        a.pts = t.ppts; a.n = t.np_pts; a.prune_gap = gap; a.mesh_vol = t.pvol;
                        LaHullTask b{};
                        b.pts = t.npts; b.n = t.nn_pts; b.prune_gap = gap; b.mesh_vol = t.nvol;
                        rh.push_back(a);
                        rwho.push_back({ci, 0});
                        rh.push_back(b);
                        rwho.push_back({ci, 1});
                    }
                    LA_TRY(hull_volumes(rh));
                    to_free.insert(to_free.end(), pts_free.begin(), pts_free.end());
                    pts_free.clear();
                    // leaf costs: the worst part is the maximum rv cost over the
                    // inherited parts and both halves; path cost = mean level cost
                    for (size_t u = 0; u < rh.size(); ++u) {
                        SRec& r = srec[rwho[u].first];
                        (rwho[u].second ? r.nhv : r.phv) = rh[u].vol;
                        if (rh[u].pruned) r.pruned = true;
                    }
                    for (size_t ri = s0; ri < srec.size(); ++ri) {
                        const SRec& r = srec[ri];
                        const Item& par = cur[r.item];
                        mesh_has[par.mesh] = 1;
                        float worst = std::max(cost_rv(r.phv, r.pvol), cost_rv(r.nhv, r.nvol));
                        for (int i = 0; i < par.np - 1; ++i) worst = std::max(worst, cost_rv(par.parts[i].hvol, par.parts[i].mvol));
                        float sum = 0.0f;
                        for (int l = 0; l < par.nl; ++l) sum += par.lc[l];
                        sum += worst;
                        Leaf L;
                        L.c = cur_c[r.item];
                        L.cut0 = par.nl == 0 ? r.k : par.cut0;
                        L.path = sum / (float)(par.nl + 1);
                        L.pruned = r.pruned;
                        leaf_recs.push_back(L);
                    }
                    return 0;
                };
                // Round 1: per cutting part, the candidates of its items with the
                // lowest bounds; round 2: the others, except those whose bound is
                // above the best leaf cost found for their cutting part. Such a
                // leaf can never be the part's argmin, so the decisions are the
                // same as scoring every leaf. A recording run (training targets:
                // every level-0 candidate's best path cost) scores every leaf,
                // since a skipped leaf would leave its first cut's cost too high.
                std::vector<int> r1, r2;
                if (leaf_bnb_on() && !ctx->record) {
                    std::vector<std::vector<int>> by_c(cutting.size());
                    for (size_t i = 0; i < cur.size(); ++i) by_c[cur_c[i]].push_back((int)i);
                    std::vector<char> first(cur.size(), 0);
                    for (auto& v : by_c) {
                        std::stable_sort(v.begin(), v.end(), [&](int a, int b) { return item_lb[a] < item_lb[b]; });
                        const int K1 = leaf_k1((int)v.size());
                        for (int j = 0; j < (int)v.size() && j < K1; ++j) first[v[j]] = 1;
                    }
                    for (size_t q = 0; q < cands.size(); ++q) (first[cands[q].item] ? r1 : r2).push_back((int)q);
                } else {
                    for (size_t q = 0; q < cands.size(); ++q) r1.push_back((int)q);
                }
                LA_TRY(run_round(r1));
                if (!r2.empty()) {
                    std::vector<float> ubc(cutting.size(), 1e30f);
                    for (const Leaf& L : leaf_recs)
                        if (!L.pruned && L.path < ubc[L.c]) ubc[L.c] = L.path;
                    std::vector<int> keep;
                    keep.reserve(r2.size());
                    for (int q : r2) {
                        const int it = cands[q].item;
                        if (mesh_has[cur[it].mesh] && item_lb[it] > ubc[cur_c[it]]) { ++bnb_skipped; continue; }
                        keep.push_back(q);
                    }
                    bnb_scored += (long long)(r1.size() + keep.size());
                    LA_TRY(run_round(keep));
                } else {
                    bnb_scored += (long long)r1.size();
                }
                for (size_t i = 0; i < cur.size(); ++i)
                    if (!mesh_has[cur[i].mesh]) {
                        leaves.push_back(cur[i]);
                        leaves_c.push_back(cur_c[i]);
                    }
                cur.clear();
                cur_c.clear();
                lap(d == 0 ? S_LEVEL0 : S_LEVELN);
                break;
            }
            LA_TRY(hull_volumes(ht));
            to_free.insert(to_free.end(), pts_free.begin(), pts_free.end());
            if (full)
                for (size_t q = 0; q < ht.size(); ++q) {
                    Item& ch = nxt[ht_who[q].first];
                    IPart& x = ch.parts[ch.np - 2 + ht_who[q].second];
                    x.hvol = ht[q].vol;
                    if (ht[q].hv) {
                        x.hv = ht[q].hv; x.ht = ht[q].ht; x.hnv = ht[q].hnv; x.hnt = ht[q].hnt;
                        store[x.store].hv = ht[q].hv;
                    }
                }
            for (Item& ch : nxt) {
                sort_item_parts(ch);
                if (ch.nl < LA_MAX_LEVELS) {
                    ch.lc[ch.nl] = cost_rv(ch.parts[ch.np - 1].hvol, ch.parts[ch.np - 1].mvol);
                    ch.nl++;
                }
            }
            if (d == 0) {
                level0_items = nxt;
                for (size_t i = 0; i < nxt.size(); ++i) l0[nxt_c[i]][nxt[i].cut0] = (int)i;
            }
            std::vector<char> has_child(M.size(), 0);
            for (const Item& ch : nxt) has_child[ch.mesh] = 1;
            for (size_t i = 0; i < cur.size(); ++i)
                if (!has_child[cur[i].mesh]) {
                    stay.push_back(cur[i]);
                    stay_c.push_back(cur_c[i]);
                }
            cur = nxt;
            cur_c = nxt_c;
            lap(d == 0 ? S_LEVEL0 : S_LEVELN);
            if (cur.empty()) break;
        }
        cur.insert(cur.end(), stay.begin(), stay.end());
        cur_c.insert(cur_c.end(), stay_c.begin(), stay_c.end());
        LA_TRY(quick_levels(P, med, total_levels, cur, cur_c, extra));
        for (size_t i = 0; i < cur.size(); ++i) {
            leaves.push_back(cur[i]);
            leaves_c.push_back(cur_c[i]);
        }
        // evaluate (cuda/la_refine.cu:418-527) and apply (cuda/la_lifecycle.cu:151-261)
        std::vector<std::vector<float>> best(cutting.size());
        for (size_t c = 0; c < cutting.size(); ++c) best[c].assign(l0[c].size(), 1e30f);
        auto consider = [&](const Item& it, int c) {
            if (it.pruned || it.nl == 0) return;
            const float pc = item_path_cost(it);
            if (it.cut0 >= 0 && it.cut0 < (int)best[c].size() && pc < best[c][it.cut0]) best[c][it.cut0] = pc;
        };
        for (size_t i = 0; i < leaves.size(); ++i) consider(leaves[i], leaves_c[i]);
        for (const Leaf& L : leaf_recs)
            if (!L.pruned && L.cut0 >= 0 && L.cut0 < (int)best[L.c].size() && L.path < best[L.c][L.cut0])
                best[L.c][L.cut0] = L.path;
        for (const Item& x : extra) consider(x, cut_begin[x.mesh] + x.src);
        if (ctx->record) LA_TRY(record_level0(cutting, rec_k, rec_pl, best));
        std::vector<int> adds_part;
        for (size_t c = 0; c < cutting.size(); ++c) {
            bool any = false;
            for (float b : best[c]) any |= b < 1e30f;
            if (!any)
                for (size_t k = 0; k < l0[c].size(); ++k)
                    if (l0[c][k] >= 0 && level0_items[l0[c][k]].nl >= 1) best[c][k] = level0_items[l0[c][k]].lc[0];
            float bc = 1e30f;
            int bi = 0;
            for (size_t k = 0; k < best[c].size(); ++k)
                if (best[c][k] < bc) { bc = best[c][k]; bi = (int)k; }
            if (P.verbose) {
                fprintf(stderr, "[la] mesh %d iter %d part %d: best cut %d cost %.6f\n", cutting[c].first, iter,
                        cutting[c].second, bi, bc);
                if (P.verbose >= 2) {
                    fprintf(stderr, "[la]   best path per first cut:");
                    for (size_t k = 0; k < best[c].size(); ++k)
                        if (best[c][k] < 1e30f) fprintf(stderr, " %zu:%.6f", k, best[c][k]);
                    fprintf(stderr, "\n");
                }
            }
            if (bi >= (int)l0[c].size() || l0[c][bi] < 0) continue;
            const Item& w = level0_items[l0[c][bi]];
            // adopt the unsorted halves of the level-0 cut: positive half replaces the part
            const IPart* hp[2] = {nullptr, nullptr};
            for (int q = 0; q < w.np; ++q) {
                if (w.parts[q].store < 0) continue;
                // store indices are consecutive per cut: positive first
                if (!hp[0] || w.parts[q].store < hp[0]->store) { hp[1] = hp[0]; hp[0] = &w.parts[q]; }
                else hp[1] = &w.parts[q];
            }
            if (!hp[0] || !hp[1]) continue;
            const int m = cutting[c].first;
            MeshState& S = M[m];
            // room for the second half first: on failure nothing has changed hands
            // (the store guard frees both halves, the part keeps its mesh)
            REQUIRE((int)S.parts.size() < LA_MAX_DECOMP, "lookahead: more than %d parts", LA_MAX_DECOMP);
            Part& old = S.parts[cutting[c].second];
            const int cc = old.cc;
            drop_part(old);
            for (int s2 = 0; s2 < 2; ++s2) {
                const IPart& x = *hp[s2];
                store[x.store].keep = true;
                Part q;
                q.v = x.v; q.t = x.t; q.nv = x.nv; q.nt = x.nt; q.mvol = x.mvol; q.hvol = x.hvol; q.haus = 0.0f;
                q.hv = x.hv; q.ht = x.ht; q.hnv = x.hnv; q.hnt = x.hnt; q.own_hull = x.hv != nullptr;
                if (q.mvol > q.hvol) q.mvol = q.hvol;  // cuda/la_lifecycle.cu:292-296
                q.cc = cc; q.own = true;
                for (int j = 0; j < 3; ++j) { q.lo[j] = x.lo[j]; q.hi[j] = x.hi[j]; }
                if (s2 == 0) old = q;
                else S.parts.push_back(q);
            }
        }
        for (const Half& hf : store)
            if (!hf.keep) {
                to_free.push_back(hf.v);
                if (hf.hv) to_free.push_back(hf.hv);
            }
        store_guard.released = true;
        const int frc = flush_frees();
        lap(S_APPLY);
        if (frc == 0) LA_TRY(bisect_snap(iter, "apply_cuts"));
        if (P.verbose) {
            const auto now = std::chrono::steady_clock::now();
            fprintf(stderr, "[la] iter %d: %.1f ms\n", iter,
                    std::chrono::duration<double, std::milli>(now - iter_t0).count());
        }
        return frc;
    }
};

}  // namespace

// Batched lookahead over n meshes (packed), each decomposed independently.
static int lookahead_run(gpu_ctx_t ctx, int n_meshes, const float* verts, const int* voff, const int* tris,
                         const int* toff, const float* hverts, const int* hvoff, const int* htris, const int* htoff,
                         const Params& P, struct gpu_result* out, int* iters_out, double* stats_out) {
    HIP_TRY(hipSetDevice(ctx->device));
    for (int m = 0; m < n_meshes; ++m) { out[m].parts = nullptr; out[m].nparts = 0; }
    REQUIRE(P.depth >= 1 && P.depth <= LA_MAX_LEVELS, "lookahead: depth must be in [1, %d]", LA_MAX_LEVELS);
    REQUIRE(P.quick_depth >= 0, "lookahead: quick_depth must be >= 0");
    REQUIRE(P.width >= 1 && P.width2 >= 1, "lookahead: width and width2 must be positive");
    REQUIRE(P.width + 4 * P.n_concave < 512, "lookahead: width + 4*n_concave_edges must be < 512");
    REQUIRE(P.max_n_cutting >= 1, "lookahead: max_n_cutting must be positive");
    // every index must name a vertex of its own mesh (the kernels trust them)
    for (int m = 0; m < n_meshes; ++m) {
        const int nvm = voff[m + 1] - voff[m], nhm = hvoff[m + 1] - hvoff[m];
        for (long long k = 3ll * toff[m]; k < 3ll * toff[m + 1]; ++k)
            REQUIRE(tris[k] >= 0 && tris[k] < nvm, "lookahead: mesh %d has a triangle index out of range", m);
        for (long long k = 3ll * htoff[m]; k < 3ll * htoff[m + 1]; ++k)
            REQUIRE(htris[k] >= 0 && htris[k] < nhm, "lookahead: hull %d has a triangle index out of range", m);
    }
    Engine E;
    E.ctx = ctx;
    E.sync_mark = ctx->host_syncs;
    // per-kernel timers for the run; detached on every exit path
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
    E.P = P;
    E.M.resize(n_meshes);
    const size_t tv = (size_t)voff[n_meshes], tt = (size_t)toff[n_meshes];
    const size_t thv = (size_t)hvoff[n_meshes], tht = (size_t)htoff[n_meshes];
    float* d_v = nullptr;
    int* d_t = nullptr;
    float* d_hv = nullptr;
    int* d_ht = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_v, 12 * tv + 16, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&d_t, 12 * tt + 16, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&d_hv, 12 * thv + 16, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&d_ht, 12 * tht + 16, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_v, verts, 12 * tv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_t, tris, 12 * tt, hipMemcpyHostToDevice, ctx->stream));
    if (thv) HIP_TRY(hipMemcpyAsync(d_hv, hverts, 12 * thv, hipMemcpyHostToDevice, ctx->stream));
    if (tht) HIP_TRY(hipMemcpyAsync(d_ht, htris, 12 * tht, hipMemcpyHostToDevice, ctx->stream));
    int rc = 0;
    std::vector<int> act;
    for (int m = 0; m < n_meshes; ++m) {
        Part p;
        p.v = d_v + 3ll * voff[m];
        p.t = d_t + 3ll * toff[m];
        p.nv = voff[m + 1] - voff[m];
        p.nt = toff[m + 1] - toff[m];
        const float* hv = verts + 3ll * voff[m];
        // FP64 volumes of the input mesh and of its hull (cuda/la_lifecycle.cu:18-81)
        auto vol = [](const float* V, const int* T, int nt) {
            if (nt <= 0) return 0.0;
            const double ax = V[0], ay = V[1], az = V[2];
            double acc = 0.0;
            for (int k = 0; k < nt; ++k) {
                const float* a = V + 3 * T[3 * k];
                const float* b = V + 3 * T[3 * k + 1];
                const float* c = V + 3 * T[3 * k + 2];
                const double x0 = a[0] - ax, y0 = a[1] - ay, z0 = a[2] - az;
                const double x1 = b[0] - ax, y1 = b[1] - ay, z1 = b[2] - az;
                const double x2 = c[0] - ax, y2 = c[1] - ay, z2 = c[2] - az;
                acc += x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
            }
            return std::fabs(acc) / 6.0;
        };
        p.mvol = vol(hv, tris + 3ll * toff[m], p.nt);
        const int hn = htoff[m + 1] - htoff[m];
        p.hv = d_hv + 3ll * hvoff[m];
        p.ht = d_ht + 3ll * htoff[m];
        p.hnv = hvoff[m + 1] - hvoff[m];
        p.hnt = hn;
        p.hvol = vol(hverts + 3ll * hvoff[m], htris + 3ll * htoff[m], hn);
        p.own = false;
        for (int j = 0; j < 3; ++j) { p.lo[j] = 1e30f; p.hi[j] = -1e30f; }
        for (int i = 0; i < p.nv; ++i)
            for (int j = 0; j < 3; ++j) {
                p.lo[j] = std::min(p.lo[j], hv[3 * i + j]);
                p.hi[j] = std::max(p.hi[j], hv[3 * i + j]);
            }
        E.M[m].parts.push_back(p);
        act.push_back(m);
    }
    E.lap(S_SETUP);
    rc = E.bisect_snap(-1, "post_init");
    for (int iter = 0; iter < P.max_iters && !act.empty() && rc == 0; ++iter) rc = E.iterate(act, iter);
    // meshes that ran out of iterations: last apply still needs components + hulls
    if (rc == 0 && !act.empty()) {
        if (P.dc_per_iter) rc = E.components(act);
        if (rc == 0) rc = E.hull_parts(act);
    }
    if (rc == 0 && P.dc_final) {
        std::vector<int> all;
        for (int m = 0; m < n_meshes; ++m) all.push_back(m);
        rc = E.components(all);
        if (rc == 0) rc = E.hull_parts(all);
    }
    if (rc == 0) {
        std::vector<int> all;
        for (int m = 0; m < n_meshes; ++m) all.push_back(m);
        rc = E.hull_parts(all, true);
        if (rc == 0 && P.merge) rc = E.merge(all);
    }
    E.lap(S_POST);
    // results: every part's mesh and hull in one batched readback
    if (rc == 0) {
        std::vector<std::pair<const void*, size_t>> src;
        for (int m = 0; m < n_meshes; ++m)
            for (Part& p : E.M[m].parts) {
                src.push_back({p.v, 12ull * p.nv});
                src.push_back({p.t, 12ull * p.nt});
                src.push_back({p.hv, p.hv ? 12ull * p.hnv : 0});
                src.push_back({p.ht, p.ht ? 12ull * p.hnt : 0});
            }
        std::vector<size_t> offs;
        int grc = 0;
        const char* hb = E.gather_to_host(src, offs, &grc);
        rc = grc;
        size_t q = 0;
        for (int m = 0; m < n_meshes && rc == 0; ++m) {
            MeshState& S = E.M[m];
            if (iters_out) iters_out[m] = S.iters;
            const int np = (int)S.parts.size();
            out[m].nparts = np;
            out[m].parts = (struct gpu_part_result*)calloc(np > 0 ? np : 1, sizeof(struct gpu_part_result));
            for (int i = 0; i < np; ++i, q += 4) {
                Part& p = S.parts[i];
                gpu_part_result& r = out[m].parts[i];
                r.nv = p.nv; r.nt = p.nt;
                r.hull_nv = p.hv ? p.hnv : 0;
                r.hull_nt = p.ht ? p.hnt : 0;
                r.mesh_vol = (float)p.mvol;
                r.hull_vol = (float)p.hvol;
                r.hausdorff = p.haus;
                r.verts = (float*)malloc(12ull * (p.nv ? p.nv : 1));
                r.tris = (int*)malloc(12ull * (p.nt ? p.nt : 1));
                r.hull_verts = (float*)malloc(12ull * (r.hull_nv ? r.hull_nv : 1));
                r.hull_tris = (int*)malloc(12ull * (r.hull_nt ? r.hull_nt : 1));
                if (src[q].second) memcpy(r.verts, hb + offs[q], src[q].second);
                if (src[q + 1].second) memcpy(r.tris, hb + offs[q + 1], src[q + 1].second);
                if (src[q + 2].second) memcpy(r.hull_verts, hb + offs[q + 2], src[q + 2].second);
                if (src[q + 3].second) memcpy(r.hull_tris, hb + offs[q + 3], src[q + 3].second);
            }
        }
    }
    // free everything the search still holds
    for (auto& S : E.M)
        for (auto& p : S.parts) {
            if (p.own && p.v) E.to_free.push_back(p.v);
            if (p.own_hull && p.hv) E.to_free.push_back(p.hv);
            p.v = nullptr;
            p.hv = nullptr;
        }
    const int frc = E.flush_frees();
    for (void* b : E.ext_bufs) (void)hipFreeAsync(b, ctx->stream);
    HIP_TRY(hipFreeAsync(d_v, ctx->stream));
    HIP_TRY(hipFreeAsync(d_t, ctx->stream));
    HIP_TRY(hipFreeAsync(d_hv, ctx->stream));
    HIP_TRY(hipFreeAsync(d_ht, ctx->stream));
    HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
    E.lap(S_OUT);
    if (stats_out) {
        for (int ph = 0; ph < PH_COUNT; ++ph)
            for (int j = 0; j < 3; ++j) stats_out[3 * ph + j] = E.stats[ph][j];
        for (int sc = 0; sc < S_COUNT; ++sc) stats_out[3 * PH_COUNT + sc] = E.sect[sc];
        (void)kt_resolve(ctx);
        for (int kc = 0; kc < KT_COUNT; ++kc)
            for (int j = 0; j < 3; ++j) stats_out[3 * PH_COUNT + S_COUNT + 3 * kc + j] = ktg.k.stats[kc][j];
        double* pr = stats_out + 3 * PH_COUNT + S_COUNT + 3 * KT_COUNT;
        pr[0] = (double)E.proposal_kept;
        pr[1] = (double)E.proposal_total;
        pr[2] = (double)E.proposal_fallbacks;
        for (int sc = 0; sc < S_COUNT; ++sc) {
            pr[3 + sc] = E.sect_gpu[sc];
            pr[3 + S_COUNT + sc] = E.sect_sync[sc];
        }
    }
    if (rc) {
        for (int m = 0; m < n_meshes; ++m) gpu_result_free(&out[m]);
    }
    return rc ? rc : frc;
}

extern "C" {

int lookahead_decompose(gpu_ctx_t ctx, const float* verts, int nv, const int* tris, int nt, const float* hull_verts,
                        int hull_nv, const int* hull_tris, int hull_nt, int max_iters, int width, int width2,
                        float threshold, int depth, int quick_depth, int max_n_cutting, int verbose, int debug,
                        int decompose_components, int decompose_components_per_iter, int n_concave_edges,
                        float concave_eps, float concave_threshold, int concave_iters, int merge_hulls,
                        struct gpu_result* out) {
    (void)debug;
    Params P{max_iters, width, width2, depth, quick_depth, max_n_cutting, verbose, threshold, decompose_components,
             decompose_components_per_iter, n_concave_edges, concave_iters, merge_hulls, concave_eps,
             concave_threshold};
    const int voff[2] = {0, nv}, toff[2] = {0, nt}, hvoff[2] = {0, hull_nv}, htoff[2] = {0, hull_nt};
    return lookahead_run(ctx, 1, verts, voff, tris, toff, hull_verts, hvoff, hull_tris, htoff, P, out, nullptr,
                         nullptr);
}

int gpu_lookahead_batch(gpu_ctx_t ctx, int n_meshes, const float* verts, const int* vert_off, const int* tris,
                        const int* tri_off, const float* hull_verts, const int* hull_vert_off, const int* hull_tris,
                        const int* hull_tri_off, int max_iters, int width, int width2, float threshold, int depth,
                        int quick_depth, int max_n_cutting, int decompose_components, int decompose_components_per_iter,
                        int n_concave_edges, float concave_eps, float concave_threshold, int concave_iters,
                        int merge_hulls, struct gpu_result* out, int* out_iters, double* out_stats) {
    Params P{max_iters, width, width2, depth, quick_depth, max_n_cutting, 0, threshold, decompose_components,
             decompose_components_per_iter, n_concave_edges, concave_iters, merge_hulls, concave_eps,
             concave_threshold};
    return lookahead_run(ctx, n_meshes, verts, vert_off, tris, tri_off, hull_verts, hull_vert_off, hull_tris,
                         hull_tri_off, P, out, out_iters, out_stats);
}

}  // extern "C"

extern "C" int gpu_lookahead_set_proposal(gpu_ctx_t ctx, gpu_mlp_t m, int top_k) {
    if (!ctx) return -1;
    REQUIRE(!m || top_k > 0, "lookahead proposal: top_k must be positive");
    ctx->proposal = m;
    ctx->proposal_topk = m ? top_k : 0;
    return 0;
}

// Level-0 candidate planes of whole meshes as the search builds them (axis
// cuts of cuda/la_axis_cut.cuh:10-19 inside the box margins, then the planes
// of sampled concave edges, cuda/la_concave.cu:48-363), for the proposal
// network's training data. Mesh m's planes go to out_planes[4 * (m * cap + j)],
// their count (at most cap) to out_count[m]; out_k[m * cap + j] is the
// candidate's index k in the search (axis cuts first).
extern "C" int gpu_lookahead_candidates(gpu_ctx_t ctx, int n_meshes, const float* verts, const int* voff,
                                        const int* tris, const int* toff, int width, float threshold, int n_concave,
                                        float concave_eps, float concave_threshold, int cap, float* out_planes,
                                        int* out_k, int* out_count) {
    HIP_TRY(hipSetDevice(ctx->device));
    REQUIRE(n_meshes >= 0 && width >= 1 && cap > 0, "lookahead_candidates: bad arguments");
    if (n_meshes == 0) return 0;
    Engine E;
    E.ctx = ctx;
    E.sync_mark = ctx->host_syncs;
    E.P = Params{};
    E.P.width = width;
    E.P.threshold = threshold;
    E.P.n_concave = n_concave;
    E.P.concave_eps = concave_eps;
    E.P.concave_threshold = concave_threshold;
    E.M.resize(n_meshes);
    const size_t tv = (size_t)voff[n_meshes], tt = (size_t)toff[n_meshes];
    float* d_v = nullptr;
    int* d_t = nullptr;
    HIP_TRY(hipMallocAsync((void**)&d_v, 12 * tv + 16, ctx->stream));
    HIP_TRY(hipMallocAsync((void**)&d_t, 12 * tt + 16, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_v, verts, 12 * tv, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(d_t, tris, 12 * tt, hipMemcpyHostToDevice, ctx->stream));
    std::vector<std::pair<int, int>> cutting;
    for (int m = 0; m < n_meshes; ++m) {
        Part p;
        p.v = d_v + 3ll * voff[m];
        p.t = d_t + 3ll * toff[m];
        p.nv = voff[m + 1] - voff[m];
        p.nt = toff[m + 1] - toff[m];
        const float* hv = verts + 3ll * voff[m];
        for (int j = 0; j < 3; ++j) { p.lo[j] = 1e30f; p.hi[j] = -1e30f; }
        for (int i = 0; i < p.nv; ++i)
            for (int j = 0; j < 3; ++j) {
                p.lo[j] = std::min(p.lo[j], hv[3 * i + j]);
                p.hi[j] = std::max(p.hi[j], hv[3 * i + j]);
            }
        E.M[m].parts.push_back(p);
        cutting.push_back({m, 0});
    }
    std::vector<std::vector<float>> ep(n_meshes);
    int rc = n_concave > 0 ? E.concave_planes(cutting, ep) : 0;
    float med = threshold * 0.25f;
    if (med > 0.015f) med = 0.015f;
    for (int m = 0; m < n_meshes && rc == 0; ++m) {
        const Part& wp = E.M[m].parts[0];
        int cnt = 0;
        auto put = [&](const float* pl, int k) {
            if (cnt >= cap) return;
            for (int j = 0; j < 4; ++j) out_planes[4 * ((size_t)m * cap + cnt) + j] = pl[j];
            out_k[(size_t)m * cap + cnt] = k;
            ++cnt;
        };
        for (int k = 0; k < width; ++k) {
            const AxisCut a = axis_cut(width, k);
            const float lov = wp.lo[a.axis] + med, hiv = wp.hi[a.axis] - med;
            if (lov >= hiv) continue;
            const float frac = (a.slice + 1.f) / (a.cpa + 1.f);
            const float pos = lov + frac * (hiv - lov);
            float pl[4] = {a.axis == 0 ? 1.f : 0.f, a.axis == 1 ? 1.f : 0.f, a.axis == 2 ? 1.f : 0.f, -pos};
            put(pl, k);
        }
        for (size_t e = 0; 4 * e + 3 < ep[m].size(); ++e) {
            const float* pl = &ep[m][4 * e];
            if (pl[0] == 0.f && pl[1] == 0.f && pl[2] == 0.f) continue;
            put(pl, width + (int)e);
        }
        out_count[m] = cnt;
    }
    (void)hipFreeAsync(d_v, ctx->stream);
    (void)hipFreeAsync(d_t, ctx->stream);
    (void)ctx_stream_sync(ctx, ctx->stream);
    return rc;
}

extern "C" int gpu_lookahead_record(gpu_ctx_t ctx, int on) {
    if (!ctx) return -1;
    delete (LaRecord*)ctx->record;
    ctx->record = on ? new LaRecord() : nullptr;
    return 0;
}

extern "C" int gpu_lookahead_record_size(gpu_ctx_t ctx, int* n_parts, int* n_cand) {
    REQUIRE(ctx->record, "lookahead record: not recording");
    const LaRecord* R = (const LaRecord*)ctx->record;
    *n_parts = R->n_parts;
    *n_cand = (int)R->target.size();
    return 0;
}

extern "C" int gpu_lookahead_record_read(gpu_ctx_t ctx, float* pts, float* planes, int* part_of, float* targets) {
    REQUIRE(ctx->record, "lookahead record: not recording");
    const LaRecord* R = (const LaRecord*)ctx->record;
    if (!R->pts.empty()) memcpy(pts, R->pts.data(), sizeof(float) * R->pts.size());
    if (!R->planes.empty()) memcpy(planes, R->planes.data(), sizeof(float) * R->planes.size());
    if (!R->part_of.empty()) memcpy(part_of, R->part_of.data(), sizeof(int) * R->part_of.size());
    if (!R->target.empty()) memcpy(targets, R->target.data(), sizeof(float) * R->target.size());
    return 0;
}
