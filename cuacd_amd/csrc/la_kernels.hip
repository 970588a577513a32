// la_kernels.hip - device work of the lookahead search: hulls (with the k-DOP
// prefilter and bound pruning), level-0 plane cuts, leaf scoring, connected
// components, concave edges. Each kernel runs over a flat task table; the
// host (lookahead.hip) builds the tables for all meshes of a batch at once.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <vector>
#include "lookahead.hpp"
#include "plane_cut.hpp"

using namespace cuacd;

namespace {

constexpr int LA_WAVE_MAX = 1024;
constexpr int LA_BLOCK_NW = 8;
constexpr int KDOP_AXES_N = 40;

// The 40 k-DOP axes: unit normals of the 80 faces of the level-1 icosphere,
// one per antipodal pair (reference cuda/kdop_const.cu:9-50 lists the same
// set). Generated on the host at first use and kept in constant memory.
__constant__ float c_kdop_axes[KDOP_AXES_N][3];

// ---------------------------------------------------------------------------
// Hull of one task's point set
// ---------------------------------------------------------------------------
// Returns false when the workspace's vertex capacity was too small and the
// task was queued on `retry` (count at retry[0], task ids after it) to be
// redone by a full-capacity kernel; the task is then left untouched.
template <int PPL, int AS, int NW>
__device__ __forceinline__ bool la_hull_item(LaHullTask& t, const HullWS<AS> w, HullBlockLDS* sb, DPool* pool,
                                             int* err, int* retry = nullptr, int task_id = -1, bool multi = false) {
    constexpr int T = 64 * NW;
    const int tid = threadIdx.x;
    const float* P = t.filt ? t.filt : t.pts;
    const int n = t.filt ? t.nf : t.n;
    if (n < 4) {
        if (tid == 0) { t.vol = 0.0; t.hnv = t.hnt = 0; t.hv = nullptr; t.ht = nullptr; }
        return true;
    }
    float xs[PPL], ys[PPL], zs[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        xs[k] = ys[k] = zs[k] = 0.0f;
        const int i = k * T + tid;
        if (i < n) {
            xs[k] = P[3ll * i];
            ys[k] = P[3ll * i + 1];
            zs[k] = P[3ll * i + 2];
        }
    }
    QFrame fr;
    if (t.use_frame) {
        fr = t.frame;
    } else {
        float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            if (k * T + tid < n) {
                mn[0] = fminf(mn[0], xs[k]); mx[0] = fmaxf(mx[0], xs[k]);
                mn[1] = fminf(mn[1], ys[k]); mx[1] = fmaxf(mx[1], ys[k]);
                mn[2] = fminf(mn[2], zs[k]); mx[2] = fmaxf(mx[2], zs[k]);
            }
        for (int j = 0; j < 3; ++j) {
            mn[j] = wave_min_f(mn[j]);
            mx[j] = wave_max_f(mx[j]);
        }
        if (NW > 1) {
            const int wv = tid >> 6;
            if (lane_id() == 0)
                for (int j = 0; j < 3; ++j) { sb->bmn[wv][j] = mn[j]; sb->bmx[wv][j] = mx[j]; }
            __syncthreads();
            for (int j = 0; j < 3; ++j)
                for (int q = 0; q < NW; ++q) {
                    mn[j] = fminf(mn[j], sb->bmn[q][j]);
                    mx[j] = fmaxf(mx[j], sb->bmx[q][j]);
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
    HullResult r;
    if constexpr (NW >= 8 && PPL * T <= 16384) {
        // multi-apex rounds (the workspace was carved or attached with their
        // scratch; the four-wave tiers keep single-apex steps and do not
        // compile them); the point index field of their keys holds 14 bits
        if (multi && n <= PPL * T) r = block_quickhull<PPL, AS, NW, true>(w, sb, n, px, py, pz);
        else r = block_quickhull<PPL, AS, NW>(w, sb, n, px, py, pz);
    } else {
        r = block_quickhull<PPL, AS, NW>(w, sb, n, px, py, pz);
    }
    if (r.status == HULL_OVERFLOW && retry) {
        if (tid == 0) {
            const int k = atomicAdd(retry, 1);
            retry[1 + k] = task_id;
        }
        return false;
    }
    if (r.status != HULL_OK) {
        if (tid == 0) {
            t.vol = 0.0;
            t.hnv = t.hnt = 0;
            t.hv = nullptr;
            t.ht = nullptr;
            if (r.status == HULL_OVERFLOW) { t.err = KERR_HULL_CAPACITY; err_or(err, KERR_HULL_CAPACITY); }
        }
        return true;
    }
    auto pos = [&](int p, float* o3) {
        o3[0] = P[3ll * p];
        o3[1] = P[3ll * p + 1];
        o3[2] = P[3ll * p + 2];
    };
    const HullCanon hc = hull_canon<AS, NW>(w, sb, r, n, pos);
    if (hc.nt < 0) {
        if (tid == 0) {
            t.vol = 0.0;
            t.hnv = t.hnt = 0;
            t.hv = nullptr;
            t.ht = nullptr;
            t.err = KERR_HULL_CAPACITY;
            err_or(err, KERR_HULL_CAPACITY);
        }
        return true;
    }
    if (tid == 0) { t.vol = hc.vol; t.ncorner = hc.nv; }
    if (!t.emit) return true;
    const int nv = hc.nv;
    const int nt = hc.nt;
    const unsigned long long vb = al16(12ull * nv);
    // emit 2: the mesh is an optional extra (built later when missing), so a
    // full heap leaves it out without an error
    if (tid == 0) sb->red[0] = (unsigned long long)heap_alloc(pool, 0, vb + 12ull * nt, t.emit == 2 ? nullptr : err);
    __syncthreads();
    float* chunk = (float*)sb->red[0];
    __syncthreads();
    if (!chunk) {
        if (tid == 0) {
            t.hv = nullptr; t.ht = nullptr; t.hnv = t.hnt = 0;
            if (t.emit != 2) t.err = KERR_HULL_HEAP_OUTPUT;
        }
        return true;
    }
    int* tris = (int*)((char*)chunk + vb);
    for (int v = tid; v < r.nvert; v += T)
        if (w.vref[v]) {
            const int p = w.vpid[v], o = w.vout[v];
            chunk[3 * o] = P[3ll * p];
            chunk[3 * o + 1] = P[3ll * p + 1];
            chunk[3 * o + 2] = P[3ll * p + 2];
        }
    hull_canon_write<AS, NW>(w, r, tris);
    if (tid == 0) { t.hv = chunk; t.ht = tris; t.hnv = nv; t.hnt = nt; }
    return true;
}

template <int PPL>
__global__ __launch_bounds__(64) void la_hull_wave(LaHullTask* tasks, const int* ids, int n, const int* dcount,
                                                   DPool* pool, int* err) {
    extern __shared__ __align__(16) char lds[];
    __shared__ HullBlockLDS sb;
    __shared__ HullBits<64 * PPL> hb;
    // ids[0, n), or the first min(*dcount, n) when the count is on the device
    if ((int)blockIdx.x >= (dcount ? min(*dcount, n) : n)) return;
    LaHullTask& t = tasks[ids[blockIdx.x]];
    const HullWS<AS_LDS> w = hull_ws_attach(hull_ws_carve<AS_LDS>(lds, 64 * PPL, 64 * PPL), &sb, &hb);
    la_hull_item<PPL, AS_LDS, 1>(t, w, &sb, pool, err);
    if (threadIdx.x == 0 && t.filt) { heap_free(pool, t.filt, err); t.filt = nullptr; }
}

// Four waves per hull, workspace in LDS: for the larger one-workgroup sets,
// where the LDS footprint allows one hull per CU anyway, all four SIMDs work.
constexpr int LA_LDS4_NW = 4;
template <int PPL>
__global__ __launch_bounds__(64 * LA_LDS4_NW) void la_hull_lds4(LaHullTask* tasks, const int* ids, int n,
                                                               const int* dcount, int multi, DPool* pool, int* err) {
    extern __shared__ __align__(16) char lds[];
    __shared__ HullBlockLDS sb;
    // ids[0, n), or the first min(*dcount, n) of them when dcount is given
    // (a retry list filled by the previous launch); the grid strides over them
    const int nn = dcount ? min(*dcount, n) : n;
    constexpr int CAP = 64 * LA_LDS4_NW * PPL;
    __shared__ HullBits<CAP> hb;
    const HullWS<AS_LDS> w = hull_ws_attach(hull_ws_carve<AS_LDS>(lds, CAP, CAP, multi != 0), &sb, &hb);
    for (int j = blockIdx.x; j < nn; j += gridDim.x) {
        __syncthreads();
        LaHullTask& t = tasks[ids[j]];
        la_hull_item<PPL, AS_LDS, LA_LDS4_NW>(t, w, &sb, pool, err, nullptr, -1, multi != 0);
        if (threadIdx.x == 0 && t.filt) { heap_free(pool, t.filt, err); t.filt = nullptr; }
    }
}

// Eight waves per hull, full capacity in LDS, multi-apex rounds when `multi`:
// the retries of the eight-wave first passes of the wave tiers (the sets
// with the most corners). Strides over its list like la_hull_lds4.
template <int PPL>
__global__ __launch_bounds__(64 * LA_BLOCK_NW) void la_hull_lds8(LaHullTask* tasks, const int* ids, int n,
                                                                const int* dcount, int multi, DPool* pool, int* err) {
    extern __shared__ __align__(16) char lds[];
    __shared__ HullBlockLDS sb;
    const int nn = dcount ? min(*dcount, n) : n;
    constexpr int CAP = 64 * LA_BLOCK_NW * PPL;
    __shared__ HullBits<CAP> hb;
    const HullWS<AS_LDS> w = hull_ws_attach(hull_ws_carve<AS_LDS>(lds, CAP, CAP, multi != 0), &sb, &hb);
    for (int j = blockIdx.x; j < nn; j += gridDim.x) {
        __syncthreads();
        LaHullTask& t = tasks[ids[j]];
        la_hull_item<PPL, AS_LDS, LA_BLOCK_NW>(t, w, &sb, pool, err, nullptr, -1, multi != 0);
        if (threadIdx.x == 0 && t.filt) { heap_free(pool, t.filt, err); t.filt = nullptr; }
    }
}

// Same, with the hull-vertex capacity reduced to vcap (a smaller LDS
// footprint, so several hulls share a CU); a set whose hull needs more
// vertices is queued on `retry` for the full-capacity kernel.
template <int PPL, int NW>
__global__ __launch_bounds__(64 * NW) void la_hull_ldsv(LaHullTask* tasks, const int* ids, int n,
                                                       const int* dcount, int vcap, int* retry, int multi,
                                                       DPool* pool, int* err) {
    extern __shared__ __align__(16) char lds[];
    __shared__ HullBlockLDS sb;
    if ((int)blockIdx.x >= (dcount ? min(*dcount, n) : n)) return;
    const int id = ids[blockIdx.x];
    LaHullTask& t = tasks[id];
    constexpr int CAP = 64 * NW * PPL;
    __shared__ HullBits<CAP> hb;
    const HullWS<AS_LDS> w = hull_ws_attach(hull_ws_carve<AS_LDS>(lds, vcap, CAP, multi != 0), &sb, &hb);
    const bool done = la_hull_item<PPL, AS_LDS, NW>(t, w, &sb, pool, err, retry, id, multi != 0);
    if (done && threadIdx.x == 0 && t.filt) { heap_free(pool, t.filt, err); t.filt = nullptr; }
}

// Workspace slot blockIdx.x of gws; ids[0, n), or the first min(*dcount, n)
// when dcount is given, strided over the grid.
// With `multi`, multi-apex rounds: their per-face marks and keys in the
// dynamic LDS (hull_ma_bytes(VC)), the faces and vertices in the global
// workspace.
template <int PPL>
__global__ __launch_bounds__(64 * LA_BLOCK_NW) void la_hull_block(LaHullTask* tasks, const int* ids, int n,
                                                                  const int* dcount, char* gws,
                                                                  unsigned long long stride, int VC, int NP,
                                                                  int multi, DPool* pool, int* err) {
    __shared__ HullBlockLDS sb;
    __shared__ HullBits<HULL_NP_MAX> hb;
    extern __shared__ __align__(16) char lds_ma[];
    const int nn = dcount ? min(*dcount, n) : n;
    HullWS<AS_GLOBAL> w =
        hull_ws_attach(hull_ws_carve<AS_GLOBAL>(gws + (unsigned long long)blockIdx.x * stride, VC, NP), &sb, &hb);
    if (multi) w = hull_ma_attach(w, lds_ma);
    for (int j = blockIdx.x; j < nn; j += gridDim.x) {
        __syncthreads();
        LaHullTask& t = tasks[ids[j]];
        la_hull_item<PPL, AS_GLOBAL, LA_BLOCK_NW>(t, w, &sb, pool, err, nullptr, -1, multi != 0);
        if (threadIdx.x == 0 && t.filt) { heap_free(pool, t.filt, err); t.filt = nullptr; }
    }
}

// k-DOP prefilter (reference cuda/kdop_hull.cuh:116-337): extreme points along
// the 40 axes -> rough inner hull -> optional bound pruning on its volume ->
// keep the points on or outside some rough-hull face (exactly, on the full
// set's grid), in input order. The filtered set keeps the full set's
// quantization frame and the input order, so its canonical hull equals the
// full set's. Three launches over the same list, one record per set between
// them: the extremes (one wave per set, passes over the points), the rough
// hull of the 80 extreme points (one wave per set: its insertion steps need
// no block barrier, and many sets share a CU), the filter (four waves).
constexpr int KD_T = 256;
constexpr int KD_NW = KD_T / 64;
constexpr int KD_NE = 2 * KDOP_AXES_N;   // extreme points
constexpr int KD_MAXPL = 2 * 128 + 8;    // rough faces kept as filter planes

struct KdopRec {
    QFrame full;                      // the full set's frame
    int status;                       // 0 filter, 1 no filter (hull the full set), 2 pruned
    int npl;                          // filter planes
    float ep[KD_NE][3];               // extreme points (axis a: max at 2a, min at 2a + 1)
    unsigned char face[KD_MAXPL][3];  // filter planes as extreme-point triples
};

// Launch 1, one wave per set: the extreme points along the 40 axes, 8 axes
// per pass over the points - max key (value, lowest index), min key (value,
// lowest index) - and the bounding box (the full set's frame) in the first
// pass. One wave per set: with four, the 16 key reductions of a pass cost as
// much as the pass itself for the typical set of about 1000 points, and the
// passes needed block barriers. min / max and the keys are independent of
// the visiting order, so the record is the same for any lane count.
__global__ __launch_bounds__(64) void la_kdop_ext_kernel(LaHullTask* tasks, const int* ids, int n,
                                                         const int* dcount, KdopRec* rec) {
    if ((int)blockIdx.x >= (dcount ? min(*dcount, n) : n)) return;
    const LaHullTask& t = tasks[ids[blockIdx.x]];
    KdopRec& R = rec[blockIdx.x];
    const int lane = threadIdx.x;
    const float* P = t.pts;
    const int np = t.n;
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int a0 = 0; a0 < KDOP_AXES_N; a0 += 8) {
        unsigned long long kmax[8], kmin[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) { kmax[q] = 0ull; kmin[q] = ~0ull; }
        for (int i = lane; i < np; i += 64) {
            const float x = P[3ll * i], y = P[3ll * i + 1], z = P[3ll * i + 2];
            if (a0 == 0) {
                mn[0] = fminf(mn[0], x); mx[0] = fmaxf(mx[0], x);
                mn[1] = fminf(mn[1], y); mx[1] = fmaxf(mx[1], y);
                mn[2] = fminf(mn[2], z); mx[2] = fmaxf(mx[2], z);
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float d = c_kdop_axes[a0 + q][0] * x + c_kdop_axes[a0 + q][1] * y + c_kdop_axes[a0 + q][2] * z;
                const unsigned long long hi = (unsigned long long)fkey(d) << 32;
                const unsigned long long km = hi | (0xFFFFFFFFu - (uint32_t)i);
                const unsigned long long kn = hi | (uint32_t)i;
                kmax[q] = km > kmax[q] ? km : kmax[q];
                kmin[q] = kn < kmin[q] ? kn : kmin[q];
            }
        }
        // lane q < 8 takes the max key of axis a0 + q, lane 8 + q its min key
        unsigned long long mine = 0ull;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const unsigned long long hi = wave_max_u64(kmax[q]);
            const unsigned long long lo = wave_min_u64(kmin[q]);
            if (lane == q) mine = hi;
            if (lane == 8 + q) mine = lo;
        }
        if (lane < 16) {
            const int q = lane & 7, isMin = lane >> 3;
            const int idx = isMin ? (int)(uint32_t)mine : (int)(0xFFFFFFFFu - (uint32_t)mine);
            const int slot = 2 * (a0 + q) + isMin;
            R.ep[slot][0] = P[3ll * idx];
            R.ep[slot][1] = P[3ll * idx + 1];
            R.ep[slot][2] = P[3ll * idx + 2];
        }
    }
    for (int j = 0; j < 3; ++j) {
        mn[j] = wave_min_f(mn[j]);
        mx[j] = wave_max_f(mx[j]);
    }
    if (lane == 0) R.full = qframe_make(mn, mx);
}

// Launch 2, one wave per set: the rough hull of the extreme points (exact, on
// a grid of their own box), its FP64 volume for the bound pruning, and the
// live rough faces as the filter's planes. A set without a usable rough hull
// is hulled whole; a pruned one keeps the rough volume as its bound.
// Vertex capacity: the 80 extreme points (a hull of them has at most 80
// corners, and the face arrays hold 2 * 80 + 8, the live faces plus one
// insertion's transient pair). A smaller workspace fits more of these
// one-wave workgroups on a CU (14 at 11.5 KB of LDS each).
constexpr int KD_VC = KD_NE;
constexpr unsigned KD_WS = 9776;  // LDS of its workspace (la_launch_kdop checks the size)
__global__ __launch_bounds__(64) void la_kdop_rough_kernel(LaHullTask* tasks, const int* ids, int n,
                                                           const int* dcount, KdopRec* rec, int2* res) {
    __shared__ HullBlockLDS sb;
    __shared__ HullBits<KD_NE> hb;
    __shared__ __align__(16) char lds[KD_WS];
    __shared__ float ep[KD_NE][3];
    __shared__ int s_npl;
    if ((int)blockIdx.x >= (dcount ? min(*dcount, n) : n)) return;
    LaHullTask& t = tasks[ids[blockIdx.x]];
    KdopRec& R = rec[blockIdx.x];
    const int tid = threadIdx.x;
    const int np = t.n;
    for (int i = tid; i < 3 * KD_NE; i += 64) (&ep[0][0])[i] = (&R.ep[0][0])[i];
    __syncthreads();
    const HullWS<AS_LDS> w = hull_ws_attach(hull_ws_carve<AS_LDS>(lds, KD_VC, KD_VC), &sb, &hb);
    float emn[3] = {1e30f, 1e30f, 1e30f}, emx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < KD_NE; ++i)
        for (int j = 0; j < 3; ++j) { emn[j] = fminf(emn[j], ep[i][j]); emx[j] = fmaxf(emx[j], ep[i][j]); }
    const QFrame rf = qframe_make(emn, emx);
    // extreme point i = k * 64 + tid in slot k
    int px[2] = {0, 0}, py[2] = {0, 0}, pz[2] = {0, 0};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = k * 64 + tid;
        if (i < KD_NE) {
            const float v[3] = {ep[i][0], ep[i][1], ep[i][2]};
            px[k] = qcoord(rf, 0, sel3(v, rf.ax[0]));
            py[k] = qcoord(rf, 1, sel3(v, rf.ax[1]));
            pz[k] = qcoord(rf, 2, sel3(v, rf.ax[2]));
        }
    }
    const QFrame full = R.full;
    HullResult r = block_quickhull<2, AS_LDS, 1>(w, &sb, KD_NE, px, py, pz);
    HullCanon rc;
    rc.nt = 0;
    if (r.status == HULL_OK) {
        auto epos = [&](int p, float* o3) { o3[0] = ep[p][0]; o3[1] = ep[p][1]; o3[2] = ep[p][2]; };
        rc = hull_canon<AS_LDS, 1>(w, &sb, r, KD_NE, epos);
    }
    if (r.status != HULL_OK || rc.nt <= 0) {
        // degenerate extremes (flat set) or no usable rough hull: no
        // prefilter, hull the full set
        if (tid == 0) {
            R.status = 1;
            res[blockIdx.x] = make_int2(np, 0);
            t.use_frame = 1; t.frame = full; t.filt = nullptr; t.nf = 0;
        }
        return;
    }
    const double rough = rc.vol;
    // bound pruning on sets above one wave's reach only (reference cuda/kdop_hull.cuh:49-115)
    if (t.prune_gap < 1e29 && np > LA_WAVE_MAX) {
        const double gap = fmax(rough - t.mesh_vol, 0.0);
        if (gap > t.prune_gap) {
            if (tid == 0) {
                R.status = 2;
                t.vol = rough;
                t.pruned = 1;
                t.nf = 0;
                t.filt = nullptr;
                res[blockIdx.x] = make_int2(0, 1);
            }
            return;
        }
    }
    // the live faces (hull_canon marks them), in face-id order
    if (tid == 0) {
        int c = 0;
        for (int f = 0; f < r.ftop && c < KD_MAXPL; ++f) {
            if (!(w.fvis[f] & 16)) continue;
            const u4v fv = w.fv[f];
            R.face[c][0] = (unsigned char)w.vpid[fv.x];
            R.face[c][1] = (unsigned char)w.vpid[fv.y];
            R.face[c][2] = (unsigned char)w.vpid[fv.z];
            ++c;
        }
        s_npl = c;
        R.npl = c;
        R.status = 0;
    }
}

// Launch 3: the exact face planes of the rough hull on the full set's grid; a
// point is dropped only when it lies strictly inside every face, i.e.
// strictly inside a solid whose vertices are points of the set, so it cannot
// be an extreme point; quantized duplicates share the decision. (Both grids
// are right-handed, so the faces keep their outward orientation.) The kept
// points, in input order, go to a scratch copy.
__global__ __launch_bounds__(KD_T) void la_kdop_filter_kernel(LaHullTask* tasks, const int* ids, int n,
                                                              const int* dcount, const KdopRec* rec, int2* res,
                                                              DPool* pool, int* err) {
    __shared__ double4 dpl[KD_MAXPL];
    __shared__ float4 fpl[KD_MAXPL];  // the same planes in FP32 ...
    __shared__ float fbd[KD_MAXPL];   // ... and a bound on that form's error
    __shared__ int s_cnt[KD_NW];
    __shared__ unsigned long long s_filt;
    if ((int)blockIdx.x >= (dcount ? min(*dcount, n) : n)) return;
    const KdopRec& R = rec[blockIdx.x];
    if (R.status != 0) return;
    LaHullTask& t = tasks[ids[blockIdx.x]];
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    const float* P = t.pts;
    const int np = t.n;
    const QFrame full = R.full;
    const int npl = R.npl;
    auto fq = [&](int p, double* o) {
        const float v[3] = {R.ep[p][0], R.ep[p][1], R.ep[p][2]};
        o[0] = (double)qcoord(full, 0, sel3(v, full.ax[0]));
        o[1] = (double)qcoord(full, 1, sel3(v, full.ax[1]));
        o[2] = (double)qcoord(full, 2, sel3(v, full.ax[2]));
    };
    for (int c = tid; c < npl; c += KD_T) {
        double a[3], b[3], cc[3];
        fq(R.face[c][0], a);
        fq(R.face[c][1], b);
        fq(R.face[c][2], cc);
        const double e1x = b[0] - a[0], e1y = b[1] - a[1], e1z = b[2] - a[2];
        const double e2x = cc[0] - a[0], e2y = cc[1] - a[1], e2z = cc[2] - a[2];
        const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
        const double pw = nx * a[0] + ny * a[1] + nz * a[2];
        dpl[c] = make_double4(nx, ny, nz, pw);
        // FP32 form: the four conversions and three roundings of the fused
        // chain each err by at most 2^-24 of T = (|nx| + |ny| + |nz|) 8192 +
        // |w| (grid coordinates are below 8192 in magnitude); 2^-20 T is four
        // times that sum
        fpl[c] = make_float4((float)nx, (float)ny, (float)nz, (float)pw);
        fbd[c] = (float)((fabs(nx) + fabs(ny) + fabs(nz)) * 8192.0 + fabs(pw)) * 0x1p-20f;
    }
    // filtered buffer: the points on or outside some rough face, in input order
    if (tid == 0) s_filt = (unsigned long long)heap_alloc(pool, 1, 12ull * (np + KD_NE), err);
    __syncthreads();
    float* filt = (float*)s_filt;
    if (!filt) {
        if (tid == 0) { t.err = KERR_KDOP_SCRATCH_OOM; res[blockIdx.x] = make_int2(-1, 0); }
        return;
    }
    int base = 0;
    for (int b = 0; b < np; b += KD_T) {
        const int i = b + tid;
        bool keep = false;
        float x = 0.f, y = 0.f, z = 0.f;
        if (i < np) {
            x = P[3ll * i]; y = P[3ll * i + 1]; z = P[3ll * i + 2];
            const float v[3] = {x, y, z};
            const double qx = (double)qcoord(full, 0, sel3(v, full.ax[0]));
            const double qy = (double)qcoord(full, 1, sel3(v, full.ax[1]));
            const double qz = (double)qcoord(full, 2, sel3(v, full.ax[2]));
            // every product and sum is an integer below 2^53: exact in FP64.
            // The FP32 form decides when it is farther from 0 than its error
            // bound; only the rest take the exact test, so the decision is
            // the exact one.
            const float fx = (float)qx, fy = (float)qy, fz = (float)qz;
            for (int f = 0; f < npl; ++f) {
                const float4 pf = fpl[f];
                const float sf = fmaf(pf.x, fx, fmaf(pf.y, fy, fmaf(pf.z, fz, -pf.w)));
                const float bf = fbd[f];
                if (sf > bf) { keep = true; break; }
                if (sf >= -bf) {
                    const double4 pl = dpl[f];
                    if (pl.x * qx + pl.y * qy + pl.z * qz - pl.w >= 0.0) { keep = true; break; }
                }
            }
        }
        const uint64_t m = ballot(keep);
        if (lane == 0) s_cnt[wv] = __popcll(m);
        __syncthreads();
        int wbase = base;
        for (int q = 0; q < wv; ++q) wbase += s_cnt[q];
        if (keep) {
            const int o = wbase + mask_rank(m);
            filt[3 * o] = x; filt[3 * o + 1] = y; filt[3 * o + 2] = z;
        }
        for (int q = 0; q < KD_NW; ++q) base += s_cnt[q];
        __syncthreads();
    }
    if (tid == 0) {
        t.filt = filt;
        t.nf = base;
        t.frame = full;
        t.use_frame = 1;
        res[blockIdx.x] = make_int2(base, 0);
    }
}

// The three launches of the prefilter over ids[0, n) (or the first *dcount).
int la_launch_kdop(gpu_ctx_t ctx, LaHullTask* d_tasks, const int* d_ids, int n, const int* dcount, int2* d_res,
                   KdopRec* d_rec, hipStream_t s) {
    if (n <= 0) return 0;
    REQUIRE(hull_ws_bytes<AS_LDS>(KD_VC, KD_VC) <= KD_WS, "k-DOP rough hull workspace exceeds its LDS");
    hipLaunchKernelGGL(la_kdop_ext_kernel, dim3((unsigned)n), dim3(64), 0, s, d_tasks, d_ids, n, dcount, d_rec);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(la_kdop_rough_kernel, dim3((unsigned)n), dim3(64), 0, s, d_tasks, d_ids, n, dcount, d_rec,
                       d_res);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(la_kdop_filter_kernel, dim3((unsigned)n), dim3(KD_T), 0, s, d_tasks, d_ids, n, dcount,
                       (const KdopRec*)d_rec, d_res, ctx->d_pool, ctx->d_err);
    HIP_TRY(hipGetLastError());
    return 0;
}

// Sets above one workgroup's reach (HULL_NP_MAX points after the prefilter)
// are hulled in chunks first: the hull of a set is the hull of the union of
// its chunks' hull corners. Chunks are contiguous runs of the set, each
// chunk's canonical corners come out in ascending point order, so the union
// is an order-preserving subsequence of the set under the set's own
// quantization frame - its canonical hull is the set's. One workgroup per
// oversized set: the chunk hulls' vertex arrays (segs[seg_off[i]..seg_off[i+1]),
// in chunk order) are concatenated into a new scratch chunk that replaces the
// task's filtered set; the chunk hulls are freed. A chunk without a hull
// (fewer than four points, or all coplanar - e.g. a run of a cut's cap
// vertices) passes all its points through (own = 0: not a heap chunk).
struct LaUnionSeg {
    const float* v;
    int n;
    int own;
};
__global__ __launch_bounds__(256) void la_union_kernel(LaHullTask* tasks, const int* ids, const int* seg_off,
                                                       const LaUnionSeg* segs, int n, DPool* pool, int* err) {
    __shared__ unsigned long long sp;
    if ((int)blockIdx.x >= n) return;
    LaHullTask& t = tasks[ids[blockIdx.x]];
    const int s0 = seg_off[blockIdx.x], s1 = seg_off[blockIdx.x + 1];
    int total = 0;
    for (int q = s0; q < s1; ++q) total += segs[q].n;
    if (threadIdx.x == 0) sp = (unsigned long long)heap_alloc(pool, 1, 12ull * total + 16, err);
    __syncthreads();
    float* dst = (float*)sp;
    if (dst) {
        int o = 0;
        for (int q = s0; q < s1; ++q) {
            const LaUnionSeg g = segs[q];
            for (int i = threadIdx.x; i < 3 * g.n; i += 256) dst[3 * o + i] = g.v[i];
            o += g.n;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = s0; q < s1; ++q)
            if (segs[q].own) heap_free(pool, (void*)segs[q].v, err);
        if (!dst) {
            t.err = KERR_KDOP_SCRATCH_OOM;
            err_or(err, KERR_KDOP_SCRATCH_OOM);
            return;
        }
        if (t.filt) heap_free(pool, t.filt, err);
        t.filt = dst;
        t.nf = total;
    }
}

// ---------------------------------------------------------------------------
// Level-0 cut: both closed halves (kept in heap 0), volumes and boxes
// ---------------------------------------------------------------------------
__device__ inline void la_block_bbox(const float* v, int nv, float* lo, float* hi, PcLDS* s, float (*scr)[6]) {
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = threadIdx.x; i < nv; i += PC_T)
        for (int j = 0; j < 3; ++j) {
            const float x = v[3 * i + j];
            mn[j] = fminf(mn[j], x);
            mx[j] = fmaxf(mx[j], x);
        }
    for (int j = 0; j < 3; ++j) { mn[j] = wave_min_f(mn[j]); mx[j] = wave_max_f(mx[j]); }
    if (lane_id() == 0)
        for (int j = 0; j < 3; ++j) { scr[threadIdx.x >> 6][j] = mn[j]; scr[threadIdx.x >> 6][3 + j] = mx[j]; }
    __syncthreads();
    for (int j = 0; j < 3; ++j) {
        float a = scr[0][j], b = scr[0][3 + j];
        for (int q = 1; q < PC_T / 64; ++q) { a = fminf(a, scr[q][j]); b = fmaxf(b, scr[q][3 + j]); }
        lo[j] = a;
        hi[j] = b;
    }
    __syncthreads();
    (void)s;
}

// at least four workgroups per CU (the caps' ear clipping is a serial chain
// on one wave, so resident workgroups, not registers, set the throughput:
// 128 VGPRs with some spills measured 149 ms per 64-mesh step against 168 ms
// at three workgroups and 215 ms at two)
__global__ __launch_bounds__(PC_T, 4) void la_cut_kernel(LaCutTask* tasks, int n, DPool* pool, int* err,
                                                         int lds_nv) {
    __shared__ PcLDS s;
    __shared__ float scr[PC_T / 64][6];
    extern __shared__ __align__(16) unsigned char pc_dyn[];
    // the candidates of one part are consecutive tasks: groups of 16 per XCD
    const int slot = (int)xcd_group_slot(blockIdx.x, gridDim.x, 16u);
    if (slot >= n) return;
    LaCutTask& T = tasks[slot];
    PcResult r = T.nv <= lds_nv ? plane_cut_block_lds(T.v, T.t, T.nv, T.nt, T.plane[0], T.plane[1], T.plane[2],
                                                      T.plane[3], pool, err, &s, pc_dyn)
                                : plane_cut_block(T.v, T.t, T.nv, T.nt, T.plane[0], T.plane[1], T.plane[2],
                                                  T.plane[3], pool, err, &s);
    float plo[3], phi[3], nlo[3], nhi[3];
    const bool ok = r.status == 0 && r.pos.v && r.neg.v;
    if (ok && r.has_box) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            plo[j] = s.boxr[j];
            phi[j] = s.boxr[3 + j];
            nlo[j] = s.boxr[6 + j];
            nhi[j] = s.boxr[9 + j];
        }
    } else if (ok) {
        la_block_bbox(r.pos.v, r.pos.nv, plo, phi, &s, scr);
        la_block_bbox(r.neg.v, r.neg.nv, nlo, nhi, &s, scr);
    }
    if (threadIdx.x == 0) {
        if (!ok) {
            heap_free(pool, r.pos.v, err);
            heap_free(pool, r.neg.v, err);
            T.pv = T.nvp = nullptr;
            T.pt = T.ntp = nullptr;
            T.pnv = T.pnt = T.nnv = T.nnt = 0;
            T.pvol = T.nvol = 0.0;
            T.status = r.status == 0 ? 2 : r.status;
        } else {
            T.pv = r.pos.v; T.pt = r.pos.t; T.pnv = r.pos.nv; T.pnt = r.pos.nt; T.pvol = r.pos.vol;
            T.nvp = r.neg.v; T.ntp = r.neg.t; T.nnv = r.neg.nv; T.nnt = r.neg.nt; T.nvol = r.neg.vol;
            for (int j = 0; j < 3; ++j) { T.plo[j] = plo[j]; T.phi[j] = phi[j]; T.nlo[j] = nlo[j]; T.nhi[j] = nhi[j]; }
            T.status = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// Leaf scoring: half-space volumes and hull point sets of both sides
// ---------------------------------------------------------------------------
// One pass over the triangles accumulates both sides' volumes, marks the
// vertices of each side and lists the triangles with vertices strictly on
// both sides (in index order); the crossing points - one per undirected edge
// with strictly opposite signs, in order of first occurrence over (triangle,
// edge) - come from that short list alone.
__global__ __launch_bounds__(PC_T) void la_score_kernel(LaScoreTask* tasks, int n, DPool* pool, int* err,
                                                        int lds_nv) {
    __shared__ PcLDS s;
    extern __shared__ __align__(16) unsigned char pc_dyn[];
    __shared__ double dsum[2][PC_T / 64];
    __shared__ int wcnt[PC_T / 64];
    // the candidates of one item's part are consecutive tasks: groups of 8 per XCD
    const int slot = (int)xcd_group_slot(blockIdx.x, gridDim.x, 8u);
    if (slot >= n) return;
    LaScoreTask& T = tasks[slot];
    const int tid = threadIdx.x, wv = tid >> 6;
    const float* V = T.v;
    const int* Tr = T.t;
    const int nv = T.nv, nt = T.nt;
    const float pa = T.plane[0], pb = T.plane[1], pc = T.plane[2], pd = T.plane[3];
    // One flag byte per vertex - bits 0-1 sign code (1 positive, 2 negative, 0
    // on the plane), bit 2 in the positive point set, bit 3 in the negative
    // one - in LDS when the mesh fits the launch's dynamic LDS, else in a heap
    // chunk of its own.
    const bool flags_in_lds = nv <= lds_nv;
    char* F = nullptr;
    if (!flags_in_lds) {
        F = (char*)pc_alloc(pool, al16(nv), &s, 1, err, 1);
        if (!F) return;
    }
    uint8_t* fl = flags_in_lds ? pc_dyn : (uint8_t*)F;
    for (int i = tid; i < nv; i += PC_T) {
        const float d = pc_side(pa, pb, pc, pd, V[3 * i], V[3 * i + 1], V[3 * i + 2]);
        fl[i] = d > PC_EPS ? 1 : (d < -PC_EPS ? 2 : 0);
    }
    __syncthreads();
    auto sgn = [&](int i) -> int {
        const int c = fl[i] & 3;
        return c == 1 ? 1 : (c == 2 ? -1 : 0);
    };
    // triangles with vertices strictly on both sides, counted first so that one
    // scratch chunk holds their list and the crossing-edge hash (every heap
    // operation synchronises the XCD's L2 with memory)
    int nx0 = 0;
    for (int k = tid; k < nt; k += PC_T) {
        const int s0 = sgn(Tr[3 * k]), s1 = sgn(Tr[3 * k + 1]), s2 = sgn(Tr[3 * k + 2]);
        nx0 += ((s0 > 0) | (s1 > 0) | (s2 > 0)) & ((s0 < 0) | (s1 < 0) | (s2 < 0));
    }
    nx0 = wave_sum_i(nx0);
    if (lane_id() == 0) wcnt[wv] = nx0;
    __syncthreads();
    nx0 = 0;
    for (int q = 0; q < PC_T / 64; ++q) nx0 += wcnt[q];
    __syncthreads();
    uint32_t H = 16;
    while ((int)H < 4 * nx0 + 2) H <<= 1;
    const unsigned long long szX = al16(4ull * nx0);
    char* A = (char*)pc_alloc(pool, szX + 12ull * H + 8ull * nx0 + 64, &s, 2, err, 1);
    if (!A) {
        if (tid == 0 && F) heap_free(pool, F, err);
        return;
    }
    int* xl = (int*)A;
    // divergence origin: the point of the plane nearest the coordinate origin
    const double n2 = (double)pa * pa + (double)pb * pb + (double)pc * pc;
    const double ox = -(double)pd * pa / n2, oy = -(double)pd * pb / n2, oz = -(double)pd * pc / n2;
    auto tet = [&](const float* a, const float* b, const float* c) {
        const double ax = a[0] - ox, ay = a[1] - oy, az = a[2] - oz;
        const double bx = b[0] - ox, by = b[1] - oy, bz = b[2] - oz;
        const double cx = c[0] - ox, cy = c[1] - oy, cz = c[2] - oz;
        return ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
    };
    // crossing point of edge (a, b), computed as the cut computes it
    auto xpt = [&](int a, int b, float* o) {
        const int lo = min(a, b), hi = max(a, b);
        const float ax = V[3 * lo], ay = V[3 * lo + 1], az = V[3 * lo + 2];
        const float bx = V[3 * hi], by = V[3 * hi + 1], bz = V[3 * hi + 2];
        const float d0 = pc_side(pa, pb, pc, pd, ax, ay, az), d1 = pc_side(pa, pb, pc, pd, bx, by, bz);
        const float t = d0 / (d0 - d1);
        o[0] = ax + t * (bx - ax);
        o[1] = ay + t * (by - ay);
        o[2] = az + t * (bz - az);
    };
    auto mark = [&](int i, int side) {
        const unsigned bits = (unsigned)side << 2;
        if ((fl[i] & bits) != bits) atomicOr((unsigned*)(fl + (i & ~3)), bits << (8 * (i & 3)));
    };
    double vp = 0.0, vn = 0.0;
    int cntp = 0, cntn = 0;
    int nx = 0;
    for (int b0 = 0; b0 < nt; b0 += PC_T) {
        const int k = b0 + tid;
        bool crossing = false;
        if (k < nt) {
            const int i0 = Tr[3 * k], i1 = Tr[3 * k + 1], i2 = Tr[3 * k + 2];
            const int s0 = sgn(i0), s1 = sgn(i1), s2 = sgn(i2);
            const float* p0 = V + 3 * i0;
            const float* p1 = V + 3 * i1;
            const float* p2 = V + 3 * i2;
            const int hp = (s0 > 0) | (s1 > 0) | (s2 > 0), hn = (s0 < 0) | (s1 < 0) | (s2 < 0);
            if (!hp || !hn) {
                int side;
                if (hn) side = 2;
                else if (hp) side = 1;
                else {
                    const float e1x = p1[0] - p0[0], e1y = p1[1] - p0[1], e1z = p1[2] - p0[2];
                    const float e2x = p2[0] - p0[0], e2y = p2[1] - p0[1], e2z = p2[2] - p0[2];
                    const float dt = pa * (e1y * e2z - e1z * e2y) + pb * (e1z * e2x - e1x * e2z) + pc * (e1x * e2y - e1y * e2x);
                    side = dt > 0 ? 1 : 2;
                }
                const double v = tet(p0, p1, p2);
                if (side == 1) { vp += v; ++cntp; } else { vn += v; ++cntn; }
                mark(i0, side); mark(i1, side); mark(i2, side);
            } else {
                crossing = true;
                const int vi[3] = {i0, i1, i2}, si[3] = {s0, s1, s2};
                int on = -1;
                for (int q = 0; q < 3; ++q) if (si[q] == 0) { on = q; break; }
                if (on >= 0) {
                    const int ov = vi[on], a = vi[(on + 1) % 3], b = vi[(on + 2) % 3], sa = si[(on + 1) % 3];
                    float x[3];
                    xpt(a, b, x);
                    const float* O = V + 3 * ov;
                    const float* A3 = V + 3 * a;
                    const float* B3 = V + 3 * b;
                    if (sa > 0) { vp += tet(O, A3, x); vn += tet(O, x, B3); }
                    else { vn += tet(O, A3, x); vp += tet(O, x, B3); }
                    ++cntp; ++cntn;
                    mark(ov, 3);
                    mark(a, sa > 0 ? 1 : 2);
                    mark(b, sa > 0 ? 2 : 1);
                } else {
                    int lone = 0;
                    for (int q = 0; q < 3; ++q)
                        if (si[q] != si[(q + 1) % 3] && si[q] != si[(q + 2) % 3]) { lone = q; break; }
                    const int lv = vi[lone], o1 = vi[(lone + 1) % 3], o2 = vi[(lone + 2) % 3];
                    float x1[3], x2[3];
                    xpt(lv, o1, x1);
                    xpt(lv, o2, x2);
                    const float* L = V + 3 * lv;
                    const float* Q1 = V + 3 * o1;
                    const float* Q2 = V + 3 * o2;
                    const double a1 = tet(L, x1, x2);
                    const double a2 = tet(Q1, x2, x1) + tet(Q1, Q2, x2);
                    const int ls = si[lone] > 0 ? 1 : 2;
                    if (ls == 1) { vp += a1; vn += a2; } else { vn += a1; vp += a2; }
                    ++cntp; ++cntn;
                    mark(lv, ls);
                    mark(o1, 3 - ls);
                    mark(o2, 3 - ls);
                }
            }
        }
        // ordered append of the crossing triangles
        const uint64_t m = ballot(crossing);
        if (lane_id() == 0) wcnt[wv] = __popcll(m);
        __syncthreads();
        int wbase = nx, tot = 0;
        for (int q = 0; q < PC_T / 64; ++q) {
            if (q < wv) wbase += wcnt[q];
            tot += wcnt[q];
        }
        if (crossing) xl[wbase + mask_rank(m)] = k;
        nx += tot;
        __syncthreads();
    }
    vp = wave_sum_d(vp);
    vn = wave_sum_d(vn);
    cntp = wave_sum_i(cntp);
    cntn = wave_sum_i(cntn);
    if (lane_id() == 0) {
        dsum[0][wv] = vp;
        dsum[1][wv] = vn;
        s.cnt[wv] = cntp;
        s.cnt[4 + wv] = cntn;
    }
    __syncthreads();
    double VP = 0.0, VN = 0.0;
    int totp = 0, totn = 0;
    for (int q = 0; q < PC_T / 64; ++q) {
        VP += dsum[0][q];
        VN += dsum[1][q];
        totp += s.cnt[q];
        totn += s.cnt[4 + q];
    }
    __syncthreads();
    // instances of the strictly crossing edges over the listed triangles
    auto strict = [&](int a, int b) {
        const int sa = sgn(a), sb = sgn(b);
        return (sa > 0 && sb < 0) || (sa < 0 && sb > 0);
    };
    auto ninst = [&](int x) {
        const int k = xl[x];
        const int a = Tr[3 * k], b = Tr[3 * k + 1], c = Tr[3 * k + 2];
        return (int)strict(a, b) + (int)strict(b, c) + (int)strict(c, a);
    };
    char* B = A + szX;  // nx == nx0
    unsigned long long* hk = (unsigned long long*)B;
    int* hmin = (int*)(B + 8ull * H);
    int* xoff = hmin + H;
    int* xnew = xoff + nx;
    const int ncross = pc_block_scan(nx, xoff, &s, ninst);
    (void)ncross;
    for (uint32_t i = tid; i < H; i += PC_T) { hk[i] = PC_EMPTY; hmin[i] = 0x7fffffff; }
    __syncthreads();
    for (int x = tid; x < nx; x += PC_T) {
        const int k = xl[x];
        int inst = xoff[x];
        for (int e = 0; e < 3; ++e) {
            const int a = Tr[3 * k + e], b = Tr[3 * k + (e == 2 ? 0 : e + 1)];
            if (!strict(a, b)) continue;
            int fresh;
            const int slot = pc_ht_insert(hk, H - 1, ((unsigned long long)min(a, b) << 32) | (unsigned)max(a, b), &fresh);
            atomicMin(&hmin[slot], inst);
            ++inst;
        }
    }
    __syncthreads();
    auto owned_new = [&](int x) {
        const int k = xl[x];
        int inst = xoff[x], c = 0;
        for (int e = 0; e < 3; ++e) {
            const int a = Tr[3 * k + e], b = Tr[3 * k + (e == 2 ? 0 : e + 1)];
            if (!strict(a, b)) continue;
            const int slot = pc_ht_find(hk, H - 1, ((unsigned long long)min(a, b) << 32) | (unsigned)max(a, b));
            if (hmin[slot] == inst) ++c;
            ++inst;
        }
        return c;
    };
    const int n_new = pc_block_scan(nx, xnew, &s, owned_new);
    // point sets: marked original vertices in index order, then the new points
    int c1 = 0, c2 = 0;
    for (int i = tid; i < nv; i += PC_T) {
        c1 += (fl[i] >> 2) & 1;
        c2 += (fl[i] >> 3) & 1;
    }
    c1 = wave_sum_i(c1);
    c2 = wave_sum_i(c2);
    if (lane_id() == 0) { s.cnt[wv] = c1; s.cnt[4 + wv] = c2; }
    __syncthreads();
    int nposo = 0, nnego = 0;
    for (int q = 0; q < PC_T / 64; ++q) { nposo += s.cnt[q]; nnego += s.cnt[4 + q]; }
    __syncthreads();
    float* pp = nullptr;
    float* npp = nullptr;
    if (totp > 0 && totn > 0) {
        // both point sets in one chunk (freed through the positive set's pointer)
        const unsigned long long pos_bytes = al16(12ull * (nposo + n_new));
        pp = (float*)pc_alloc(pool, pos_bytes + 12ull * (nnego + n_new) + 16, &s, 3, err, 1);
        npp = pp ? (float*)((char*)pp + pos_bytes) : nullptr;
        if (pp) {
            // originals, stable by index: per-chunk ballot compaction
            int bp = 0, bn = 0;
            for (int b0 = 0; b0 < nv; b0 += PC_T) {
                const int i = b0 + tid;
                const int f1 = i < nv ? (fl[i] >> 2) & 1 : 0, f2 = i < nv ? (fl[i] >> 3) & 1 : 0;
                const uint64_t mp = ballot(f1), mn_ = ballot(f2);
                if (lane_id() == 0) { s.cnt[wv] = __popcll(mp); s.cnt[4 + wv] = __popcll(mn_); }
                __syncthreads();
                int op = bp, on_ = bn;
                for (int q = 0; q < wv; ++q) { op += s.cnt[q]; on_ += s.cnt[4 + q]; }
                if (f1) {
                    const int o = op + mask_rank(mp);
                    pp[3 * o] = V[3 * i]; pp[3 * o + 1] = V[3 * i + 1]; pp[3 * o + 2] = V[3 * i + 2];
                }
                if (f2) {
                    const int o = on_ + mask_rank(mn_);
                    npp[3 * o] = V[3 * i]; npp[3 * o + 1] = V[3 * i + 1]; npp[3 * o + 2] = V[3 * i + 2];
                }
                for (int q = 0; q < PC_T / 64; ++q) { bp += s.cnt[q]; bn += s.cnt[4 + q]; }
                __syncthreads();
            }
            for (int x = tid; x < nx; x += PC_T) {
                const int k = xl[x];
                int inst = xoff[x], c = 0;
                for (int e = 0; e < 3; ++e) {
                    const int a = Tr[3 * k + e], b = Tr[3 * k + (e == 2 ? 0 : e + 1)];
                    if (!strict(a, b)) continue;
                    const int slot =
                        pc_ht_find(hk, H - 1, ((unsigned long long)min(a, b) << 32) | (unsigned)max(a, b));
                    if (hmin[slot] == inst) {
                        float q3[3];
                        xpt(a, b, q3);
                        const int j = xnew[x] + c;
                        float* dp = pp + 3ll * (nposo + j);
                        float* dn = npp + 3ll * (nnego + j);
                        dp[0] = q3[0]; dp[1] = q3[1]; dp[2] = q3[2];
                        dn[0] = q3[0]; dn[1] = q3[1]; dn[2] = q3[2];
                        ++c;
                    }
                    ++inst;
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        T.pvol = VP / 6.0;
        T.nvol = VN / 6.0;
        T.npt = totp;
        T.nnt = totn;
        T.ppts = pp;
        T.npts = npp;
        T.np_pts = pp ? nposo + n_new : 0;
        T.nn_pts = npp ? nnego + n_new : 0;
        void* const f[2] = {A, F};
        heap_free_n(pool, f, err);
    }
}

// ---------------------------------------------------------------------------
// Connected components (reference cuda/postprocess.cuh:139-527)
// ---------------------------------------------------------------------------
__device__ inline int uf_find(int* par, int x) {
    for (;;) {
        const int p = __atomic_load_n(&par[x], __ATOMIC_RELAXED);
        if (p == x) return x;
        x = p;
    }
}
__device__ inline void uf_union(int* par, int a, int b) {
    for (;;) {
        a = uf_find(par, a);
        b = uf_find(par, b);
        if (a == b) return;
        const int hi = max(a, b), lo = min(a, b);
        if (atomicCAS(&par[hi], hi, lo) == hi) return;
    }
}

__global__ __launch_bounds__(PC_T) void la_comp_kernel(LaCompTask* tasks, int n, DPool* pool, int* err) {
    __shared__ PcLDS s;
    __shared__ double dred[PC_T / 64];
    __shared__ float scr[PC_T / 64][6];
    if ((int)blockIdx.x >= n) return;
    LaCompTask& T = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    const float* V = T.v;
    const int* Tr = T.t;
    const int nv = T.nv, nt = T.nt;
    if (nv <= 0 || nt <= 0) {
        if (tid == 0) T.n_comp = 0;
        return;
    }
    char* A = (char*)pc_alloc(pool, 12ull * nv + 64, &s, 1, err, 1);
    if (!A) {
        if (tid == 0) { T.n_comp = 0; err_or(err, KERR_DC_OOM); }
        return;
    }
    int* par = (int*)A;
    int* cid = par + nv;   // used flag, then component of a root
    int* lidx = cid + nv;  // local index inside its component
    for (int i = tid; i < nv; i += PC_T) { par[i] = i; cid[i] = 0; }
    __syncthreads();
    for (int k = tid; k < nt; k += PC_T) {
        const int a = Tr[3 * k], b = Tr[3 * k + 1], c = Tr[3 * k + 2];
        cid[a] = 1; cid[b] = 1; cid[c] = 1;
        uf_union(par, a, b);
        uf_union(par, b, c);
    }
    __syncthreads();
    for (int i = tid; i < nv; i += PC_T) par[i] = uf_find(par, i);
    __syncthreads();
    // components numbered by their lowest vertex
    const int ncomp = pc_block_scan(nv, lidx, &s, [&](int i) { return (cid[i] && par[i] == i) ? 1 : 0; });
    if (ncomp <= 1) {
        if (tid == 0) { T.n_comp = 0; heap_free(pool, A, err); }
        return;
    }
    // component id of every used vertex (roots first, then members)
    for (int i = tid; i < nv; i += PC_T)
        if (cid[i] && par[i] == i) cid[i] = 1 + min(lidx[i], LA_MAX_COMP - 1);
    __syncthreads();
    for (int i = tid; i < nv; i += PC_T)
        if (cid[i] && par[i] != i) cid[i] = -(1 + min(lidx[par[i]], LA_MAX_COMP - 1));
    __syncthreads();
    for (int i = tid; i < nv; i += PC_T) if (cid[i] < 0) cid[i] = -cid[i];
    __syncthreads();
    const int nc = min(ncomp, LA_MAX_COMP);
    for (int c = 0; c < nc; ++c) {
        const int cv = pc_block_scan(nv, lidx, &s, [&](int i) { return cid[i] == c + 1 ? 1 : 0; });
        const int ct = pc_block_scan(nt, nullptr, &s, [&](int k) { return cid[Tr[3 * k]] == c + 1 ? 1 : 0; });
        const unsigned long long vb = al16(12ull * cv);
        float* chunk = (float*)pc_alloc(pool, vb + 12ull * ct, &s, 2, err, 0);
        if (!chunk) {
            if (tid == 0) { T.out[c].v = nullptr; T.out[c].nv = 0; T.out[c].nt = 0; T.out[c].svol = 0.0; }
            continue;
        }
        int* ot = (int*)((char*)chunk + vb);
        for (int i = tid; i < nv; i += PC_T)
            if (cid[i] == c + 1) {
                const int o = lidx[i];
                chunk[3 * o] = V[3 * i]; chunk[3 * o + 1] = V[3 * i + 1]; chunk[3 * o + 2] = V[3 * i + 2];
            }
        __syncthreads();
        // triangles in input order (ballot compaction per 256-chunk)
        int base = 0;
        for (int b0 = 0; b0 < nt; b0 += PC_T) {
            const int k = b0 + tid;
            const bool in = k < nt && cid[Tr[3 * k]] == c + 1;
            const uint64_t m = ballot(in);
            if (lane_id() == 0) s.cnt[tid >> 6] = __popcll(m);
            __syncthreads();
            int o = base;
            for (int q = 0; q < (tid >> 6); ++q) o += s.cnt[q];
            if (in) {
                o += mask_rank(m);
                ot[3 * o] = lidx[Tr[3 * k]]; ot[3 * o + 1] = lidx[Tr[3 * k + 1]]; ot[3 * o + 2] = lidx[Tr[3 * k + 2]];
            }
            for (int q = 0; q < PC_T / 64; ++q) base += s.cnt[q];
            __syncthreads();
        }
        // FP64 signed volume anchored at the component's first vertex
        double acc = 0.0;
        const double ax = chunk[0], ay = chunk[1], az = chunk[2];
        for (int k = tid; k < ct; k += PC_T) {
            const int a = ot[3 * k], b = ot[3 * k + 1], d = ot[3 * k + 2];
            const double x0 = chunk[3 * a] - ax, y0 = chunk[3 * a + 1] - ay, z0 = chunk[3 * a + 2] - az;
            const double x1 = chunk[3 * b] - ax, y1 = chunk[3 * b + 1] - ay, z1 = chunk[3 * b + 2] - az;
            const double x2 = chunk[3 * d] - ax, y2 = chunk[3 * d + 1] - ay, z2 = chunk[3 * d + 2] - az;
            acc += x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
        }
        acc = wave_sum_d(acc);
        if (lane_id() == 0) dred[tid >> 6] = acc;
        float lo[3], hi[3];
        la_block_bbox(chunk, cv, lo, hi, &s, scr);
        double sv = 0.0;
        for (int q = 0; q < PC_T / 64; ++q) sv += dred[q];
        if (tid == 0) {
            LaCompOut& o = T.out[c];
            o.v = chunk; o.t = ot; o.nv = cv; o.nt = ct; o.svol = sv / 6.0;
            for (int j = 0; j < 3; ++j) { o.lo[j] = lo[j]; o.hi[j] = hi[j]; }
        }
        __syncthreads();
    }
    if (tid == 0) {
        // drop inner shells (negative signed volume) when inner and outer mix
        int n_inner = 0;
        for (int c = 0; c < nc; ++c) n_inner += T.out[c].v && T.out[c].svol < 0.0;
        int keep = nc;
        if (n_inner > 0 && n_inner < nc) {
            keep = 0;
            for (int c = 0; c < nc; ++c) {
                if (T.out[c].v && T.out[c].svol < 0.0) {
                    heap_free(pool, T.out[c].v, err);
                } else {
                    if (keep != c) T.out[keep] = T.out[c];
                    ++keep;
                }
            }
        }
        T.n_comp = keep;
        heap_free(pool, A, err);
    }
}

// ---------------------------------------------------------------------------
// Concave edges (reference cuda/la_concave.cu:48-363)
// ---------------------------------------------------------------------------
__device__ inline bool la_unit_normal(const float* V, const int* Tr, int f, float* n) {
    const int a = Tr[3 * f], b = Tr[3 * f + 1], c = Tr[3 * f + 2];
    const float e1x = V[3 * b] - V[3 * a], e1y = V[3 * b + 1] - V[3 * a + 1], e1z = V[3 * b + 2] - V[3 * a + 2];
    const float e2x = V[3 * c] - V[3 * a], e2y = V[3 * c + 1] - V[3 * a + 1], e2z = V[3 * c + 2] - V[3 * a + 2];
    float x = e1y * e2z - e1z * e2y, y = e1z * e2x - e1x * e2z, z = e1x * e2y - e1y * e2x;
    const float len = sqrtf(x * x + y * y + z * z);
    if (len < 1e-12f) return false;
    n[0] = x / len; n[1] = y / len; n[2] = z / len;
    return true;
}

__global__ __launch_bounds__(PC_T) void la_edges_kernel(LaEdgeTask* tasks, int n, float cos_pos, float cos_neg,
                                                        int pos_mode, int neg_mode, DPool* pool, int* err) {
    __shared__ PcLDS s;
    if ((int)blockIdx.x >= n) return;
    LaEdgeTask& T = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    const float* V = T.v;
    const int* Tr = T.t;
    const int nt = T.nt;
    if (T.nv <= 0 || nt < 2) {
        if (tid == 0) T.count = 0;
        return;
    }
    uint32_t H = 16;
    while ((int)H < 6 * nt + 2) H <<= 1;
    char* B = (char*)pc_alloc(pool, 16ull * H + 64, &s, 1, err, 1);
    if (!B) {
        if (tid == 0) T.count = 0;
        return;
    }
    unsigned long long* hk = (unsigned long long*)B;
    int* f0 = (int*)(B + 8ull * H);
    int* f1 = f0 + H;
    for (uint32_t i = tid; i < H; i += PC_T) { hk[i] = PC_EMPTY; f0[i] = 0x7fffffff; f1[i] = 0x7fffffff; }
    if (tid == 0) s.cnt[0] = 0;
    __syncthreads();
    auto key = [&](int k, int e) {
        const int a = Tr[3 * k + e], b = Tr[3 * k + (e == 2 ? 0 : e + 1)];
        return ((unsigned long long)min(a, b) << 32) | (unsigned)max(a, b);
    };
    for (int k = tid; k < nt; k += PC_T)
        for (int e = 0; e < 3; ++e) {
            int fresh;
            const int slot = pc_ht_insert(hk, H - 1, key(k, e), &fresh);
            atomicMin(&f0[slot], k);
        }
    __syncthreads();
    for (int k = tid; k < nt; k += PC_T)
        for (int e = 0; e < 3; ++e) {
            const int slot = pc_ht_find(hk, H - 1, key(k, e));
            if (f0[slot] != k) atomicMin(&f1[slot], k);
        }
    __syncthreads();
    for (int k = tid; k < nt; k += PC_T)
        for (int e = 0; e < 3; ++e) {
            const unsigned long long kk = key(k, e);
            // the edge's first face handles it, once (first matching edge slot of that face)
            bool first = true;
            for (int e2 = 0; e2 < e; ++e2) first &= key(k, e2) != kk;
            if (!first) continue;
            const int slot = pc_ht_find(hk, H - 1, kk);
            if (f0[slot] != k || f1[slot] == 0x7fffffff) continue;
            const int g = f1[slot];
            float n0[3], n1[3];
            if (!la_unit_normal(V, Tr, k, n0) || !la_unit_normal(V, Tr, g, n1)) continue;
            const int v0 = (int)(kk >> 32), v1 = (int)(kk & 0xFFFFFFFFu);
            float ca = n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2];
            ca = fmaxf(-1.0f, fminf(1.0f, ca));
            int vo = -1;
            for (int q = 0; q < 3; ++q) {
                const int vv = Tr[3 * k + q];
                if (vv != v0 && vv != v1) { vo = vv; break; }
            }
            if (vo < 0) continue;
            const float dx = V[3 * vo] - V[3 * v0], dy = V[3 * vo + 1] - V[3 * v0 + 1], dz = V[3 * vo + 2] - V[3 * v0 + 2];
            const float d = dx * n1[0] + dy * n1[1] + dz * n1[2];
            // dihedral = pi + acos(ca) if d > 0 else pi - acos(ca); concave iff dihedral > threshold,
            // decided on the cosine (acos is decreasing) with the host's per-branch constants
            bool concave;
            if (d > 0.0f) concave = pos_mode == 1 || (pos_mode == 2 && ca < cos_pos);
            else concave = neg_mode == 1 || (neg_mode == 2 && ca > cos_neg);
            if (!concave) continue;
            const int j = atomicAdd(&s.cnt[0], 1);
            if (j < T.cap) {
                LaEdgeRec& r = T.out[j];
                r.lo = v0; r.hi = v1; r.f0 = k; r.f1 = g;
                for (int q = 0; q < 3; ++q) {
                    r.n0[q] = n0[q];
                    r.n1[q] = n1[q];
                    r.mid[q] = (V[3 * v0 + q] + V[3 * v1 + q]) * 0.5f;
                }
                r.pad = 0;
            }
        }
    __syncthreads();
    if (tid == 0) {
        T.count = s.cnt[0];
        heap_free(pool, B, err);
    }
}

__global__ __launch_bounds__(PC_T) void la_bbox_kernel(LaBoxTask* tasks, int n) {
    __shared__ PcLDS s;
    __shared__ float scr[PC_T / 64][6];
    if ((int)blockIdx.x >= n) return;
    LaBoxTask& T = tasks[blockIdx.x];
    float lo[3], hi[3];
    la_block_bbox(T.v, T.nv, lo, hi, &s, scr);
    if (threadIdx.x == 0)
        for (int j = 0; j < 3; ++j) { T.lo[j] = lo[j]; T.hi[j] = hi[j]; }
}

__global__ __launch_bounds__(256) void la_segcopy_kernel(const LaSeg* segs, int n) {
    if ((int)blockIdx.x >= n) return;
    const LaSeg g = segs[blockIdx.x];
    const uint32_t* src = (const uint32_t*)g.src;
    uint32_t* dst = (uint32_t*)g.dst;
    const unsigned long long words = g.bytes / 4;
    for (unsigned long long i = threadIdx.x; i < words; i += 256) dst[i] = src[i];
}

__global__ __launch_bounds__(PC_T) void la_gather_kernel(LaGatherTask* tasks, int n) {
    __shared__ int wc[PC_T / 64];
    if ((int)blockIdx.x >= n) return;
    LaGatherTask& T = tasks[blockIdx.x];
    const int tid = threadIdx.x;
    for (int i = tid; i < 3 * T.n1; i += PC_T) T.dv[i] = T.v1[i];
    for (int i = tid; i < 3 * T.n2; i += PC_T) T.dv[3 * T.n1 + i] = T.v2[i];
    if (!T.dt) return;
    const float pa = T.plane[0], pb = T.plane[1], pc = T.plane[2], pd = T.plane[3];
    auto on_plane = [&](const float* V, int i) {
        const float s = pa * V[3 * i] + pb * V[3 * i + 1] + pc * V[3 * i + 2] - pd;
        return fabsf(s) <= 1e-3f;
    };
    int base = 0;
    for (int part = 0; part < 2; ++part) {
        const float* V = part ? T.v2 : T.v1;
        const int* Tr = part ? T.t2 : T.t1;
        const int m = part ? T.m2 : T.m1;
        const int off = part ? T.n1 : 0;
        for (int b0 = 0; b0 < m; b0 += PC_T) {
            const int k = b0 + tid;
            bool keep = false;
            int a = 0, b = 0, c = 0;
            if (k < m) {
                a = Tr[3 * k]; b = Tr[3 * k + 1]; c = Tr[3 * k + 2];
                keep = !(T.has_plane && on_plane(V, a) && on_plane(V, b) && on_plane(V, c));
            }
            const uint64_t msk = ballot(keep);
            if (lane_id() == 0) wc[tid >> 6] = __popcll(msk);
            __syncthreads();
            int o = base;
            for (int q = 0; q < (tid >> 6); ++q) o += wc[q];
            if (keep) {
                o += mask_rank(msk);
                T.dt[3 * o] = a + off; T.dt[3 * o + 1] = b + off; T.dt[3 * o + 2] = c + off;
            }
            for (int q = 0; q < PC_T / 64; ++q) base += wc[q];
            __syncthreads();
        }
    }
    if (tid == 0) T.nt_out = base;
}

__global__ __launch_bounds__(256) void la_free_kernel(void** ptrs, int n, DPool* pool, int* err) {
    __shared__ unsigned long long lds[DH_BATCH_LDS / 2];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    heap_free_batch(pool, i < n ? ptrs[i] : nullptr, err, (uint32_t*)lds);
}

// host helpers --------------------------------------------------------------------
template <class T> T* dev_tmp(gpu_ctx_t ctx, size_t n) {
    void* p = nullptr;
    if (hipMallocAsync(&p, n * sizeof(T) + 16, ctx->stream) != hipSuccess) return nullptr;
    return (T*)p;
}

int kdop_axes_ready = -1;

int ensure_kdop_axes(gpu_ctx_t ctx) {
    if (kdop_axes_ready == ctx->device) return 0;
    // level-1 icosphere (12 vertices + 30 edge midpoints on the unit sphere)
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<std::array<double, 3>> v = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                                            {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    for (auto& p : v) {
        const double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (auto& c : p) c /= l;
    }
    const int F[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    auto mid = [&](int a, int b) {
        std::array<double, 3> m{v[a][0] + v[b][0], v[a][1] + v[b][1], v[a][2] + v[b][2]};
        const double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        for (auto& c : m) c /= l;
        for (size_t i = 0; i < v.size(); ++i)
            if (std::fabs(v[i][0] - m[0]) + std::fabs(v[i][1] - m[1]) + std::fabs(v[i][2] - m[2]) < 1e-12) return (int)i;
        v.push_back(m);
        return (int)v.size() - 1;
    };
    std::vector<std::array<double, 3>> normals;
    for (auto& f : F) {
        const int a = f[0], b = f[1], c = f[2];
        const int ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
        const int sub[4][3] = {{a, ab, ca}, {b, bc, ab}, {c, ca, bc}, {ab, bc, ca}};
        for (auto& s3 : sub) {
            const auto &p = v[s3[0]], &q = v[s3[1]], &r = v[s3[2]];
            const double ux = q[0] - p[0], uy = q[1] - p[1], uz = q[2] - p[2];
            const double wx = r[0] - p[0], wy = r[1] - p[1], wz = r[2] - p[2];
            std::array<double, 3> nn{uy * wz - uz * wy, uz * wx - ux * wz, ux * wy - uy * wx};
            const double l = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            for (auto& c2 : nn) c2 /= l;
            bool dup = false;
            for (auto& o : normals)
                if (std::fabs(o[0] + nn[0]) + std::fabs(o[1] + nn[1]) + std::fabs(o[2] + nn[2]) < 1e-9) dup = true;
            if (!dup) normals.push_back(nn);
        }
    }
    REQUIRE(normals.size() == KDOP_AXES_N, "k-DOP axis construction gave %d axes", (int)normals.size());
    float ax[KDOP_AXES_N][3];
    for (int i = 0; i < KDOP_AXES_N; ++i)
        for (int j = 0; j < 3; ++j) ax[i][j] = (float)normals[i][j];
    HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(c_kdop_axes), ax, sizeof(ax), 0, hipMemcpyHostToDevice));
    kdop_axes_ready = ctx->device;
    return 0;
}

using WaveFn = void (*)(LaHullTask*, const int*, int, const int*, DPool*, int*);

// Launch-shape knobs of the hull tiers (they choose kernels and capacities,
// never results: every tier computes the same canonical hull). Process-wide;
// initial values from the environment, changed at run time through
// gpu_set_tuning (the autotuner).
//   0 wide_min: wave-tier sets above this many points run four waves per hull
//     (CUACD_HULL_WIDE_MIN, 0 disables; default 256)
//   1 vcap8: first-pass vertex capacity of the four-wave tiers, in eighths of
//     their point capacity (CUACD_HULL_VCAP8, 1..8; default 3)
//   2 half_cap: reduced-capacity first passes on/off (CUACD_HULL_HALFCAP; default 1)
//   3 kdop_min: sets above this many points are k-DOP prefiltered
//     (CUACD_KDOP_MIN, >= 4; default 512: re-tuned on the 64-mesh search
//     step with the current tiers, 1.71-1.73 s against 1.76-1.80 s at 256)
//   4 chunk_above: prefiltered sets above this many points are hulled in
//     chunks of half that size first (CUACD_HULL_CHUNK_ABOVE, 1024..16384;
//     default 16384, the most one workgroup hulls)
//   5 multi: multi-apex rounds in the eight-wave LDS tiers (1025-4096 points;
//     CUACD_HULL_MULTI, 0/1; default 1)
//   6 mlp_fused: the proposal network's inference runs its point layers and
//     pooling as one kernel per candidate instead of layer-by-layer GEMMs
//     (CUACD_MLP_FUSED, 0/1; default 1; bit-identical scores)
//   7 ldsv8_vcap: first-pass vertex capacity of the eight-wave LDS tier
//     (1025-2048 points; CUACD_HULL_LDSV8_VCAP, 64..1024, default 1024).
//     At 448 or less with multi-apex rounds the workspace fits two
//     workgroups per CU; hulls with more corners are redone in the HBM tier
//     (the same canonical hull). On SDF surface parts a quarter of the sets
//     have more than 448 corners, and their redo costs more than the second
//     workgroup gains, hence the default
//   8 ldsb8_vcap: vertex capacity of an LDS first pass for the 2049-4096-point
//     sets (CUACD_HULL_LDSB8_VCAP, 0 = none: straight to the HBM tier; else
//     64..1024; default 1024); hulls with more corners are redone in the HBM
//     tier
//   9 cut_lds_verts: plane cuts of meshes with up to this many vertices keep
//     their per-vertex state in LDS and write no per-triangle scratch
//     (CUACD_CUT_LDS_VERTS, 0 = every cut on the heap-scratch path, else
//     256..110592; default 98304); bit-identical halves either way
//   10 lpt: each hull tier's sets are launched largest first (CUACD_HULL_LPT,
//     0/1; default 0): the longest insertion chains start on the first
//     workgroups and the short ones fill in behind them. On the 64-mesh
//     search step it measured slower (2.40 s against 2.31 s): the tiers
//     already overlap on side streams, and task order keeps a mesh's sets
//     together
//   11 b4_split: percent of the 1025-2048-point sets (the smallest) hulled by
//     the HBM-workspace kernel on the context's stream beside the LDS tier's
//     stream (CUACD_HULL_B4_SPLIT, 0..90; default 0): that kernel needs 5 KB
//     of LDS, so its workgroups fit on the CUs the LDS tier fills, whose
//     VALUs are mostly idle. On the 64-mesh search step 25 and 40 measured
//     slower (2.43 s and 2.50 s against 2.27 s)
//   12 order: the four-wave and wave tiers of a hull call wait for the first
//     passes of the eight-wave LDS tiers (CUACD_HULL_ORDER, 0/1; default 1).
//     An eight-wave LDS workgroup needs a whole CU's LDS, so it only starts on
//     a CU that no smaller workgroup holds; launched beside the smaller tiers
//     it waits for such CUs, and its launch stretches far beyond its work (on
//     the 64-mesh search step the 1025-2048-point tier was in flight 583 ms
//     for 178 ms of work per 256 workgroup slots; ordered, 322 ms)
//   13 w8: first passes of the wave tiers of 513-1024 points (bit 0) and
//     257-512 points (bit 1) on eight waves with multi-apex rounds instead of
//     four waves with single-apex steps; bits 2 (513-1024 points) and 3
//     (257-512 points): the same for their full-capacity retries
//     (CUACD_HULL_W8, 0..15; default 5). On the 64-mesh search step a
//     513-1024-point hull takes 0.38 ms on eight waves against 0.69 ms on
//     four (two workgroups per CU against three), and the step went from
//     2.05 s to 1.96 s; the 257-512-point tier packs better on four waves
//   14 multi_hbm: multi-apex rounds in the HBM-workspace tiers (their
//     per-face marks in LDS; CUACD_HULL_MULTI_HBM, 0/1;
//     default 1). These tiers hold the hulls with the most corners, so the
//     rounds pay most there: on the 256-mesh scorer workload the 4097-8192-
//     point tier went from 260 ms to 67 ms per step (14.8k -> 27.0k cuts/s),
//     and the 64-mesh search step from 1.98 s to 1.83 s, although the
//     kernels spill (224 and 608 bytes per lane at 4096 and 8192 points)
constexpr int LA_TUNE_N = 15;
int g_tune[LA_TUNE_N];
bool g_tune_init = false;

int tune_clamp(int key, int v) {
    switch (key) {
        case 0: return std::max(0, v);
        case 1: return std::min(8, std::max(1, v));
        case 2: return v != 0;
        case 3: return std::max(4, v);
        case 4: return std::min(HULL_NP_MAX, std::max(1024, v));
        case 7: return std::min(1024, std::max(64, v / 32 * 32));
        case 8: return v <= 0 ? 0 : std::min(1024, std::max(64, v / 32 * 32));
        case 9: return v <= 0 ? 0 : std::min(PC_LDS_VERTS_MAX, std::max(256, v));
        case 11: return std::min(90, std::max(0, v));
        case 13: return std::min(15, std::max(0, v));
        default: return v != 0;
    }
}

void tune_init() {
    if (g_tune_init) return;
    const char* env[LA_TUNE_N] = {"CUACD_HULL_WIDE_MIN",   "CUACD_HULL_VCAP8", "CUACD_HULL_HALFCAP",
                                  "CUACD_KDOP_MIN",        "CUACD_HULL_CHUNK_ABOVE", "CUACD_HULL_MULTI",
                                  "CUACD_MLP_FUSED",       "CUACD_HULL_LDSV8_VCAP",  "CUACD_HULL_LDSB8_VCAP",
                                  "CUACD_CUT_LDS_VERTS",   "CUACD_HULL_LPT",         "CUACD_HULL_B4_SPLIT",
                                  "CUACD_HULL_ORDER",      "CUACD_HULL_W8",          "CUACD_HULL_MULTI_HBM"};
    const int def[LA_TUNE_N] = {256, 3, 1, 512, HULL_NP_MAX, 1, 1, 1024, 1024, 98304, 0, 0, 1, 5, 1};
    for (int k = 0; k < LA_TUNE_N; ++k) {
        const char* e = getenv(env[k]);
        g_tune[k] = tune_clamp(k, e ? atoi(e) : def[k]);
    }
    g_tune_init = true;
}

int hull_wide_min() { tune_init(); return g_tune[0]; }
int hull_vcap_eighths() { tune_init(); return g_tune[1]; }
int hull_half_cap() { tune_init(); return g_tune[2]; }
int hull_multi() { tune_init(); return g_tune[5]; }
int hull_multi_hbm() { tune_init(); return g_tune[14]; }
int hull_ldsv8_vcap() { tune_init(); return g_tune[7]; }
int hull_ldsb8_vcap() { tune_init(); return g_tune[8]; }

// One hull call: its size tiers are independent, so they run on separate
// streams - the latency-bound large-set tiers (a few hulls of thousands of
// points, each a long sequential insertion chain) beside the
// throughput-bound small-set ones, instead of after them. Buffers are
// allocated on the context's stream, which each side stream waits for before
// its launches (fork), and freed there after the side streams have joined
// back. Retry passes read their task counts on the device, so no launch waits
// for the host.
struct HullCall {
    gpu_ctx_t ctx;
    std::vector<void*> tmp;
    bool used[3] = {false, false, false};
    bool joined = false;
    // a caller's arena instead of stream-ordered allocations (a captured
    // launch sequence must not allocate): bump allocation, 256-byte aligned
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    explicit HullCall(gpu_ctx_t c) : ctx(c) {}
    ~HullCall() {
        if (!joined) (void)join();
    }
    // stream k (-1: the context's stream); a side stream is first made to
    // wait for everything issued on the context's stream so far
    int fork(int k, hipStream_t* s) {
        if (k < 0) {
            *s = ctx->stream;
            return 0;
        }
        HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
        HIP_TRY(hipStreamWaitEvent(ctx->aux[k], ctx->ev_fork, 0));
        used[k] = true;
        *s = ctx->aux[k];
        return 0;
    }
    template <class T> T* alloc(size_t n) {
        if (arena) {
            const size_t b = (n * sizeof(T) + 255) / 256 * 256;
            if (arena_used + b > arena_cap) return nullptr;
            T* p = (T*)(arena + arena_used);
            arena_used += b;
            return p;
        }
        T* p = dev_tmp<T>(ctx, n);
        if (p) tmp.push_back(p);
        return p;
    }
    int join() {
        joined = true;
        for (int k = 0; k < 3; ++k)
            if (used[k]) {
                HIP_TRY(hipEventRecord(ctx->ev_join[k], ctx->aux[k]));
                HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join[k], 0));
                used[k] = false;
            }
        for (void* p : tmp) HIP_TRY(hipFreeAsync(p, ctx->stream));
        tmp.clear();
        return 0;
    }
};

// grid of a retry pass (its count is only known on the device)
int retry_grid(gpu_ctx_t ctx, size_t cap) {
    return (int)std::min(cap, (size_t)std::max(1, ctx->n_cu));
}

// Zero n ints (retry counters) with a kernel: such clears may sit inside a
// captured graph, where memset nodes are avoided.
__global__ void la_zero_ints_kernel(int* p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

int zero_ints(gpu_ctx_t ctx, int* p, int n, hipStream_t s) {
    hipLaunchKernelGGL(la_zero_ints_kernel, dim3((unsigned)std::max(1, std::min(64, (n + 255) / 256))), dim3(256), 0,
                       s, p, n);
    HIP_TRY(hipGetLastError());
    return 0;
}

// One wave tier (sets of up to 64 * ppl points). d_ids holds nids task ids,
// or - with dcnt given - up to nids of them, *dcnt on the device (the grid is
// sized for nids and the kernels read the count).
int launch_wave_ids(HullCall& hc, int sk, LaHullTask* d_tasks, const int* d_ids, int nids, int ppl,
                    double bytes, const int* dcnt = nullptr) {
    gpu_ctx_t ctx = hc.ctx;
    if (nids <= 0) return 0;
    const int kid = ppl == 1 ? KT_HULL_W1 : ppl == 2 ? KT_HULL_W2 : ppl == 4 ? KT_HULL_W4 : ppl == 8 ? KT_HULL_W8
                                                                                                     : KT_HULL_W16;
    const int wm = hull_wide_min();
    if (wm > 0 && ppl >= 8 && 64 * ppl > wm && hull_half_cap()) {
        // first pass with a reduced vertex capacity (several hulls per CU),
        // then the sets whose hulls did not fit, at full capacity
        using CapFn = void (*)(LaHullTask*, const int*, int, const int*, int, int*, int, DPool*, int*);
        using FullFn = void (*)(LaHullTask*, const int*, int, const int*, int, DPool*, int*);
        // the w8 knob (bit 0: 513-1024 points, bit 1: 257-512 points) runs
        // the first pass on eight waves with multi-apex rounds instead of
        // four waves with single-apex steps
        const bool w8 = (g_tune[13] >> (ppl == 16 ? 0 : 1)) & 1;
        const int nw = w8 ? LA_BLOCK_NW : LA_LDS4_NW;
        CapFn kc = w8 ? (ppl == 8 ? la_hull_ldsv<1, LA_BLOCK_NW> : la_hull_ldsv<2, LA_BLOCK_NW>)
                      : (ppl == 8 ? la_hull_ldsv<2, LA_LDS4_NW> : la_hull_ldsv<4, LA_LDS4_NW>);
        // bits 2 (513-1024 points) and 3 (257-512 points): the full-capacity
        // retries on eight waves with multi-apex rounds
        const bool w8f = (g_tune[13] >> (ppl == 16 ? 2 : 3)) & 1;
        const int nwf = w8f ? LA_BLOCK_NW : LA_LDS4_NW;
        FullFn kf = w8f ? (ppl == 8 ? la_hull_lds8<1> : la_hull_lds8<2>)
                        : (ppl == 8 ? la_hull_lds4<2> : la_hull_lds4<4>);
        // the canonical hulls of these sets have few corners (mostly under a
        // quarter of the points), so 3/8 of the points fits nearly all
        const int vcap = (64 * ppl * hull_vcap_eighths()) / 8;
        // four waves take single-apex steps: on these sets (257-1024 points)
        // the larger LDS footprint of multi-apex rounds costs more occupancy
        // than four candidates per round save
        const int multi = w8 ? hull_multi() : 0;
        const int multi_full = w8f ? hull_multi() : 0;
        const unsigned ldc = hull_ws_bytes<AS_LDS>(vcap, 64 * ppl, multi != 0);
        const unsigned ldf = hull_ws_bytes<AS_LDS>(64 * ppl, 64 * ppl, multi_full != 0);
        HIP_TRY(hipFuncSetAttribute((const void*)kc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldc));
        HIP_TRY(hipFuncSetAttribute((const void*)kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldf));
        int* d_retry = hc.alloc<int>((size_t)nids + 1);
        REQUIRE(d_retry, "lookahead: id list allocation failed");
        if (int rc = zero_ints(ctx, d_retry, 1, ctx->stream)) return rc;
        hipStream_t s;
        if (int rc = hc.fork(sk, &s)) return rc;
        hipEvent_t t0 = kt_begin(ctx, s);
        hipLaunchKernelGGL(kc, dim3((unsigned)(size_t)nids), dim3(64 * nw), ldc, s, d_tasks, d_ids,
                           (int)(size_t)nids, dcnt, vcap, d_retry, multi, ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
        kt_end(ctx, kid, t0, bytes, s);
        hipEvent_t t1 = kt_begin(ctx, s);
        hipLaunchKernelGGL(kf, dim3((unsigned)retry_grid(ctx, (size_t)nids)), dim3(64 * nwf), ldf, s, d_tasks,
                           d_retry + 1, (int)(size_t)nids, d_retry, multi_full, ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
        kt_end(ctx, ppl == 8 ? KT_HULL_R8 : KT_HULL_R16, t1, 0.0, s);
        return 0;
    }
    if (wm > 0 && ppl >= 4 && 64 * ppl > wm) {
        // same capacity (64 * ppl points and vertices), four waves
        using FullFn = void (*)(LaHullTask*, const int*, int, const int*, int, DPool*, int*);
        FullFn k = ppl == 4 ? la_hull_lds4<1> : ppl == 8 ? la_hull_lds4<2> : la_hull_lds4<4>;
        const int multi = 0;
        const unsigned lds = hull_ws_bytes<AS_LDS>(64 * ppl, 64 * ppl, multi != 0);
        if (lds > 65536u)
            HIP_TRY(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipStream_t s;
        if (int rc = hc.fork(sk, &s)) return rc;
        hipEvent_t t0 = kt_begin(ctx, s);
        hipLaunchKernelGGL(k, dim3((unsigned)(size_t)nids), dim3(64 * LA_LDS4_NW), lds, s, d_tasks, d_ids,
                           (int)(size_t)nids, dcnt, multi, ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
        kt_end(ctx, kid, t0, bytes, s);
        return 0;
    }
    WaveFn k = ppl == 1 ? la_hull_wave<1> : ppl == 2 ? la_hull_wave<2> : ppl == 4 ? la_hull_wave<4>
             : ppl == 8 ? la_hull_wave<8> : la_hull_wave<16>;
    const unsigned lds = hull_ws_bytes<AS_LDS>(64 * ppl, 64 * ppl);
    if (lds > 65536u)
        HIP_TRY(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipStream_t s;
    if (int rc = hc.fork(sk, &s)) return rc;
    hipEvent_t t0 = kt_begin(ctx, s);
    hipLaunchKernelGGL(k, dim3((unsigned)(size_t)nids), dim3(64), lds, s, d_tasks, d_ids, (int)(size_t)nids, dcnt,
                       ctx->d_pool, ctx->d_err);
    HIP_TRY(hipGetLastError());
    kt_end(ctx, kid, t0, bytes, s);
    return 0;
}

using BlockFn = void (*)(LaHullTask*, const int*, int, const int*, char*, unsigned long long, int, int, int,
                         DPool*, int*);

// Multi-apex rounds for an HBM-workspace kernel of vc vertices (knob
// multi_hbm; their LDS scratch, 3 bytes per face and 15.5 KB, fits up to
// 16384 vertices): the flag, and the dynamic LDS it needs, set on the kernel.
int block_multi(gpu_ctx_t ctx, BlockFn k, int vc, int* multi, unsigned* lds) {
    *multi = (hull_multi_hbm() && vc <= 16384) ? 1 : 0;
    *lds = *multi ? hull_ma_bytes(vc) : 0u;
    if (*lds > 0) HIP_TRY(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)*lds));
    return 0;
}

// The context's stream (and so every side stream forked from it later)
// waits for the recorded first passes of the eight-wave LDS tiers.
int wait_tiers(gpu_ctx_t ctx, hipEvent_t e8, hipEvent_t e4) {
    if (e8) HIP_TRY(hipStreamWaitEvent(ctx->stream, e8, 0));
    if (e4) HIP_TRY(hipStreamWaitEvent(ctx->stream, e4, 0));
    return 0;
}

// One block tier (sets above 1024 points, at most maxn). With dcnt given,
// d_ids holds up to n ids, *dcnt on the device, and maxn is the tier's bound.
int launch_block_ids(HullCall& hc, int sk, LaHullTask* d_tasks, const int* d_ids, int n, int maxn,
                     double bytes, const int* dcnt = nullptr, bool hbm_only = false,
                     hipEvent_t after_first = nullptr) {
    gpu_ctx_t ctx = hc.ctx;
    if (n <= 0) return 0;
    constexpr int T = 64 * LA_BLOCK_NW;
    const bool lds4 = !hbm_only && maxn <= 4 * T && hull_half_cap();
    const bool lds8 = !hbm_only && !lds4 && maxn <= 8 * T && hull_half_cap() && hull_ldsb8_vcap() > 0;
    if (lds4 || lds8) {
        // first pass with the workspace in LDS (a bounded number of hull
        // vertices), then the sets whose hulls did not fit, with the workspace
        // in HBM
        using CapFn = void (*)(LaHullTask*, const int*, int, const int*, int, int*, int, DPool*, int*);
        CapFn kc = lds4 ? la_hull_ldsv<4, LA_BLOCK_NW> : la_hull_ldsv<8, LA_BLOCK_NW>;
        BlockFn kr = lds4 ? la_hull_block<4> : la_hull_block<8>;
        const int multi = hull_multi();
        const int vc = lds4 ? hull_ldsv8_vcap() : hull_ldsb8_vcap();
        const unsigned ldc = hull_ws_bytes<AS_LDS>(vc, (lds4 ? 4 : 8) * T, multi != 0);
        HIP_TRY(hipFuncSetAttribute((const void*)kc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldc));
        int* d_retry = hc.alloc<int>((size_t)n + 1);
        REQUIRE(d_retry, "lookahead: id list allocation failed");
        if (int rc = zero_ints(ctx, d_retry, 1, ctx->stream)) return rc;
        const int np = (maxn + 63) / 64 * 64;
        const unsigned long long stride = (hull_ws_bytes<AS_GLOBAL>(np, np) + 255ull) & ~255ull;
        const int g = retry_grid(ctx, (size_t)n);
        char* gws = hc.alloc<char>(stride * (size_t)g + 256);
        REQUIRE(gws, "lookahead: hull workspace allocation failed");
        hipStream_t s;
        if (int rc = hc.fork(sk, &s)) return rc;
        hipEvent_t t0 = kt_begin(ctx, s);
        hipLaunchKernelGGL(kc, dim3((unsigned)n), dim3(T), ldc, s, d_tasks, d_ids, n, dcnt, vc, d_retry, multi,
                           ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
        kt_end(ctx, lds4 ? KT_HULL_B4 : KT_HULL_B8, t0, bytes, s);
        if (after_first) HIP_TRY(hipEventRecord(after_first, s));
        int mr;
        unsigned lr;
        if (int rc = block_multi(ctx, kr, np, &mr, &lr)) return rc;
        hipEvent_t t1 = kt_begin(ctx, s);
        hipLaunchKernelGGL(kr, dim3((unsigned)g), dim3(T), lr, s, d_tasks, d_retry + 1, n, (const int*)d_retry, gws,
                           stride, np, np, mr, ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
        kt_end(ctx, lds4 ? KT_HULL_RB4 : KT_HULL_RB8, t1, 0.0, s);
        return 0;
    }
    BlockFn k;
    int kid;
    if (maxn <= 4 * T) { k = la_hull_block<4>; kid = KT_HULL_B4; }
    else if (maxn <= 8 * T) { k = la_hull_block<8>; kid = KT_HULL_B8; }
    else if (maxn <= 16 * T) { k = la_hull_block<16>; kid = KT_HULL_B16; }
    else { k = la_hull_block<32>; kid = KT_HULL_B32; }
    const int np = (maxn + 63) / 64 * 64;
    const int vc = np < 32704 ? np : 32704;
    const unsigned long long stride = (hull_ws_bytes<AS_GLOBAL>(vc, np) + 255ull) & ~255ull;
    int mk;
    unsigned lk;
    if (int rc = block_multi(ctx, k, vc, &mk, &lk)) return rc;
    if (dcnt) {
        // device count: one workspace per CU, the grid strides over the list
        const int g = retry_grid(ctx, (size_t)n);
        char* gws = hc.alloc<char>(stride * (size_t)g + 256);
        REQUIRE(gws, "lookahead: hull workspace allocation failed");
        hipStream_t s;
        if (int rc = hc.fork(sk, &s)) return rc;
        hipEvent_t t0 = kt_begin(ctx, s);
        hipLaunchKernelGGL(k, dim3((unsigned)g), dim3(T), lk, s, d_tasks, d_ids, n, dcnt, gws, stride, vc, np, mk,
                           ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
        kt_end(ctx, kid, t0, bytes, s);
        return 0;
    }
    // bounded workspace: launch in slices of at most 2048 tasks
    const int slice = 2048;
    const int g = std::min(slice, n);
    char* gws = hc.alloc<char>(stride * (size_t)g + 256);
    REQUIRE(gws, "lookahead: hull workspace allocation failed");
    hipStream_t s;
    if (int rc = hc.fork(sk, &s)) return rc;
    for (int b = 0; b < n; b += slice) {
        const int cnt = std::min(slice, n - b);
        hipEvent_t t0 = kt_begin(ctx, s);
        hipLaunchKernelGGL(k, dim3(cnt), dim3(T), lk, s, d_tasks, d_ids + b, cnt, (const int*)nullptr, gws, stride,
                           vc, np, mk, ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
        kt_end(ctx, kid, t0, bytes * (double)cnt / (double)n, s);
    }
    return 0;
}

int ppl_class(int n) { return n <= 64 ? 1 : n <= 128 ? 2 : n <= 256 ? 4 : n <= 512 ? 8 : 16; }

// Launch streams by tier: the large-set global-workspace tiers (longest
// single-hull latency) first and on their own stream, then the 1025-2048
// and 513-1024 LDS tiers on one stream each, the small-set tiers last.
// `pre`: the sets of the call that skip the k-DOP prefilter, launched while
// it runs on the context's stream.
int launch_by_size(HullCall& hc, LaHullTask* d_tasks, const std::vector<std::pair<int, int>>& idn_in, bool pre) {
    // largest sets first (a stable order: ties keep the task order)
    tune_init();
    std::vector<std::pair<int, int>> idn_sorted;
    if (g_tune[10]) {
        idn_sorted = idn_in;
        std::stable_sort(idn_sorted.begin(), idn_sorted.end(),
                         [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.second > b.second; });
    }
    const std::vector<std::pair<int, int>>& idn = g_tune[10] ? idn_sorted : idn_in;
    std::vector<int> cls[5];
    double cbytes[5] = {0, 0, 0, 0, 0};
    std::vector<int> b4, b8, b16, b32, b4h;
    std::vector<std::pair<int, int>> b4n;
    int m4 = 0, m8 = 0, m16 = 0, m32 = 0, m4h = 0;
    double y4 = 0, y8 = 0, y16 = 0, y32 = 0, y4h = 0;
    for (auto& q : idn) {
        const int n = q.second;
        const double y = 12.0 * n + 8.0;  // points read, volume written
        if (n <= LA_WAVE_MAX) {
            const int p = ppl_class(n);
            const int c = p == 1 ? 0 : p == 2 ? 1 : p == 4 ? 2 : p == 8 ? 3 : 4;
            cls[c].push_back(q.first);
            cbytes[c] += y;
        } else if (n <= 2048) { b4n.push_back(q); }
        else if (n <= 4096) { b8.push_back(q.first); m8 = std::max(m8, n); y8 += y; }
        else if (n <= 8192) { b16.push_back(q.first); m16 = std::max(m16, n); y16 += y; }
        else { b32.push_back(q.first); m32 = std::max(m32, n); y32 += y; }
    }
    // the 1025-2048-point tier, less the smallest b4_split percent of its sets
    // (ties by task order), which go to the HBM-workspace kernel
    {
        const size_t nh = b4n.size() * (size_t)g_tune[11] / 100;
        std::vector<int> pos(b4n.size());
        for (size_t j = 0; j < pos.size(); ++j) pos[j] = (int)j;
        std::stable_sort(pos.begin(), pos.end(), [&](int a, int b) { return b4n[a].second < b4n[b].second; });
        std::vector<char> to_hbm(b4n.size(), 0);
        for (size_t j = 0; j < nh; ++j) to_hbm[pos[j]] = 1;
        for (size_t j = 0; j < b4n.size(); ++j) {
            const int id = b4n[j].first, n = b4n[j].second;
            const double y = 12.0 * n + 8.0;
            if (to_hbm[j]) { b4h.push_back(id); m4h = std::max(m4h, n); y4h += y; }
            else { b4.push_back(id); m4 = std::max(m4, n); y4 += y; }
        }
    }
    // every tier's id list in one upload, before any launch of this call on
    // the context's stream (a pageable copy queued behind a running tier
    // would hold the host until that tier finishes)
    gpu_ctx_t ctx = hc.ctx;
    const std::vector<int>* lists[10] = {&b32, &b16, &b8, &b4, &cls[4], &cls[3], &cls[2], &cls[1], &cls[0], &b4h};
    size_t off[10], total = 0;
    for (int k = 0; k < 10; ++k) { off[k] = total; total += lists[k]->size(); }
    if (total == 0) return 0;
    std::vector<int> all;
    all.reserve(total);
    for (int k = 0; k < 10; ++k) all.insert(all.end(), lists[k]->begin(), lists[k]->end());
    int* d_all = hc.alloc<int>(total);
    REQUIRE(d_all, "lookahead: id list allocation failed");
    HIP_TRY(hipMemcpyAsync(d_all, all.data(), 4 * total, hipMemcpyHostToDevice, ctx->stream));
    auto L = [&](int k) { return d_all + off[k]; };
    auto N = [&](int k) { return (int)lists[k]->size(); };
    // with the order knob, the eight-wave LDS first passes are recorded and
    // the smaller tiers wait for them
    hipEvent_t e8 = g_tune[12] && N(2) > 0 ? ctx->ev_tier[0] : nullptr;
    hipEvent_t e4 = g_tune[12] && N(3) > 0 ? ctx->ev_tier[1] : nullptr;
    if (int rc = launch_block_ids(hc, 0, d_tasks, L(0), N(0), m32, y32)) return rc;
    if (int rc = launch_block_ids(hc, 0, d_tasks, L(1), N(1), m16, y16)) return rc;
    if (int rc = launch_block_ids(hc, 0, d_tasks, L(2), N(2), m8, y8, nullptr, false, e8)) return rc;
    if (int rc = launch_block_ids(hc, 1, d_tasks, L(3), N(3), m4, y4, nullptr, false, e4)) return rc;
    if (int rc = launch_block_ids(hc, -1, d_tasks, L(9), N(9), m4h, y4h, nullptr, true)) return rc;
    if (int rc = wait_tiers(ctx, e8, e4)) return rc;
    const int ppls[5] = {1, 2, 4, 8, 16};
    if (int rc = launch_wave_ids(hc, 2, d_tasks, L(4), N(4), ppls[4], cbytes[4])) return rc;
    for (int c = 3; c >= 0; --c)
        if (int rc = launch_wave_ids(hc, pre ? 2 : -1, d_tasks, L(8 - c), N(8 - c), ppls[c], cbytes[c])) return rc;
    return 0;
}

}  // namespace

namespace cuacd {

// Hulls of n tasks. Host copy h_tasks gives the point counts; sets above the
// wave tier are prefiltered (k-DOP) first, which needs one readback.
static int kdop_min_points() {
    tune_init();
    return g_tune[3];
}

// Oversized sets (more than chunk_above points after the prefilter, `over`
// as (task, points)): hulls of chunks of chunk_above / 2 points, then the
// union of their corners replaces the set (la_union_kernel), repeated while a
// union is still too large; the shrunken sets join `fin`.
static int hull_oversized(gpu_ctx_t ctx, LaHullTask* d_tasks, std::vector<std::pair<int, int>> over,
                          std::vector<std::pair<int, int>>& fin) {
    tune_init();
    const int limit = g_tune[4], chunk = limit / 2;
    while (!over.empty()) {
        // the sets' current point arrays and frames (written by the prefilter)
        std::vector<LaHullTask> par(over.size());
        for (size_t i = 0; i < over.size(); ++i)
            HIP_TRY(hipMemcpyAsync(&par[i], d_tasks + over[i].first, sizeof(LaHullTask), hipMemcpyDeviceToHost,
                                   ctx->stream));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        std::vector<LaHullTask> sub;
        std::vector<int> sub_of;
        for (size_t i = 0; i < over.size(); ++i) {
            const LaHullTask& p = par[i];
            const float* P = p.filt ? p.filt : p.pts;
            const int n = over[i].second;
            for (int o = 0; o < n; o += chunk) {
                LaHullTask c{};
                c.pts = P + 3ll * o;
                c.n = std::min(chunk, n - o);
                c.emit = 1;
                c.prune_gap = 1e30;
                c.frame = p.frame;
                c.use_frame = 1;
                sub.push_back(c);
                sub_of.push_back((int)i);
            }
        }
        ctx->chunk_subtasks += (long long)sub.size();
        int* d_ids = dev_tmp<int>(ctx, over.size());
        int* d_off = dev_tmp<int>(ctx, over.size() + 1);
        LaHullTask* d_sub = dev_tmp<LaHullTask>(ctx, sub.size());
        LaUnionSeg* d_segs = dev_tmp<LaUnionSeg>(ctx, sub.size());
        REQUIRE(d_ids && d_off && d_sub && d_segs, "lookahead: chunked hull buffers allocation failed");
        HIP_TRY(hipMemcpyAsync(d_sub, sub.data(), sizeof(LaHullTask) * sub.size(), hipMemcpyHostToDevice,
                               ctx->stream));
        {
            HullCall hc(ctx);
            std::vector<std::pair<int, int>> idn;
            for (size_t k = 0; k < sub.size(); ++k) idn.push_back({(int)k, sub[k].n});
            if (int rc = launch_by_size(hc, d_sub, idn, false)) return rc;
            if (int rc = hc.join()) return rc;
        }
        HIP_TRY(hipMemcpyAsync(sub.data(), d_sub, sizeof(LaHullTask) * sub.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        std::vector<LaUnionSeg> segs;
        std::vector<int> ids, off(1, 0), tot(over.size(), 0);
        for (size_t i = 0, k = 0; i < over.size(); ++i) {
            ids.push_back(over[i].first);
            for (; k < sub.size() && sub_of[k] == (int)i; ++k) {
                REQUIRE(sub[k].hv || sub[k].err == 0, "lookahead: chunk hull failed (error %d)", sub[k].err);
                if (sub[k].hv) {
                    segs.push_back({sub[k].hv, sub[k].hnv, 1});
                    tot[i] += sub[k].hnv;
                } else {
                    segs.push_back({sub[k].pts, sub[k].n, 0});
                    tot[i] += sub[k].n;
                }
            }
            off.push_back((int)segs.size());
        }
        HIP_TRY(hipMemcpyAsync(d_ids, ids.data(), 4 * ids.size(), hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(d_off, off.data(), 4 * off.size(), hipMemcpyHostToDevice, ctx->stream));
        if (!segs.empty())
            HIP_TRY(hipMemcpyAsync(d_segs, segs.data(), sizeof(LaUnionSeg) * segs.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
        hipLaunchKernelGGL(la_union_kernel, dim3((unsigned)over.size()), dim3(256), 0, ctx->stream, d_tasks, d_ids,
                           d_off, d_segs, (int)over.size(), ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipFreeAsync(d_ids, ctx->stream));
        HIP_TRY(hipFreeAsync(d_off, ctx->stream));
        HIP_TRY(hipFreeAsync(d_sub, ctx->stream));
        HIP_TRY(hipFreeAsync(d_segs, ctx->stream));
        std::vector<std::pair<int, int>> next;
        for (size_t i = 0; i < over.size(); ++i) {
            if (tot[i] <= limit) {
                fin.push_back({over[i].first, tot[i]});
            } else if (tot[i] < over[i].second) {
                next.push_back({over[i].first, tot[i]});  // shrinking: chunk again
            } else {
                // no chunk lost a point: every point is a corner
                REQUIRE(tot[i] <= HULL_NP_MAX, "lookahead: hull set of %d points with %d chunk corners exceeds %d",
                        over[i].second, tot[i], HULL_NP_MAX);
                fin.push_back({over[i].first, tot[i]});
            }
        }
        over.swap(next);
    }
    return 0;
}

int la_launch_hull(gpu_ctx_t ctx, LaHullTask* d_tasks, const LaHullTask* h_tasks, int n) {
    if (n <= 0) return 0;
    if (int rc = ensure_kdop_axes(ctx)) return rc;
    std::vector<std::pair<int, int>> small;
    std::vector<int> big;
    const int kmin = std::min(kdop_min_points(), LA_WAVE_MAX);
    for (int i = 0; i < n; ++i) {
        if (h_tasks[i].n <= kmin) small.push_back({i, h_tasks[i].n});
        else big.push_back(i);
    }
    HullCall hc(ctx);
    // the sets that skip the prefilter run beside it
    if (int rc = launch_by_size(hc, d_tasks, small, true)) return rc;
    if (!big.empty()) {
        REQUIRE(big.size() < (1u << 31), "too many hull tasks");
        int* d_ids = hc.alloc<int>(big.size());
        int2* d_res = hc.alloc<int2>(big.size());
        KdopRec* d_rec = hc.alloc<KdopRec>(big.size());
        REQUIRE(d_ids && d_res && d_rec, "lookahead: k-DOP buffers allocation failed");
        HIP_TRY(hipMemcpyAsync(d_ids, big.data(), 4 * big.size(), hipMemcpyHostToDevice, ctx->stream));
        double kbytes = 0;
        for (int i : big) kbytes += 12.0 * h_tasks[i].n;  // points read (the filtered copy is extra)
        hipEvent_t t0 = kt_begin(ctx);
        if (int rc = la_launch_kdop(ctx, d_tasks, d_ids, (int)big.size(), nullptr, d_res, d_rec, ctx->stream))
            return rc;
        kt_end(ctx, KT_KDOP, t0, kbytes);
        std::vector<int2> res(big.size());
        HIP_TRY(hipMemcpyAsync(res.data(), d_res, sizeof(int2) * big.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx_stream_sync(ctx, ctx->stream));
        std::vector<std::pair<int, int>> fin, over;
        for (size_t i = 0; i < big.size(); ++i) {
            if (res[i].y) continue;          // pruned: volume is the rough bound
            if (res[i].x < 0) continue;      // allocation failure, flagged in the task
            if (res[i].x > g_tune[4]) over.push_back({big[i], res[i].x});
            else fin.push_back({big[i], res[i].x});
        }
        if (!over.empty())
            if (int rc = hull_oversized(ctx, d_tasks, over, fin)) return rc;
        if (int rc = launch_by_size(hc, d_tasks, fin, false)) return rc;
    }
    return hc.join();
}

// ---- device-sized hull call -------------------------------------------------
// Tier index of a set of m points (after the prefilter): 0-4 the wave tiers
// of 64, 128, 256, 512, 1024 points, 5-8 the block tiers up to 2048, 4096,
// 8192 and 16384 points - the classes of launch_by_size.
__device__ __forceinline__ int la_tier_of(int m) {
    if (m <= LA_WAVE_MAX) return m <= 64 ? 0 : m <= 128 ? 1 : m <= 256 ? 2 : m <= 512 ? 3 : 4;
    return m <= 2048 ? 5 : m <= 4096 ? 6 : m <= 8192 ? 7 : 8;
}

// Without a chunk table every oversized set is left unhulled.
__global__ void la_hdev_mark_over_kernel(int* counters) {
    if (threadIdx.x == 0) counters[LA_HDEV_FAIL] += counters[LA_HDEV_OVER];
}

// The sets above kmin points go to the prefilter's list (any order: every
// set is hulled on its own).
__global__ void la_hdev_prelist_kernel(const LaHullTask* tasks, int n, int kmin, int* counters, int* big) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (tasks[i].n > kmin) big[atomicAdd(&counters[LA_HDEV_BIG], 1)] = i;
}

// Every set with points into its tier's list: the small ones by their size,
// the prefiltered ones by their filtered size (pruned sets and failed filters
// drop out as in la_launch_hull); a set above `limit` is counted as oversized.
__global__ void la_hdev_bin_kernel(const LaHullTask* tasks, int n, int kmin, int limit, const int* big,
                                   const int2* res, int* counters, int* lists, int* over) {
    const int stride = gridDim.x * blockDim.x, g = blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = g; i < n; i += stride) {
        const int m = tasks[i].n;
        if (m <= 0 || m > kmin) continue;
        const int t = la_tier_of(m);
        lists[(size_t)t * n + atomicAdd(&counters[1 + t], 1)] = i;
    }
    const int nb = counters[LA_HDEV_BIG];
    for (int j = g; j < nb; j += stride) {
        const int2 r = res[j];
        if (r.y || r.x < 0) continue;
        if (r.x > limit) {
            over[atomicAdd(&counters[LA_HDEV_OVER], 1)] = big[j];
            continue;
        }
        const int t = la_tier_of(r.x);
        lists[(size_t)t * n + atomicAdd(&counters[1 + t], 1)] = big[j];
    }
}

// Oversized sets: chunk tasks of chunk points each (contiguous in the table).
__global__ void la_hdev_chunk_kernel(const LaHullTask* tasks, const int* over, int* counters, int chunk,
                                     LaHullTask* sub, int cap, int2* range) {
    const int no = counters[LA_HDEV_OVER];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < no; j += gridDim.x * blockDim.x) {
        const LaHullTask& p = tasks[over[j]];
        const float* P = p.filt ? p.filt : p.pts;
        const int m = p.filt ? p.nf : p.n;
        const int k = (m + chunk - 1) / chunk;
        const int f = atomicAdd(&counters[LA_HDEV_SUBN], k);
        if (f + k > cap) {
            atomicAdd(&counters[LA_HDEV_FAIL], 1);
            range[j] = make_int2(-1, 0);
            continue;
        }
        for (int q = 0; q < k; ++q) {
            LaHullTask c{};
            c.pts = P + 3ll * q * chunk;
            c.n = min(chunk, m - q * chunk);
            c.emit = 1;
            c.prune_gap = 1e30;
            c.frame = p.frame;
            c.use_frame = 1;
            sub[f + q] = c;
        }
        range[j] = make_int2(f, k);
    }
}

// Every chunk task empty (n = 0) before a replay fills the table.
__global__ void la_hdev_clear_kernel(LaHullTask* sub, int cap) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < cap; i += gridDim.x * blockDim.x) sub[i].n = 0;
}

// One workgroup per oversized set: the union of its chunks' corners (a chunk
// without a hull passes its points through) replaces the set's filtered
// points, as la_union_kernel does; the set then goes to its tier's list, or
// counts as unhulled when it is still above the limit.
__global__ __launch_bounds__(256) void la_hdev_union_kernel(LaHullTask* tasks, const int* over, int* counters,
                                                            const int2* range, const LaHullTask* sub, int limit,
                                                            int n, int* lists, DPool* pool, int* err) {
    __shared__ unsigned long long sp;
    __shared__ int sfail;
    const int no = counters[LA_HDEV_OVER];
    for (int j = blockIdx.x; j < no; j += gridDim.x) {
        __syncthreads();
        const int2 rg = range[j];
        if (rg.x < 0) continue;
        LaHullTask& t = tasks[over[j]];
        int total = 0, bad = 0;
        for (int q = rg.x; q < rg.x + rg.y; ++q) {
            total += sub[q].hv ? sub[q].hnv : sub[q].n;
            bad |= !sub[q].hv && sub[q].err != 0;
        }
        if (threadIdx.x == 0) {
            sfail = bad || total > limit;
            sp = sfail ? 0ull : (unsigned long long)heap_alloc(pool, 1, 12ull * total + 16, err);
        }
        __syncthreads();
        float* dst = (float*)sp;
        if (dst) {
            int o = 0;
            for (int q = rg.x; q < rg.x + rg.y; ++q) {
                const float* v = sub[q].hv ? sub[q].hv : sub[q].pts;
                const int m = sub[q].hv ? sub[q].hnv : sub[q].n;
                for (int i = threadIdx.x; i < 3 * m; i += 256) dst[3 * o + i] = v[i];
                o += m;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int q = rg.x; q < rg.x + rg.y; ++q)
                if (sub[q].hv) heap_free(pool, sub[q].hv, err);
            if (sfail) {
                atomicAdd(&counters[LA_HDEV_FAIL], 1);
            } else if (!dst) {
                t.err = KERR_KDOP_SCRATCH_OOM;
                err_or(err, KERR_KDOP_SCRATCH_OOM);
            } else {
                if (t.filt) heap_free(pool, t.filt, err);
                t.filt = dst;
                t.nf = total;
                const int tr = la_tier_of(total);
                lists[(size_t)tr * n + atomicAdd(&counters[1 + tr], 1)] = over[j];
            }
        }
    }
}

size_t la_hull_dev_work_bytes(int n) {
    const size_t nn = (size_t)std::max(n, 1);
    return 256 + (4 * nn + 255) / 256 * 256 + (8 * nn + 255) / 256 * 256 +
           (4 * nn * LA_HDEV_TIERS + 255) / 256 * 256 + (4 * nn + 255) / 256 * 256 + (8 * nn + 255) / 256 * 256 +
           (sizeof(KdopRec) * nn + 255) / 256 * 256;
}

size_t la_hull_dev_arena_bytes(gpu_ctx_t ctx, int n) {
    // retry lists of the four tiers with a first pass, and one global
    // workspace per CU for each block tier at its largest set size
    const size_t nn = (size_t)std::max(n, 1), g = (size_t)retry_grid(ctx, nn);
    size_t b = 4 * ((4 * (nn + 1) + 255) / 256 * 256);
    for (int np = 2048; np <= HULL_NP_MAX; np *= 2) {
        const int vc = np < 32704 ? np : 32704;
        b += ((hull_ws_bytes<AS_GLOBAL>(vc, np) + 255ull) & ~255ull) * g + 512;
    }
    return b;
}

int la_launch_hull_dev(gpu_ctx_t ctx, LaHullTask* d_tasks, int n, char* work, char* arena, size_t arena_bytes,
                       const LaHullDevChunks* chunks, bool prefilter) {
    if (n <= 0) return 0;
    if (int rc = ensure_kdop_axes(ctx)) return rc;
    tune_init();
    const size_t nn = (size_t)n;
    int* counters = (int*)work;
    int* big = (int*)(work + 256);
    int2* res = (int2*)(work + 256 + (4 * nn + 255) / 256 * 256);
    int* lists = (int*)((char*)res + (8 * nn + 255) / 256 * 256);
    int* over = (int*)((char*)lists + (4 * nn * LA_HDEV_TIERS + 255) / 256 * 256);
    int2* range = (int2*)((char*)over + (4 * nn + 255) / 256 * 256);
    KdopRec* krec = (KdopRec*)((char*)range + (8 * nn + 255) / 256 * 256);
    // without the prefilter every set is binned by its own size (the chunk
    // hulls of oversized sets, as hull_oversized hulls them)
    const int kmin = prefilter ? std::min(kdop_min_points(), LA_WAVE_MAX) : 0x7fffffff;
    const unsigned gb = (unsigned)std::max(1, std::min(1024, (n + 255) / 256));
    if (int rc = zero_ints(ctx, counters, LA_HDEV_NCOUNT, ctx->stream)) return rc;
    if (prefilter) {
        hipLaunchKernelGGL(la_hdev_prelist_kernel, dim3(gb), dim3(256), 0, ctx->stream, (const LaHullTask*)d_tasks,
                           n, kmin, counters, big);
        HIP_TRY(hipGetLastError());
        if (int rc = la_launch_kdop(ctx, d_tasks, (const int*)big, n, (const int*)(counters + LA_HDEV_BIG), res, krec,
                                    ctx->stream))
            return rc;
    }
    hipLaunchKernelGGL(la_hdev_bin_kernel, dim3(gb), dim3(256), 0, ctx->stream, (const LaHullTask*)d_tasks, n, kmin,
                       g_tune[4], (const int*)big, (const int2*)res, counters, lists, over);
    HIP_TRY(hipGetLastError());
    if (chunks && chunks->cap > 0) {
        // oversized sets: chunk hulls (a nested device-sized call), then unions
        const int cap = chunks->cap;
        const unsigned gc = (unsigned)std::max(1, std::min(1024, (cap + 255) / 256));
        hipLaunchKernelGGL(la_hdev_clear_kernel, dim3(gc), dim3(256), 0, ctx->stream, chunks->sub, cap);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(la_hdev_chunk_kernel, dim3(gb), dim3(256), 0, ctx->stream, (const LaHullTask*)d_tasks,
                           (const int*)over, counters, g_tune[4] / 2, chunks->sub, cap, range);
        HIP_TRY(hipGetLastError());
        if (int rc = la_launch_hull_dev(ctx, chunks->sub, cap, chunks->work, chunks->arena, chunks->arena_bytes,
                                        nullptr, false))
            return rc;
        hipLaunchKernelGGL(la_hdev_union_kernel, dim3((unsigned)std::max(1, std::min(n, 1024))), dim3(256), 0,
                           ctx->stream, d_tasks, (const int*)over, counters, (const int2*)range,
                           (const LaHullTask*)chunks->sub, g_tune[4], n, lists, ctx->d_pool, ctx->d_err);
        HIP_TRY(hipGetLastError());
    } else {
        hipLaunchKernelGGL(la_hdev_mark_over_kernel, dim3(1), dim3(64), 0, ctx->stream, counters);
        HIP_TRY(hipGetLastError());
    }
    HullCall hc(ctx);
    hc.arena = arena;
    hc.arena_cap = arena_bytes;
    auto L = [&](int t) { return (const int*)(lists + (size_t)t * nn); };
    auto C = [&](int t) { return (const int*)(counters + 1 + t); };
    // the launch_by_size order and streams: block tiers (largest first) on
    // side streams 0 and 1, the 1024-point wave tier on stream 2, the small
    // wave tiers on the context's stream
    hipEvent_t e8 = g_tune[12] ? ctx->ev_tier[0] : nullptr;
    hipEvent_t e4 = g_tune[12] ? ctx->ev_tier[1] : nullptr;
    if (int rc = launch_block_ids(hc, 0, d_tasks, L(8), n, HULL_NP_MAX, 0.0, C(8))) return rc;
    if (int rc = launch_block_ids(hc, 0, d_tasks, L(7), n, 8192, 0.0, C(7))) return rc;
    if (int rc = launch_block_ids(hc, 0, d_tasks, L(6), n, 4096, 0.0, C(6), false, e8)) return rc;
    if (int rc = launch_block_ids(hc, 1, d_tasks, L(5), n, 2048, 0.0, C(5), false, e4)) return rc;
    if (int rc = wait_tiers(ctx, e8, e4)) return rc;
    const int ppls[5] = {1, 2, 4, 8, 16};
    if (int rc = launch_wave_ids(hc, 2, d_tasks, L(4), n, ppls[4], 0.0, C(4))) return rc;
    for (int c = 3; c >= 0; --c)
        if (int rc = launch_wave_ids(hc, -1, d_tasks, L(c), n, ppls[c], 0.0, C(c))) return rc;
    return hc.join();
}

int la_launch_cut(gpu_ctx_t ctx, LaCutTask* d_tasks, int n, int max_nv) {
    if (n <= 0) return 0;
    int lds_nv;
    size_t lds;
    if (int rc = pc_lds_launch(ctx, (const void*)la_cut_kernel, max_nv, &lds_nv, &lds)) return rc;
    hipLaunchKernelGGL(la_cut_kernel, dim3(n), dim3(PC_T), lds, ctx->stream, d_tasks, n, ctx->d_pool, ctx->d_err,
                       lds_nv);
    HIP_TRY(hipGetLastError());
    return 0;
}

int la_launch_score(gpu_ctx_t ctx, LaScoreTask* d_tasks, int n, int max_nv) {
    if (n <= 0) return 0;
    // one flag byte per vertex: cut_lds_verts() bounds it like the cut's state
    const int lds_nv = std::min(std::max(max_nv, 0), cut_lds_verts());
    const size_t lds = lds_nv > 0 ? ((size_t)lds_nv + 15) / 16 * 16 : 0;
    if (lds > 0)
        HIP_TRY(hipFuncSetAttribute((const void*)la_score_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
    hipLaunchKernelGGL(la_score_kernel, dim3(n), dim3(PC_T), lds, ctx->stream, d_tasks, n, ctx->d_pool, ctx->d_err,
                       lds_nv);
    HIP_TRY(hipGetLastError());
    return 0;
}

int la_launch_components(gpu_ctx_t ctx, LaCompTask* d_tasks, int n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(la_comp_kernel, dim3(n), dim3(PC_T), 0, ctx->stream, d_tasks, n, ctx->d_pool, ctx->d_err);
    HIP_TRY(hipGetLastError());
    return 0;
}

int la_launch_edges(gpu_ctx_t ctx, LaEdgeTask* d_tasks, int n, float cos_pos, float cos_neg, int pos_mode,
                    int neg_mode) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(la_edges_kernel, dim3(n), dim3(PC_T), 0, ctx->stream, d_tasks, n, cos_pos, cos_neg, pos_mode,
                       neg_mode, ctx->d_pool, ctx->d_err);
    HIP_TRY(hipGetLastError());
    return 0;
}

int la_launch_bbox(gpu_ctx_t ctx, LaBoxTask* d_tasks, int n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(la_bbox_kernel, dim3(n), dim3(PC_T), 0, ctx->stream, d_tasks, n);
    HIP_TRY(hipGetLastError());
    return 0;
}

int la_launch_segcopy(gpu_ctx_t ctx, const LaSeg* d_segs, int n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(la_segcopy_kernel, dim3(n), dim3(256), 0, ctx->stream, d_segs, n);
    HIP_TRY(hipGetLastError());
    return 0;
}

int la_launch_gather(gpu_ctx_t ctx, LaGatherTask* d_tasks, int n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(la_gather_kernel, dim3(n), dim3(PC_T), 0, ctx->stream, d_tasks, n);
    HIP_TRY(hipGetLastError());
    return 0;
}

int mlp_fused_on() {
    tune_init();
    return g_tune[6];
}

int cut_lds_verts() {
    tune_init();
    return g_tune[9];
}

int pc_lds_launch(gpu_ctx_t ctx, const void* kernel, int max_nv, int* lds_nv, size_t* lds_bytes) {
    (void)ctx;
    const int cap = cut_lds_verts();
    *lds_nv = std::min(std::max(max_nv, 0), cap);
    *lds_bytes = *lds_nv > 0 ? (size_t)pc_lds_bytes(*lds_nv) : 0;
    if (*lds_bytes > 0)
        HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)*lds_bytes));
    return 0;
}

int la_launch_free(gpu_ctx_t ctx, void** d_ptrs, int n) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(la_free_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_ptrs, n, ctx->d_pool,
                       ctx->d_err);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace cuacd

extern "C" {

int gpu_set_tuning(gpu_ctx_t ctx, int key, int value) {
    REQUIRE(key >= 0 && key < LA_TUNE_N, "unknown tuning key %d", key);
    tune_init();
    g_tune[key] = tune_clamp(key, value);
    return 0;
}

int gpu_get_tuning(gpu_ctx_t ctx, int key, int* value) {
    REQUIRE(key >= 0 && key < LA_TUNE_N, "unknown tuning key %d", key);
    tune_init();
    *value = g_tune[key];
    return 0;
}

}  // extern "C"

extern "C" {

// Plain device copy of `bytes` (a multiple of 4) through la_segcopy_kernel in
// 1 MiB segments, `reps` times: the known-size copy that calibrates HBM
// traffic counters. *ms = average GPU milliseconds per copy.
int gpu_copy_probe(gpu_ctx_t ctx, size_t bytes, int reps, float* ms) {
    REQUIRE(bytes >= 4 && bytes % 4 == 0 && reps >= 1, "copy_probe: bad arguments");
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t seg = 1u << 20;
    const size_t n = (bytes + seg - 1) / seg;
    char* src = nullptr;
    char* dst = nullptr;
    LaSeg* d_segs = nullptr;
    HIP_TRY(hipMalloc((void**)&src, bytes));
    HIP_TRY(hipMalloc((void**)&dst, bytes));
    HIP_TRY(hipMalloc((void**)&d_segs, sizeof(LaSeg) * n));
    std::vector<LaSeg> h(n);
    for (size_t i = 0; i < n; ++i)
        h[i] = LaSeg{src + i * seg, dst + i * seg, (unsigned long long)std::min(seg, bytes - i * seg)};
    HIP_TRY(hipMemcpy(d_segs, h.data(), sizeof(LaSeg) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(src, 1, bytes));
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, ctx->stream));
    int rc = 0;
    for (int r = 0; r < reps && rc == 0; ++r) rc = la_launch_segcopy(ctx, d_segs, (int)n);
    HIP_TRY(hipEventRecord(b, ctx->stream));
    HIP_TRY(hipEventSynchronize(b));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, a, b));
    *ms = t / (float)reps;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(src);
    (void)hipFree(dst);
    (void)hipFree(d_segs);
    return rc;
}

}  // extern "C"
