// hausdorff.hpp - sampled directed Hausdorff distance, one workgroup per task.
//
// Reference semantics (cuda/hausdorff.cuh:271-1478): samples on the source
// surface, exact nearest distance to the target's triangles, maximum over the
// samples; the concavity metric is sqrt(max(d2(A->B), d2(B->A))).
//
// Two samplers:
//   * reference sampler: res = max(int(2000*area), 1000) samples spread by
//     area, per-triangle count int(res/area*a_t) with the reference's
//     fallback for tiny triangles, barycentric points from Wang hashes of
//     (t*65537 + k) (cuda/hausdorff.cuh:380-382, 432-466, 516-558);
//   * fixed-count stratified sampler: K samples, sample i at cumulative-area
//     position (i + r_i)/K of the surface, barycentric point from hashes of i.
//
// Nearest-triangle queries run on a BVH designed for wave64:
//   * triangles are counting-sorted by a 12-bit Morton cell of their centroid
//     (LDS histogram + scan), so no comparison sort is needed;
//   * the tree is implicit: level 0 groups 4 consecutive sorted triangles into
//     a leaf, every level above groups 8 nodes, so child/parent are index
//     arithmetic and there are no pointers;
//   * each lane answers one sample: a greedy descent to the nearest-looking
//     leaf gives an upper bound, then a stackless in-order walk of the
//     implicit tree prunes every box farther than the current best.
// Distances are fp32 (closest point on triangle by Voronoi regions).
#pragma once
#include "wave.hpp"
#include "hull.hpp"

namespace cuacd {

constexpr int HD_T = 256;                 // threads per task workgroup
constexpr int HD_CELLS = 4096;            // 16^3 Morton cells for the counting sort
constexpr int HD_RESOLUTION = 2000;       // reference sampler density (cuda/hausdorff.cuh:27)
constexpr int HD_MIN_SAMPLES = 1000;      // cuda/hausdorff.cuh:28
constexpr int HD_MAX_LEVELS = 12;

// Thomas Wang's 32-bit integer hash (the reference sampler's hash,
// cuda/hausdorff.cuh:37-44).
__host__ __device__ __forceinline__ uint32_t wang_hash(uint32_t s) {
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float v3dot(V3 a, V3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// Squared distance from p to triangle abc: closest point by Voronoi regions
// (Ericson, Real-Time Collision Detection, 5.1.5).
__device__ __forceinline__ float pt_tri_d2(V3 p, V3 a, V3 b, V3 c) {
    const V3 ab = v3sub(b, a), ac = v3sub(c, a), ap = v3sub(p, a);
    const float d1 = v3dot(ab, ap), d2 = v3dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return v3dot(ap, ap);
    const V3 bp = v3sub(p, b);
    const float d3 = v3dot(ab, bp), d4 = v3dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return v3dot(bp, bp);
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        const float v = d1 / (d1 - d3);
        const V3 q = {ap.x - v * ab.x, ap.y - v * ab.y, ap.z - v * ab.z};
        return v3dot(q, q);
    }
    const V3 cp = v3sub(p, c);
    const float d5 = v3dot(ab, cp), d6 = v3dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return v3dot(cp, cp);
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        const float w = d2 / (d2 - d6);
        const V3 q = {ap.x - w * ac.x, ap.y - w * ac.y, ap.z - w * ac.z};
        return v3dot(q, q);
    }
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        const float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        const V3 bc = v3sub(c, b);
        const V3 q = {bp.x - w * bc.x, bp.y - w * bc.y, bp.z - w * bc.z};
        return v3dot(q, q);
    }
    const float den = va + vb + vc;
    if (!(den > 0.0f)) {  // degenerate triangle: nearest vertex
        return fminf(v3dot(ap, ap), fminf(v3dot(bp, bp), v3dot(cp, cp)));
    }
    const float inv = 1.0f / den;
    const float v = vb * inv, w = vc * inv;
    const V3 q = {ap.x - ab.x * v - ac.x * w, ap.y - ab.y * v - ac.y * w, ap.z - ab.z * v - ac.z * w};
    return v3dot(q, q);
}

__device__ __forceinline__ float tri_area(V3 a, V3 b, V3 c) {
#pragma clang fp contract(off)
    const V3 ab = v3sub(b, a), ac = v3sub(c, a);
    const float nx = ab.y * ac.z - ab.z * ac.y;
    const float ny = ab.z * ac.x - ab.x * ac.z;
    const float nz = ab.x * ac.y - ab.y * ac.x;
    return 0.5f * sqrtf(nx * nx + ny * ny + nz * nz);
}

// Mesh view: float xyz vertices, int triangles.
struct MeshView {
    const float* v;
    const int* t;
    int nv, nt;
    __device__ __forceinline__ V3 vert(int i) const { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }
    __device__ __forceinline__ void tri(int k, V3& a, V3& b, V3& c) const {
        const int i0 = t[3 * k], i1 = t[3 * k + 1], i2 = t[3 * k + 2];
        a = vert(i0);
        b = vert(i1);
        c = vert(i2);
    }
};

// Implicit BVH over a target mesh (global memory).
struct HdBvh {
    f4v* tri;    // [nt_pad*3] sorted triangles: a, b, c (w unused)
    f4v* box;    // [2*nbox] lo, hi per node, levels concatenated (level 0 = leaves)
    int nt;
    int top;     // index of the top level (count 1)
    asptr<AS_LDS, int> off;  // [HD_MAX_LEVELS] level offsets (LDS: indexed per lane)
    asptr<AS_LDS, int> cnt;  // [HD_MAX_LEVELS] level node counts
};

__host__ __device__ inline int hd_bvh_levels(int nt, int* off, int* cnt, int* nbox) {
    int n = (nt + 3) / 4, L = 0, o = 0;
    if (n < 1) n = 1;
    for (;;) {
        if (off) { off[L] = o; cnt[L] = n; }
        o += n;
        if (n == 1) break;
        n = (n + 7) / 8;
        ++L;
    }
    if (nbox) *nbox = o;
    return L;
}

// Bytes of global workspace for one task: target BVH + source prefix.
__host__ __device__ inline unsigned long long hd_ws_bytes(int nt_target, int nt_source) {
    int nbox = 0;
    hd_bvh_levels(nt_target, nullptr, nullptr, &nbox);
    const unsigned long long ntp = (unsigned long long)((nt_target + 3) / 4 * 4);
    return ntp * 48ull + (unsigned long long)nbox * 32ull + ((unsigned long long)(nt_source + 1) * 4ull + 255ull) / 256ull * 256ull + 256ull;
}

struct HdLDS {
    int hist[HD_CELLS];
    float red[HD_T / 64][8];
    double dred[HD_T / 64];
    int ired[HD_T / 64];
    int scan[HD_T];
    int loff[HD_MAX_LEVELS];
    int lcnt[HD_MAX_LEVELS];
};

__device__ __forceinline__ uint32_t morton4(uint32_t v) {  // spread 4 bits
    v &= 15u;
    v = (v | (v << 4)) & 0x0C3u;
    v = (v | (v << 2)) & 0x249u;
    return v;
}

// Build the BVH of target mesh m into workspace (T = HD_T threads).
__device__ inline HdBvh hd_build(const MeshView& m, char* ws, HdLDS* s) {
    const int tid = threadIdx.x;
    HdBvh b;
    b.nt = m.nt;
    int nbox = 0;
    {
        int lo[HD_MAX_LEVELS], lc[HD_MAX_LEVELS];
        b.top = hd_bvh_levels(m.nt, lo, lc, &nbox);
        if (tid == 0)
            for (int L = 0; L <= b.top; ++L) { s->loff[L] = lo[L]; s->lcnt[L] = lc[L]; }
    }
    b.off = (asptr<AS_LDS, int>)s->loff;
    b.cnt = (asptr<AS_LDS, int>)s->lcnt;
    const int ntp = (m.nt + 3) / 4 * 4;
    b.tri = (f4v*)ws;
    b.box = (f4v*)(ws + (unsigned long long)ntp * 48ull);
    // bbox of the vertices
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = tid; i < m.nv; i += HD_T) {
        const V3 p = m.vert(i);
        mn[0] = fminf(mn[0], p.x); mx[0] = fmaxf(mx[0], p.x);
        mn[1] = fminf(mn[1], p.y); mx[1] = fmaxf(mx[1], p.y);
        mn[2] = fminf(mn[2], p.z); mx[2] = fmaxf(mx[2], p.z);
    }
    const int wv = tid >> 6;
    for (int j = 0; j < 3; ++j) {
        mn[j] = wave_min_f(mn[j]);
        mx[j] = wave_max_f(mx[j]);
    }
    if (lane_id() == 0)
        for (int j = 0; j < 3; ++j) { s->red[wv][j] = mn[j]; s->red[wv][3 + j] = mx[j]; }
    for (int c = tid; c < HD_CELLS; c += HD_T) s->hist[c] = 0;
    __syncthreads();
    for (int j = 0; j < 3; ++j) {
        mn[j] = s->red[0][j];
        mx[j] = s->red[0][3 + j];
        for (int q = 1; q < HD_T / 64; ++q) {
            mn[j] = fminf(mn[j], s->red[q][j]);
            mx[j] = fmaxf(mx[j], s->red[q][3 + j]);
        }
    }
    float sc[3];
    for (int j = 0; j < 3; ++j) sc[j] = (mx[j] > mn[j]) ? 16.0f / (mx[j] - mn[j]) : 0.0f;
    auto cell_of = [&](int k) {
        V3 a, bb, c;
        m.tri(k, a, bb, c);
        const float cx = (a.x + bb.x + c.x) * (1.0f / 3.0f), cy = (a.y + bb.y + c.y) * (1.0f / 3.0f),
                    cz = (a.z + bb.z + c.z) * (1.0f / 3.0f);
        const int qx = min(15, max(0, (int)((cx - mn[0]) * sc[0])));
        const int qy = min(15, max(0, (int)((cy - mn[1]) * sc[1])));
        const int qz = min(15, max(0, (int)((cz - mn[2]) * sc[2])));
        return (int)(morton4(qx) | (morton4(qy) << 1) | (morton4(qz) << 2));
    };
    for (int k = tid; k < m.nt; k += HD_T) atomicAdd(&s->hist[cell_of(k)], 1);
    __syncthreads();
    // exclusive scan of the 4096 counters: 16 per thread
    {
        constexpr int PER = HD_CELLS / HD_T;
        int loc[PER], sum = 0;
        for (int q = 0; q < PER; ++q) { loc[q] = s->hist[tid * PER + q]; sum += loc[q]; }
        int incl = wave_incl_scan(sum);
        if (lane_id() == 63) s->ired[wv] = incl;
        __syncthreads();
        int base = 0;
        for (int q = 0; q < wv; ++q) base += s->ired[q];
        int run = base + incl - sum;
        for (int q = 0; q < PER; ++q) { s->hist[tid * PER + q] = run; run += loc[q]; }
    }
    __syncthreads();
    for (int k = tid; k < m.nt; k += HD_T) {
        const int pos = atomicAdd(&s->hist[cell_of(k)], 1);
        V3 a, bb, c;
        m.tri(k, a, bb, c);
        b.tri[3 * pos + 0] = mk_f4(a.x, a.y, a.z, 0.0f);
        b.tri[3 * pos + 1] = mk_f4(bb.x, bb.y, bb.z, 0.0f);
        b.tri[3 * pos + 2] = mk_f4(c.x, c.y, c.z, 0.0f);
    }
    __syncthreads();  // slot 0 is written by whichever thread scattered it
    // pad the last leaf with copies of the first triangle (never nearer than it)
    for (int k = m.nt + tid; k < ntp; k += HD_T)
        for (int e = 0; e < 3; ++e) b.tri[3 * k + e] = b.tri[e];
    __syncthreads();
    // leaf boxes
    for (int j = tid; j < b.cnt[0]; j += HD_T) {
        f4v lo = mk_f4(1e30f, 1e30f, 1e30f, 0.0f), hi = mk_f4(-1e30f, -1e30f, -1e30f, 0.0f);
        for (int q = 4 * j; q < 4 * j + 4 && q < ntp; ++q)
            for (int e = 0; e < 3; ++e) {
                const f4v p = b.tri[3 * q + e];
                lo.x = fminf(lo.x, p.x); lo.y = fminf(lo.y, p.y); lo.z = fminf(lo.z, p.z);
                hi.x = fmaxf(hi.x, p.x); hi.y = fmaxf(hi.y, p.y); hi.z = fmaxf(hi.z, p.z);
            }
        b.box[2 * j] = lo;
        b.box[2 * j + 1] = hi;
    }
    __syncthreads();
    for (int L = 1; L <= b.top; ++L) {
        for (int j = tid; j < b.cnt[L]; j += HD_T) {
            f4v lo = mk_f4(1e30f, 1e30f, 1e30f, 0.0f), hi = mk_f4(-1e30f, -1e30f, -1e30f, 0.0f);
            const int c0 = 8 * j, c1 = min(8 * j + 8, b.cnt[L - 1]);
            for (int c = c0; c < c1; ++c) {
                const f4v l = b.box[2 * (b.off[L - 1] + c)], h = b.box[2 * (b.off[L - 1] + c) + 1];
                lo.x = fminf(lo.x, l.x); lo.y = fminf(lo.y, l.y); lo.z = fminf(lo.z, l.z);
                hi.x = fmaxf(hi.x, h.x); hi.y = fmaxf(hi.y, h.y); hi.z = fmaxf(hi.z, h.z);
            }
            b.box[2 * (b.off[L] + j)] = lo;
            b.box[2 * (b.off[L] + j) + 1] = hi;
        }
        __syncthreads();
    }
    return b;
}

__device__ __forceinline__ float box_d2(const HdBvh& b, int L, int i, V3 p) {
    const f4v lo = b.box[2 * (b.off[L] + i)], hi = b.box[2 * (b.off[L] + i) + 1];
    const float dx = fmaxf(fmaxf(lo.x - p.x, 0.0f), p.x - hi.x);
    const float dy = fmaxf(fmaxf(lo.y - p.y, 0.0f), p.y - hi.y);
    const float dz = fmaxf(fmaxf(lo.z - p.z, 0.0f), p.z - hi.z);
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

__device__ __forceinline__ float leaf_d2(const HdBvh& b, int j, V3 p, float best) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = 4 * j + q;
        const f4v a = b.tri[3 * k], bb = b.tri[3 * k + 1], c = b.tri[3 * k + 2];
        best = fminf(best, pt_tri_d2(p, {a.x, a.y, a.z}, {bb.x, bb.y, bb.z}, {c.x, c.y, c.z}));
    }
    return best;
}

// Exact squared distance from p to the nearest target triangle.
__device__ inline float hd_nearest_d2(const HdBvh& b, V3 p) {
    // greedy descent for an upper bound
    int L = b.top, i = 0;
    while (L > 0) {
        const int c0 = 8 * i, c1 = min(8 * i + 8, b.cnt[L - 1]);
        int bi = c0;
        float bd = 3.4e38f;
        for (int c = c0; c < c1; ++c) {
            const float d = box_d2(b, L - 1, c, p);
            if (d < bd) { bd = d; bi = c; }
        }
        --L;
        i = bi;
    }
    float best = leaf_d2(b, i, p, 3.4e38f);
    const int greedy_leaf = i;
    // stackless in-order walk of the implicit tree with pruning
    L = b.top;
    i = 0;
    for (;;) {
        const float d = box_d2(b, L, i, p);
        if (d < best) {
            if (L == 0) {
                if (i != greedy_leaf) best = leaf_d2(b, i, p, best);
            } else {
                --L;
                i *= 8;
                continue;
            }
        }
        // advance to the next sibling, climbing while the group is exhausted
        for (;;) {
            if (L == b.top) return best;
            if ((i & 7) != 7 && i + 1 < b.cnt[L]) {
                ++i;
                break;
            }
            ++L;
            i >>= 3;
        }
    }
}

// One directed Hausdorff task: samples on the source surface, nearest
// distances to the target surface.
struct HdTask {
    const float* sv;  // source vertices / triangles (samples are drawn here)
    const int* st;
    const float* tv;  // target vertices / triangles (distances are measured to it)
    const int* tt;
    int snv, snt, tnv, tnt;
    unsigned long long ws_off;  // byte offset of the task's workspace
    int mode;                   // 0 reference sampler, 1 stratified K samples
    int K;
    float* out_d2;              // max squared nearest distance
    int* out_ns;                // samples used (may be null)
    float* out_samples;         // optional sample dump (K*3 floats)
    // filled by the preparation kernel for the query kernel
    int ns = 0;                 // samples drawn
    float area = 0.0f;          // source area (fp32 as the sampler uses it)
};

}  // namespace cuacd
