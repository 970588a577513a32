// plane_cut.hpp - cut a closed triangle mesh by a plane into two closed meshes.
//
// Semantics follow the reference plane cut (cuda/plane_cut.cuh:167-1562):
//   * vertex sign with eps 1e-6; an edge crosses when its endpoints' signs
//     differ with on-plane counted as positive (269-318);
//   * a crossing edge with an on-plane endpoint reuses that vertex, otherwise
//     the intersection point p_a + t (p_b - p_a), t = d_a / (d_a - d_b) (548-566,
//     125-135);
//   * triangles split by lone vertex or on-plane vertex; triangles lying in
//     the plane go to the side their normal faces (582-652);
//   * the cut boundary is capped on both sides: boundary loops, orientation
//     by signed area, nesting by ray casting (odd depth = hole), holes bridged
//     into their enclosing loop from the hole's rightmost vertex, ear clipping
//     with collinear ears emitted (1008-1308);
//   * no crossing edge: the mesh goes whole to its side, or its triangles are
//     split by side when vertices lie on both (331-474);
//   * a cap that cannot be built returns the mesh unsplit (whole mesh on the
//     positive side), which callers treat as a failed cut (1339-1368).
//
// The implementation is this library's own, for one 256-thread workgroup:
// intersection vertices are deduplicated through an open-addressing hash
// table in HBM (no sort), per-triangle outputs are placed by block prefix
// scans (deterministic triangle order), boundary edges are the on-plane
// directed edges without a twin (hash set), loops come from pointer jumping,
// and one wave64 runs nesting, bridging and ear clipping with lane-parallel
// inner tests. Each side's FP64 volume is computed on the way out. Every
// choice is order-independent of thread timing: new vertices are numbered in
// (triangle, edge) order of first occurrence and boundary edges are listed in
// (positive triangle, edge) order, so a cut is reproducible bit for bit.
#pragma once
#include "dheap.hpp"
#include "hausdorff.hpp"

// Geometry below is evaluated exactly as written (no FMA contraction), so
// that a serial C evaluation of the same expressions gives the same bits.
#pragma clang fp contract(off)

namespace cuacd {

constexpr int PC_T = 256;
constexpr float PC_EPS = 1e-6f;
constexpr int PC_MAX_LOOPS = 256;
constexpr unsigned long long PC_EMPTY = ~0ull;
// most vertices of an LDS-resident cut (its dynamic LDS stays under 152 KiB)
constexpr int PC_LDS_VERTS_MAX = 110592;

struct PcMesh {
    float* v;   // heap chunk (verts, then tris) - free with heap_free(v)
    int* t;
    int nv, nt;
    double vol;
};

struct PcResult {
    PcMesh pos, neg;
    int status;  // 0 cut, 1 cap failure (mesh returned whole on pos), 2 one side empty, <0 error
    int has_box;  // 1: PcLDS::boxr holds both halves' vertex bounds (LDS-resident cut, status 0)
};

struct PcLDS {
    int ired[PC_T / 64];
    double dred[PC_T / 64];
    int cnt[8];
    int wtot[PC_T / 64][4];
    float box[PC_T / 64][12];
    float boxr[12];  // positive side min xyz, max xyz; negative side min xyz, max xyz
    void* ptr[4];
    int n_loops;
    int fail;
    int loop_start[PC_MAX_LOOPS + 1];
    int loop_parent[PC_MAX_LOOPS];
    int loop_hole[PC_MAX_LOOPS];
    float loop_area[PC_MAX_LOOPS];
};

__device__ __forceinline__ float pc_side(float pa, float pb, float pc, float pd, float x, float y, float z) {
    return fmaf(pa, x, fmaf(pb, y, fmaf(pc, z, pd)));
}

__device__ __forceinline__ uint32_t pc_hash(unsigned long long k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k;
}

// insert key; returns slot. *fresh = 1 if this call inserted it.
__device__ __forceinline__ int pc_ht_insert(unsigned long long* keys, uint32_t mask, unsigned long long key,
                                            int* fresh) {
    uint32_t h = pc_hash(key) & mask;
    for (;;) {
        const unsigned long long old = atomicCAS(&keys[h], PC_EMPTY, key);
        if (old == PC_EMPTY) { *fresh = 1; return (int)h; }
        if (old == key) { *fresh = 0; return (int)h; }
        h = (h + 1) & mask;
    }
}
__device__ __forceinline__ int pc_ht_find(const unsigned long long* keys, uint32_t mask, unsigned long long key) {
    uint32_t h = pc_hash(key) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long k = keys[h];
        if (k == key) return (int)h;
        if (k == PC_EMPTY) return -1;
        h = (h + 1) & mask;
    }
    return -1;
}

// Block-wide exclusive scan over n elements of f(i) (int). Contiguous chunks.
template <class F>
__device__ int pc_block_scan(int n, int* out, PcLDS* s, F f) {
    const int tid = threadIdx.x, wv = tid >> 6;
    const int per = (n + PC_T - 1) / PC_T;
    const int b = min(n, tid * per), e = min(n, b + per);
    int sum = 0;
    for (int i = b; i < e; ++i) sum += f(i);
    const int incl = wave_incl_scan(sum);
    if (lane_id() == 63) s->ired[wv] = incl;
    __syncthreads();
    int base = 0, tot = 0;
    for (int q = 0; q < PC_T / 64; ++q) {
        if (q < wv) base += s->ired[q];
        tot += s->ired[q];
    }
    int run = base + incl - sum;
    if (out)
        for (int i = b; i < e; ++i) {
            out[i] = run;
            run += f(i);
        }
    __syncthreads();
    return tot;
}

__device__ __forceinline__ void* pc_alloc(DPool* pool, unsigned long long bytes, PcLDS* s, int slot, int* err,
                                          int heap_id) {
    if (threadIdx.x == 0) s->ptr[slot] = heap_alloc(pool, heap_id, bytes ? bytes : 16, err);
    __syncthreads();
    void* p = s->ptr[slot];
    __syncthreads();
    return p;
}

__device__ __forceinline__ unsigned long long al16(unsigned long long x) { return (x + 15ull) & ~15ull; }

// 2D helpers for the cap (projection drops the plane normal's dominant axis)
struct P2 {
    float u, v;
};
__device__ __forceinline__ float cross2(P2 a, P2 b, P2 c) { return (b.u - a.u) * (c.v - a.v) - (b.v - a.v) * (c.u - a.u); }

// Whole-mesh copy to one side (no crossing edges or failed cap).
__device__ inline PcMesh pc_copy_mesh(const float* V, const int* T, int nv, int nt, DPool* pool, PcLDS* s, int* err) {
    PcMesh m{nullptr, nullptr, 0, 0, 0.0};
    const unsigned long long vb = al16(12ull * nv), tb = al16(12ull * nt);
    float* chunk = (float*)pc_alloc(pool, vb + tb, s, 0, err, 0);
    if (!chunk) return m;
    int* t = (int*)((char*)chunk + vb);
    for (int i = threadIdx.x; i < 3 * nv; i += PC_T) chunk[i] = V[i];
    for (int i = threadIdx.x; i < 3 * nt; i += PC_T) t[i] = T[i];
    m.v = chunk;
    m.t = t;
    m.nv = nv;
    m.nt = nt;
    return m;
}

// FP64 volume of a mesh chunk (anchored at vertex 0), block-uniform.
__device__ inline double pc_volume(const float* v, const int* t, int nt, PcLDS* s) {
    double acc = 0.0;
    if (nt > 0) {
        const double ax = v[0], ay = v[1], az = v[2];
        for (int k = threadIdx.x; k < nt; k += PC_T) {
            const int a = t[3 * k], b = t[3 * k + 1], c = t[3 * k + 2];
            const double x0 = v[3 * a] - ax, y0 = v[3 * a + 1] - ay, z0 = v[3 * a + 2] - az;
            const double x1 = v[3 * b] - ax, y1 = v[3 * b + 1] - ay, z1 = v[3 * b + 2] - az;
            const double x2 = v[3 * c] - ax, y2 = v[3 * c + 1] - ay, z2 = v[3 * c + 2] - az;
            acc += x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
        }
    }
    acc = wave_sum_d(acc);
    if (lane_id() == 0) s->dred[threadIdx.x >> 6] = acc;
    __syncthreads();
    double r = 0.0;
    for (int q = 0; q < PC_T / 64; ++q) r += s->dred[q];
    __syncthreads();
    return fabs(r) / 6.0;
}

// Closed point-in-triangle test, either orientation (reference cuda/plane_cut.cuh:101-109).
__device__ __forceinline__ bool pc_pt_in_tri(P2 X, P2 A, P2 B, P2 C) {
    const float d1 = cross2(A, B, X), d2 = cross2(B, C, X), d3 = cross2(C, A, X);
    return !((d1 < -1e-10f || d2 < -1e-10f || d3 < -1e-10f) && (d1 > 1e-10f || d2 > 1e-10f || d3 > 1e-10f));
}

// Ear-clip one weakly simple polygon (indices into all-vertex space, CCW in
// 2D), executed by wave 0 (reference ear test, cuda/plane_cut.cuh:1195-1290):
// cross < -1e-9 is not an ear; |cross| <= 1e-9 is clipped as a zero-area ear;
// otherwise an ear has no other live vertex (farther than 1e-6 from its
// corners) in its closed triangle, tested lane-parallel. The walk moves on to
// the next vertex after a clip. After a lap without an ear, the least reflex
// vertex is clipped when it is reflex only by rounding (a sliver: cross above
// -1e-6 of the squared extent); otherwise the cap fails (-1). Writes n - 2
// triangles (vertex ids) to out_tris and returns that count.
__device__ inline int pc_earclip(const int* poly, int n, const P2* uv, int* prv, int* nxt, uint8_t* unused,
                                 int* out_tris) {
    (void)unused;
    const int lane = lane_id();
    if (n < 3) return 0;
    float mnu = 3.4e38f, mxu = -3.4e38f, mnv = 3.4e38f, mxv = -3.4e38f;
    for (int i = lane; i < n; i += 64) {
        mnu = fminf(mnu, uv[i].u); mxu = fmaxf(mxu, uv[i].u);
        mnv = fminf(mnv, uv[i].v); mxv = fmaxf(mxv, uv[i].v);
        prv[i] = i == 0 ? n - 1 : i - 1;
        nxt[i] = i == n - 1 ? 0 : i + 1;
    }
    mnu = wave_min_f(mnu); mxu = wave_max_f(mxu); mnv = wave_min_f(mnv); mxv = wave_max_f(mxv);
    const float ext = fmaxf(mxu - mnu, mxv - mnv);
    const float sliver = -1e-6f * ext * ext;
    // up to 64 R vertices: lane l keeps vertices l + 64 k in registers with a
    // removed bit each (a clipped vertex is exactly one whose predecessor no
    // longer links to it), so the ear test reads no LDS
    constexpr int R = 8;
    const bool regs = n <= 64 * R;
    float ru[R], rv[R];
    uint32_t rem = 0;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        const int i = lane + 64 * k;
        ru[k] = 0.0f;
        rv[k] = 0.0f;
        if (regs && i < n) { ru[k] = uv[i].u; rv[k] = uv[i].v; }
    }
    __threadfence_block();
    auto conv = [&](int p, int c, int q) { return cross2(uv[p], uv[c], uv[q]); };
    int nt = 0, remaining = n, cur = 0, lap = 0;
    while (remaining > 3) {
        int p = prv[cur], q = nxt[cur];
        const float cr = conv(p, cur, q);
        bool ear;
        if (lap >= remaining) {
            // a lap without an ear: least reflex vertex, first in ring order from cur
            int best = cur, c = cur;
            float bc = cr;
            for (int k = 0; k < remaining; ++k) {
                const float x = conv(prv[c], c, nxt[c]);
                if (x > bc) { bc = x; best = c; }
                c = nxt[c];
            }
            if (bc < sliver) return -1;
            cur = best;
            p = prv[cur];
            q = nxt[cur];
            ear = true;
        } else if (cr < -1e-9f) {
            ear = false;
        } else if (cr <= 1e-9f) {
            ear = true;
        } else {
            const P2 A = uv[p], B = uv[cur], C = uv[q];
            bool bad = false;
            auto inside = [&](const P2& X) {
                return fabsf(X.u - A.u) + fabsf(X.v - A.v) > 1e-6f && fabsf(X.u - B.u) + fabsf(X.v - B.v) > 1e-6f &&
                       fabsf(X.u - C.u) + fabsf(X.v - C.v) > 1e-6f && pc_pt_in_tri(X, A, B, C);
            };
            if (regs) {
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    const int i = lane + 64 * k;
                    if (i >= n || ((rem >> k) & 1u) || i == p || i == cur || i == q) continue;
                    if (inside(P2{ru[k], rv[k]})) bad = true;
                }
            } else {
                for (int i = lane; i < n; i += 64) {
                    if (i == p || i == cur || i == q) continue;
                    if (nxt[prv[i]] != i) continue;  // removed
                    if (inside(uv[i])) bad = true;
                }
            }
            ear = !__any(bad);
        }
        if (ear) {
            if (lane == 0) {
                out_tris[3 * nt] = poly[p];
                out_tris[3 * nt + 1] = poly[cur];
                out_tris[3 * nt + 2] = poly[q];
                nxt[p] = q;
                prv[q] = p;
            }
            if (lane == (cur & 63)) rem |= 1u << (cur >> 6);
            __threadfence_block();
            ++nt;
            --remaining;
            cur = q;
            lap = 0;
        } else {
            cur = q;
            ++lap;
        }
    }
    if (remaining == 3) {
        const int p = prv[cur], q = nxt[cur];
        if (lane == 0) {
            out_tris[3 * nt] = poly[p];
            out_tris[3 * nt + 1] = poly[cur];
            out_tris[3 * nt + 2] = poly[q];
        }
        ++nt;
    }
    __threadfence_block();
    return nt;
}

// Scratch of the cap builder, carved from one heap chunk (pc_cap_carve).
struct PcCapWork {
    unsigned long long* ek;  // [H] directed boundary-candidate edges (hash set), then the head table
    int2* be;                // [E] directed boundary edges (a, b) of the positive side
    int* succ;               // [E] successor edge
    int* lab;                // [E] loop label (min edge index in the loop)
    int* rank;               // [E] distance from the loop head
    int* jmp;                // [E] pointer-jumping scratch
    int* jl;                 // [E]
    int* jr;                 // [E]
    int* tmp;                // [E] in-degree / predecessor / loop id of a head
    int* poly;               // [E] loop-major boundary vertices
    P2* uv;                  // [E] their 2D coordinates
    int* wid;                // [W] working polygon (vertex ids)
    P2* wuv;                 // [W] its 2D coordinates
    int* prv;                // [W] ear-clip ring
    int* nxt;                // [W]
    int* capbuf;             // [3W] cap triangles
    uint8_t* refl;           // [W] reflex flags
    int* extra;              // [n_extra] caller's per-triangle scratch
    uint32_t H;
};

// Bytes of the cap builder's chunk for E boundary-candidate edges (W: outer
// loop + every hole + 2 bridge vertices per hole) and n_extra ints.
__device__ __forceinline__ uint32_t pc_cap_hash_size(int E) {
    uint32_t H = 16;
    while ((int)H < 2 * E + 2) H <<= 1;
    return H;
}
__device__ __forceinline__ unsigned long long pc_cap_bytes(unsigned long long E, uint32_t H, unsigned long long n_extra) {
    const unsigned long long W = E + 2ull * PC_MAX_LOOPS + 4ull;
    return 8ull * H + 8ull * E + 4ull * E * 7 + 4ull * E + 8ull * E + W * (4ull + 8ull + 4ull + 4ull + 12ull) +
           al16(W) + 4ull * n_extra + 256ull;
}
__device__ inline PcCapWork pc_cap_carve(char* C, unsigned long long E, uint32_t H) {
    const unsigned long long W = E + 2ull * PC_MAX_LOOPS + 4ull;
    PcCapWork w;
    w.H = H;
    w.ek = (unsigned long long*)C;
    w.be = (int2*)(C + 8ull * H);
    w.succ = (int*)(w.be + E);
    w.lab = w.succ + E;
    w.rank = w.lab + E;
    w.jmp = w.rank + E;
    w.jl = w.jmp + E;
    w.jr = w.jl + E;
    w.tmp = w.jr + E;
    w.poly = w.tmp + E;
    w.uv = (P2*)(w.poly + E);
    w.wid = (int*)(w.uv + E);
    w.wuv = (P2*)(w.wid + W);
    w.prv = (int*)(w.wuv + W);
    w.nxt = w.prv + W;
    w.capbuf = w.nxt + W;
    w.refl = (uint8_t*)(w.capbuf + 3 * W);
    w.extra = (int*)(w.refl + al16(W));
    return w;
}

// Cap polygons of the positive side's boundary edges be[0, nb) (directed
// edges of the positive side with no twin, in (positive triangle, edge)
// order): successor links, loops by pointer jumping, loop orientation and
// nesting, hole bridging and ear clipping (reference cuda/plane_cut.cuh:
// 1008-1308). Vertex ids below nv index V, the others newv. Writes the cap
// triangles to w.capbuf and returns their count, or -1 when the cut cannot
// be capped (open boundary, too many loops, a hole with no bridge, a polygon
// ear clipping cannot finish). Block-uniform; the hash set w.ek is reused as
// the head table (start vertex -> lowest edge index).
__device__ inline int pc_cap_from_edges(const PcCapWork& w, int nb, const float* __restrict__ V, int nv,
                                        const float* newv, float pa, float pb, float pc, PcLDS* s) {
    const int tid = threadIdx.x;
    int2* be = w.be;
    int* succ = w.succ;
    int* lab = w.lab;
    int* rank = w.rank;
    int* jmp = w.jmp;
    int* jl = w.jl;
    int* jr = w.jr;
    int* tmp = w.tmp;
    int* poly = w.poly;
    P2* uv = w.uv;
    int* wid = w.wid;
    P2* wuv = w.wuv;
    int* prv = w.prv;
    int* nxt = w.nxt;
    int* capbuf = w.capbuf;
    uint8_t* refl = w.refl;
    int* hkey = (int*)w.ek;
    int* hval = hkey + w.H;
    if (tid == 0) s->fail = 0;
    __syncthreads();
    // successor of edge j: the boundary edge starting where j ends. Start
    // vertices are unique on a manifold cut; scatter head[start] = edge
    // (atomic min keeps it deterministic), non-manifold touches fixed below.
    for (uint32_t i = tid; i < w.H; i += PC_T) { hkey[i] = -1; hval[i] = 0x7fffffff; }
    __syncthreads();
    for (int j = tid; j < nb; j += PC_T) {
        const int v = be[j].x;
        uint32_t h = pc_hash((unsigned long long)(unsigned)v) & (w.H - 1);
        for (;;) {
            const int old = atomicCAS(&hkey[h], -1, v);
            if (old == -1 || old == v) break;
            h = (h + 1) & (w.H - 1);
        }
        atomicMin(&hval[h], j);
    }
    __syncthreads();
    for (int j = tid; j < nb; j += PC_T) {
        const int v = be[j].y;
        uint32_t h = pc_hash((unsigned long long)(unsigned)v) & (w.H - 1);
        int e = -1;
        for (uint32_t probe = 0; probe < w.H; ++probe) {
            const int k = hkey[h];
            if (k == v) { e = hval[h]; break; }
            if (k == -1) break;
            h = (h + 1) & (w.H - 1);
        }
        succ[j] = e;
    }
    __syncthreads();
    // non-manifold vertices (two loops touching): give the second out-edge
    // of a vertex to the second in-edge so every edge has a distinct successor
    for (int j = tid; j < nb; j += PC_T) tmp[j] = 0;
    __syncthreads();
    for (int j = tid; j < nb; j += PC_T)
        if (succ[j] >= 0) atomicAdd(&tmp[succ[j]], 1);
    __syncthreads();
    if (tid == 0) {
        int bad = 0;
        for (int j = 0; j < nb && !bad; ++j) {
            if (succ[j] < 0) { bad = 1; break; }
            if (tmp[succ[j]] > 1) {
                // reassign: find another edge starting at the same vertex not yet taken
                const int v = be[succ[j]].x;
                for (int q = 0; q < nb; ++q)
                    if (be[q].x == v && tmp[q] == 0) {
                        tmp[succ[j]]--;
                        succ[j] = q;
                        tmp[q] = 1;
                        break;
                    }
            }
        }
        for (int j = 0; j < nb && !bad; ++j)
            if (tmp[j] != 1) bad = 1;
        s->fail = bad;
    }
    __syncthreads();
    if (s->fail) return -1;  // open boundary: the cut cannot be capped
    // loop labels by pointer jumping: label = min edge index on the cycle
    for (int j = tid; j < nb; j += PC_T) { lab[j] = j; jmp[j] = succ[j]; }
    __syncthreads();
    for (int step = 1; step < nb; step <<= 1) {
        for (int j = tid; j < nb; j += PC_T) { jl[j] = min(lab[j], lab[jmp[j]]); jr[j] = jmp[jmp[j]]; }
        __syncthreads();
        for (int j = tid; j < nb; j += PC_T) { lab[j] = jl[j]; jmp[j] = jr[j]; }
        __syncthreads();
    }
    // rank along the loop from its head: walk backwards via doubling on the
    // predecessor-free formulation: rank = number of steps from head
    for (int j = tid; j < nb; j += PC_T) { rank[j] = (j == lab[j]) ? 0 : 1; jmp[j] = j; }
    __syncthreads();
    // predecessor pointers
    for (int j = tid; j < nb; j += PC_T) tmp[succ[j]] = j;
    __syncthreads();
    for (int j = tid; j < nb; j += PC_T) jmp[j] = (j == lab[j]) ? j : tmp[j];
    __syncthreads();
    for (int step = 1; step < nb; step <<= 1) {
        for (int j = tid; j < nb; j += PC_T) { jl[j] = rank[j] + (jmp[j] != j ? rank[jmp[j]] : 0); jr[j] = jmp[jmp[j]]; }
        __syncthreads();
        for (int j = tid; j < nb; j += PC_T) { rank[j] = jl[j]; jmp[j] = jr[j]; }
        __syncthreads();
    }
    // loop table: heads are edges with lab == j
    if (tid == 0) {
        int nl = 0, off = 0;
        for (int j = 0; j < nb; ++j)
            if (lab[j] == j) {
                if (nl >= PC_MAX_LOOPS) { s->fail = 1; break; }
                s->loop_start[nl] = off;
                tmp[j] = nl;  // head -> loop id
                ++nl;
                off = off;  // sizes filled below
            }
        s->n_loops = nl;
    }
    __syncthreads();
    if (s->fail) return -1;
    const int nl = s->n_loops;
    // loop sizes -> starts
    if (tid < nl) s->loop_parent[tid] = 0;
    __syncthreads();
    for (int j = tid; j < nb; j += PC_T) atomicAdd(&s->loop_parent[tmp[lab[j]] ], 1);
    __syncthreads();
    if (tid == 0) {
        int o = 0;
        for (int l = 0; l < nl; ++l) { s->loop_start[l] = o; o += s->loop_parent[l]; }
        s->loop_start[nl] = o;
    }
    __syncthreads();
    // cap polygons wind opposite to the positive side's boundary: a cap
    // edge runs b -> a for boundary edge (a, b); lay loops out reversed
    for (int j = tid; j < nb; j += PC_T) {
        const int l = tmp[lab[j]];
        const int sz = s->loop_start[l + 1] - s->loop_start[l];
        poly[s->loop_start[l] + (sz - 1 - rank[j])] = be[j].x;
    }
    // projection drops the dominant axis of the plane normal
    const float anx = fabsf(pa), any_ = fabsf(pb), anz = fabsf(pc);
    const int pu = (anx >= any_ && anx >= anz) ? 1 : 0;
    const int pv = (anx >= any_ && anx >= anz) ? 2 : ((any_ >= anz) ? 2 : 1);
    __syncthreads();
    for (int j = tid; j < nb; j += PC_T) {
        const int v = poly[j];
        const float* P3 = v < nv ? V + 3 * v : newv + 3 * (v - nv);
        uv[j] = P2{pu == 0 ? P3[0] : P3[1], pv == 1 ? P3[1] : P3[2]};
    }
    __syncthreads();
    // ---- wave 0: orientation, nesting, bridging, ear clipping ------------
    if (tid < 64) {
        const int lane = tid;
        // signed areas: one FP64 pass per loop in vertex order on lane 0 (a
        // fixed summation order keeps orientation and nesting reproducible)
        if (lane == 0)
            for (int l = 0; l < nl; ++l) {
                const int b0 = s->loop_start[l], sz = s->loop_start[l + 1] - b0;
                double a = 0.0;
                for (int i = 0; i < sz; ++i) {
                    const P2 p = uv[b0 + i], q = uv[b0 + (i + 1 == sz ? 0 : i + 1)];
                    a += (double)p.u * (double)q.v - (double)q.u * (double)p.v;
                }
                s->loop_area[l] = (float)(a * 0.5);
            }
        __threadfence_block();
        // nesting: parent = smallest enclosing loop (ray casting from loop i's first vertex)
        for (int l = 0; l < nl; ++l) {
            const P2 q = uv[s->loop_start[l]];
            int par = -1;
            float best = 3.4e38f;
            for (int o = 0; o < nl; ++o) {
                if (o == l) continue;
                const int b0 = s->loop_start[o], sz = s->loop_start[o + 1] - b0;
                if (sz < 3) continue;
                int cr = 0;
                for (int i = lane; i < sz; i += 64) {
                    const P2 a = uv[b0 + i], b = uv[b0 + (i + 1 == sz ? 0 : i + 1)];
                    if ((a.v <= q.v) == (b.v <= q.v)) continue;
                    const float t = (q.v - a.v) / (b.v - a.v);
                    if (a.u + t * (b.u - a.u) > q.u) ++cr;
                }
                cr = wave_sum_i(cr);
                const float ar = fabsf(s->loop_area[o]);
                if ((cr & 1) && ar < best) { best = ar; par = o; }
            }
            if (lane == 0) s->loop_parent[l] = par;
        }
        __threadfence_block();
        if (lane == 0) {
            int bad = 0;
            for (int l = 0; l < nl && !bad; ++l) {
                int d = 0;
                for (int p = s->loop_parent[l]; p >= 0; p = s->loop_parent[p])
                    if (++d > nl) { bad = 1; break; }
                if (bad) break;
                s->loop_hole[l] = d & 1;
                if (!(d & 1) && d >= 2) s->loop_parent[l] = -1;  // island inside a hole
            }
            s->fail = bad;
        }
        __threadfence_block();
        int ncap = 0;
        if (!s->fail) {
            // orient: outer loops CCW, holes CW (in 2D)
            for (int l = 0; l < nl; ++l) {
                const int b0 = s->loop_start[l], sz = s->loop_start[l + 1] - b0;
                const bool want_ccw = !s->loop_hole[l];
                if ((s->loop_area[l] > 0) != want_ccw) {
                    for (int i = lane; i < sz / 2; i += 64) {
                        const int x = poly[b0 + i];
                        poly[b0 + i] = poly[b0 + sz - 1 - i];
                        poly[b0 + sz - 1 - i] = x;
                        const P2 y = uv[b0 + i];
                        uv[b0 + i] = uv[b0 + sz - 1 - i];
                        uv[b0 + sz - 1 - i] = y;
                    }
                }
            }
            __threadfence_block();
            // each outer loop with its holes becomes one simple polygon in the
            // working arrays (capacity W: all loops plus 2 bridge vertices per hole)
            for (int o = 0; o < nl && !s->fail; ++o) {
                if (s->loop_hole[o]) continue;
                const int ob = s->loop_start[o], osz = s->loop_start[o + 1] - ob;
                if (osz < 3) continue;
                int* pid = wid;
                P2* puv = wuv;
                int n = osz;
                for (int i = lane; i < osz; i += 64) { pid[i] = poly[ob + i]; puv[i] = uv[ob + i]; }
                __threadfence_block();
                // holes of this outer loop, by decreasing max-u
                for (;;) {
                    int hbest = -1;
                    float hmax = -3.4e38f;
                    for (int h = 0; h < nl; ++h) {
                        if (!s->loop_hole[h] || s->loop_parent[h] != o) continue;
                        if (s->loop_area[h] == 3.4e38f) continue;  // already bridged
                        const int hb = s->loop_start[h], hs = s->loop_start[h + 1] - hb;
                        float m = -3.4e38f;
                        for (int i = lane; i < hs; i += 64) m = fmaxf(m, uv[hb + i].u);
                        m = wave_max_f(m);
                        if (m > hmax) { hmax = m; hbest = h; }
                    }
                    if (hbest < 0) break;
                    const int hb = s->loop_start[hbest], hs = s->loop_start[hbest + 1] - hb;
                    // rightmost hole vertex
                    uint64_t key = 0;
                    for (int i = lane; i < hs; i += 64) {
                        const uint64_t kk = ((uint64_t)fkey(uv[hb + i].u) << 32) | (uint32_t)(0xFFFFFFFFu - i);
                        key = kk > key ? kk : key;
                    }
                    key = wave_max_u64(key);
                    const int mi = (int)(0xFFFFFFFFu - (uint32_t)key);
                    const P2 M = uv[hb + mi];
                    // nearest polygon edge hit by the ray +u from M
                    uint64_t ek2 = ~0ull;
                    for (int i = lane; i < n; i += 64) {
                        const P2 a = puv[i], b = puv[i + 1 == n ? 0 : i + 1];
                        const float dv = b.v - a.v;
                        if (fabsf(dv) < 1e-12f) continue;
                        const float t = (M.v - a.v) / dv;
                        if (t < -1e-7f || t > 1.0f + 1e-7f) continue;
                        const float x = a.u + t * (b.u - a.u) - M.u;
                        if (x <= 1e-12f) continue;
                        const uint64_t kk = ((uint64_t)fkey(x) << 32) | (uint32_t)i;
                        ek2 = kk < ek2 ? kk : ek2;
                    }
                    ek2 = wave_min_u64(ek2);
                    if (ek2 == ~0ull) { if (lane == 0) s->fail = 1; __threadfence_block(); break; }
                    const int be_ = (int)(uint32_t)ek2;
                    const int bn = be_ + 1 == n ? 0 : be_ + 1;
                    int pp = puv[be_].u >= puv[bn].u ? be_ : bn;
                    {
                        // Visibility (Eberly's hole bridging): if a polygon vertex lies
                        // inside the triangle (M, I, P), with I the ray hit, the bridge
                        // M-P would cross an edge; bridge to the vertex of smallest
                        // angle from the ray instead. Every such bridge is visible.
                        const P2 a = puv[be_], b = puv[bn];
                        const P2 I{a.u + (M.v - a.v) / (b.v - a.v) * (b.u - a.u), M.v};
                        const P2 Pp = puv[pp];
                        // boundary points count as inside (a vertex on the bridge must be
                        // bridged to instead), within a tolerance of the triangle's size;
                        // P itself competes, ties go to the nearest
                        const float L = fmaxf(fmaxf(fabsf(Pp.u - M.u), fabsf(Pp.v - M.v)), I.u - M.u);
                        const float tol = 1e-5f * L * L;
                        uint32_t vt = 0xFFFFFFFFu;
                        uint64_t vk = ~0ull;
                        for (int i = lane; i < n; i += 64) {
                            const P2 X = puv[i];
                            if (X.u <= M.u) continue;
                            const float d1 = cross2(M, I, X), d2 = cross2(I, Pp, X), d3 = cross2(Pp, M, X);
                            const bool in = (d1 >= -tol && d2 >= -tol && d3 >= -tol) ||
                                            (d1 <= tol && d2 <= tol && d3 <= tol);
                            if (!in) continue;
                            // smallest angle to the ray = smallest |dv| / du (du > 0), then nearest
                            const uint32_t k1 = fkey(fabsf(X.v - M.v) / (X.u - M.u));
                            const uint64_t k2 = ((uint64_t)fkey(X.u - M.u) << 32) | (uint32_t)i;
                            if (k1 < vt || (k1 == vt && k2 < vk)) { vt = k1; vk = k2; }
                        }
                        const uint32_t vtm = (uint32_t)wave_min_u64(vt);
                        vk = wave_min_u64(vt == vtm ? vk : ~0ull);
                        if (vk != ~0ull) pp = (int)(uint32_t)vk;
                    }
                    {
                        // an earlier bridge may have duplicated the chosen vertex: take
                        // the lowest copy whose interior wedge contains M
                        const P2 Q = puv[pp];
                        int best = 0x7fffffff;
                        for (int i = lane; i < n; i += 64) {
                            if (puv[i].u != Q.u || puv[i].v != Q.v) continue;
                            const P2 A = puv[i == 0 ? n - 1 : i - 1], Bn = puv[i + 1 == n ? 0 : i + 1];
                            bool in;
                            if (cross2(A, Q, Bn) > 0.0f) in = cross2(Q, Bn, M) >= 0.0f && cross2(Q, M, A) >= 0.0f;
                            else in = cross2(Q, A, M) <= 0.0f || cross2(Q, M, Bn) <= 0.0f;
                            if (in) best = min(best, i);
                        }
                        best = (int)wave_min_u64((uint64_t)(uint32_t)best);
                        if (best != 0x7fffffff) pp = best;
                    }
                    // splice: poly[0..pp], hole from mi around back to mi, poly[pp], poly[pp+1..]
                    const int nn = n + hs + 2;
                    // shift the tail right by hs + 2 (from the end, lane 0 serial for safety)
                    if (lane == 0) {
                        for (int i = n - 1; i > pp; --i) { pid[i + hs + 2] = pid[i]; puv[i + hs + 2] = puv[i]; }
                        for (int j = 0; j <= hs; ++j) {
                            const int src = hb + (mi + j) % hs;
                            pid[pp + 1 + j] = poly[src];
                            puv[pp + 1 + j] = uv[src];
                        }
                        pid[pp + hs + 2] = pid[pp];
                        puv[pp + hs + 2] = puv[pp];
                        s->loop_area[hbest] = 3.4e38f;  // mark bridged
                    }
                    __threadfence_block();
                    n = nn;
                }
                if (s->fail) break;
                const int got = pc_earclip(pid, n, puv, prv, nxt, refl, capbuf + 3 * ncap);
                if (got != n - 2) { if (lane == 0) s->fail = 1; __threadfence_block(); break; }
                ncap += got;
            }
        }
        if (lane == 0) s->cnt[4] = ncap;
    }
    __syncthreads();
    if (s->fail) return -1;
    return s->cnt[4];
}

// Cut mesh (V, T) by plane pa x + pb y + pc z + pd = 0. All PC_T threads call.
// Output meshes are allocated from the output heap (heap 0); scratch from heap 1.
__device__ inline PcResult plane_cut_block(const float* __restrict__ V, const int* __restrict__ T, int nv, int nt,
                                           float pa, float pb, float pc, float pd, DPool* pool, int* err,
                                           PcLDS* s) {
    const int tid = threadIdx.x;
    PcResult R;
    R.pos = PcMesh{nullptr, nullptr, 0, 0, 0.0};
    R.neg = PcMesh{nullptr, nullptr, 0, 0, 0.0};
    R.status = 0;
    R.has_box = 0;
    if (nv <= 0 || nt <= 0) { R.status = 2; return R; }

    // ---- stage A: signs, per-triangle crossing and output counts ------------
    const unsigned long long szA = al16(nv) + 20ull * nt;
    char* A = (char*)pc_alloc(pool, szA, s, 1, err, 1);
    if (!A) { R.status = -1; return R; }
    int8_t* sg = (int8_t*)A;
    int* tc = (int*)(A + al16(nv));    // per tri: pos count | neg count << 8 | crossing instances << 16
    int* toff = tc + nt;               // scratch for scans
    int* toff2 = toff + nt;
    int* toff3 = toff2 + nt;
    int* tnew = toff3 + nt;            // per tri: first new-vertex slot it owns
    if (tid == 0) { s->cnt[0] = 0; s->cnt[1] = 0; s->fail = 0; }
    for (int i = tid; i < nv; i += PC_T) {
        const float d = pc_side(pa, pb, pc, pd, V[3 * i], V[3 * i + 1], V[3 * i + 2]);
        sg[i] = d > PC_EPS ? 1 : (d < -PC_EPS ? -1 : 0);
    }
    __syncthreads();
    int any_pos = 0, any_neg = 0;
    for (int k = tid; k < nt; k += PC_T) {
        const int i0 = T[3 * k], i1 = T[3 * k + 1], i2 = T[3 * k + 2];
        const int s0 = sg[i0], s1 = sg[i1], s2 = sg[i2];
        const int e0 = s0 >= 0, e1 = s1 >= 0, e2 = s2 >= 0;
        const int ncross = (e0 != e1) + (e1 != e2) + (e2 != e0);
        const int hp = (s0 > 0) | (s1 > 0) | (s2 > 0), hn = (s0 < 0) | (s1 < 0) | (s2 < 0);
        int np_ = 0, nn_ = 0;
        if (!hp || !hn) {
            if (hn) nn_ = 1;
            else if (hp) np_ = 1;
            else {
                // all three on the plane: side by the triangle normal
                const float e1x = V[3 * i1] - V[3 * i0], e1y = V[3 * i1 + 1] - V[3 * i0 + 1], e1z = V[3 * i1 + 2] - V[3 * i0 + 2];
                const float e2x = V[3 * i2] - V[3 * i0], e2y = V[3 * i2 + 1] - V[3 * i0 + 1], e2z = V[3 * i2 + 2] - V[3 * i0 + 2];
                const float dt = pa * (e1y * e2z - e1z * e2y) + pb * (e1z * e2x - e1x * e2z) + pc * (e1x * e2y - e1y * e2x);
                if (dt > 0) np_ = 1;
                else nn_ = 1;
            }
        } else if (s0 == 0 || s1 == 0 || s2 == 0) {
            np_ = 1;  // one vertex on the plane, the others on opposite sides
            nn_ = 1;
        } else {
            const int lone_pos = (s0 > 0) + (s1 > 0) + (s2 > 0) == 1;
            np_ = lone_pos ? 1 : 2;
            nn_ = lone_pos ? 2 : 1;
        }
        any_pos |= (s0 > 0) | (s1 > 0) | (s2 > 0);
        any_neg |= (s0 < 0) | (s1 < 0) | (s2 < 0);
        tc[k] = np_ | (nn_ << 8) | (ncross << 16);
    }
    if (__any(any_pos) && lane_id() == 0) atomicOr(&s->cnt[0], 1);
    if (__any(any_neg) && lane_id() == 0) atomicOr(&s->cnt[0], 2);
    __syncthreads();
    const int sides = s->cnt[0];
    const int n_cross_inst = pc_block_scan(nt, toff, s, [&](int k) { return (tc[k] >> 16) & 0xFF; });
    const int npt = pc_block_scan(nt, toff2, s, [&](int k) { return tc[k] & 0xFF; });
    const int nnt = pc_block_scan(nt, toff3, s, [&](int k) { return (tc[k] >> 8) & 0xFF; });

    if (n_cross_inst == 0) {
        // no crossing edge: one side, or disjoint pieces on both sides
        if (!(sides & 1) || !(sides & 2)) {
            PcMesh m = pc_copy_mesh(V, T, nv, nt, pool, s, err);
            m.vol = m.v ? pc_volume(m.v, m.t, m.nt, s) : 0.0;
            if (sides & 2 && !(sides & 1)) R.neg = m;
            else R.pos = m;
            R.status = 2;
            if (tid == 0) heap_free(pool, A, err);
            return R;
        }
    }

    // ---- stage B: intersection vertices via hash, split triangles ---------------
    // Crossing-edge instances are numbered in (triangle, edge) order; the
    // lowest instance of an undirected edge owns it, and owners of edges with
    // both ends off the plane number the new vertices in instance order. The
    // output is therefore the same on every run and equals the serial order.
    uint32_t H = 16;
    while ((int)H < 2 * n_cross_inst + 2) H <<= 1;
    const int maxnew = n_cross_inst;
    const int nall_cap = nv + maxnew;
    const unsigned long long szB = 8ull * H + 8ull * H + al16(12ull * maxnew) + 12ull * npt + 12ull * nnt +
                                   8ull * nall_cap + 64;
    char* B = (char*)pc_alloc(pool, szB, s, 2, err, 1);
    if (!B) {
        if (tid == 0) heap_free(pool, A, err);
        R.status = -1;
        return R;
    }
    unsigned long long* hk = (unsigned long long*)B;
    int* hmin = (int*)(B + 8ull * H);       // lowest instance of the edge
    int* hid = hmin + H;                    // vertex id of the edge's crossing point
    float* newv = (float*)(B + 16ull * H);
    int* ptri = (int*)((char*)newv + al16(12ull * maxnew));
    int* ntri = ptri + 3 * npt;
    int* remap = ntri + 3 * nnt;  // [nall_cap] per side, reused
    int* remap2 = remap + nall_cap;
    for (uint32_t i = tid; i < H; i += PC_T) { hk[i] = PC_EMPTY; hmin[i] = 0x7fffffff; }
    __syncthreads();
    // pass 1: insert undirected crossing edges, keep the lowest instance
    for (int k = tid; k < nt; k += PC_T) {
        if (!((tc[k] >> 16) & 0xFF)) continue;
        const int vi[3] = {T[3 * k], T[3 * k + 1], T[3 * k + 2]};
        int inst = toff[k];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int a = vi[e], b = vi[e == 2 ? 0 : e + 1];
            if ((sg[a] >= 0) == (sg[b] >= 0)) continue;
            const int lo = min(a, b), hi = max(a, b);
            int fresh = 0;
            const int slot = pc_ht_insert(hk, H - 1, ((unsigned long long)lo << 32) | (unsigned)hi, &fresh);
            atomicMin(&hmin[slot], inst);
            ++inst;
        }
    }
    __syncthreads();
    // pass 2: owners of off-plane edges get consecutive new-vertex slots
    auto owned_new = [&](int k, int* first_slot_out) {
        const int vi[3] = {T[3 * k], T[3 * k + 1], T[3 * k + 2]};
        int inst = toff[k], c = 0;
        for (int e = 0; e < 3; ++e) {
            const int a = vi[e], b = vi[e == 2 ? 0 : e + 1];
            if ((sg[a] >= 0) == (sg[b] >= 0)) continue;
            const int lo = min(a, b), hi = max(a, b);
            const int slot = pc_ht_find(hk, H - 1, ((unsigned long long)lo << 32) | (unsigned)hi);
            if (hmin[slot] == inst) {
                if (sg[a] != 0 && sg[b] != 0) {
                    if (first_slot_out) {
                        const int j = *first_slot_out + c;
                        const float ax = V[3 * lo], ay = V[3 * lo + 1], az = V[3 * lo + 2];
                        const float bx = V[3 * hi], by = V[3 * hi + 1], bz = V[3 * hi + 2];
                        const float d0 = pc_side(pa, pb, pc, pd, ax, ay, az);
                        const float d1 = pc_side(pa, pb, pc, pd, bx, by, bz);
                        const float t = d0 / (d0 - d1);
                        newv[3 * j] = ax + t * (bx - ax);
                        newv[3 * j + 1] = ay + t * (by - ay);
                        newv[3 * j + 2] = az + t * (bz - az);
                        hid[slot] = nv + j;
                    }
                    ++c;
                } else if (first_slot_out) {
                    hid[slot] = sg[lo] == 0 ? lo : hi;
                }
            }
            ++inst;
        }
        return c;
    };
    const int n_new = pc_block_scan(nt, tnew, s, [&](int k) { return ((tc[k] >> 16) & 0xFF) ? owned_new(k, nullptr) : 0; });
    for (int k = tid; k < nt; k += PC_T)
        if ((tc[k] >> 16) & 0xFF) {
            int first = tnew[k];
            owned_new(k, &first);
        }
    __syncthreads();
    const int nall = nv + n_new;
    auto isect = [&](int a, int b) {
        const int lo = min(a, b), hi = max(a, b);
        const int slot = pc_ht_find(hk, H - 1, ((unsigned long long)lo << 32) | (unsigned)hi);
        return slot >= 0 ? hid[slot] : -1;
    };
    // pass 3: split triangles into their precomputed slots
    for (int k = tid; k < nt; k += PC_T) {
        const int i0 = T[3 * k], i1 = T[3 * k + 1], i2 = T[3 * k + 2];
        const int s0 = sg[i0], s1 = sg[i1], s2 = sg[i2];
        int* P = ptri + 3 * toff2[k];
        int* N = ntri + 3 * toff3[k];
        const int npk = tc[k] & 0xFF, nnk = (tc[k] >> 8) & 0xFF;
        const int hp = (s0 > 0) | (s1 > 0) | (s2 > 0), hn = (s0 < 0) | (s1 < 0) | (s2 < 0);
        if (!hp || !hn) {
            int* D = npk ? P : N;
            D[0] = i0; D[1] = i1; D[2] = i2;
            continue;
        }
        const int vi[3] = {i0, i1, i2}, si[3] = {s0, s1, s2};
        int on = -1;
        for (int q = 0; q < 3; ++q)
            if (si[q] == 0) { on = q; break; }
        if (on >= 0) {
            const int ov = vi[on], a = vi[(on + 1) % 3], b = vi[(on + 2) % 3], sa = si[(on + 1) % 3];
            const int x = isect(a, b);
            if (sa > 0) {
                P[0] = ov; P[1] = a; P[2] = x;
                N[0] = ov; N[1] = x; N[2] = b;
            } else {
                N[0] = ov; N[1] = a; N[2] = x;
                P[0] = ov; P[1] = x; P[2] = b;
            }
        } else {
            int lone = 0;
            for (int q = 0; q < 3; ++q)
                if (si[q] != si[(q + 1) % 3] && si[q] != si[(q + 2) % 3]) { lone = q; break; }
            const int lv = vi[lone], o1 = vi[(lone + 1) % 3], o2 = vi[(lone + 2) % 3];
            const int x1 = isect(lv, o1), x2 = isect(lv, o2);
            int* L1 = si[lone] > 0 ? P : N;
            int* L2 = si[lone] > 0 ? N : P;
            L1[0] = lv; L1[1] = x1; L1[2] = x2;
            L2[0] = o1; L2[1] = x2; L2[2] = x1;
            L2[3] = o1; L2[4] = o2; L2[5] = x2;
        }
        (void)nnk;
    }
    __syncthreads();
    if (tid == 0) heap_free(pool, A, err);  // signs and counts done

    // ---- stage C: boundary edges of the positive side --------------------------
    // A boundary edge (a, b) of the positive side has its twin (b, a) in the
    // negative side, so both ends are used by both sides: mark side usage and
    // only consider edges between doubly-used vertices.
    for (int i = tid; i < nall; i += PC_T) remap[i] = 0;
    __syncthreads();
    // mark vertices appearing on both sides: exactly the plane vertices of a closed cut
    for (int k = tid; k < npt; k += PC_T)
        for (int e = 0; e < 3; ++e) atomicOr(&remap[ptri[3 * k + e]], 1);
    for (int k = tid; k < nnt; k += PC_T)
        for (int e = 0; e < 3; ++e) atomicOr(&remap[ntri[3 * k + e]], 2);
    __syncthreads();
    const int nbe = pc_block_scan(npt, nullptr, s, [&](int k) {
        int c = 0;
        for (int e = 0; e < 3; ++e)
            c += (remap[ptri[3 * k + e]] == 3 && remap[ptri[3 * k + (e == 2 ? 0 : e + 1)]] == 3) ? 1 : 0;
        return c;
    });
    int n_cap = 0;
    int* cap = nullptr;
    char* Cbuf = nullptr;
    if (nbe > 0) {
        const uint32_t H2 = pc_cap_hash_size(nbe);
        Cbuf = (char*)pc_alloc(pool, pc_cap_bytes(nbe, H2, npt), s, 3, err, 1);
        if (!Cbuf) {
            if (tid == 0) heap_free(pool, B, err);
            R.status = -1;
            return R;
        }
        const PcCapWork cw = pc_cap_carve(Cbuf, nbe, H2);
        unsigned long long* ek = cw.ek;
        int2* be = cw.be;
        int* boff = cw.extra;  // [npt] first boundary edge of each pos triangle
        for (uint32_t i = tid; i < H2; i += PC_T) ek[i] = PC_EMPTY;
        if (tid == 0) s->cnt[3] = 0;
        __syncthreads();
        for (int k = tid; k < npt; k += PC_T)
            for (int e = 0; e < 3; ++e) {
                const int a = ptri[3 * k + e], b = ptri[3 * k + (e == 2 ? 0 : e + 1)];
                if (remap[a] == 3 && remap[b] == 3) {
                    int fresh;
                    pc_ht_insert(ek, H2 - 1, ((unsigned long long)a << 32) | (unsigned)b, &fresh);
                }
            }
        __syncthreads();
        // boundary = on-plane directed edge whose twin is absent, listed in
        // (positive triangle, edge) order
        auto is_bnd = [&](int k, int e) {
            const int a = ptri[3 * k + e], b = ptri[3 * k + (e == 2 ? 0 : e + 1)];
            return remap[a] == 3 && remap[b] == 3 &&
                   pc_ht_find(ek, H2 - 1, ((unsigned long long)b << 32) | (unsigned)a) < 0;
        };
        const int nb = pc_block_scan(npt, boff, s, [&](int k) {
            return (int)is_bnd(k, 0) + (int)is_bnd(k, 1) + (int)is_bnd(k, 2);
        });
        for (int k = tid; k < npt; k += PC_T) {
            int j = boff[k];
            for (int e = 0; e < 3; ++e)
                if (is_bnd(k, e)) be[j++] = make_int2(ptri[3 * k + e], ptri[3 * k + (e == 2 ? 0 : e + 1)]);
        }
        __syncthreads();
        // ek is free again: the cap builder reuses it as its head table
        const int got_cap = pc_cap_from_edges(cw, nb, V, nv, newv, pa, pb, pc, s);
        if (got_cap < 0) {
            if (tid == 0) { heap_free(pool, Cbuf, err); heap_free(pool, B, err); }
            R.pos = pc_copy_mesh(V, T, nv, nt, pool, s, err);
            R.pos.vol = R.pos.v ? pc_volume(R.pos.v, R.pos.t, R.pos.nt, s) : 0.0;
            R.status = 1;
            return R;
        }
        n_cap = got_cap;
        cap = cw.capbuf;
    }

    // ---- stage D: per-side compaction and output -------------------------------
    // Cap triangles wind counter-clockwise in the 2D projection, i.e. their 3D
    // normal points along +x, -y or +z for the dropped axis x, y or z. The
    // positive side's cap must face the negative side (normal . plane < 0), so
    // one flip decision per side keeps every cap triangle (zero-area ears
    // included) consistent with its neighbours.
    bool flip_pos = false;
    {
        const float anx = fabsf(pa), any_ = fabsf(pb), anz = fabsf(pc);
        const float ccw_dot = (anx >= any_ && anx >= anz) ? pa : ((any_ >= anz) ? -pb : pc);
        flip_pos = ccw_dot > 0.0f;
    }
    for (int side = 0; side < 2; ++side) {
        const int* st = side == 0 ? ptri : ntri;
        const int sn = side == 0 ? npt : nnt;
        for (int i = tid; i < nall; i += PC_T) remap2[i] = 0;
        __syncthreads();
        for (int k = tid; k < sn; k += PC_T)
            for (int e = 0; e < 3; ++e) remap2[st[3 * k + e]] = 1;
        for (int k = tid; k < n_cap; k += PC_T)
            for (int e = 0; e < 3; ++e) remap2[cap[3 * k + e]] = 1;
        __syncthreads();
        const int onv = pc_block_scan(nall, remap, s, [&](int i) { return remap2[i]; });
        const int ont = sn + n_cap;
        PcMesh m{nullptr, nullptr, 0, 0, 0.0};
        if (ont > 0 && onv > 0) {
            const unsigned long long vb = al16(12ull * onv), tb = al16(12ull * ont);
            float* chunk = (float*)pc_alloc(pool, vb + tb, s, 0, err, 0);
            if (chunk) {
                int* ot = (int*)((char*)chunk + vb);
                for (int i = tid; i < nall; i += PC_T)
                    if (remap2[i]) {
                        const float* P3 = i < nv ? V + 3 * i : newv + 3 * (i - nv);
                        const int o = remap[i];
                        chunk[3 * o] = P3[0];
                        chunk[3 * o + 1] = P3[1];
                        chunk[3 * o + 2] = P3[2];
                    }
                for (int k = tid; k < sn; k += PC_T)
                    for (int e = 0; e < 3; ++e) ot[3 * k + e] = remap[st[3 * k + e]];
                const bool flip = side == 0 ? flip_pos : !flip_pos;
                for (int k = tid; k < n_cap; k += PC_T) {
                    const int a = cap[3 * k], b = cap[3 * k + 1], c = cap[3 * k + 2];
                    int* o = ot + 3 * (sn + k);
                    o[0] = remap[a];
                    o[1] = remap[flip ? c : b];
                    o[2] = remap[flip ? b : c];
                }
                m.v = chunk;
                m.t = ot;
                m.nv = onv;
                m.nt = ont;
            }
        }
        __syncthreads();
        m.vol = m.v ? pc_volume(m.v, m.t, m.nt, s) : 0.0;
        if (side == 0) R.pos = m;
        else R.neg = m;
    }
    if (tid == 0) {
        void* const f[2] = {Cbuf, B};
        heap_free_n(pool, f, err);
    }
    __syncthreads();
    if (!R.pos.v || !R.neg.v) R.status = 2;
    return R;
}

// ---------------------------------------------------------------------------
// LDS-resident cut: the same output as plane_cut_block, bit for bit, with the
// per-vertex state in LDS and no per-triangle scratch at all.
//
// plane_cut_block keeps per-triangle counts, three prefix-sum arrays, the
// split triangle lists and two vertex remaps in heap scratch: tens of bytes
// per triangle, written and re-read over several passes - at hundreds of
// concurrent cuts that working set cannot stay in L2, so it streams through
// HBM. Here each of the four waves owns a contiguous quarter of the triangle
// list and walks it 64 triangles at a time; output positions come from a
// wave prefix sum plus the wave's base (the totals of the waves before it),
// so every triangle is re-derived from its indices and the vertex signs
// whenever a pass needs it, in the same (triangle, piece) order. LDS holds one
// flag byte per vertex (sign code, used by the positive / negative side) and,
// per 64 vertices and side, the used mask and the exclusive count of used
// vertices before the block, from which a vertex's output index is one
// popcount. Heap scratch is left for what scales with the cut, not the mesh:
// the crossing-edge hash, the new vertices and the cap builder.
// ---------------------------------------------------------------------------

// Dynamic LDS bytes of plane_cut_block_lds for meshes of up to nv vertices.
__host__ __device__ constexpr unsigned long long pc_lds_bytes(long long nv) {
    return ((unsigned long long)nv + 15ull) / 16ull * 16ull + 24ull * (((unsigned long long)nv + 63ull) / 64ull);
}

// One triangle as the LDS-resident cut sees it: vertex ids and signs, kept in
// scalars (no private arrays, so nothing is placed in scratch memory).
struct PcTri {
    int v0, v1, v2, s0, s1, s2;
};
// Up to two pieces of a triangle on one side of the cut, vertex ids in order.
struct PcPieces {
    int n, a0, a1, a2, b0, b1, b2;
};

__device__ inline PcResult plane_cut_block_lds(const float* __restrict__ V, const int* __restrict__ T, int nv, int nt,
                                               float pa, float pb, float pc, float pd, DPool* pool, int* err,
                                               PcLDS* s, unsigned char* dyn) {
    constexpr int NW = PC_T / 64;
    const int tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    PcResult R;
    R.pos = PcMesh{nullptr, nullptr, 0, 0, 0.0};
    R.neg = PcMesh{nullptr, nullptr, 0, 0, 0.0};
    R.status = 0;
    R.has_box = 0;
    if (nv <= 0 || nt <= 0) { R.status = 2; return R; }
    const int nb64 = (nv + 63) >> 6;
    uint8_t* fl = dyn;                                                         // [nv] flags
    unsigned long long* um = (unsigned long long*)(dyn + al16(nv));           // [2][nb64] used masks
    int* up = (int*)(um + 2 * nb64);                                           // [2][nb64] used before the block
    // wave wv owns triangles [tb, te), whole multiples of 64 but the last
    const int tch = (nt + NW * 64 - 1) / (NW * 64) * 64;
    const int tb = min(nt, wv * tch), te = min(nt, tb + tch);

    // flag byte: bits 0-1 sign code (1 positive, 2 negative, 0 on the plane),
    // bit 2 used by the positive side, bit 3 used by the negative side
    auto sgn = [&](int i) -> int {
        const int c = fl[i] & 3;
        return c == 1 ? 1 : (c == 2 ? -1 : 0);
    };
    auto use = [&](int i, unsigned bits) {
        if ((fl[i] & bits) != bits) atomicOr((unsigned*)(fl + (i & ~3)), bits << (8 * (i & 3)));
    };
    auto load = [&](int k) {
        PcTri t;
        t.v0 = T[3 * k];
        t.v1 = T[3 * k + 1];
        t.v2 = T[3 * k + 2];
        t.s0 = sgn(t.v0);
        t.s1 = sgn(t.v1);
        t.s2 = sgn(t.v2);
        return t;
    };
    // an edge crosses when its ends' signs differ with on-plane counted positive
    auto xe = [](int sa, int sb) { return (sa >= 0) != (sb >= 0); };
    auto ncross = [&](const PcTri& t) { return (int)xe(t.s0, t.s1) + (int)xe(t.s1, t.s2) + (int)xe(t.s2, t.s0); };
    // np | nn << 2 | crossing instances << 4 | has positive << 7 | has negative << 8
    auto code = [&](const PcTri& t) -> int {
        const int hp = (t.s0 > 0) | (t.s1 > 0) | (t.s2 > 0), hn = (t.s0 < 0) | (t.s1 < 0) | (t.s2 < 0);
        int np_ = 0, nn_ = 0;
        if (!hp || !hn) {
            if (hn) nn_ = 1;
            else if (hp) np_ = 1;
            else {
                // all three on the plane: side by the triangle normal
                const int i0 = t.v0, i1 = t.v1, i2 = t.v2;
                const float e1x = V[3 * i1] - V[3 * i0], e1y = V[3 * i1 + 1] - V[3 * i0 + 1], e1z = V[3 * i1 + 2] - V[3 * i0 + 2];
                const float e2x = V[3 * i2] - V[3 * i0], e2y = V[3 * i2 + 1] - V[3 * i0 + 1], e2z = V[3 * i2 + 2] - V[3 * i0 + 2];
                const float dt = pa * (e1y * e2z - e1z * e2y) + pb * (e1z * e2x - e1x * e2z) + pc * (e1x * e2y - e1y * e2x);
                if (dt > 0) np_ = 1;
                else nn_ = 1;
            }
        } else if (t.s0 == 0 || t.s1 == 0 || t.s2 == 0) {
            np_ = 1;  // one vertex on the plane, the others on opposite sides
            nn_ = 1;
        } else {
            const int lone_pos = (t.s0 > 0) + (t.s1 > 0) + (t.s2 > 0) == 1;
            np_ = lone_pos ? 1 : 2;
            nn_ = lone_pos ? 2 : 1;
        }
        return np_ | (nn_ << 2) | (ncross(t) << 4) | (hp << 7) | (hn << 8);
    };
    auto block_sum = [&](int v) -> int {
        v = wave_sum_i(v);
        __syncthreads();
        if (lane == 0) s->ired[wv] = v;
        __syncthreads();
        int t = 0;
        for (int q = 0; q < NW; ++q) t += s->ired[q];
        __syncthreads();
        return t;
    };

    // ---- stage A: signs; per-side triangle counts, crossing instances, used vertices
    for (int i = tid; i < nv; i += PC_T) {
        const float d = pc_side(pa, pb, pc, pd, V[3 * i], V[3 * i + 1], V[3 * i + 2]);
        fl[i] = d > PC_EPS ? 1 : (d < -PC_EPS ? 2 : 0);
    }
    if (tid == 0) { s->cnt[0] = 0; s->fail = 0; }
    // the wave's special triangles (crossing an edge, or with two vertices on
    // the plane), in triangle order from sl[tb]: the passes that only concern
    // them walk this list instead of the whole triangle range
    int* sl = (int*)pc_alloc(pool, 4ull * nt + 16, s, 1, err, 1);
    if (!sl) { R.status = -1; return R; }
    int nsl = 0;  // wave-uniform
    {
        int cnp = 0, cnn = 0, cnc = 0, anyp = 0, anyn = 0;
        for (int k0 = tb; k0 < te; k0 += 64) {
            const int k = k0 + lane;
            PcTri t{0, 0, 0, 0, 0, 0};
            int c = 0;
            if (k < te) {
                t = load(k);
                c = code(t);
            }
            const bool special = k < te && (((c >> 4) & 7) || ((t.s0 == 0) + (t.s1 == 0) + (t.s2 == 0) >= 2));
            const unsigned long long sm = ballot(special);
            if (special) sl[tb + nsl + mask_rank(sm)] = k;
            nsl += (int)__popcll(sm);
            if (k >= te) continue;
            const int hp = (c >> 7) & 1, hn = (c >> 8) & 1;
            // a side uses the vertices its pieces of this triangle reference:
            // a whole triangle all three, a split one each off-plane vertex on
            // its own side and an on-plane vertex on both
            if (!hp || !hn) {
                const unsigned bits = (c & 3) ? 4u : 8u;
                use(t.v0, bits);
                use(t.v1, bits);
                use(t.v2, bits);
            } else {
                use(t.v0, t.s0 > 0 ? 4u : (t.s0 < 0 ? 8u : 12u));
                use(t.v1, t.s1 > 0 ? 4u : (t.s1 < 0 ? 8u : 12u));
                use(t.v2, t.s2 > 0 ? 4u : (t.s2 < 0 ? 8u : 12u));
            }
            cnp += c & 3;
            cnn += (c >> 2) & 3;
            cnc += (c >> 4) & 7;
            anyp |= hp;
            anyn |= hn;
        }
        cnp = wave_sum_i(cnp);
        cnn = wave_sum_i(cnn);
        cnc = wave_sum_i(cnc);
        if (lane == 0) { s->wtot[wv][0] = cnp; s->wtot[wv][1] = cnn; s->wtot[wv][2] = cnc; }
        if (__any(anyp) && lane == 0) atomicOr(&s->cnt[0], 1);
        if (__any(anyn) && lane == 0) atomicOr(&s->cnt[0], 2);
    }
    __syncthreads();
    const int sides = s->cnt[0];
    int bnp = 0, bnn = 0, bnc = 0, npt = 0, nnt = 0, n_cross_inst = 0;
    for (int q = 0; q < NW; ++q) {
        if (q < wv) { bnp += s->wtot[q][0]; bnn += s->wtot[q][1]; bnc += s->wtot[q][2]; }
        npt += s->wtot[q][0];
        nnt += s->wtot[q][1];
        n_cross_inst += s->wtot[q][2];
    }
    __syncthreads();
    if (n_cross_inst == 0 && (!(sides & 1) || !(sides & 2))) {
        // no crossing edge and one side only: the mesh goes whole to it
        if (tid == 0) heap_free(pool, sl, err);
        PcMesh m = pc_copy_mesh(V, T, nv, nt, pool, s, err);
        m.vol = m.v ? pc_volume(m.v, m.t, m.nt, s) : 0.0;
        if (sides & 2 && !(sides & 1)) R.neg = m;
        else R.pos = m;
        R.status = 2;
        return R;
    }

    // ---- stage B: crossing edges (lowest instance owns), new vertices -------
    uint32_t H = 16;
    while ((int)H < 2 * n_cross_inst + 2) H <<= 1;
    const unsigned long long szB = 16ull * H + al16(12ull * n_cross_inst) + 64;
    char* B = (char*)pc_alloc(pool, szB, s, 2, err, 1);
    if (!B) {
        if (tid == 0) heap_free(pool, sl, err);
        R.status = -1;
        return R;
    }
    unsigned long long* hk = (unsigned long long*)B;
    int* hmin = (int*)(B + 8ull * H);  // lowest instance of the edge
    int* hid = hmin + H;               // vertex id of the edge's crossing point
    float* newv = (float*)(B + 16ull * H);
    for (uint32_t i = tid; i < H; i += PC_T) { hk[i] = PC_EMPTY; hmin[i] = 0x7fffffff; }
    __syncthreads();
    auto ekey = [](int a, int b) {
        return ((unsigned long long)min(a, b) << 32) | (unsigned)max(a, b);
    };
    {
        auto ins = [&](int a, int sa, int b, int sb, int& inst) {
            if (!xe(sa, sb)) return;
            int fresh = 0;
            const int slot = pc_ht_insert(hk, H - 1, ekey(a, b), &fresh);
            atomicMin(&hmin[slot], inst);
            ++inst;
        };
        int base = bnc;
        for (int j0 = 0; j0 < nsl; j0 += 64) {
            const int j = j0 + lane;
            PcTri t{0, 0, 0, 0, 0, 0};
            int nc = 0;
            if (j < nsl) {
                t = load(sl[tb + j]);
                nc = ncross(t);
            }
            const int incl = wave_incl_scan(nc);
            if (nc) {
                int inst = base + incl - nc;
                ins(t.v0, t.s0, t.v1, t.s1, inst);
                ins(t.v1, t.s1, t.v2, t.s2, inst);
                ins(t.v2, t.s2, t.v0, t.s0, inst);
            }
            base += read_lane(incl, 63);
        }
    }
    __syncthreads();
    // owners of off-plane edges number the new vertices in instance order; the
    // owner of an edge with an on-plane end records that vertex
    auto own_edge = [&](int a, int sa, int b, int sb, int& inst, int* slot_next, int& c) {
        if (!xe(sa, sb)) return;
        const int lo = min(a, b), hi = max(a, b);
        const int slot = pc_ht_find(hk, H - 1, ekey(a, b));
        if (hmin[slot] == inst) {
            if (sa != 0 && sb != 0) {
                if (slot_next) {
                    const int j = *slot_next + c;
                    const float ax = V[3 * lo], ay = V[3 * lo + 1], az = V[3 * lo + 2];
                    const float bx = V[3 * hi], by = V[3 * hi + 1], bz = V[3 * hi + 2];
                    const float d0 = pc_side(pa, pb, pc, pd, ax, ay, az);
                    const float d1 = pc_side(pa, pb, pc, pd, bx, by, bz);
                    const float tt = d0 / (d0 - d1);
                    newv[3 * j] = ax + tt * (bx - ax);
                    newv[3 * j + 1] = ay + tt * (by - ay);
                    newv[3 * j + 2] = az + tt * (bz - az);
                    hid[slot] = nv + j;
                }
                ++c;
            } else if (slot_next) {
                hid[slot] = sgn(lo) == 0 ? lo : hi;
            }
        }
        ++inst;
    };
    auto owned_new = [&](const PcTri& t, int inst, int* slot_next) {
        int c = 0;
        own_edge(t.v0, t.s0, t.v1, t.s1, inst, slot_next, c);
        own_edge(t.v1, t.s1, t.v2, t.s2, inst, slot_next, c);
        own_edge(t.v2, t.s2, t.v0, t.s0, inst, slot_next, c);
        return c;
    };
    {
        int base = bnc, cnew = 0;
        for (int j0 = 0; j0 < nsl; j0 += 64) {
            const int j = j0 + lane;
            PcTri t{0, 0, 0, 0, 0, 0};
            int nc = 0;
            if (j < nsl) {
                t = load(sl[tb + j]);
                nc = ncross(t);
            }
            const int incl = wave_incl_scan(nc);
            if (nc) cnew += owned_new(t, base + incl - nc, nullptr);
            base += read_lane(incl, 63);
        }
        cnew = wave_sum_i(cnew);
        if (lane == 0) s->wtot[wv][3] = cnew;
    }
    __syncthreads();
    int bnew = 0, n_new = 0;
    for (int q = 0; q < NW; ++q) {
        if (q < wv) bnew += s->wtot[q][3];
        n_new += s->wtot[q][3];
    }
    {
        int base = bnc, nbase = bnew;
        for (int j0 = 0; j0 < nsl; j0 += 64) {
            const int j = j0 + lane;
            PcTri t{0, 0, 0, 0, 0, 0};
            int nc = 0;
            if (j < nsl) {
                t = load(sl[tb + j]);
                nc = ncross(t);
            }
            const int incl = wave_incl_scan(nc);
            const int inst0 = base + incl - nc;
            const int c = nc ? owned_new(t, inst0, nullptr) : 0;
            const int nincl = wave_incl_scan(c);
            if (nc) {
                int first = nbase + nincl - c;
                owned_new(t, inst0, &first);
            }
            base += read_lane(incl, 63);
            nbase += read_lane(nincl, 63);
        }
    }
    // used masks and the used counts before each 64-vertex block, per side
    for (int b = wv; b < nb64; b += NW) {
        const int i = b * 64 + lane;
        const int f = i < nv ? fl[i] : 0;
        const unsigned long long mp = ballot((f & 4) != 0), mn = ballot((f & 8) != 0);
        if (lane == 0) {
            um[b] = mp;
            um[nb64 + b] = mn;
        }
    }
    __syncthreads();
    const int onv_pos = pc_block_scan(nb64, up, s, [&](int b) { return (int)__popcll(um[b]); });
    const int onv_neg = pc_block_scan(nb64, up + nb64, s, [&](int b) { return (int)__popcll(um[nb64 + b]); });
    // output index of a vertex on side 0 (positive) or 1 (negative); new
    // vertices are used by both sides and follow the original ones
    auto rmap = [&](int side, int i) -> int {
        if (i >= nv) return (side ? onv_neg : onv_pos) + (i - nv);
        const int b = i >> 6;
        const unsigned long long m = um[side * nb64 + b];
        return up[side * nb64 + b] + (int)__popcll(m & ((1ull << (i & 63)) - 1ull));
    };
    auto bused = [&](int v) { return v >= nv || (fl[v] & 12) == 12; };
    auto isect = [&](int a, int b) {
        const int slot = pc_ht_find(hk, H - 1, ekey(a, b));
        return slot >= 0 ? hid[slot] : -1;
    };
    // the triangle's pieces on one side (0 positive, 1 negative), in the
    // order of that side's triangle list
    auto pieces = [&](const PcTri& t, int c, int side) -> PcPieces {
        PcPieces r{0, 0, 0, 0, 0, 0, 0};
        const int hp = (c >> 7) & 1, hn = (c >> 8) & 1;
        if (!hp || !hn) {
            if (((c & 3) ? 0 : 1) != side) return r;
            r.n = 1; r.a0 = t.v0; r.a1 = t.v1; r.a2 = t.v2;
            return r;
        }
        // rotate so that the first vertex is the on-plane one, else the lone one
        const int on = t.s0 == 0 ? 0 : (t.s1 == 0 ? 1 : (t.s2 == 0 ? 2 : -1));
        if (on >= 0) {
            const int ov = on == 0 ? t.v0 : (on == 1 ? t.v1 : t.v2);
            const int a = on == 0 ? t.v1 : (on == 1 ? t.v2 : t.v0);
            const int b = on == 0 ? t.v2 : (on == 1 ? t.v0 : t.v1);
            const int sa = on == 0 ? t.s1 : (on == 1 ? t.s2 : t.s0);
            const int x = isect(a, b);
            r.n = 1;
            r.a0 = ov;
            if ((sa > 0) == (side == 0)) { r.a1 = a; r.a2 = x; }
            else { r.a1 = x; r.a2 = b; }
            return r;
        }
        const int lone = (t.s0 != t.s1 && t.s0 != t.s2) ? 0 : ((t.s1 != t.s2 && t.s1 != t.s0) ? 1
                                                                 : ((t.s2 != t.s0 && t.s2 != t.s1) ? 2 : 0));
        const int lv = lone == 0 ? t.v0 : (lone == 1 ? t.v1 : t.v2);
        const int o1 = lone == 0 ? t.v1 : (lone == 1 ? t.v2 : t.v0);
        const int o2 = lone == 0 ? t.v2 : (lone == 1 ? t.v0 : t.v1);
        const int sl = lone == 0 ? t.s0 : (lone == 1 ? t.s1 : t.s2);
        const int x1 = isect(lv, o1), x2 = isect(lv, o2);
        if ((sl > 0) == (side == 0)) {
            r.n = 1; r.a0 = lv; r.a1 = x1; r.a2 = x2;
            return r;
        }
        r.n = 2;
        r.a0 = o1; r.a1 = x2; r.a2 = x1;
        r.b0 = o1; r.b1 = o2; r.b2 = x2;
        return r;
    };

    // ---- stage C: boundary edges of the positive side, caps ---------------------
    // Only an edge between vertices used by both sides can be on the boundary:
    // on-plane vertices and new ones, so a triangle contributes only if it is
    // split or has two vertices on the plane.
    auto may_bnd = [&](const PcTri& t, int c) {
        return ((c >> 7) & (c >> 8) & 1) || ((t.s0 == 0) + (t.s1 == 0) + (t.s2 == 0) >= 2);
    };
    auto both2 = [&](int a, int b) { return (int)(bused(a) && bused(b)); };
    int nbe_l = 0;
    for (int j = lane; j < nsl; j += 64) {
        const PcTri t = load(sl[tb + j]);
        const int c = code(t);
        if (!may_bnd(t, c)) continue;
        const PcPieces q = pieces(t, c, 0);
        if (q.n >= 1) nbe_l += both2(q.a0, q.a1) + both2(q.a1, q.a2) + both2(q.a2, q.a0);
        if (q.n == 2) nbe_l += both2(q.b0, q.b1) + both2(q.b1, q.b2) + both2(q.b2, q.b0);
    }
    const int nbe = block_sum(nbe_l);
    int n_cap = 0;
    const int* cap = nullptr;
    char* Cbuf = nullptr;
    if (nbe > 0) {
        const uint32_t H2 = pc_cap_hash_size(nbe);
        Cbuf = (char*)pc_alloc(pool, pc_cap_bytes(nbe, H2, 0), s, 3, err, 1);
        if (!Cbuf) {
            if (tid == 0) { heap_free(pool, B, err); heap_free(pool, sl, err); }
            R.status = -1;
            return R;
        }
        const PcCapWork cw = pc_cap_carve(Cbuf, nbe, H2);
        for (uint32_t i = tid; i < H2; i += PC_T) cw.ek[i] = PC_EMPTY;
        __syncthreads();
        auto dkey = [](int a, int b) { return ((unsigned long long)a << 32) | (unsigned)b; };
        auto put = [&](int a, int b) {
            if (bused(a) && bused(b)) {
                int fresh;
                pc_ht_insert(cw.ek, H2 - 1, dkey(a, b), &fresh);
            }
        };
        for (int j = lane; j < nsl; j += 64) {
            const PcTri t = load(sl[tb + j]);
            const int c = code(t);
            if (!may_bnd(t, c)) continue;
            const PcPieces q = pieces(t, c, 0);
            if (q.n >= 1) { put(q.a0, q.a1); put(q.a1, q.a2); put(q.a2, q.a0); }
            if (q.n == 2) { put(q.b0, q.b1); put(q.b1, q.b2); put(q.b2, q.b0); }
        }
        __syncthreads();
        // boundary = such a directed edge whose twin is absent, listed in
        // (positive triangle, piece, edge) order
        auto bnd = [&](int a, int b) {
            return bused(a) && bused(b) && pc_ht_find(cw.ek, H2 - 1, dkey(b, a)) < 0;
        };
        auto n_bnd = [&](const PcPieces& q) {
            int m = 0;
            if (q.n >= 1) m += (int)bnd(q.a0, q.a1) + (int)bnd(q.a1, q.a2) + (int)bnd(q.a2, q.a0);
            if (q.n == 2) m += (int)bnd(q.b0, q.b1) + (int)bnd(q.b1, q.b2) + (int)bnd(q.b2, q.b0);
            return m;
        };
        int cb = 0;
        for (int j = lane; j < nsl; j += 64) {
            const PcTri t = load(sl[tb + j]);
            const int c = code(t);
            if (may_bnd(t, c)) cb += n_bnd(pieces(t, c, 0));
        }
        cb = wave_sum_i(cb);
        if (lane == 0) s->wtot[wv][0] = cb;
        __syncthreads();
        int base = 0, nb = 0;
        for (int q = 0; q < NW; ++q) {
            if (q < wv) base += s->wtot[q][0];
            nb += s->wtot[q][0];
        }
        for (int j0 = 0; j0 < nsl; j0 += 64) {
            const int j1 = j0 + lane;
            PcPieces q{0, 0, 0, 0, 0, 0, 0};
            int m = 0;
            if (j1 < nsl) {
                const PcTri t = load(sl[tb + j1]);
                const int c = code(t);
                if (may_bnd(t, c)) {
                    q = pieces(t, c, 0);
                    m = n_bnd(q);
                }
            }
            const int incl = wave_incl_scan(m);
            if (m) {
                int j = base + incl - m;
                auto emit = [&](int a, int b) {
                    if (bnd(a, b)) cw.be[j++] = make_int2(a, b);
                };
                if (q.n >= 1) { emit(q.a0, q.a1); emit(q.a1, q.a2); emit(q.a2, q.a0); }
                if (q.n == 2) { emit(q.b0, q.b1); emit(q.b1, q.b2); emit(q.b2, q.b0); }
            }
            base += read_lane(incl, 63);
        }
        __syncthreads();
        const int got_cap = pc_cap_from_edges(cw, nb, V, nv, newv, pa, pb, pc, s);
        if (got_cap < 0) {
            if (tid == 0) { heap_free(pool, Cbuf, err); heap_free(pool, B, err); heap_free(pool, sl, err); }
            R.pos = pc_copy_mesh(V, T, nv, nt, pool, s, err);
            R.pos.vol = R.pos.v ? pc_volume(R.pos.v, R.pos.t, R.pos.nt, s) : 0.0;
            R.status = 1;
            return R;
        }
        n_cap = got_cap;
        cap = cw.capbuf;
    }

    // ---- stage D: both sides written in place --------------------------------
    // Cap triangles wind counter-clockwise in the 2D projection; the positive
    // side's cap must face the negative side (as in plane_cut_block).
    bool flip_pos;
    {
        const float anx = fabsf(pa), any_ = fabsf(pb), anz = fabsf(pc);
        const float ccw_dot = (anx >= any_ && anx >= anz) ? pa : ((any_ >= anz) ? -pb : pc);
        flip_pos = ccw_dot > 0.0f;
    }
    const int onv_p = onv_pos + n_new, onv_n = onv_neg + n_new;
    const int ont_p = npt + n_cap, ont_n = nnt + n_cap;
    float* chp = nullptr;
    float* chn = nullptr;
    if (ont_p > 0 && onv_p > 0) chp = (float*)pc_alloc(pool, al16(12ull * onv_p) + al16(12ull * ont_p), s, 0, err, 0);
    if (ont_n > 0 && onv_n > 0) chn = (float*)pc_alloc(pool, al16(12ull * onv_n) + al16(12ull * ont_n), s, 1, err, 0);
    int* otp = chp ? (int*)((char*)chp + al16(12ull * onv_p)) : nullptr;
    int* otn = chn ? (int*)((char*)chn + al16(12ull * onv_n)) : nullptr;
    // vertices, and each side's bounding box on the way (min / max are exact,
    // so the order does not matter)
    float px0 = 1e30f, py0 = 1e30f, pz0 = 1e30f, px1 = -1e30f, py1 = -1e30f, pz1 = -1e30f;
    float nx0 = 1e30f, ny0 = 1e30f, nz0 = 1e30f, nx1 = -1e30f, ny1 = -1e30f, nz1 = -1e30f;
    auto grow = [&](int side, float x, float y, float z) {
        if (side == 0) {
            px0 = fminf(px0, x); py0 = fminf(py0, y); pz0 = fminf(pz0, z);
            px1 = fmaxf(px1, x); py1 = fmaxf(py1, y); pz1 = fmaxf(pz1, z);
        } else {
            nx0 = fminf(nx0, x); ny0 = fminf(ny0, y); nz0 = fminf(nz0, z);
            nx1 = fmaxf(nx1, x); ny1 = fmaxf(ny1, y); nz1 = fmaxf(nz1, z);
        }
    };
    for (int i = tid; i < nv; i += PC_T) {
        const int f = fl[i];
        if (!(f & 12)) continue;
        const float x = V[3 * i], y = V[3 * i + 1], z = V[3 * i + 2];
        if ((f & 4) && chp) {
            const int o = rmap(0, i);
            chp[3 * o] = x; chp[3 * o + 1] = y; chp[3 * o + 2] = z;
            grow(0, x, y, z);
        }
        if ((f & 8) && chn) {
            const int o = rmap(1, i);
            chn[3 * o] = x; chn[3 * o + 1] = y; chn[3 * o + 2] = z;
            grow(1, x, y, z);
        }
    }
    for (int q = tid; q < n_new; q += PC_T) {
        const float x = newv[3 * q], y = newv[3 * q + 1], z = newv[3 * q + 2];
        if (chp) { const int o = onv_pos + q; chp[3 * o] = x; chp[3 * o + 1] = y; chp[3 * o + 2] = z; grow(0, x, y, z); }
        if (chn) { const int o = onv_neg + q; chn[3 * o] = x; chn[3 * o + 1] = y; chn[3 * o + 2] = z; grow(1, x, y, z); }
    }
    px0 = wave_min_f(px0); py0 = wave_min_f(py0); pz0 = wave_min_f(pz0);
    px1 = wave_max_f(px1); py1 = wave_max_f(py1); pz1 = wave_max_f(pz1);
    nx0 = wave_min_f(nx0); ny0 = wave_min_f(ny0); nz0 = wave_min_f(nz0);
    nx1 = wave_max_f(nx1); ny1 = wave_max_f(ny1); nz1 = wave_max_f(nz1);
    if (lane == 0) {
        float* q = s->box[wv];
        q[0] = px0; q[1] = py0; q[2] = pz0; q[3] = px1; q[4] = py1; q[5] = pz1;
        q[6] = nx0; q[7] = ny0; q[8] = nz0; q[9] = nx1; q[10] = ny1; q[11] = nz1;
    }
    auto put_pieces = [&](int side, const PcPieces& q, int* o) {
        o[0] = rmap(side, q.a0);
        o[1] = rmap(side, q.a1);
        o[2] = rmap(side, q.a2);
        if (q.n == 2) {
            o[3] = rmap(side, q.b0);
            o[4] = rmap(side, q.b1);
            o[5] = rmap(side, q.b2);
        }
    };
    {
        int bp = bnp, bn = bnn;
        for (int k0 = tb; k0 < te; k0 += 64) {
            const int k = k0 + lane;
            PcTri t{0, 0, 0, 0, 0, 0};
            int c = 0;
            if (k < te) {
                t = load(k);
                c = code(t);
            }
            const int np_ = c & 3, nn_ = (c >> 2) & 3;
            const int incl = wave_incl_scan(np_ | (nn_ << 16));
            if (np_ && otp) put_pieces(0, pieces(t, c, 0), otp + 3 * (bp + (incl & 0xFFFF) - np_));
            if (nn_ && otn) put_pieces(1, pieces(t, c, 1), otn + 3 * (bn + (incl >> 16) - nn_));
            const int tot = read_lane(incl, 63);
            bp += tot & 0xFFFF;
            bn += tot >> 16;
        }
    }
    for (int k = tid; k < n_cap; k += PC_T) {
        const int a = cap[3 * k], b = cap[3 * k + 1], c = cap[3 * k + 2];
        if (otp) {
            int* o = otp + 3 * (npt + k);
            o[0] = rmap(0, a);
            o[1] = rmap(0, flip_pos ? c : b);
            o[2] = rmap(0, flip_pos ? b : c);
        }
        if (otn) {
            int* o = otn + 3 * (nnt + k);
            o[0] = rmap(1, a);
            o[1] = rmap(1, flip_pos ? b : c);
            o[2] = rmap(1, flip_pos ? c : b);
        }
    }
    __syncthreads();
    if (tid < 12) {
        const bool mx = (tid % 6) >= 3;
        float a = s->box[0][tid];
        for (int w = 1; w < NW; ++w) a = mx ? fmaxf(a, s->box[w][tid]) : fminf(a, s->box[w][tid]);
        s->boxr[tid] = a;
    }
    if (chp) R.pos = PcMesh{chp, otp, onv_p, ont_p, pc_volume(chp, otp, ont_p, s)};
    if (chn) R.neg = PcMesh{chn, otn, onv_n, ont_n, pc_volume(chn, otn, ont_n, s)};
    if (tid == 0) {
        void* const f[3] = {Cbuf, B, sl};
        heap_free_n(pool, f, err);
    }
    __syncthreads();
    if (!R.pos.v || !R.neg.v) R.status = 2;
    else R.has_box = 1;
    return R;
}

}  // namespace cuacd
