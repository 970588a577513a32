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

struct PcMesh {
    float* v;   // heap chunk (verts, then tris) - free with heap_free(v)
    int* t;
    int nv, nt;
    double vol;
};

struct PcResult {
    PcMesh pos, neg;
    int status;  // 0 cut, 1 cap failure (mesh returned whole on pos), 2 one side empty, <0 error
};

struct PcLDS {
    int ired[PC_T / 64];
    double dred[PC_T / 64];
    int cnt[8];
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
            for (int i = lane; i < n; i += 64) {
                if (i == p || i == cur || i == q) continue;
                if (nxt[prv[i]] != i) continue;  // removed
                const P2 X = uv[i];
                if (fabsf(X.u - A.u) + fabsf(X.v - A.v) > 1e-6f && fabsf(X.u - B.u) + fabsf(X.v - B.v) > 1e-6f &&
                    fabsf(X.u - C.u) + fabsf(X.v - C.v) > 1e-6f && pc_pt_in_tri(X, A, B, C))
                    bad = true;
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
        uint32_t H2 = 16;
        while ((int)H2 < 2 * nbe + 2) H2 <<= 1;
        // E = boundary edges, W = working polygon capacity (outer loop + every
        // hole + 2 bridge vertices per hole)
        const unsigned long long E = (unsigned long long)nbe, W = E + 2ull * PC_MAX_LOOPS + 4ull;
        const unsigned long long szC = 8ull * H2 + 8ull * E + 4ull * E * 7 + 4ull * E + 8ull * E +
                                       W * (4ull + 8ull + 4ull + 4ull + 12ull) + al16(W) + 4ull * npt + 256ull;
        Cbuf = (char*)pc_alloc(pool, szC, s, 3, err, 1);
        if (!Cbuf) {
            if (tid == 0) heap_free(pool, B, err);
            R.status = -1;
            return R;
        }
        unsigned long long* ek = (unsigned long long*)Cbuf;
        int2* be = (int2*)(Cbuf + 8ull * H2);  // [E] directed boundary edges (a, b) of the pos side
        int* succ = (int*)(be + E);            // [E] successor edge
        int* lab = succ + E;                   // [E] loop label (min edge index in the loop)
        int* rank = lab + E;                   // [E] distance from the loop head
        int* jmp = rank + E;                   // [E] pointer-jumping scratch
        int* jl = jmp + E;                     // [E]
        int* jr = jl + E;                      // [E]
        int* tmp = jr + E;                     // [E] in-degree / predecessor / loop id of a head
        int* poly = tmp + E;                   // [E] loop-major boundary vertices
        P2* uv = (P2*)(poly + E);              // [E] their 2D coordinates
        int* wid = (int*)(uv + E);             // [W] working polygon (vertex ids)
        P2* wuv = (P2*)(wid + W);              // [W] its 2D coordinates
        int* prv = (int*)(wuv + W);            // [W] ear-clip ring
        int* nxt = prv + W;                    // [W]
        int* capbuf = nxt + W;                 // [3W] cap triangles
        uint8_t* refl = (uint8_t*)(capbuf + 3 * W);  // [W] reflex flags
        int* boff = (int*)(refl + al16(W));          // [npt] first boundary edge of each pos triangle
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
        // successor of edge j: the boundary edge starting where j ends. Start
        // vertices are unique on a manifold cut; scatter head[start] = edge
        // (atomic min keeps it deterministic), non-manifold touches fixed below.
        for (int i = tid; i < nall; i += PC_T) remap2[i] = 0x7fffffff;
        __syncthreads();
        for (int j = tid; j < nb; j += PC_T) atomicMin(&remap2[be[j].x], j);
        __syncthreads();
        for (int j = tid; j < nb; j += PC_T) {
            const int h = remap2[be[j].y];
            succ[j] = h == 0x7fffffff ? -1 : h;
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
        if (s->fail) {
            // open boundary: the cut cannot be capped
            if (tid == 0) { heap_free(pool, Cbuf, err); heap_free(pool, B, err); }
            R.pos = pc_copy_mesh(V, T, nv, nt, pool, s, err);
            R.pos.vol = R.pos.v ? pc_volume(R.pos.v, R.pos.t, R.pos.nt, s) : 0.0;
            R.status = 1;
            return R;
        }
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
        if (s->fail) {
            if (tid == 0) { heap_free(pool, Cbuf, err); heap_free(pool, B, err); }
            R.pos = pc_copy_mesh(V, T, nv, nt, pool, s, err);
            R.pos.vol = R.pos.v ? pc_volume(R.pos.v, R.pos.t, R.pos.nt, s) : 0.0;
            R.status = 1;
            return R;
        }
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
        if (s->fail) {
            if (tid == 0) { heap_free(pool, Cbuf, err); heap_free(pool, B, err); }
            R.pos = pc_copy_mesh(V, T, nv, nt, pool, s, err);
            R.pos.vol = R.pos.v ? pc_volume(R.pos.v, R.pos.t, R.pos.nt, s) : 0.0;
            R.status = 1;
            return R;
        }
        n_cap = s->cnt[4];
        cap = capbuf;
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
        if (Cbuf) heap_free(pool, Cbuf, err);
        heap_free(pool, B, err);
    }
    __syncthreads();
    if (!R.pos.v || !R.neg.v) R.status = 2;
    return R;
}

}  // namespace cuacd
