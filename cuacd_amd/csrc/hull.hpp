// hull.hpp - exact convex hull of a point cloud on one wave64.
//
// Semantics follow the reference D&C hull (cuda/hull_dandc.cuh): the fp32
// input is quantized onto an integer grid (span/10216 per unit, axes permuted
// to (median, max, min) extent, cuda/hull_dandc.cuh:1351-1411), the hull is
// computed EXACTLY on that grid, quantized duplicates resolve to the lowest
// input index (cuda/hull_dandc.cuh:1101-1104, 1494-1509), and the output mesh
// uses the ORIGINAL fp32 points looked up by index (1268-1274).
//
// The algorithm is this library's own and is built for a 64-lane wave:
//   * points live in registers, PPL per lane (point i sits in lane i%64,
//     slot i/64), so "farthest conflict point" is one 64-bit wave max;
//   * faces (plane, vertices, neighbours) live in LDS; the region visible from
//     a new point is found by a wave-parallel BFS over face adjacency, the
//     horizon is extracted lane-parallel, and the cone of new faces is stitched
//     with per-vertex start/end maps (no serial ring walk);
//   * only conflict points of deleted faces are re-tested, against the new
//     faces (quickhull outside-set reassignment).
// Orientation predicate: fp32 plane filter with a proven error bound, and the
// exact FP64 determinant (all terms < 2^46 on the +-5109 grid, so exact) only
// where the filter is ambiguous.
#pragma once
#include "wave.hpp"
#include "kerr.h"

namespace cuacd {

constexpr float QUANT_STEPS = 10216.0f;  // grid span (cuda/hull_dandc.cuh:1374-1376)

// Permuted quantization frame: q.x <- median axis, q.y <- max axis, q.z <- min axis.
struct QFrame {
    int ax[3];
    float cen[3];
    float inv[3];
};

// Frame from the cloud's AABB, step for step as the reference presort does it
// (cuda/hull_dandc.cuh:1365-1395), in IEEE fp32 (correctly rounded divisions).
__device__ __forceinline__ QFrame qframe_make(const float mn[3], const float mx[3]) {
    float span[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    int maxAx = (span[0] >= span[1]) ? ((span[0] >= span[2]) ? 0 : 2) : ((span[1] >= span[2]) ? 1 : 2);
    int minAx = (span[0] <= span[1]) ? ((span[0] <= span[2]) ? 0 : 2) : ((span[1] <= span[2]) ? 1 : 2);
    if (minAx == maxAx) minAx = (maxAx + 1) % 3;
    int medAx = 3 - maxAx - minAx;
    float sc[3];
    for (int i = 0; i < 3; ++i) sc[i] = span[i] / QUANT_STEPS;
    if (((medAx + 1) % 3) != maxAx)
        for (int i = 0; i < 3; ++i) sc[i] = -sc[i];
    QFrame f;
    f.ax[0] = medAx; f.ax[1] = maxAx; f.ax[2] = minAx;
    for (int j = 0; j < 3; ++j) {
        int a = f.ax[j];
        f.cen[j] = (mn[a] + mx[a]) * 0.5f;
        f.inv[j] = (sc[a] != 0.0f) ? 1.0f / sc[a] : 0.0f;
    }
    return f;
}

__device__ __forceinline__ int qcoord(const QFrame& f, int j, float v) {
    return (int)((v - f.cen[j]) * f.inv[j]);
}

// ---------------------------------------------------------------------------
// LDS workspace
// ---------------------------------------------------------------------------
struct HullWS {
    float4* fpl;     // [FC] unit outward normal + offset, grid space
    ushort4* fv;     // [FC] v0,v1,v2 (vertex ids, CCW seen from outside), .w = birth/death stamp
    ushort4* fn;     // [FC] neighbour across edge (v0v1, v1v2, v2v0), .w = alive flag
    int* fstamp;     // [FC] BFS visit stamp
    uint8_t* fvis;   // [FC] visibility from the current apex
    uint16_t* ffree; // [FC] free face ids
    uint16_t* vis;   // [FC] visible faces of the current insertion
    short4* vq;      // [VC] grid coordinates of hull vertices
    int* vpid;       // [VC] local point id of each hull vertex
    uint16_t* smap;  // [VC] horizon edge starting at vertex
    uint16_t* emap;  // [VC] horizon edge ending at vertex
    uint16_t* hface; // [VC] face beyond horizon edge i
    uint8_t* hedge;  // [VC] edge slot of edge i inside hface
    uint16_t* ha;    // [VC] horizon edge i = ha -> hb
    uint16_t* hb;    // [VC]
    uint16_t* hnew;  // [VC] new face built on horizon edge i
    uint16_t* vout;  // [VC] output index of a hull vertex
    uint8_t* vref;   // [VC] vertex referenced by a live face
    uint32_t* bm;    // [64] bitmap over local point ids (for output ordering)
    int* wpre;       // [64] prefix popcounts of bm
    int* ctr;        // [8] shared counters
};

// FC = 2*VC + 8 faces bound the live faces of a hull with VC vertices plus the
// transient pair created by one insertion (nhz <= nvis + 2).
__host__ __device__ constexpr int hull_fc(int vc) { return 2 * vc + 8; }

__host__ __device__ constexpr unsigned hull_ws_bytes(int vc) {
    // keep every sub-array 16-byte aligned
    return (unsigned)(((unsigned)hull_fc(vc) * (16u + 8u + 8u + 4u) + 15u) / 16u * 16u
                    + ((unsigned)hull_fc(vc) * (1u + 2u + 2u) + 15u) / 16u * 16u
                    + (unsigned)vc * 8u + (unsigned)vc * 4u
                    + ((unsigned)vc * (2u * 6u + 1u + 1u) + 15u) / 16u * 16u
                    + 64u * 4u + 64u * 4u + 8u * 4u + 64u);
}

__device__ __forceinline__ HullWS hull_ws_carve(char* p, int vc) {
    HullWS w;
    const int fc = hull_fc(vc);
    auto take = [&](unsigned bytes) { char* r = p; p += (bytes + 15u) / 16u * 16u; return r; };
    w.fpl = (float4*)take(fc * 16u);
    w.fv = (ushort4*)take(fc * 8u);
    w.fn = (ushort4*)take(fc * 8u);
    w.fstamp = (int*)take(fc * 4u);
    w.fvis = (uint8_t*)take(fc);
    w.ffree = (uint16_t*)take(fc * 2u);
    w.vis = (uint16_t*)take(fc * 2u);
    w.vq = (short4*)take(vc * 8u);
    w.vpid = (int*)take(vc * 4u);
    w.smap = (uint16_t*)take(vc * 2u);
    w.emap = (uint16_t*)take(vc * 2u);
    w.hface = (uint16_t*)take(vc * 2u);
    w.hedge = (uint8_t*)take(vc);
    w.ha = (uint16_t*)take(vc * 2u);
    w.hb = (uint16_t*)take(vc * 2u);
    w.hnew = (uint16_t*)take(vc * 2u);
    w.vout = (uint16_t*)take(vc * 2u);
    w.vref = (uint8_t*)take(vc);
    w.bm = (uint32_t*)take(64u * 4u);
    w.wpre = (int*)take(64u * 4u);
    w.ctr = (int*)take(8u * 4u);
    return w;
}

// ---------------------------------------------------------------------------
// Predicates
// ---------------------------------------------------------------------------
// Filter bound (grid units). With |q_i| <= 5109 and a unit normal rounded once
// from an FP64 normalisation, |t_f - t| <= eps32*(|q|_inf*sum|n_i| + |w|)
// + 3 fma roundings of magnitude <= 2*8850 -> < 4.5e-3. 1/64 leaves 3.4x margin.
constexpr float HULL_FILTER_EPS = 1.0f / 64.0f;

__device__ __forceinline__ double orient_exact(short4 a, short4 b, short4 c, int qx, int qy, int qz) {
    // ((b-a) x (c-a)) . (q-a); every product < 2^29 * 2^14, exact in FP64.
    double bx = b.x - a.x, by = b.y - a.y, bz = b.z - a.z;
    double cx = c.x - a.x, cy = c.y - a.y, cz = c.z - a.z;
    double dx = qx - a.x, dy = qy - a.y, dz = qz - a.z;
    double nx = by * cz - bz * cy;
    double ny = bz * cx - bx * cz;
    double nz = bx * cy - by * cx;
    return nx * dx + ny * dy + nz * dz;
}

// >0 iff q strictly above face f; *tf receives the fp32 signed distance.
__device__ __forceinline__ bool face_above(const HullWS& w, int f, int qx, int qy, int qz, float* tf) {
    float4 pl = w.fpl[f];
    float t = fmaf(pl.x, (float)qx, fmaf(pl.y, (float)qy, fmaf(pl.z, (float)qz, pl.w)));
    *tf = t;
    if (t > HULL_FILTER_EPS) return true;
    if (t < -HULL_FILTER_EPS) return false;
    ushort4 v = w.fv[f];
    return orient_exact(w.vq[v.x], w.vq[v.y], w.vq[v.z], qx, qy, qz) > 0.0;
}

__device__ __forceinline__ void face_set_plane(const HullWS& w, int f, short4 a, short4 b, short4 c) {
    int bx = b.x - a.x, by = b.y - a.y, bz = b.z - a.z;
    int cx = c.x - a.x, cy = c.y - a.y, cz = c.z - a.z;
    // exact in int32: |components| < 2^29
    int nx = by * cz - bz * cy;
    int ny = bz * cx - bx * cz;
    int nz = bx * cy - by * cx;
    double dnx = nx, dny = ny, dnz = nz;
    double len = sqrt(dnx * dnx + dny * dny + dnz * dnz);
    double il = len > 0.0 ? 1.0 / len : 0.0;
    double ux = dnx * il, uy = dny * il, uz = dnz * il;
    double ww = -(ux * a.x + uy * a.y + uz * a.z);
    w.fpl[f] = make_float4((float)ux, (float)uy, (float)uz, (float)ww);
}

// ---------------------------------------------------------------------------
// Quickhull on one wave.
// ---------------------------------------------------------------------------
enum HullStatus { HULL_OK = 0, HULL_DEGENERATE = 1, HULL_OVERFLOW = 2 };

struct HullResult {
    int status;
    int nvert;  // vertices inserted (vids 0..nvert-1)
    int ftop;   // face ids in use are [0, ftop)
};

// px/py/pz: grid coords of local points i = k*64 + lane (valid iff i < n).
// VC: vertex capacity of the workspace (>= hull vertex count).
template <int PPL>
__device__ HullResult wave_quickhull(const HullWS& w, int n, int VC, const int (&px)[PPL],
                                     const int (&py)[PPL], const int (&pz)[PPL]) {
    const int lane = lane_id();
    const int FC = hull_fc(VC);
    int conf[PPL];
    float dist[PPL];
    HullResult res{HULL_OK, 0, 0};

    auto fetch = [&](int ip, int& x, int& y, int& z) {
        const int own = ip & 63, slot = ip >> 6;
#pragma unroll
        for (int k = 0; k < PPL; ++k)
            if (k == slot) {
                x = read_lane(px[k], own);
                y = read_lane(py[k], own);
                z = read_lane(pz[k], own);
            }
    };

    // ---- initial simplex --------------------------------------------------
    uint64_t key = ~0ull;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * 64 + lane;
        if (i < n) {
            uint64_t kk = ((uint64_t)(py[k] + 8192) << 49) | ((uint64_t)(px[k] + 8192) << 35) |
                          ((uint64_t)(pz[k] + 8192) << 21) | (uint64_t)i;
            key = kk < key ? kk : key;
        }
    }
    key = wave_min_u64(key);
    const int i0 = (int)(key & 0x1FFFFF);
    int x0, y0, z0;
    fetch(i0, x0, y0, z0);

    key = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * 64 + lane;
        if (i < n) {
            int dx = px[k] - x0, dy = py[k] - y0, dz = pz[k] - z0;
            uint64_t d2 = (uint64_t)(dx * dx) + (uint64_t)(dy * dy) + (uint64_t)(dz * dz);
            uint64_t kk = (d2 << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)i);
            key = kk > key ? kk : key;
        }
    }
    key = wave_max_u64(key);
    if ((key >> 32) == 0) { res.status = HULL_DEGENERATE; return res; }
    int i1 = (int)(0xFFFFFFFFu - (uint32_t)key);
    int x1, y1, z1;
    fetch(i1, x1, y1, z1);

    const int ex = x1 - x0, ey = y1 - y0, ez = z1 - z0;
    key = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * 64 + lane;
        if (i < n) {
            int dx = px[k] - x0, dy = py[k] - y0, dz = pz[k] - z0;
            double cx = (double)(ey * dz - ez * dy), cy = (double)(ez * dx - ex * dz),
                   cz = (double)(ex * dy - ey * dx);
            float c2 = (float)(cx * cx + cy * cy + cz * cz);
            uint64_t kk = ((uint64_t)__float_as_uint(c2) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)i);
            key = kk > key ? kk : key;
        }
    }
    key = wave_max_u64(key);
    if ((key >> 32) == 0) { res.status = HULL_DEGENERATE; return res; }
    int i2 = (int)(0xFFFFFFFFu - (uint32_t)key);
    int x2, y2, z2;
    fetch(i2, x2, y2, z2);

    const short4 s0 = make_short4(x0, y0, z0, 0);
    const short4 s1 = make_short4(x1, y1, z1, 0);
    const short4 s2 = make_short4(x2, y2, z2, 0);
    key = 0;
    double o_lane = 0.0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * 64 + lane;
        if (i < n) {
            double o = orient_exact(s0, s1, s2, px[k], py[k], pz[k]);
            float ao = (float)fabs(o);
            uint64_t kk = ((uint64_t)__float_as_uint(ao) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)i);
            if (kk > key) { key = kk; o_lane = o; }
        }
    }
    {
        uint64_t km = wave_max_u64(key);
        // the winning lane holds the exact signed orientation of the apex
        uint64_t win = ballot(key == km && km != 0);
        if ((km >> 32) == 0) { res.status = HULL_DEGENERATE; return res; }
        int wl = __builtin_ffsll((long long)win) - 1;
        double o3 = __shfl(o_lane, wl);
        key = km;
        if (o3 < 0.0) {  // make (p0,p1,p2,p3) positively oriented
            int t = i1; i1 = i2; i2 = t;
        }
    }
    const int i3 = (int)(0xFFFFFFFFu - (uint32_t)key);
    {
        const int ids[4] = {i0, i1, i2, i3};
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int x, y, z;
            fetch(ids[v], x, y, z);
            if (lane == 0) {
                w.vq[v] = make_short4(x, y, z, 0);
                w.vpid[v] = ids[v];
            }
        }
    }
    __syncthreads();
    // faces (0,2,1) (0,1,3) (1,2,3) (0,3,2); neighbours derived from shared edges
    if (lane < 4) {
        const ushort4 fvv[4] = {make_ushort4(0, 2, 1, 0), make_ushort4(0, 1, 3, 0),
                                make_ushort4(1, 2, 3, 0), make_ushort4(0, 3, 2, 0)};
        const ushort4 fnn[4] = {make_ushort4(3, 2, 1, 1), make_ushort4(0, 2, 3, 1),
                                make_ushort4(0, 3, 1, 1), make_ushort4(1, 2, 0, 1)};
        ushort4 fvl = fvv[0], fnl = fnn[0];
        if (lane == 1) { fvl = fvv[1]; fnl = fnn[1]; }
        if (lane == 2) { fvl = fvv[2]; fnl = fnn[2]; }
        if (lane == 3) { fvl = fvv[3]; fnl = fnn[3]; }
        w.fv[lane] = fvl;
        w.fn[lane] = fnl;
        w.fstamp[lane] = 0;
        w.fvis[lane] = 0;
        face_set_plane(w, lane, w.vq[fvl.x], w.vq[fvl.y], w.vq[fvl.z]);
    }
    if (lane == 0) { w.ctr[0] = 0; w.ctr[1] = 0; }
    __syncthreads();

    // ---- assign every other point to the face it is farthest above ----------
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * 64 + lane;
        conf[k] = -1;
        dist[k] = 0.0f;
        if (i < n) {
            if (i == i0 || i == i1 || i == i2 || i == i3) {
                conf[k] = -2;
            } else {
                for (int f = 0; f < 4; ++f) {
                    float t;
                    if (face_above(w, f, px[k], py[k], pz[k], &t) && (conf[k] < 0 || t > dist[k])) {
                        conf[k] = f;
                        dist[k] = t;
                    }
                }
            }
        }
    }

    int nvert = 4, ftop = 4, nfree = 0;
    for (int it = 1;; ++it) {
        // ---- select the farthest conflict point (ties: lowest index) -------
        uint64_t sk = 0;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            if (conf[k] >= 0) {
                uint64_t kk = ((uint64_t)fkey(dist[k]) << 32) |
                              (uint64_t)(0xFFFFFFFFu - (uint32_t)(k * 64 + lane));
                sk = kk > sk ? kk : sk;
            }
        }
        sk = wave_max_u64(sk);
        if (sk == 0) break;
        const int ip = (int)(0xFFFFFFFFu - (uint32_t)sk);
        if (nvert >= VC) { res.status = HULL_OVERFLOW; break; }
        int qx, qy, qz, f0 = 0;
        {
            const int own = ip & 63, slot = ip >> 6;
#pragma unroll
            for (int k = 0; k < PPL; ++k)
                if (k == slot) {
                    qx = read_lane(px[k], own);
                    qy = read_lane(py[k], own);
                    qz = read_lane(pz[k], own);
                    f0 = read_lane(conf[k], own);
                    if (lane == own) conf[k] = -2;
                }
        }
        const int vid = nvert++;
        if (lane == 0) {
            w.vq[vid] = make_short4(qx, qy, qz, 0);
            w.vpid[vid] = ip;
            w.fstamp[f0] = it;
            w.fvis[f0] = 1;
            w.vis[0] = (uint16_t)f0;
            w.ctr[0] = 1;  // nvis
            w.ctr[1] = 0;  // nhz
        }
        __syncthreads();

        // ---- BFS over the faces visible from the apex --------------------
        int head = 0, nvis = 1;
        while (head < nvis) {
            const int cnt = min(nvis - head, 64);
            if (lane < cnt) {
                const int f = w.vis[head + lane];
                const ushort4 nb = w.fn[f];
                const int g3[3] = {nb.x, nb.y, nb.z};
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const int g = g3[e];
                    if (atomicExch(&w.fstamp[g], it) != it) {
                        float t;
                        const bool v = face_above(w, g, qx, qy, qz, &t);
                        w.fvis[g] = v ? 1 : 0;
                        if (v) {
                            int s = atomicAdd(&w.ctr[0], 1);
                            w.vis[s] = (uint16_t)g;
                        }
                    }
                }
            }
            head += cnt;
            __syncthreads();
            nvis = uniform(w.ctr[0]);
        }

        // ---- horizon edges (visible face -> non-visible neighbour) ---------
        for (int b = 0; b < nvis; b += 64) {
            const int j = b + lane;
            if (j < nvis) {
                const int f = w.vis[j];
                const ushort4 fvl = w.fv[f];
                const ushort4 nb = w.fn[f];
                const int vv[3] = {fvl.x, fvl.y, fvl.z};
                const int gg[3] = {nb.x, nb.y, nb.z};
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const int g = gg[e];
                    if (!w.fvis[g]) {
                        const ushort4 gn = w.fn[g];
                        const int ge = (gn.x == f) ? 0 : ((gn.y == f) ? 1 : 2);
                        int s = atomicAdd(&w.ctr[1], 1);
                        if (s < VC) {
                            w.ha[s] = (uint16_t)vv[e];
                            w.hb[s] = (uint16_t)vv[(e + 1) % 3];
                            w.hface[s] = (uint16_t)g;
                            w.hedge[s] = (uint8_t)ge;
                        }
                    }
                }
            }
        }
        __syncthreads();
        const int nhz = uniform(w.ctr[1]);
        if (nhz > VC || nhz < 3) { res.status = HULL_OVERFLOW; break; }

        // ---- face ids for the cone: reuse visible ids, then free list, then top
        const int extra = nhz > nvis ? nhz - nvis : 0;
        const int from_free = min(extra, nfree);
        const int from_top = extra - from_free;
        if (ftop + from_top > FC) { res.status = HULL_OVERFLOW; break; }
        for (int b = 0; b < max(nhz, nvis); b += 64) {
            const int j = b + lane;
            if (j < nhz) {
                int id;
                if (j < nvis) id = w.vis[j];
                else if (j - nvis < from_free) id = w.ffree[nfree - 1 - (j - nvis)];
                else id = ftop + (j - nvis - from_free);
                w.hnew[j] = (uint16_t)id;
                w.smap[w.ha[j]] = (uint16_t)j;
                w.emap[w.hb[j]] = (uint16_t)j;
            } else if (j < nvis) {
                // surplus visible faces die and go back to the free list
                const int f = w.vis[j];
                w.ffree[nfree - from_free + (j - nhz)] = (uint16_t)f;
                ushort4 t = w.fn[f];
                t.w = 0;
                w.fn[f] = t;
                ushort4 u = w.fv[f];
                u.w = (uint16_t)it;
                w.fv[f] = u;
            }
        }
        nfree = nfree - from_free + (nvis > nhz ? nvis - nhz : 0);
        ftop += from_top;
        __syncthreads();

        // ---- stitch the cone ---------------------------------------------
        const short4 sp = make_short4(qx, qy, qz, 0);
        for (int b = 0; b < nhz; b += 64) {
            const int j = b + lane;
            if (j < nhz) {
                const int id = w.hnew[j];
                const int a = w.ha[j], c = w.hb[j];
                const int g = w.hface[j];
                w.fv[id] = make_ushort4(a, c, vid, (uint16_t)it);
                w.fn[id] = make_ushort4(g, w.hnew[w.smap[c]], w.hnew[w.emap[a]], 1);
                w.fstamp[id] = it;
                w.fvis[id] = 0;
                face_set_plane(w, id, w.vq[a], w.vq[c], sp);
            }
        }
        __syncthreads();
        for (int b = 0; b < nhz; b += 64) {
            const int j = b + lane;
            if (j < nhz) {
                const int g = w.hface[j];
                const int ge = w.hedge[j];
                const uint16_t id = w.hnew[j];
                uint16_t* slotp = ((uint16_t*)&w.fn[g]) + ge;
                *slotp = id;
            }
        }
        __syncthreads();

        // ---- re-home conflict points of the faces that died ----------------
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            const int c = conf[k];
            if (c >= 0 && w.fv[c].w == (uint16_t)it) {
                int best = -1;
                float bd = 0.0f;
                for (int j = 0; j < nhz; ++j) {
                    const int id = w.hnew[j];
                    float t;
                    if (face_above(w, id, px[k], py[k], pz[k], &t) && (best < 0 || t > bd)) {
                        best = id;
                        bd = t;
                    }
                }
                conf[k] = best;
                dist[k] = bd;
            }
        }
    }
    res.nvert = nvert;
    res.ftop = ftop;
    return res;
}

// After wave_quickhull: mark referenced vertices and give them output indices in
// ascending local-point order. Returns the output vertex count (wave-uniform).
__device__ inline int hull_rank_vertices(const HullWS& w, const HullResult& r) {
    const int lane = lane_id();
    for (int v = lane; v < r.nvert; v += 64) w.vref[v] = 0;
    w.bm[lane] = 0;
    __syncthreads();
    for (int f = lane; f < r.ftop; f += 64) {
        const ushort4 nb = w.fn[f];
        if (nb.w) {
            const ushort4 v = w.fv[f];
            w.vref[v.x] = 1;
            w.vref[v.y] = 1;
            w.vref[v.z] = 1;
        }
    }
    __syncthreads();
    for (int v = lane; v < r.nvert; v += 64)
        if (w.vref[v]) {
            const int p = w.vpid[v];
            atomicOr(&w.bm[p >> 5], 1u << (p & 31));
        }
    __syncthreads();
    const int c = __popc(w.bm[lane]);
    const int incl = wave_incl_scan(c);
    w.wpre[lane] = incl - c;
    __syncthreads();
    for (int v = lane; v < r.nvert; v += 64)
        if (w.vref[v]) {
            const int p = w.vpid[v];
            w.vout[v] = (uint16_t)(w.wpre[p >> 5] + __popc(w.bm[p >> 5] & ((1u << (p & 31)) - 1u)));
        }
    __syncthreads();
    return uniform(__shfl(incl, 63));
}

__device__ inline int hull_count_faces(const HullWS& w, const HullResult& r) {
    int c = 0;
    for (int f = lane_id(); f < r.ftop; f += 64) c += w.fn[f].w ? 1 : 0;
    return wave_sum_i(c);
}

// Write triangles (output vertex indices) in face-id order; returns count.
__device__ inline int hull_write_tris(const HullWS& w, const HullResult& r, int* __restrict__ out_tris) {
    const int lane = lane_id();
    int base = 0;
    for (int b = 0; b < r.ftop; b += 64) {
        const int f = b + lane;
        const bool alive = f < r.ftop && w.fn[f].w;
        const uint64_t m = ballot(alive);
        if (alive) {
            const int t = base + mask_rank(m);
            const ushort4 v = w.fv[f];
            out_tris[3 * t + 0] = w.vout[v.x];
            out_tris[3 * t + 1] = w.vout[v.y];
            out_tris[3 * t + 2] = w.vout[v.z];
        }
        base += __popcll(m);
    }
    return base;
}

// FP64 divergence-theorem volume over live faces, original fp32 coordinates
// of the point with local id p fetched by `pos(p, xyz)`. Anchored at the first
// hull vertex to avoid cancellation. Wave-uniform result.
template <class PosFn>
__device__ inline double hull_volume_f64(const HullWS& w, const HullResult& r, PosFn pos) {
    // anchor: first vertex of the lowest live face id (wave-uniform)
    int f0 = 0;
    for (int b = 0; b < r.ftop; b += 64) {
        const uint64_t m = ballot(b + lane_id() < r.ftop && w.fn[b + lane_id()].w);
        if (m) { f0 = b + __builtin_ffsll((long long)m) - 1; break; }
    }
    float a[3];
    pos(w.vpid[w.fv[f0].x], a);
    const double ax = a[0], ay = a[1], az = a[2];
    double acc = 0.0;
    for (int f = lane_id(); f < r.ftop; f += 64) {
        if (!w.fn[f].w) continue;
        const ushort4 v = w.fv[f];
        float p0[3], p1[3], p2[3];
        pos(w.vpid[v.x], p0);
        pos(w.vpid[v.y], p1);
        pos(w.vpid[v.z], p2);
        const double x0 = p0[0] - ax, y0 = p0[1] - ay, z0 = p0[2] - az;
        const double x1 = p1[0] - ax, y1 = p1[1] - ay, z1 = p1[2] - az;
        const double x2 = p2[0] - ax, y2 = p2[1] - ay, z2 = p2[2] - az;
        acc += x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
    }
    return wave_sum_d(acc) / 6.0;
}

}  // namespace cuacd
