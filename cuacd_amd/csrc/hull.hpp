// hull.hpp - exact convex hull of a point cloud, cooperative over one workgroup.
//
// Semantics follow the reference D&C hull (cuda/hull_dandc.cuh): the fp32
// input is quantized onto an integer grid (span/10216 per unit, axes permuted
// to (median, max, min) extent, cuda/hull_dandc.cuh:1351-1411), the hull is
// computed EXACTLY on that grid, quantized duplicates resolve to the lowest
// input index (cuda/hull_dandc.cuh:1101-1104, 1494-1509), and the output mesh
// uses the ORIGINAL fp32 points looked up by index (1268-1274).
//
// The algorithm is this library's own and is built for 64-lane waves:
//   * points live in registers, PPL per thread (local point i sits in thread
//     i % T, slot i / T, T = threads of the group), so "farthest conflict
//     point" is one 64-bit max reduction;
//   * faces (plane, vertices, neighbours) live in a workspace that is LDS for
//     small hulls (one wave, AS = 3) and global memory for big ones (several
//     waves, AS = 1); the region visible from a new apex is found by a
//     cooperative BFS over face adjacency, the horizon is extracted in
//     parallel and the cone of new faces is stitched with per-vertex start/end
//     maps (no serial ring walk);
//   * only conflict points of deleted faces are re-tested, against the new
//     faces (quickhull outside-set reassignment).
// Orientation predicate: fp32 plane filter with a proven error bound, and the
// exact FP64 determinant (all terms < 2^46 on the +-5109 grid, so exact) only
// where the filter is ambiguous.
#pragma once
#include "wave.hpp"
#include "kerr.h"

namespace cuacd {

template <int AS, class T> using asptr = __attribute__((address_space(AS))) T*;

// clang vector types (not HIP's wrapper classes) so they load and store
// through address-space-qualified pointers as single ds/global ops
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned short u4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));
constexpr int AS_GLOBAL = 1;
constexpr int AS_LDS = 3;

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
        const int a = f.ax[j];
        const float lo = a == 0 ? mn[0] : (a == 1 ? mn[1] : mn[2]);
        const float hi = a == 0 ? mx[0] : (a == 1 ? mx[1] : mx[2]);
        const float s = a == 0 ? sc[0] : (a == 1 ? sc[1] : sc[2]);
        f.cen[j] = (lo + hi) * 0.5f;
        f.inv[j] = (s != 0.0f) ? 1.0f / s : 0.0f;
    }
    return f;
}

__device__ __forceinline__ int qcoord(const QFrame& f, int j, float v) {
    return (int)((v - f.cen[j]) * f.inv[j]);
}

// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------
template <int AS> struct HullWS {
    asptr<AS, f4v> fpl;     // [FC] unit outward normal + offset, grid space
    asptr<AS, u4v> fv;     // [FC] v0,v1,v2 (CCW seen from outside), .w = birth/death stamp
    asptr<AS, u4v> fn;     // [FC] neighbour across edge (v0v1, v1v2, v2v0), .w = alive flag
    asptr<AS, int> fstamp;     // [FC] BFS visit stamp
    asptr<AS, uint8_t> fvis;   // [FC] visibility from the current apex
    asptr<AS, uint16_t> ffree; // [FC] free face ids
    asptr<AS, uint16_t> vis;   // [FC] visible faces of the current insertion
    asptr<AS, s4v> vq;      // [VC] grid coordinates of hull vertices
    asptr<AS, int> vpid;       // [VC] local point id of each hull vertex
    asptr<AS, uint16_t> smap;  // [VC] horizon edge starting at vertex
    asptr<AS, uint16_t> emap;  // [VC] horizon edge ending at vertex
    asptr<AS, uint16_t> hface; // [VC] face beyond horizon edge i
    asptr<AS, uint8_t> hedge;  // [VC] edge slot of edge i inside hface
    asptr<AS, uint16_t> ha;    // [VC] horizon edge i = ha -> hb
    asptr<AS, uint16_t> hb;    // [VC]
    asptr<AS, uint16_t> hnew;  // [VC] new face built on horizon edge i
    asptr<AS, uint16_t> vout;  // [VC] output index of a hull vertex
    asptr<AS, uint8_t> vref;   // [VC] vertex referenced by a live face / corner flag
    asptr<AS, int> cv;         // [2*VC] canonical form: facet list head, triangle offset per c0
    // Always in LDS, whatever AS is: these are updated with atomics and then
    // read back; global atomics execute in L2 while plain loads may hit a
    // stale line of the CU's L1, so atomic-then-load state stays in LDS.
    asptr<AS_LDS, uint32_t> bm; // [NP/32] bitmap over local point ids (output order)
    asptr<AS_LDS, int> wpre;    // [NP/32] prefix popcounts of bm
    asptr<AS_LDS, int> ctr;     // [8] counters: 0 nvis, 1 nhz
    // multi-apex scratch (LDS workspaces with several waves; null otherwise)
    asptr<AS_LDS, uint32_t> mfm;    // [FC/2] per face, 16 bits (face f in word f/2, bits 16 (f&1) up):
                                    // bit r visited by candidate r, bit 8+r visible from it
    asptr<AS_LDS, uint32_t> mfmax;  // [FC] per face: best conflict key (quantised distance, point index);
                                    // only for up to 8192 vertices (hull_ma_layout)
    asptr<AS_LDS, unsigned long long> msk;  // [MA_W] per wave share: best (conflict key << 32 | face)
    asptr<AS_LDS, uint8_t> mfcone;  // [FC] candidate whose cone replaced the face this round
    asptr<AS_LDS, uint16_t> mlist;  // [MA_W][MA_CAP][5]: visible faces, ha, hb, hface, hnew (ranks >= 1)
    asptr<AS_LDS, uint8_t> mhedge;  // [MA_W][MA_CAP]
    asptr<AS_LDS, int> mcand;       // [MA_W][MA_CI] candidate table, then [MA_W] conflict rows, misc
    asptr<AS_LDS, f4v> mcpl;        // [MA_W][MA_CPL] planes of each accepted cone's faces, horizon order
    int VC, FC, NP;
};

// Candidate lists of the multi-apex rounds: candidate 0 uses the workspace's
// own arrays, the others LDS scratch; with a workspace in global memory both
// are addressed through generic pointers.
template <int AS, class T> struct MaList { using type = T*; };
template <class T> struct MaList<AS_LDS, T> { using type = asptr<AS_LDS, T>; };

// Multi-apex rounds: at most MA_W candidates, MA_CAP visible faces and
// horizon edges for each candidate but the first (which uses the main arrays).
// The planes of a cone of at most MA_CPL faces are also kept in cone order, so
// that the points of the faces it replaced scan them without an id lookup.
constexpr int MA_W = 8, MA_CAP = 128, MA_CI = 12, MA_CPL = 32;

// LDS scratch of the multi-apex rounds for fc faces, carved from p (16-byte
// aligned sub-arrays; p == nullptr only sizes it). An LDS workspace carries
// it after its own arrays; a workspace in global memory takes it from the
// kernel's dynamic LDS (hull_ma_attach).
template <int AS>
__host__ __device__ __forceinline__ unsigned hull_ma_layout(char* p, unsigned fc, HullWS<AS>* w) {
    unsigned off = 0;
    auto take = [&](unsigned bytes) {
        char* r = p ? p + off : nullptr;
        off += (bytes + 15u) / 16u * 16u;
        return r;
    };
    w->mfm = (asptr<AS_LDS, uint32_t>)take((fc + 1u) / 2u * 4u);
    w->msk = (asptr<AS_LDS, unsigned long long>)take(MA_W * 8u);
    // per-face keys while they fit (2 + 4 + 1 bytes per face); above 8192
    // vertices (3 bytes per face) the rounds reduce keys per wave share
    if (fc <= 2u * 8192u + 8u) w->mfmax = (asptr<AS_LDS, uint32_t>)take(fc * 4u);
    w->mfcone = (asptr<AS_LDS, uint8_t>)take(fc);
    w->mlist = (asptr<AS_LDS, uint16_t>)take(MA_W * MA_CAP * 5 * 2u);
    w->mhedge = (asptr<AS_LDS, uint8_t>)take(MA_W * MA_CAP);
    w->mcand = (asptr<AS_LDS, int>)take((MA_W * MA_CI + MA_W + 16) * 4u);
    w->mcpl = (asptr<AS_LDS, f4v>)take(MA_W * MA_CPL * 16u);
    return off;
}
__host__ __device__ inline unsigned hull_ma_bytes(int vc) {
    HullWS<AS_LDS> w{};
    return hull_ma_layout<AS_LDS>(nullptr, 2u * (unsigned)vc + 8u, &w);
}
template <int AS> __device__ __forceinline__ HullWS<AS> hull_ma_attach(HullWS<AS> w, char* lds) {
    hull_ma_layout<AS>(lds, (unsigned)w.FC, &w);
    return w;
}

// FC = 2*VC + 8 faces bound the live faces of a hull with VC vertices plus the
// transient pair created by one insertion (nhz <= nvis + 2). Face and vertex
// ids are 16-bit, so VC <= 32760.
__host__ __device__ constexpr int hull_fc(int vc) { return 2 * vc + 8; }

// Carves the workspace from `p` (16-byte aligned sub-arrays) and returns its
// size; with p == nullptr it only sizes it. One routine for both, so the size
// requested and the carve can never disagree. np = point capacity.
template <int AS>
__host__ __device__ __forceinline__ unsigned hull_ws_layout(char* p, int vc, int np, HullWS<AS>* out,
                                                            bool multi = false) {
    HullWS<AS> w{};
    const unsigned fc = (unsigned)hull_fc(vc), v = (unsigned)vc, nw = (unsigned)(np + 31) / 32u;
    unsigned off = 0;
    auto take = [&](unsigned bytes) {
        char* r = p ? p + off : nullptr;
        off += (bytes + 15u) / 16u * 16u;
        return r;
    };
    w.fpl = (asptr<AS, f4v>)take(fc * 16u);
    w.fv = (asptr<AS, u4v>)take(fc * 8u);
    w.fn = (asptr<AS, u4v>)take(fc * 8u);
    w.fstamp = (asptr<AS, int>)take(fc * 4u);
    w.fvis = (asptr<AS, uint8_t>)take(fc);
    w.ffree = (asptr<AS, uint16_t>)take(fc * 2u);
    w.vis = (asptr<AS, uint16_t>)take(fc * 2u);
    w.vq = (asptr<AS, s4v>)take(v * 8u);
    w.vpid = (asptr<AS, int>)take(v * 4u);
    w.smap = (asptr<AS, uint16_t>)take(v * 2u);
    w.emap = (asptr<AS, uint16_t>)take(v * 2u);
    w.hface = (asptr<AS, uint16_t>)take(v * 2u);
    w.hedge = (asptr<AS, uint8_t>)take(v);
    w.ha = (asptr<AS, uint16_t>)take(v * 2u);
    w.hb = (asptr<AS, uint16_t>)take(v * 2u);
    w.hnew = (asptr<AS, uint16_t>)take(v * 2u);
    w.vout = (asptr<AS, uint16_t>)take(v * 2u);
    w.vref = (asptr<AS, uint8_t>)take(v);
    w.cv = (asptr<AS, int>)take(v * 8u);
    if (multi && AS == AS_LDS) off += hull_ma_layout(p ? p + off : nullptr, fc, &w);
    (void)nw;  // bm / wpre / ctr live in HullBlockLDS (hull_ws_attach)
    w.VC = vc;
    w.FC = (int)fc;
    w.NP = np;
    if (out) *out = w;
    return off;
}

template <int AS> __host__ __device__ inline unsigned hull_ws_bytes(int vc, int np, bool multi = false) {
    return hull_ws_layout<AS>(nullptr, vc, np, nullptr, multi);
}

template <int AS> __device__ __forceinline__ HullWS<AS> hull_ws_carve(char* p, int vc, int np, bool multi = false) {
    HullWS<AS> w;
    hull_ws_layout<AS>(p, vc, np, &w, multi);
    return w;
}

__device__ __forceinline__ s4v mk_s4(int x, int y, int z, int w) {
    s4v r = {(short)x, (short)y, (short)z, (short)w};
    return r;
}
__device__ __forceinline__ u4v mk_u4(int x, int y, int z, int w) {
    u4v r = {(unsigned short)x, (unsigned short)y, (unsigned short)z, (unsigned short)w};
    return r;
}
__device__ __forceinline__ f4v mk_f4(float x, float y, float z, float w) {
    f4v r = {x, y, z, w};
    return r;
}
__device__ __forceinline__ float sel3(const float* v, int i) { return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]); }

constexpr int HULL_NP_MAX = 16384;  // largest point set one workgroup hulls

// Small block-wide scratch that always lives in LDS (reductions, broadcasts,
// counters). The output-order bitmap is HullBits, sized per kernel.
struct HullBlockLDS {
    unsigned long long red[16];
    double dred[16];
    int bc[8];
    int ctr[8];
    float bmn[16][3];  // per-wave AABB partials
    float bmx[16][3];
};

// Output-order bitmap (one bit per input point) and its per-word prefix
// counts, in LDS, for point sets of up to NPMAX points: a one-wave tier of
// 256 points needs 64 bytes where the 16384-point bound takes 4 KB, and LDS
// is what limits the small tiers' workgroups per CU.
template <int NPMAX> struct HullBits {
    static_assert(NPMAX > 0 && NPMAX <= HULL_NP_MAX, "HullBits capacity");
    uint32_t bm[(NPMAX + 31) / 32];
    int wpre[(NPMAX + 31) / 32];
};

template <int AS, int NPMAX>
__device__ __forceinline__ HullWS<AS> hull_ws_attach(HullWS<AS> w, HullBlockLDS* sb, HullBits<NPMAX>* hb) {
    w.bm = (asptr<AS_LDS, uint32_t>)hb->bm;
    w.wpre = (asptr<AS_LDS, int>)hb->wpre;
    w.ctr = (asptr<AS_LDS, int>)sb->ctr;
    return w;
}

// ---------------------------------------------------------------------------
// Predicates
// ---------------------------------------------------------------------------
// Filter bound (grid units). The stored plane is n' = s*n + e, w = -fl(n'.a),
// with n the exact integer normal, s ~ 1/|n| and |e_i| <= 2 eps s|n_i|. For
// |q_i| <= 5109 and |q - a|_inf <= 10218 the computed value satisfies
// |t_f - s*T| <= eps*|n'|_1*(2*10218 + 9*5109) < 7e-3 (T = exact orientation,
// eps = 2^-24, |n'|_1 <= sqrt(3)). 1/64 leaves a 2.2x margin, so |t_f| > 1/64
// decides the sign of T.
constexpr float HULL_FILTER_EPS = 1.0f / 64.0f;

__device__ __forceinline__ double orient_exact(s4v a, s4v b, s4v c, int qx, int qy, int qz) {
    // ((b-a) x (c-a)) . (q-a); every product < 2^29 * 2^14, exact in FP64
    // (also under fma contraction: each fma adds exact integers below 2^53).
    double bx = b.x - a.x, by = b.y - a.y, bz = b.z - a.z;
    double cx = c.x - a.x, cy = c.y - a.y, cz = c.z - a.z;
    double dx = qx - a.x, dy = qy - a.y, dz = qz - a.z;
    double nx = by * cz - bz * cy;
    double ny = bz * cx - bx * cz;
    double nz = bx * cy - by * cx;
    return nx * dx + ny * dy + nz * dz;
}

// true iff q is strictly above face f; *tf receives the fp32 signed distance.
template <int AS>
__device__ __forceinline__ bool face_above(const HullWS<AS> w, int f, int qx, int qy, int qz, float* tf) {
    const f4v pl = w.fpl[f];
    const float t = fmaf(pl.x, (float)qx, fmaf(pl.y, (float)qy, fmaf(pl.z, (float)qz, pl.w)));
    *tf = t;
    if (t > HULL_FILTER_EPS) return true;
    if (t < -HULL_FILTER_EPS) return false;
    const u4v v = w.fv[f];
    return orient_exact(w.vq[v.x], w.vq[v.y], w.vq[v.z], qx, qy, qz) > 0.0;
}

template <int AS>
__device__ __forceinline__ f4v face_set_plane(const HullWS<AS> w, int f, s4v a, s4v b, s4v c) {
    const int bx = b.x - a.x, by = b.y - a.y, bz = b.z - a.z;
    const int cx = c.x - a.x, cy = c.y - a.y, cz = c.z - a.z;
    // exact in int32: |components| < 2^29
    const float nx = (float)(by * cz - bz * cy);
    const float ny = (float)(bz * cx - bx * cz);
    const float nz = (float)(bx * cy - by * cx);
    const float l2 = fmaf(nx, nx, fmaf(ny, ny, nz * nz));
    const float il = l2 > 0.0f ? rsqrtf(l2) : 0.0f;
    const float ux = nx * il, uy = ny * il, uz = nz * il;
    const float ww = -fmaf(ux, (float)a.x, fmaf(uy, (float)a.y, uz * (float)a.z));
    const f4v pl = mk_f4(ux, uy, uz, ww);
    w.fpl[f] = pl;
    return pl;
}

template <class P> __device__ __forceinline__ int atom_exch(P p, int v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class P> __device__ __forceinline__ int atom_add(P p, int v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
template <class P> __device__ __forceinline__ void atom_or(P p, uint32_t v) {
    __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Block-wide max of a 64-bit key (every thread gets the result).
template <int NW> __device__ __forceinline__ uint64_t block_max_u64(uint64_t v, HullBlockLDS* s) {
    v = wave_max_u64(v);
    if (NW == 1) return v;
    const int wv = threadIdx.x >> 6;
    if (lane_id() == 0) s->red[wv] = v;
    __syncthreads();
    uint64_t r = s->red[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) r = s->red[i] > r ? s->red[i] : r;
    __syncthreads();
    return r;
}
template <int NW> __device__ __forceinline__ uint64_t block_min_u64(uint64_t v, HullBlockLDS* s) {
    v = wave_min_u64(v);
    if (NW == 1) return v;
    const int wv = threadIdx.x >> 6;
    if (lane_id() == 0) s->red[wv] = v;
    __syncthreads();
    uint64_t r = s->red[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) r = s->red[i] < r ? s->red[i] : r;
    __syncthreads();
    return r;
}
template <int NW> __device__ __forceinline__ int block_sum_i(int v, HullBlockLDS* s) {
    v = wave_sum_i(v);
    if (NW == 1) return v;
    const int wv = threadIdx.x >> 6;
    if (lane_id() == 0) s->red[wv] = (unsigned long long)(unsigned)v;
    __syncthreads();
    int r = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) r += (int)(unsigned)s->red[i];
    __syncthreads();
    return r;
}
template <int NW> __device__ __forceinline__ double block_sum_d(double v, HullBlockLDS* s) {
    v = wave_sum_d(v);
    if (NW == 1) return v;
    const int wv = threadIdx.x >> 6;
    if (lane_id() == 0) s->dred[wv] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < NW; ++i) r += s->dred[i];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Quickhull
// ---------------------------------------------------------------------------
enum HullStatus { HULL_OK = 0, HULL_DEGENERATE = 1, HULL_OVERFLOW = 2 };

struct HullResult {
    int status;
    int nvert;  // vertices inserted (vids 0..nvert-1)
    int ftop;   // face ids in use are [0, ftop)
};

// Cooperative exact quickhull over NW waves (blockDim.x == 64*NW).
// px/py/pz: grid coords of local points i = k*T + tid (valid iff i < n), T = 64*NW.
template <int PPL, int AS, int NW, bool MULTI = false>
__device__ __forceinline__ HullResult block_quickhull(const HullWS<AS> w, HullBlockLDS* sb, int n, const int (&px)[PPL],
                                      const int (&py)[PPL], const int (&pz)[PPL]) {
    constexpr int T = 64 * NW;
    const int tid = threadIdx.x;
    const int VC = w.VC, FC = w.FC;
    int conf[PPL];
    float dist[PPL];
    HullResult res{HULL_OK, 0, 0};

    // coordinates of local point ip, broadcast to the whole block
    // Slot `slot` of a per-lane array, selected with masks: written as
    // `if (k == slot) v = a[k]` the compiler turns the loop into an indexed
    // access and moves the whole array to scratch memory.
    auto pick = [&](const int (&a)[PPL], int slot) {
        int v = 0;
#pragma unroll
        for (int k = 0; k < PPL; ++k) v |= a[k] & -(int)(k == slot);
        return v;
    };
    auto fetch = [&](int ip, int& x, int& y, int& z) {
        const int own = ip % T, slot = ip / T;
        if (NW == 1) {
            x = read_lane(pick(px, slot), own);
            y = read_lane(pick(py, slot), own);
            z = read_lane(pick(pz, slot), own);
        } else {
            if (tid == own) { sb->bc[0] = pick(px, slot); sb->bc[1] = pick(py, slot); sb->bc[2] = pick(pz, slot); }
            __syncthreads();
            x = sb->bc[0]; y = sb->bc[1]; z = sb->bc[2];
            __syncthreads();
        }
    };

    // ---- initial simplex --------------------------------------------------
    uint64_t key = ~0ull;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * T + tid;
        if (i < n) {
            uint64_t kk = ((uint64_t)(py[k] + 8192) << 49) | ((uint64_t)(px[k] + 8192) << 35) |
                          ((uint64_t)(pz[k] + 8192) << 21) | (uint64_t)i;
            key = kk < key ? kk : key;
        }
    }
    key = block_min_u64<NW>(key, sb);
    const int i0 = (int)(key & 0x1FFFFF);
    int x0, y0, z0;
    fetch(i0, x0, y0, z0);

    key = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * T + tid;
        if (i < n) {
            int dx = px[k] - x0, dy = py[k] - y0, dz = pz[k] - z0;
            uint64_t d2 = (uint64_t)(dx * dx) + (uint64_t)(dy * dy) + (uint64_t)(dz * dz);
            uint64_t kk = (d2 << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)i);
            key = kk > key ? kk : key;
        }
    }
    key = block_max_u64<NW>(key, sb);
    if ((key >> 32) == 0) { res.status = HULL_DEGENERATE; return res; }
    int i1 = (int)(0xFFFFFFFFu - (uint32_t)key);
    int x1, y1, z1;
    fetch(i1, x1, y1, z1);

    const int ex = x1 - x0, ey = y1 - y0, ez = z1 - z0;
    key = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * T + tid;
        if (i < n) {
            int dx = px[k] - x0, dy = py[k] - y0, dz = pz[k] - z0;
            double cx = (double)(ey * dz - ez * dy), cy = (double)(ez * dx - ex * dz),
                   cz = (double)(ex * dy - ey * dx);
            float c2 = (float)(cx * cx + cy * cy + cz * cz);
            uint64_t kk = ((uint64_t)__float_as_uint(c2) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)i);
            key = kk > key ? kk : key;
        }
    }
    key = block_max_u64<NW>(key, sb);
    if ((key >> 32) == 0) { res.status = HULL_DEGENERATE; return res; }
    int i2 = (int)(0xFFFFFFFFu - (uint32_t)key);
    int x2, y2, z2;
    fetch(i2, x2, y2, z2);

    const s4v s0 = mk_s4(x0, y0, z0, 0);
    const s4v s1 = mk_s4(x1, y1, z1, 0);
    const s4v s2 = mk_s4(x2, y2, z2, 0);
    key = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * T + tid;
        if (i < n) {
            double o = orient_exact(s0, s1, s2, px[k], py[k], pz[k]);
            float ao = (float)fabs(o);
            uint64_t kk = ((uint64_t)__float_as_uint(ao) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)i);
            key = kk > key ? kk : key;
        }
    }
    key = block_max_u64<NW>(key, sb);
    if ((key >> 32) == 0) { res.status = HULL_DEGENERATE; return res; }
    const int i3 = (int)(0xFFFFFFFFu - (uint32_t)key);
    int x3, y3, z3;
    fetch(i3, x3, y3, z3);
    if (orient_exact(s0, s1, s2, x3, y3, z3) < 0.0) {  // make (p0,p1,p2,p3) positively oriented
        int t = i1; i1 = i2; i2 = t;
        t = x1; x1 = x2; x2 = t;
        t = y1; y1 = y2; y2 = t;
        t = z1; z1 = z2; z2 = t;
    }
    if (tid == 0) {
        w.vq[0] = mk_s4(x0, y0, z0, 0); w.vpid[0] = i0;
        w.vq[1] = mk_s4(x1, y1, z1, 0); w.vpid[1] = i1;
        w.vq[2] = mk_s4(x2, y2, z2, 0); w.vpid[2] = i2;
        w.vq[3] = mk_s4(x3, y3, z3, 0); w.vpid[3] = i3;
    }
    __syncthreads();
    // faces (0,2,1) (0,1,3) (1,2,3) (0,3,2); neighbours from the shared edges
    if (tid < 4) {
        u4v fvl, fnl;
        if (tid == 0) { fvl = mk_u4(0, 2, 1, 0); fnl = mk_u4(3, 2, 1, 1); }
        else if (tid == 1) { fvl = mk_u4(0, 1, 3, 0); fnl = mk_u4(0, 2, 3, 1); }
        else if (tid == 2) { fvl = mk_u4(1, 2, 3, 0); fnl = mk_u4(0, 3, 1, 1); }
        else { fvl = mk_u4(0, 3, 2, 0); fnl = mk_u4(1, 2, 0, 1); }
        w.fv[tid] = fvl;
        w.fn[tid] = fnl;
        w.fstamp[tid] = 0;
        w.fvis[tid] = 0;
        face_set_plane(w, tid, w.vq[fvl.x], w.vq[fvl.y], w.vq[fvl.z]);
    }
    __syncthreads();

    // ---- assign every other point to the face it is farthest above ----------
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
        const int i = k * T + tid;
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
    if constexpr (MULTI) {
        // ---- multi-apex rounds ---------------------------------------------
        // Each round takes the farthest conflict point of every face, keeps
        // the best face of each wave's share of the faces (ids interleaved
        // over the waves; at most NW
        // candidates, the first being the overall best), and lets wave r find
        // the visible region of candidate r. Two candidates conflict when one's
        // visible region meets the other's visited faces (its visible region
        // and their neighbours); a conflict-free set of candidates (greedy by
        // rank) is inserted at once. For such a set, inserting them together
        // gives the hull of inserting them one after another: a candidate sees
        // none of another's cone faces, since a walk over visible faces from
        // its region to such a cone face would cross a horizon face of the
        // other, which would then be in its own visible region. Candidates may
        // be non-extreme points (their distances are quantised); the exact
        // incremental construction still ends at the unique hull.
        static_assert(NW > 1 && NW <= MA_W, "multi-apex rounds need several waves");
        // a conflict key is the face's best point: its quantised distance
        // above the high bits, its (inverted) local index in the low ones
        static_assert(PPL * T <= 16384, "multi-apex keys hold point indices below 16384");
        constexpr uint32_t IMASK =
            PPL * T <= 2048 ? 0x7FFu : (PPL * T <= 4096 ? 0xFFFu : (PPL * T <= 8192 ? 0x1FFFu : 0x3FFFu));
        using L16 = typename MaList<AS, uint16_t>::type;
        using L8 = typename MaList<AS, uint8_t>::type;
        // the 16 mark bits of face f in its word
        auto mw = [&](int f) { return w.mfm + (f >> 1); };
        auto msh = [](int f) { return (uint32_t)(f & 1) << 4; };
        const int wv = tid >> 6, lane = tid & 63;
        asptr<AS_LDS, int> cand = w.mcand;
        asptr<AS_LDS, int> crow = w.mcand + MA_W * MA_CI;
        asptr<AS_LDS, int> misc = w.mcand + MA_W * MA_CI + MA_W;
        auto lvis = [&](int r) { return w.mlist + (r * 5 + 0) * MA_CAP; };
        auto lha = [&](int r) { return w.mlist + (r * 5 + 1) * MA_CAP; };
        auto lhb = [&](int r) { return w.mlist + (r * 5 + 2) * MA_CAP; };
        auto lhf = [&](int r) { return w.mlist + (r * 5 + 3) * MA_CAP; };
        auto lhn = [&](int r) { return w.mlist + (r * 5 + 4) * MA_CAP; };
        auto wsync = [&]() {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
        };
        if (tid < MA_W) w.msk[tid] = 0ull;
        __syncthreads();
        for (int it = 1;; ++it) {
            if constexpr (PPL * T <= 8192) {
                // A: clear the per-face marks and best keys
                for (int f = tid; f < ftop; f += T) w.mfmax[f] = 0u;
                for (int q = tid; q < (ftop + 1) >> 1; q += T) w.mfm[q] = 0u;
                if (tid < MA_W) crow[tid] = 0;
                __syncthreads();
                // B: best conflict key per face (quantised distance, lowest index)
#pragma unroll
                for (int k = 0; k < PPL; ++k)
                    if (conf[k] >= 0) {
                        const uint32_t key = (fkey(dist[k]) & ~IMASK) | (IMASK - (uint32_t)(k * T + tid));
                        __hip_atomic_fetch_max(&w.mfmax[conf[k]], key, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                    }
                __syncthreads();
                // C: candidates - the best face of each wave's faces (face
                // ids interleaved over the waves, so every wave has faces)
                uint64_t bk = 0;
                for (int f = wv + NW * lane; f < ftop; f += NW * 64) {
                    const uint32_t k32 = w.mfmax[f];
                    const uint64_t v = ((uint64_t)k32 << 32) | (uint32_t)f;
                    if (k32 && v > bk) bk = v;
                }
                bk = wave_max_u64(bk);
                if (lane == 0) w.msk[wv] = bk;
                __syncthreads();
            } else {
                // A: clear the per-face marks. B (same interval: it reads no
                // shared state): the same candidates without per-face keys -
                // the best conflict point over the faces of each share (face
                // f is in share f % NW), as (key << 32 | face): per lane and
                // share, then over the wave, then over the waves
                for (int q = tid; q < (ftop + 1) >> 1; q += T) w.mfm[q] = 0u;
                if (tid < MA_W) crow[tid] = 0;
                uint64_t sk[NW];
#pragma unroll
                for (int q = 0; q < NW; ++q) sk[q] = 0;
#pragma unroll
                for (int k = 0; k < PPL; ++k)
                    if (conf[k] >= 0) {
                        const uint32_t key = (fkey(dist[k]) & ~IMASK) | (IMASK - (uint32_t)(k * T + tid));
                        const uint64_t v = ((uint64_t)key << 32) | (uint32_t)conf[k];
                        const int sh = conf[k] % NW;
#pragma unroll
                        for (int q = 0; q < NW; ++q)
                            if (q == sh && v > sk[q]) sk[q] = v;
                    }
#pragma unroll
                for (int q = 0; q < NW; ++q) {
                    sk[q] = wave_max_u64(sk[q]);
                    if (lane == 0 && sk[q])
                        __hip_atomic_fetch_max(&w.msk[q], (unsigned long long)sk[q], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                __syncthreads();
            }
            uint64_t rk[NW];
            int ncand = 0;
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                rk[q] = w.msk[q];
                ncand += rk[q] != 0;
            }
            if (ncand == 0) break;
            // the owners of the candidate points publish them at their rank,
            // in descending key order (keys differ: they hold distinct face ids)
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                if (!rk[q]) continue;
                const int ip = (int)(IMASK - ((uint32_t)(rk[q] >> 32) & IMASK));
                if (tid == ip % T) {
                    int rank = 0;
#pragma unroll
                    for (int o = 0; o < NW; ++o) rank += rk[o] > rk[q];
                    const int slot = ip / T;
                    cand[rank * MA_CI + 0] = (int)(uint32_t)(rk[q] >> 32);
                    cand[rank * MA_CI + 1] = (int)(uint32_t)rk[q];
                    cand[rank * MA_CI + 2] = ip;
                    cand[rank * MA_CI + 3] = pick(px, slot);
                    cand[rank * MA_CI + 4] = pick(py, slot);
                    cand[rank * MA_CI + 5] = pick(pz, slot);
                }
            }
            __syncthreads();
            // the share maxima were read before the barrier above: clear them
            // for the next round (several barriers before its first update)
            if (tid < MA_W) w.msk[tid] = 0ull;
            // D: wave r finds candidate r's visible region and, on the way, its
            // horizon: every visible face tests its three neighbours at once
            // (visit marks, planes and back links loaded together); the first
            // visitor of a visible neighbour queues it, and every neighbour
            // not visible gives a horizon edge (its test is deterministic, so
            // every lane that reaches it agrees)
            if (wv < ncand) {
                const int r = wv;
                const int f0 = cand[r * MA_CI + 1];
                const int qx = cand[r * MA_CI + 3], qy = cand[r * MA_CI + 4], qz = cand[r * MA_CI + 5];
                const float fqx = (float)qx, fqy = (float)qy, fqz = (float)qz;
                L16 V = r == 0 ? (L16)w.vis : (L16)lvis(r);
                const int vcap = r == 0 ? FC : MA_CAP;
                L16 HA = r == 0 ? (L16)w.ha : (L16)lha(r);
                L16 HB = r == 0 ? (L16)w.hb : (L16)lhb(r);
                L16 HF = r == 0 ? (L16)w.hface : (L16)lhf(r);
                L8 HE = r == 0 ? (L8)w.hedge : (L8)(w.mhedge + r * MA_CAP);
                const int hcap = r == 0 ? VC : MA_CAP;
                const uint32_t vb = 1u << r, sbit = 1u << (8 + r);
                if (lane == 0) {
                    __hip_atomic_fetch_or(mw(f0), (vb | sbit) << msh(f0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    V[0] = (uint16_t)f0;
                }
                wsync();
                int head = 0, nv = 1, nh = 0;
                bool over = false;
                while (head < nv) {
                    const int cnt = min(nv - head, 64);
                    const bool act = lane < cnt;
                    int f = 0;
                    u4v fvl = {0, 0, 0, 0}, nb = {0, 0, 0, 0};
                    if (act) {
                        f = V[head + lane];
                        fvl = w.fv[f];
                        nb = w.fn[f];
                    }
                    const int gg[3] = {nb.x, nb.y, nb.z};
                    const int vv[3] = {fvl.x, fvl.y, fvl.z};
                    uint32_t old[3] = {vb, vb, vb};
                    bool vis[3] = {false, false, false};
                    int back[3] = {0, 0, 0};
                    if (act) {
                        f4v pl[3];
                        u4v gn[3];
#pragma unroll
                        for (int e = 0; e < 3; ++e) {
                            old[e] = __hip_atomic_fetch_or(mw(gg[e]), vb << msh(gg[e]), __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_WORKGROUP) >> msh(gg[e]);
                            pl[e] = w.fpl[gg[e]];
                            gn[e] = w.fn[gg[e]];
                        }
#pragma unroll
                        for (int e = 0; e < 3; ++e) {
                            // face_above on the loaded plane
                            const float t = fmaf(pl[e].x, fqx, fmaf(pl[e].y, fqy, fmaf(pl[e].z, fqz, pl[e].w)));
                            bool above = t > HULL_FILTER_EPS;
                            if (!above && t >= -HULL_FILTER_EPS) {
                                const u4v gv = w.fv[gg[e]];
                                above = orient_exact(w.vq[gv.x], w.vq[gv.y], w.vq[gv.z], qx, qy, qz) > 0.0;
                            }
                            vis[e] = above;
                            back[e] = (gn[e].x == f) ? 0 : ((gn[e].y == f) ? 1 : 2);
                        }
                    }
                    head += cnt;
                    int add = 0;
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const bool ap = act && vis[e] && !(old[e] & vb);
                        if (ap)
                            __hip_atomic_fetch_or(mw(gg[e]), sbit << msh(gg[e]), __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
                        const uint64_t m = ballot(ap);
                        const int pos = nv + add + mask_rank(m);
                        if (ap) {
                            if (pos < vcap) V[pos] = (uint16_t)gg[e];
                            else over = true;
                        }
                        add += __popcll(m);
                    }
#pragma unroll
                    for (int e = 0; e < 3; ++e) {
                        const bool h = act && !vis[e];
                        const uint64_t m = ballot(h);
                        const int pos = nh + mask_rank(m);
                        if (h && pos < hcap) {
                            HA[pos] = (uint16_t)vv[e];
                            HB[pos] = (uint16_t)vv[e == 2 ? 0 : e + 1];
                            HF[pos] = (uint16_t)gg[e];
                            HE[pos] = (uint8_t)back[e];
                        }
                        nh += __popcll(m);
                    }
                    nv += add;
                    wsync();
                    if (nv > vcap) { over = true; nv = vcap; break; }
                }
                over = ballot(over) != 0;
                if (lane == 0) {
                    cand[r * MA_CI + 6] = nv;
                    cand[r * MA_CI + 7] = nh;
                    cand[r * MA_CI + 8] = (!over && nh >= 3 && nh <= hcap) ? 1 : 0;
                }
            }
            __syncthreads();
            // E: conflicts - candidate r's visible faces against the others' visited faces
            for (int f = tid; f < ftop; f += T) {
                const uint32_t m = (w.mfm[f >> 1] >> msh(f)) & 0xFFFFu;
                uint32_t sv = (m >> 8) & 0xFFu;
                const uint32_t vis_by = m & 0xFFu;
                while (sv) {
                    const int r = __ffs(sv) - 1;
                    sv &= sv - 1;
                    const uint32_t o = vis_by & ~(1u << r);
                    if (o) __hip_atomic_fetch_or(&crow[r], (int)o, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
            }
            __syncthreads();
            // F: wave 0 accepts greedily by rank and assigns vertex and face
            // ids: lane r loads candidate r's sizes and conflict row, the rank
            // walk runs on scalar broadcasts, then each accepted lane writes
            // its cone's extra face ids (from the free list, then the top)
            if (wv == 0) {
                const int nc = __builtin_amdgcn_readfirstlane(ncand);
                int c_nv = 0, c_nh = 0, c_ok = 0, c_row = 0;
                if (lane < nc) {
                    c_nv = cand[lane * MA_CI + 6];
                    c_nh = cand[lane * MA_CI + 7];
                    c_ok = cand[lane * MA_CI + 8];
                    c_row = crow[lane];
                }
                uint32_t acc = 0, blocked = 0;  // blocked: visited the visible faces of an accepted one
                int nacc = 0, st = HULL_OK;
                int nf = __builtin_amdgcn_readfirstlane(nfree), ft = __builtin_amdgcn_readfirstlane(ftop);
                const int nv0 = __builtin_amdgcn_readfirstlane(nvert);
                int my_vid = 0, my_free = 0, my_ff = 0, my_top = 0;
                for (int r = 0; r < nc; ++r) {
                    const int nv = read_lane(c_nv, r), nh = read_lane(c_nh, r);
                    const uint32_t row = (uint32_t)read_lane(c_row, r);
                    if (!read_lane(c_ok, r)) {
                        if (r == 0) { st = HULL_OVERFLOW; break; }
                        continue;
                    }
                    if ((row & acc) || ((blocked >> r) & 1u)) continue;
                    const int extra = nh > nv ? nh - nv : 0;
                    const int from_free = min(extra, nf);
                    const int from_top = extra - from_free;
                    if (nv0 + nacc >= VC || ft + from_top > FC) {
                        if (r == 0) { st = HULL_OVERFLOW; break; }
                        continue;
                    }
                    if (lane == r) {
                        my_vid = nv0 + nacc;
                        my_free = nf;
                        my_ff = from_free;
                        my_top = ft;
                    }
                    nf -= from_free;
                    ft += from_top;
                    ++nacc;
                    acc |= 1u << r;
                    blocked |= row;
                }
                // the surplus visible faces of the accepted cones go back to the free list
                int my_push = 0;
                for (int r = 0; r < nc; ++r)
                    if ((acc >> r) & 1u) {
                        const int nv = read_lane(c_nv, r), nh = read_lane(c_nh, r);
                        if (lane == r) my_push = nf;
                        nf += nv > nh ? nv - nh : 0;
                    }
                if (lane < nc && ((acc >> lane) & 1u)) {
                    L16 HN = lane == 0 ? (L16)w.hnew : (L16)lhn(lane);
                    const int extra = c_nh > c_nv ? c_nh - c_nv : 0;
                    for (int k = 0; k < extra; ++k)
                        HN[c_nv + k] = k < my_ff ? w.ffree[my_free - 1 - k] : (uint16_t)(my_top + k - my_ff);
                    cand[lane * MA_CI + 9] = my_vid;
                    cand[lane * MA_CI + 10] = my_push;
                }
                if (lane == 0) {
                    misc[0] = (int)acc;
                    misc[1] = nacc;
                    misc[2] = ft;
                    misc[3] = nf;
                    misc[4] = st;
                }
            }
            __syncthreads();
            const uint32_t acc = (uint32_t)misc[0];
            if (misc[4] != HULL_OK) { res.status = HULL_OVERFLOW; break; }
            // G: wave r builds the cone of accepted candidate r
            if (wv < ncand && ((acc >> wv) & 1u)) {
                const int r = wv;
                const int nv = cand[r * MA_CI + 6], nh = cand[r * MA_CI + 7];
                const int vid = cand[r * MA_CI + 9], push = cand[r * MA_CI + 10];
                const int qx = cand[r * MA_CI + 3], qy = cand[r * MA_CI + 4], qz = cand[r * MA_CI + 5];
                L16 V = r == 0 ? (L16)w.vis : (L16)lvis(r);
                L16 HA = r == 0 ? (L16)w.ha : (L16)lha(r);
                L16 HB = r == 0 ? (L16)w.hb : (L16)lhb(r);
                L16 HF = r == 0 ? (L16)w.hface : (L16)lhf(r);
                L8 HE = r == 0 ? (L8)w.hedge : (L8)(w.mhedge + r * MA_CAP);
                L16 HN = r == 0 ? (L16)w.hnew : (L16)lhn(r);
                if (lane == 0) {
                    w.vq[vid] = mk_s4(qx, qy, qz, 0);
                    w.vpid[vid] = cand[r * MA_CI + 2];
                }
                const int mx = max(nh, nv);
                for (int j = lane; j < mx; j += 64) {
                    if (j < nv) w.mfcone[V[j]] = (uint8_t)r;
                    if (j < nh) {
                        if (j < nv) HN[j] = V[j];
                    } else {
                        const int f = V[j];
                        w.ffree[push + (j - nh)] = (uint16_t)f;
                        u4v t = w.fn[f];
                        t.w = 0;
                        w.fn[f] = t;
                        u4v u = w.fv[f];
                        u.w = (uint16_t)it;
                        w.fv[f] = u;
                    }
                }
                wsync();
                const s4v sp = mk_s4(qx, qy, qz, 0);
                for (int j = lane; j < nh; j += 64) {
                    const int a = HA[j], c = HB[j];
                    int nx = 0, pv = 0;  // cone faces across (c, apex) and (apex, a)
                    for (int k = 0; k < nh; ++k) {
                        if (HA[k] == c) nx = k;
                        if (HB[k] == a) pv = k;
                    }
                    const int id = HN[j];
                    w.fv[id] = mk_u4(a, c, vid, (uint16_t)it);
                    w.fn[id] = mk_u4(HF[j], HN[nx], HN[pv], 1);
                    const f4v pl = face_set_plane(w, id, w.vq[a], w.vq[c], sp);
                    if (nh <= MA_CPL) w.mcpl[r * MA_CPL + j] = pl;
                }
                wsync();
                for (int j = lane; j < nh; j += 64) {
                    asptr<AS, uint16_t> slotp = ((asptr<AS, uint16_t>)&w.fn[HF[j]]) + HE[j];
                    *slotp = HN[j];
                }
            }
            nvert += misc[1];
            ftop = misc[2];
            nfree = misc[3];
            __syncthreads();
            // H: accepted apexes leave the conflict sets; points of replaced
            // faces move to the cone that replaced them
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                if (q >= ncand || !((acc >> q) & 1u)) continue;
                const int ip = cand[q * MA_CI + 2];
                const int own = ip % T, slot = ip / T;
#pragma unroll
                for (int k = 0; k < PPL; ++k) {
                    const int m = -(int)(tid == own && k == slot);
                    conf[k] = (conf[k] & ~m) | (-2 & m);
                }
            }
            // the lane's points whose face was replaced this round, as a bit
            // mask: the wave then loops as often as its busiest lane needs,
            // not once per point slot that any lane needs
            uint32_t need = 0;
#pragma unroll
            for (int k = 0; k < PPL; ++k) {
                const int c = conf[k];
                if (c >= 0 && w.fv[c].w == (uint16_t)it) need |= 1u << k;
            }
            while (need) {
                const int k = __ffs(need) - 1;
                need &= need - 1;
                const int c = pick(conf, k);
                const int qx = pick(px, k), qy = pick(py, k), qz = pick(pz, k);
                const int r = w.mfcone[c];
                const int nh = cand[r * MA_CI + 7];
                L16 HN = r == 0 ? (L16)w.hnew : (L16)lhn(r);
                // the cone's planes in cone order when it kept them (the same
                // test as face_above, on the same plane values), else by id
                const bool cp = nh <= MA_CPL;
                const asptr<AS_LDS, f4v> CP = w.mcpl + r * MA_CPL;
                const float fx = (float)qx, fy = (float)qy, fz = (float)qz;
                int bj = -1;
                float bd = 0.0f;
                for (int j = 0; j < nh; ++j) {
                    const f4v pl = cp ? CP[j] : w.fpl[HN[j]];
                    const float t = fmaf(pl.x, fx, fmaf(pl.y, fy, fmaf(pl.z, fz, pl.w)));
                    bool above = t > HULL_FILTER_EPS;
                    if (!above && t >= -HULL_FILTER_EPS) {
                        const u4v v = w.fv[HN[j]];
                        above = orient_exact(w.vq[v.x], w.vq[v.y], w.vq[v.z], qx, qy, qz) > 0.0;
                    }
                    if (above && (bj < 0 || t > bd)) {
                        bj = j;
                        bd = t;
                    }
                }
                const int best = bj >= 0 ? (int)HN[bj] : -1;
#pragma unroll
                for (int kk = 0; kk < PPL; ++kk)
                    if (kk == k) {
                        conf[kk] = best;
                        dist[kk] = bd;
                    }
            }
        }
        res.nvert = nvert;
        res.ftop = ftop;
        __syncthreads();
        return res;
    }
    for (int it = 1;; ++it) {
        // ---- select the farthest conflict point (ties: lowest index) -------
        uint64_t sk = 0;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
            if (conf[k] >= 0) {
                uint64_t kk = ((uint64_t)fkey(dist[k]) << 32) |
                              (uint64_t)(0xFFFFFFFFu - (uint32_t)(k * T + tid));
                sk = kk > sk ? kk : sk;
            }
        }
        sk = block_max_u64<NW>(sk, sb);
        if (sk == 0) break;
        const int ip = (int)(0xFFFFFFFFu - (uint32_t)sk);
        if (nvert >= VC) { res.status = HULL_OVERFLOW; break; }
        const int vid = nvert++;
        int qx, qy, qz, f0 = 0;
        {
            const int own = ip % T, slot = ip / T;
            if (NW == 1) {
                qx = read_lane(pick(px, slot), own);
                qy = read_lane(pick(py, slot), own);
                qz = read_lane(pick(pz, slot), own);
                f0 = read_lane(pick(conf, slot), own);
            } else {
                if (tid == own) {
                    sb->bc[0] = pick(px, slot); sb->bc[1] = pick(py, slot); sb->bc[2] = pick(pz, slot);
                    sb->bc[3] = pick(conf, slot);
                }
                __syncthreads();
                qx = sb->bc[0]; qy = sb->bc[1]; qz = sb->bc[2]; f0 = sb->bc[3];
            }
            // the apex leaves the conflict sets
#pragma unroll
            for (int k = 0; k < PPL; ++k) {
                const int m = -(int)(tid == own && k == slot);
                conf[k] = (conf[k] & ~m) | (-2 & m);
            }
        }
        if (tid == 0) {
            w.vq[vid] = mk_s4(qx, qy, qz, 0);
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
            const int cnt = min(nvis - head, T);
            if (tid < cnt) {
                const int f = w.vis[head + tid];
                const u4v nb = w.fn[f];
                const int g3[3] = {nb.x, nb.y, nb.z};
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    const int g = g3[e];
                    if (atom_exch(&w.fstamp[g], it) != it) {
                        float t;
                        const bool v = face_above(w, g, qx, qy, qz, &t);
                        w.fvis[g] = v ? 1 : 0;
                        if (v) {
                            const int s = atom_add(&w.ctr[0], 1);
                            w.vis[s] = (uint16_t)g;
                        }
                    }
                }
            }
            head += cnt;
            __syncthreads();
            nvis = w.ctr[0];
            if (NW == 1) nvis = uniform(nvis);
            // waves run independently: nobody may append to ctr[0] in the next
            // round before every wave has read this round's count
            if (NW > 1) __syncthreads();
        }

        // ---- horizon edges (visible face -> non-visible neighbour) ---------
        for (int j = tid; j < nvis; j += T) {
            const int f = w.vis[j];
            const u4v fvl = w.fv[f];
            const u4v nb = w.fn[f];
            const int vv[3] = {fvl.x, fvl.y, fvl.z};
            const int gg[3] = {nb.x, nb.y, nb.z};
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const int g = gg[e];
                if (!w.fvis[g]) {
                    const u4v gn = w.fn[g];
                    const int ge = (gn.x == f) ? 0 : ((gn.y == f) ? 1 : 2);
                    const int s = atom_add(&w.ctr[1], 1);
                    if (s < VC) {
                        w.ha[s] = (uint16_t)vv[e];
                        w.hb[s] = (uint16_t)vv[e == 2 ? 0 : e + 1];
                        w.hface[s] = (uint16_t)g;
                        w.hedge[s] = (uint8_t)ge;
                    }
                }
            }
        }
        __syncthreads();
        int nhz = w.ctr[1];
        if (NW == 1) nhz = uniform(nhz);
        if (nhz > VC || nhz < 3) { res.status = HULL_OVERFLOW; break; }

        // ---- face ids for the cone: reuse visible ids, then free list, then top
        const int extra = nhz > nvis ? nhz - nvis : 0;
        const int from_free = min(extra, nfree);
        const int from_top = extra - from_free;
        if (ftop + from_top > FC) { res.status = HULL_OVERFLOW; break; }
        const int mx = max(nhz, nvis);
        for (int j = tid; j < mx; j += T) {
            if (j < nhz) {
                int id;
                if (j < nvis) id = w.vis[j];
                else if (j - nvis < from_free) id = w.ffree[nfree - 1 - (j - nvis)];
                else id = ftop + (j - nvis - from_free);
                w.hnew[j] = (uint16_t)id;
                w.smap[w.ha[j]] = (uint16_t)j;
                w.emap[w.hb[j]] = (uint16_t)j;
            } else {
                // surplus visible faces die and go back to the free list
                const int f = w.vis[j];
                w.ffree[nfree - from_free + (j - nhz)] = (uint16_t)f;
                u4v t = w.fn[f];
                t.w = 0;
                w.fn[f] = t;
                u4v u = w.fv[f];
                u.w = (uint16_t)it;
                w.fv[f] = u;
            }
        }
        nfree = nfree - from_free + (nvis > nhz ? nvis - nhz : 0);
        ftop += from_top;
        __syncthreads();

        // ---- stitch the cone ---------------------------------------------
        const s4v sp = mk_s4(qx, qy, qz, 0);
        for (int j = tid; j < nhz; j += T) {
            const int id = w.hnew[j];
            const int a = w.ha[j], c = w.hb[j];
            const int g = w.hface[j];
            w.fv[id] = mk_u4(a, c, vid, (uint16_t)it);
            w.fn[id] = mk_u4(g, w.hnew[w.smap[c]], w.hnew[w.emap[a]], 1);
            w.fstamp[id] = it;
            w.fvis[id] = 0;
            face_set_plane(w, id, w.vq[a], w.vq[c], sp);
        }
        __syncthreads();
        for (int j = tid; j < nhz; j += T) {
            const int g = w.hface[j];
            const int ge = w.hedge[j];
            asptr<AS, uint16_t> slotp = ((asptr<AS, uint16_t>)&w.fn[g]) + ge;
            *slotp = w.hnew[j];
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
    __syncthreads();
    return res;
}

// After block_quickhull: mark referenced vertices and give them output
// indices in ascending local-point order. Returns the output vertex count.
template <int AS, int NW>
__device__ __forceinline__ int hull_rank_vertices(const HullWS<AS> w, HullBlockLDS* sb, const HullResult& r, int n) {
    constexpr int T = 64 * NW;
    const int tid = threadIdx.x;
    const int nw = (n + 31) >> 5;
    for (int v = tid; v < r.nvert; v += T) w.vref[v] = 0;
    for (int i = tid; i < nw; i += T) w.bm[i] = 0;
    __syncthreads();
    for (int f = tid; f < r.ftop; f += T) {
        const u4v nb = w.fn[f];
        if (nb.w) {
            const u4v v = w.fv[f];
            w.vref[v.x] = 1;
            w.vref[v.y] = 1;
            w.vref[v.z] = 1;
        }
    }
    __syncthreads();
    for (int v = tid; v < r.nvert; v += T)
        if (w.vref[v]) {
            const int p = w.vpid[v];
            atom_or(&w.bm[p >> 5], 1u << (p & 31));
        }
    __syncthreads();
    // exclusive prefix popcount over the bitmap words, 64 words per wave step
    int carry = 0;
    for (int b = 0; b < nw; b += 64) {
        int c = 0;
        if (tid < 64 && b + tid < nw) c = __popc(w.bm[b + tid]);
        const int incl = wave_incl_scan(c);
        if (tid < 64 && b + tid < nw) w.wpre[b + tid] = carry + incl - c;
        carry += __shfl(incl, 63);
    }
    if (NW > 1) {
        if (tid == 0) sb->bc[7] = carry;
        __syncthreads();
        carry = sb->bc[7];
        __syncthreads();
    } else {
        __syncthreads();
    }
    for (int v = tid; v < r.nvert; v += T)
        if (w.vref[v]) {
            const int p = w.vpid[v];
            w.vout[v] = (uint16_t)(w.wpre[p >> 5] + __popc(w.bm[p >> 5] & ((1u << (p & 31)) - 1u)));
        }
    __syncthreads();
    return carry;
}

template <int AS, int NW>
__device__ __forceinline__ int hull_count_faces(const HullWS<AS> w, HullBlockLDS* sb, const HullResult& r) {
    int c = 0;
    for (int f = threadIdx.x; f < r.ftop; f += 64 * NW) c += w.fn[f].w ? 1 : 0;
    return block_sum_i<NW>(c, sb);
}

// Write triangles (output vertex indices) in face-id order; returns count.
template <int AS, int NW>
__device__ __forceinline__ int hull_write_tris(const HullWS<AS> w, HullBlockLDS* sb, const HullResult& r,
                                      int* __restrict__ out_tris) {
    const int tid = threadIdx.x;
    int base = 0;
    for (int b = 0; b < r.ftop; b += 64 * NW) {
        const int f = b + tid;
        const bool alive = f < r.ftop && w.fn[f].w;
        const uint64_t m = ballot(alive);
        int wbase = 0;
        if (NW > 1) {
            const int wv = tid >> 6;
            if (lane_id() == 0) sb->bc[wv] = __popcll(m);
            __syncthreads();
            for (int i = 0; i < wv; ++i) wbase += sb->bc[i];
        }
        if (alive) {
            const int t = base + wbase + mask_rank(m);
            const u4v v = w.fv[f];
            out_tris[3 * t + 0] = w.vout[v.x];
            out_tris[3 * t + 1] = w.vout[v.y];
            out_tris[3 * t + 2] = w.vout[v.z];
        }
        if (NW > 1) {
            int tot = 0;
            for (int i = 0; i < NW; ++i) tot += sb->bc[i];
            base += tot;
            __syncthreads();
        } else {
            base += __popcll(m);
        }
    }
    return base;
}

// FP64 divergence-theorem volume over live faces, original fp32 coordinates
// of the point with local id p fetched by `pos(p, xyz)`. Anchored at the first
// vertex of the lowest live face to avoid cancellation. Block-uniform result.
template <int AS, int NW, class PosFn>
__device__ __forceinline__ double hull_volume_f64(const HullWS<AS> w, HullBlockLDS* sb, const HullResult& r, PosFn pos) {
    int f0 = 0;
    for (int b = 0; b < r.ftop; b += 64) {
        const uint64_t m = ballot(b + lane_id() < r.ftop && w.fn[b + lane_id()].w);
        if (m) { f0 = b + __builtin_ffsll((long long)m) - 1; break; }
    }
    float a[3];
    pos(w.vpid[w.fv[f0].x], a);
    const double ax = a[0], ay = a[1], az = a[2];
    double acc = 0.0;
    for (int f = threadIdx.x; f < r.ftop; f += 64 * NW) {
        if (!w.fn[f].w) continue;
        const u4v v = w.fv[f];
        float p0[3], p1[3], p2[3];
        pos(w.vpid[v.x], p0);
        pos(w.vpid[v.y], p1);
        pos(w.vpid[v.z], p2);
        const double x0 = p0[0] - ax, y0 = p0[1] - ay, z0 = p0[2] - az;
        const double x1 = p1[0] - ax, y1 = p1[1] - ay, z1 = p1[2] - az;
        const double x2 = p2[0] - ax, y2 = p2[1] - ay, z2 = p2[2] - az;
        acc += x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
    }
    return block_sum_d<NW>(acc, sb) / 6.0;
}

// ---------------------------------------------------------------------------
// Canonical form of a finished hull
// ---------------------------------------------------------------------------
// Where grid points are coplanar, the faces the quickhull leaves (and the
// non-extreme points it keeps as vertices, inside a facet or on an edge)
// depend on the insertion order. The canonical form depends on the point set
// alone:
//   * facets = maximal planar faces (adjacent faces merged when exactly
//     coplanar on the grid);
//   * corners = extreme points: facet boundary vertices where the boundary
//     turns (exact grid cross product); output vertices = the corners in
//     ascending local point order;
//   * each facet is the fan (c0, c_k, c_k+1), k = 1..m-2, of its corners in
//     counter-clockwise order (seen from outside) starting at its corner of
//     lowest output index c0; facets are listed by (c0, c1);
//   * volume: FP64 divergence terms of the fan triangles (original fp32
//     coordinates, anchored at output vertex 0, no fused multiply-add),
//     summed in a fixed order: the triangle list in canonical order is cut
//     into 64 contiguous chunks of ceil(nt / 64), each summed in order, then
//     the 64 partial sums in order - so the value does not depend on how the
//     hull was built.
// Scratch reused after block_quickhull, per face f: fvis = coplanar-edge bits
// (bit 3: facet leader, bit 4: live face), fstamp = facet label (lowest face id of the facet),
// then a leader's triangle offset; fpl.xyz = successor of each boundary edge
// (3*face + slot, -1 inside a facet), fpl.w = a leader's edge ending at c0;
// fv.w = a leader's boundary face, then its corner count; ffree = a leader's
// c1; vis = next leader in the list of facets sharing c0; fn (no longer
// needed once the boundaries are traced) = the triangles' volume terms. Per
// vertex: vref = corner flag, vout = output index, cv = list head and
// triangle offset per c0.
struct HullCanon {
    int nv, nt;
    double vol;
};

template <int AS> __device__ __forceinline__ int canon_ld(asptr<AS, int> p) {
    if constexpr (AS == AS_LDS) return *p;
    else return __hip_atomic_load((int*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int AS> __device__ __forceinline__ int canon_exch(asptr<AS, int> p, int v) {
    if constexpr (AS == AS_LDS) return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else return __hip_atomic_exchange((int*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int fpl_geti(f4v v, int j) {
    return __float_as_int(j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w)));
}
__device__ __forceinline__ int u4_get(u4v v, int j) { return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w)); }

template <int NW> __device__ __forceinline__ bool block_any(bool p, HullBlockLDS* sb) {
    if (NW == 1) return ballot(p) != 0ull;
    return block_sum_i<NW>(p ? 1 : 0, sb) != 0;
}

// Builds the canonical form's vertex set (vref/vout), per-facet fans and
// offsets, and its volume. Block-uniform result; nt = 0 on a topology fault.
template <int AS, int NW, class PosFn>
__device__ HullCanon hull_canon(const HullWS<AS> w, HullBlockLDS* sb, const HullResult& r, int n, PosFn pos) {
    constexpr int T = 64 * NW;
    const int tid = threadIdx.x;
    const int ftop = r.ftop, FC = w.FC;
    HullCanon out{0, 0, 0.0};
    auto alive = [&](int f) { return w.fn[f].w != 0; };
    auto opp_of = [&](int g, int a, int b) {
        const u4v gv = w.fv[g];
        return (gv.x != a && gv.x != b) ? (int)gv.x : ((gv.y != a && gv.y != b) ? (int)gv.y : (int)gv.z);
    };
    // A: exact coplanarity of each face with its neighbour across each edge
    for (int f = tid; f < ftop; f += T) {
        w.fstamp[f] = f;
        if (!alive(f)) { w.fvis[f] = 0; continue; }
        const u4v v = w.fv[f], nb = w.fn[f];
        const s4v qa = w.vq[v.x], qb = w.vq[v.y], qc = w.vq[v.z];
        int bits = 0;
        for (int e = 0; e < 3; ++e) {
            const int a = u4_get(v, e), b = u4_get(v, e == 2 ? 0 : e + 1);
            const s4v q = w.vq[opp_of(u4_get(nb, e), a, b)];
            if (orient_exact(qa, qb, qc, q.x, q.y, q.z) == 0.0) bits |= 1 << e;
        }
        w.fvis[f] = (uint8_t)(bits | 16);
    }
    for (int v = tid; v < r.nvert; v += T) w.vref[v] = 0;
    __syncthreads();
    // B: facet labels = lowest face id over coplanar adjacency (min propagation
    // with pointer jumping; a round that changes nothing ends it)
    for (int round = 0; round < FC; ++round) {
        bool changed = false;
        for (int f = tid; f < ftop; f += T) {
            if (!alive(f)) continue;
            const int bits = w.fvis[f] & 7;
            if (!bits) continue;
            const u4v nb = w.fn[f];
            int l = w.fstamp[f];
            const int l0 = l;
            for (int e = 0; e < 3; ++e)
                if (bits & (1 << e)) l = min(l, w.fstamp[u4_get(nb, e)]);
            l = min(l, w.fstamp[l]);
            if (l < l0) { w.fstamp[f] = l; changed = true; }
        }
        if (!block_any<NW>(changed, sb)) break;
        __syncthreads();
    }
    __syncthreads();
    // C: successor of every boundary edge inside its facet, corner flags, a
    // boundary face per facet (any one: the walks cover the whole cycle)
    int fault = 0;
    for (int f = tid; f < ftop; f += T) {
        if (!alive(f)) continue;
        const int bits = w.fvis[f] & 7;
        const u4v v = w.fv[f];
        f4v sc = w.fpl[f];
        int succ[3] = {-1, -1, -1};
        for (int e = 0; e < 3; ++e) {
            if (bits & (1 << e)) continue;
            // rotate around the edge's end vertex b, staying inside the facet
            const int b = u4_get(v, e == 2 ? 0 : e + 1);
            int h = f, j = e == 2 ? 0 : e + 1;
            int guard = 0;
            while (((w.fvis[h] >> j) & 1) && guard < FC) {
                const int g = u4_get(w.fn[h], j);
                const u4v gv = w.fv[g];
                j = gv.x == b ? 0 : (gv.y == b ? 1 : 2);
                h = g;
                ++guard;
            }
            if (guard >= FC) { fault = 1; break; }
            succ[e] = 3 * h + j;
            const int a = u4_get(v, e), c = u4_get(w.fv[h], j == 2 ? 0 : j + 1);
            const s4v qa = w.vq[a], qb = w.vq[b], qc = w.vq[c];
            const int ux = qb.x - qa.x, uy = qb.y - qa.y, uz = qb.z - qa.z;
            const int vx = qc.x - qb.x, vy = qc.y - qb.y, vz = qc.z - qb.z;
            if (uy * vz - uz * vy != 0 || uz * vx - ux * vz != 0 || ux * vy - uy * vx != 0) w.vref[b] = 1;
            asptr<AS, uint16_t> fw = (asptr<AS, uint16_t>)(w.fv + w.fstamp[f]);
            fw[3] = (uint16_t)f;
        }
        sc.x = __int_as_float(succ[0]);
        sc.y = __int_as_float(succ[1]);
        sc.z = __int_as_float(succ[2]);
        w.fpl[f] = sc;
    }
    if (block_any<NW>(fault != 0, sb)) {
        out.nt = -1;
        return out;
    }
    __syncthreads();
    // D: output indices of the corners, ascending local point order
    {
        const int nw = (n + 31) >> 5;
        for (int i = tid; i < nw; i += T) w.bm[i] = 0;
        for (int v = tid; v < r.nvert; v += T) { w.cv[2 * v] = -1; w.cv[2 * v + 1] = 0; }
        __syncthreads();
        for (int v = tid; v < r.nvert; v += T)
            if (w.vref[v]) {
                const int p = w.vpid[v];
                atom_or(&w.bm[p >> 5], 1u << (p & 31));
            }
        __syncthreads();
        int carry = 0;
        for (int b = 0; b < nw; b += 64) {
            int c = 0;
            if (tid < 64 && b + tid < nw) c = __popc(w.bm[b + tid]);
            const int incl = wave_incl_scan(c);
            if (tid < 64 && b + tid < nw) w.wpre[b + tid] = carry + incl - c;
            carry += __shfl(incl, 63);
        }
        if (tid == 0) sb->bc[7] = carry;
        __syncthreads();
        out.nv = sb->bc[7];
        for (int v = tid; v < r.nvert; v += T)
            if (w.vref[v]) {
                const int p = w.vpid[v];
                const int o = w.wpre[p >> 5] + __popc(w.bm[p >> 5] & ((1u << (p & 31)) - 1u));
                w.vout[v] = (uint16_t)o;
                if (o == 0) sb->bc[6] = v;
            }
        __syncthreads();
    }
    if (out.nv < 4) return out;
    float an[3];
    pos(w.vpid[sb->bc[6]], an);
    const double ax = an[0], ay = an[1], az = an[2];
    auto succ_of = [&](int e) { return fpl_geti(w.fpl[e / 3], e % 3); };
    auto end_of = [&](int e) { const int h = e / 3, j = e % 3; return u4_get(w.fv[h], j == 2 ? 0 : j + 1); };
    // E: per facet (its leader's thread): corners, c0, c1
    auto live = [&](int f) { return (w.fvis[f] & 16) != 0; };
    for (int f = tid; f < ftop; f += T) {
        if (!live(f) || w.fstamp[f] != f) continue;
        const int h0 = u4_get(w.fv[f], 3);
        const int b0 = w.fvis[h0] & 7;
        const int j0 = !(b0 & 1) ? 0 : (!(b0 & 2) ? 1 : 2);
        const int start = 3 * h0 + j0;
        int m = 0, best = 0x7fffffff, ebest = start;
        int e = start, guard = 0;
        do {
            const int b = end_of(e);
            if (w.vref[b]) {
                ++m;
                if (w.vout[b] < best) { best = w.vout[b]; ebest = e; }
            }
            e = succ_of(e);
        } while (e != start && ++guard < 3 * FC);
        int c1 = -1;
        e = succ_of(ebest);
        guard = 0;
        while (e != ebest && ++guard < 3 * FC) {
            const int b = end_of(e);
            if (w.vref[b]) { c1 = w.vout[b]; break; }
            e = succ_of(e);
        }
        w.fvis[f] = (uint8_t)(w.fvis[f] | 8);
        f4v sc = w.fpl[f];
        sc.w = __int_as_float(ebest);
        w.fpl[f] = sc;
        asptr<AS, uint16_t> fw = (asptr<AS, uint16_t>)(w.fv + f);
        fw[3] = (uint16_t)m;
        w.ffree[f] = (uint16_t)(c1 < 0 ? 0 : c1);
        w.vis[f] = (uint16_t)canon_exch<AS>(w.cv + 2 * best, f);
    }
    __syncthreads();
    // F: triangle offsets: per c0 bucket (scan over output vertices), then by c1 inside it
    auto tris_of = [&](int L) { return max(u4_get(w.fv[L], 3) - 2, 0); };
    {
        int carry = 0;
        for (int b = 0; b < out.nv; b += T) {
            const int o = b + tid;
            int c = 0;
            if (o < out.nv)
                for (int L = canon_ld<AS>(w.cv + 2 * o), guard = 0; L >= 0 && L != 0xFFFF && guard < FC; ++guard) {
                    c += tris_of(L);
                    L = w.vis[L];
                }
            int incl = wave_incl_scan(c);
            int wbase = 0, tot = __shfl(incl, 63);
            if (NW > 1) {
                const int wv = tid >> 6;
                if (lane_id() == 63) sb->bc[wv] = incl;
                __syncthreads();
                tot = 0;
                for (int q = 0; q < NW; ++q) {
                    if (q < wv) wbase += sb->bc[q];
                    tot += sb->bc[q];
                }
                __syncthreads();
            }
            if (o < out.nv) w.cv[2 * o + 1] = carry + wbase + incl - c;
            carry += tot;
        }
        out.nt = carry;
    }
    __syncthreads();
    for (int f = tid; f < ftop; f += T) {
        if (!(w.fvis[f] & 8)) continue;
        const int c0 = w.vout[end_of(fpl_geti(w.fpl[f], 3))];
        const int c1 = w.ffree[f];
        int off = w.cv[2 * c0 + 1];
        for (int L = canon_ld<AS>(w.cv + 2 * c0), guard = 0; L >= 0 && L != 0xFFFF && guard < FC; ++guard) {
            if (L != f && w.ffree[L] < c1) off += tris_of(L);
            L = w.vis[L];
        }
        w.fstamp[f] = off;
    }
    __syncthreads();
    // V: volume terms at the triangles' canonical positions, fixed-order sum
    asptr<AS, double> term = (asptr<AS, double>)w.fn;
    for (int f = tid; f < ftop; f += T) {
#pragma clang fp contract(off)
        if (!(w.fvis[f] & 8)) continue;
        const int ebest = fpl_geti(w.fpl[f], 3);
        float p0[3];
        pos(w.vpid[end_of(ebest)], p0);
        const double x0 = (double)p0[0] - ax, y0 = (double)p0[1] - ay, z0 = (double)p0[2] - az;
        int t = w.fstamp[f];
        int prev = -1, e = succ_of(ebest), guard = 0;
        while (e != ebest && ++guard < 3 * FC) {
            const int b = end_of(e);
            if (w.vref[b]) {
                if (prev >= 0) {
                    float p1[3], p2[3];
                    pos(w.vpid[prev], p1);
                    pos(w.vpid[b], p2);
                    const double x1 = (double)p1[0] - ax, y1 = (double)p1[1] - ay, z1 = (double)p1[2] - az;
                    const double x2 = (double)p2[0] - ax, y2 = (double)p2[1] - ay, z2 = (double)p2[2] - az;
                    term[t++] = x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
                }
                prev = b;
            }
            e = succ_of(e);
        }
    }
    __syncthreads();
    if (tid < 64) {
        const int chunk = (out.nt + 63) / 64;
        const int b = min(out.nt, tid * chunk), e = min(out.nt, b + chunk);
        double part = 0.0;
        for (int t = b; t < e; ++t) part += term[t];
        double acc = 0.0;
        for (int l = 0; l < 64; ++l) acc += __shfl(part, l);
        if (tid == 0) sb->dred[0] = acc / 6.0;
    }
    __syncthreads();
    out.vol = sb->dred[0];
    __syncthreads();
    return out;
}

// Writes the canonical triangles (output vertex indices) after hull_canon.
template <int AS, int NW>
__device__ void hull_canon_write(const HullWS<AS> w, const HullResult& r, int* __restrict__ out_tris) {
    constexpr int T = 64 * NW;
    const int FC = w.FC;
    auto succ_of = [&](int e) { return fpl_geti(w.fpl[e / 3], e % 3); };
    auto end_of = [&](int e) { const int h = e / 3, j = e % 3; return u4_get(w.fv[h], j == 2 ? 0 : j + 1); };
    for (int f = threadIdx.x; f < r.ftop; f += T) {
        if (!(w.fvis[f] & 8)) continue;
        const int ebest = fpl_geti(w.fpl[f], 3);
        const int c0 = w.vout[end_of(ebest)];
        int t = w.fstamp[f];
        int prev = -1, e = succ_of(ebest), guard = 0;
        while (e != ebest && ++guard < 3 * FC) {
            const int b = end_of(e);
            if (w.vref[b]) {
                const int o = w.vout[b];
                if (prev >= 0) {
                    out_tris[3 * t] = c0;
                    out_tris[3 * t + 1] = prev;
                    out_tris[3 * t + 2] = o;
                    ++t;
                }
                prev = o;
            }
            e = succ_of(e);
        }
    }
}

}  // namespace cuacd
