/* hull_oracle.c - CPU oracle for the exact convex hull and its volume.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load this library, and only as the checker or the CPU
 * baseline; the product path never calls it.
 *
 * Restates the reference semantics of the D&C hull:
 *   - quantization onto the integer grid: cuda/hull_dandc.cuh:1351-1411
 *     (AABB, axis order (median, max, min), scale span/10216 with the sign flip
 *     when (med+1)%3 != max, centre (min+max)/2, truncating float->int);
 *   - the hull is the exact hull of the quantized points, duplicates resolve
 *     to the lowest input index (cuda/hull_dandc.cuh:1101-1104, 1494-1509);
 *   - output vertices are the original fp32 points (1268-1274);
 *   - volume: divergence theorem over the hull triangles (tests/bench_cpu_hull.cpp:87-111,
 *     cuacd/__init__.py:46-52), here in FP64.
 * The algorithm is a plain incremental hull in input order with exact int64
 * orientation tests (a different construction from the GPU quickhull).
 *
 * Pinned by: brute-force facet enumeration on small clouds (orc_hull_bruteforce),
 * closed forms (cube 8/12, tetrahedron 4/4, unit-cube volume 1), and scipy's
 * Qhull vertex sets on the same integer points (tests/test_oracle_hull.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <omp.h>
#include <string.h>

/* ---- quantization (cuda/hull_dandc.cuh:1351-1411) ------------------------ */
void orc_quantize(const float* pts, int n, int* q /* n*3: (med, max, min) axes */) {
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j) {
            float v = pts[3 * i + j];
            if (v < mn[j]) mn[j] = v;
            if (v > mx[j]) mx[j] = v;
        }
    float span[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
    int maxAx = (span[0] >= span[1]) ? ((span[0] >= span[2]) ? 0 : 2) : ((span[1] >= span[2]) ? 1 : 2);
    int minAx = (span[0] <= span[1]) ? ((span[0] <= span[2]) ? 0 : 2) : ((span[1] <= span[2]) ? 1 : 2);
    if (minAx == maxAx) minAx = (maxAx + 1) % 3;
    int medAx = 3 - maxAx - minAx;
    float sc[3];
    for (int j = 0; j < 3; ++j) sc[j] = span[j] / 10216.0f;
    if (((medAx + 1) % 3) != maxAx)
        for (int j = 0; j < 3; ++j) sc[j] = -sc[j];
    int ax[3] = {medAx, maxAx, minAx};
    float cen[3], inv[3];
    for (int j = 0; j < 3; ++j) {
        cen[j] = (mn[ax[j]] + mx[ax[j]]) * 0.5f;
        inv[j] = (sc[ax[j]] != 0.0f) ? 1.0f / sc[ax[j]] : 0.0f;
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j) {
            float d = pts[3 * i + ax[j]] - cen[j];
            float s = d * inv[j];
            q[3 * i + j] = (int)s;
        }
}

static int64_t orient64(const int* a, const int* b, const int* c, const int* p) {
    int64_t bx = b[0] - a[0], by = b[1] - a[1], bz = b[2] - a[2];
    int64_t cx = c[0] - a[0], cy = c[1] - a[1], cz = c[2] - a[2];
    int64_t dx = p[0] - a[0], dy = p[1] - a[1], dz = p[2] - a[2];
    int64_t nx = by * cz - bz * cy, ny = bz * cx - bx * cz, nz = bx * cy - by * cx;
    return nx * dx + ny * dy + nz * dz;
}

typedef struct {
    int v[3];
    int alive;
} OFace;

/* Exact hull of n quantized points (q: n*3). Writes is_vertex[n] (0/1) and the
 * live faces (indices into the input, outward CCW) to faces (cap = 2n+8
 * triangles). Returns the face count, 0 for a degenerate (flat) input, -1 on
 * allocation failure. */
/* directed edge (a, b) -> face id, -1 when none; open addressing, entries are
 * never removed (a dead edge is set to -1), so lookups stay valid */
typedef struct { uint64_t* k; int* v; uint64_t mask; } OEdgeMap;
static uint64_t oem_hash(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
    return k;
}
static int* oem_slot(OEdgeMap* m, int a, int b, int create) {
    uint64_t key = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
    uint64_t h = oem_hash(key) & m->mask;
    for (;;) {
        if (m->k[h] == key) return &m->v[h];
        if (m->k[h] == ~0ull) {
            if (!create) return NULL;
            m->k[h] = key;
            m->v[h] = -1;
            return &m->v[h];
        }
        h = (h + 1) & m->mask;
    }
}
static int oem_get(OEdgeMap* m, int a, int b) {
    int* s = oem_slot(m, a, b, 0);
    return s ? *s : -1;
}

/* One incremental construction with room for cap faces created in all
 * (dead ones included); -2 when that is not enough. */
static int orc_hull_cap(const int* q, int n, unsigned char* is_vertex, int* faces, int face_cap, int cap);

/* Exact hull of integer points by incremental insertion in input order
 * (strictly visible faces only, so points on a face are not inserted).
 * The number of faces created over all insertions depends on the order and
 * can exceed any fixed multiple of n on lattices with many coplanar points,
 * so the face store grows until the construction fits. */
int orc_hull(const int* q, int n, unsigned char* is_vertex, int* faces, int face_cap) {
    int64_t cap = 8 * (int64_t)n + 64;
    for (;;) {
        int r = orc_hull_cap(q, n, is_vertex, faces, face_cap, (int)cap);
        if (r != -2) return r;
        cap *= 2;
        if (cap > (1 << 28)) return -1;
    }
}

static int orc_hull_cap(const int* q, int n, unsigned char* is_vertex, int* faces, int face_cap, int cap) {
    memset(is_vertex, 0, (size_t)n);
    if (n < 4) return 0;
    /* initial tetrahedron: first points (input order) that span 0D..3D */
    int i0 = 0, i1 = -1, i2 = -1, i3 = -1;
    for (int i = 1; i < n && i1 < 0; ++i)
        if (q[3 * i] != q[0] || q[3 * i + 1] != q[1] || q[3 * i + 2] != q[2]) i1 = i;
    if (i1 < 0) return 0;
    for (int i = i1 + 1; i < n && i2 < 0; ++i) {
        int64_t ex = q[3 * i1] - q[0], ey = q[3 * i1 + 1] - q[1], ez = q[3 * i1 + 2] - q[2];
        int64_t dx = q[3 * i] - q[0], dy = q[3 * i + 1] - q[1], dz = q[3 * i + 2] - q[2];
        if (ey * dz - ez * dy || ez * dx - ex * dz || ex * dy - ey * dx) i2 = i;
    }
    if (i2 < 0) return 0;
    for (int i = i2 + 1; i < n && i3 < 0; ++i)
        if (orient64(q + 3 * i0, q + 3 * i1, q + 3 * i2, q + 3 * i) != 0) i3 = i;
    if (i3 < 0) return 0;
    if (orient64(q + 3 * i0, q + 3 * i1, q + 3 * i2, q + 3 * i3) < 0) {
        int t = i1; i1 = i2; i2 = t;
    }
    OFace* F = (OFace*)malloc(sizeof(OFace) * (size_t)cap);
    OEdgeMap em;
    uint64_t H = 64;
    while (H < 2ull * 3ull * (uint64_t)cap + 2) H <<= 1;
    em.k = (uint64_t*)malloc(8ull * H);
    em.v = (int*)malloc(4ull * H);
    em.mask = H - 1;
    int* hz = (int*)malloc(sizeof(int) * 2 * (size_t)cap);
    int* vis = (int*)malloc(sizeof(int) * (size_t)cap);
    if (!F || !em.k || !em.v || !hz || !vis) { free(F); free(em.k); free(em.v); free(hz); free(vis); return -1; }
    for (uint64_t k = 0; k < H; ++k) em.k[k] = ~0ull;
    int nf = 0;
#define ADDF(a, b, c)                                                                     \
    do {                                                                                  \
        F[nf].v[0] = (a); F[nf].v[1] = (b); F[nf].v[2] = (c); F[nf].alive = 1;           \
        *oem_slot(&em, (a), (b), 1) = nf; *oem_slot(&em, (b), (c), 1) = nf;               \
        *oem_slot(&em, (c), (a), 1) = nf; ++nf;                                           \
    } while (0)
    ADDF(i0, i2, i1);
    ADDF(i0, i1, i3);
    ADDF(i1, i2, i3);
    ADDF(i0, i3, i2);
    int failed = 0;
    /* live face ids (the face array keeps dead faces so edge-map ids stay valid) */
    int* L = (int*)malloc(sizeof(int) * (size_t)cap);
    int nL = 4;
    for (int k = 0; k < 4; ++k) L[k] = k;
    for (int p = 0; p < n && !failed; ++p) {
        if (p == i0 || p == i1 || p == i2 || p == i3) continue;
        int nvis = 0;
        for (int li = 0; li < nL; ++li) {
            int f = L[li];
            if (!F[f].alive) continue;
            if (orient64(q + 3 * F[f].v[0], q + 3 * F[f].v[1], q + 3 * F[f].v[2], q + 3 * p) > 0) {
                vis[nvis++] = f;
                F[f].alive = 2; /* marked visible */
            }
        }
        if (!nvis) continue;
        int nh = 0;
        for (int k = 0; k < nvis; ++k) {
            const OFace* f = &F[vis[k]];
            for (int e = 0; e < 3; ++e) {
                int a = f->v[e], b = f->v[(e + 1) % 3];
                int g = oem_get(&em, b, a);
                if (g < 0 || F[g].alive != 2) { hz[2 * nh] = a; hz[2 * nh + 1] = b; ++nh; }
            }
        }
        for (int k = 0; k < nvis; ++k) {
            OFace* f = &F[vis[k]];
            f->alive = 0;
            for (int e = 0; e < 3; ++e) {
                int a = f->v[e], b = f->v[(e + 1) % 3];
                int* sl = oem_slot(&em, a, b, 0);
                if (sl && *sl == vis[k]) *sl = -1;
            }
        }
        if (nf + nh > cap) { failed = 1; break; }
        int w = 0;
        for (int li = 0; li < nL; ++li)
            if (F[L[li]].alive) L[w++] = L[li];
        for (int k = 0; k < nh; ++k) { L[w++] = nf; ADDF(hz[2 * k], hz[2 * k + 1], p); }
        nL = w;
    }
    free(L);
#undef ADDF
    int out = 0;
    if (!failed) {
        for (int f = 0; f < nf; ++f) {
            if (!F[f].alive) continue;
            if (out < face_cap) {
                faces[3 * out] = F[f].v[0];
                faces[3 * out + 1] = F[f].v[1];
                faces[3 * out + 2] = F[f].v[2];
            }
            is_vertex[F[f].v[0]] = is_vertex[F[f].v[1]] = is_vertex[F[f].v[2]] = 1;
            ++out;
        }
    }
    free(F); free(em.k); free(em.v); free(hz); free(vis);
    return failed ? -2 : out;
}

/* Canonical form of the exact hull (the GPU library emits the same):
 *   facets = maximal planar faces (adjacent faces merged when exactly
 *   coplanar); corners = facet boundary vertices where the boundary turns;
 *   each facet is the fan (c0, c_k, c_k+1) over its corners in CCW order from
 *   its lowest-index corner c0; facets ordered by (c0, c1).
 * Writes is_vertex[n] = corner flags and the triangles (input indices);
 * returns the triangle count, 0 for a flat input, -1 on failure. */
static int ocn_find(int* uf, int x) {
    while (uf[x] != x) { uf[x] = uf[uf[x]]; x = uf[x]; }
    return x;
}
typedef struct { int c0, c1, f, e, m; } OFacet;
static int ofacet_cmp(const void* a, const void* b) {
    const OFacet* x = (const OFacet*)a;
    const OFacet* y = (const OFacet*)b;
    if (x->c0 != y->c0) return x->c0 < y->c0 ? -1 : 1;
    if (x->c1 != y->c1) return x->c1 < y->c1 ? -1 : 1;
    return 0;
}
int orc_hull_canon(const int* q, int n, unsigned char* is_vertex, int* out, int cap) {
    memset(is_vertex, 0, (size_t)(n > 0 ? n : 1));
    if (n < 4) return 0;
    int fcap = 2 * n + 8;
    int* F = (int*)malloc(sizeof(int) * 3 * (size_t)fcap);
    unsigned char* tmpv = (unsigned char*)malloc((size_t)n);
    if (!F || !tmpv) { free(F); free(tmpv); return -1; }
    int nf = orc_hull(q, n, tmpv, F, fcap);
    free(tmpv);
    if (nf <= 0) { free(F); return nf; }
    /* directed edge -> face */
    OEdgeMap em;
    uint64_t H = 64;
    while (H < 2ull * 3ull * (uint64_t)nf + 2) H <<= 1;
    em.k = (uint64_t*)malloc(8ull * H);
    em.v = (int*)malloc(4ull * H);
    em.mask = H - 1;
    int* uf = (int*)malloc(sizeof(int) * (size_t)nf);
    unsigned char* cop = (unsigned char*)malloc((size_t)nf);
    int* succ = (int*)malloc(sizeof(int) * 3 * (size_t)nf);
    unsigned char* corner = (unsigned char*)calloc((size_t)n, 1);
    OFacet* fl = (OFacet*)malloc(sizeof(OFacet) * (size_t)nf);
    int* start = (int*)malloc(sizeof(int) * (size_t)nf);
    if (!em.k || !em.v || !uf || !cop || !succ || !corner || !fl || !start) {
        free(F); free(em.k); free(em.v); free(uf); free(cop); free(succ); free(corner); free(fl); free(start);
        return -1;
    }
    for (uint64_t k = 0; k < H; ++k) em.k[k] = ~0ull;
    for (int f = 0; f < nf; ++f)
        for (int e = 0; e < 3; ++e) *oem_slot(&em, F[3 * f + e], F[3 * f + (e + 1) % 3], 1) = f;
    int fail = 0;
    for (int f = 0; f < nf; ++f) {
        uf[f] = f;
        cop[f] = 0;
        for (int e = 0; e < 3; ++e) {
            int a = F[3 * f + e], b = F[3 * f + (e + 1) % 3];
            int g = oem_get(&em, b, a);
            if (g < 0) { fail = 1; continue; }
            int o = F[3 * g];
            if (o == a || o == b) o = F[3 * g + 1];
            if (o == a || o == b) o = F[3 * g + 2];
            if (orient64(q + 3 * F[3 * f], q + 3 * F[3 * f + 1], q + 3 * F[3 * f + 2], q + 3 * o) == 0)
                cop[f] |= (unsigned char)(1 << e);
        }
    }
    for (int f = 0; f < nf && !fail; ++f)
        for (int e = 0; e < 3; ++e)
            if (cop[f] & (1 << e)) {
                int g = oem_get(&em, F[3 * f + (e + 1) % 3], F[3 * f + e]);
                int a = ocn_find(uf, f), b = ocn_find(uf, g);
                if (a != b) { if (a < b) uf[b] = a; else uf[a] = b; }
            }
    /* boundary successors (rotate around the end vertex inside the facet), corners */
    for (int f = 0; f < nf && !fail; ++f)
        for (int e = 0; e < 3; ++e) {
            succ[3 * f + e] = -1;
            if (cop[f] & (1 << e)) continue;
            int b = F[3 * f + (e + 1) % 3];
            int h = f, j = (e + 1) % 3, guard = 0;
            while ((cop[h] >> j) & 1) {
                int c = F[3 * h + (j + 1) % 3];
                int g = oem_get(&em, c, b);
                if (g < 0 || ++guard > nf) { fail = 1; break; }
                j = F[3 * g] == b ? 0 : (F[3 * g + 1] == b ? 1 : 2);
                h = g;
            }
            if (fail) break;
            succ[3 * f + e] = 3 * h + j;
            int a = F[3 * f + e], c = F[3 * h + (j + 1) % 3];
            int64_t ux = q[3 * b] - q[3 * a], uy = q[3 * b + 1] - q[3 * a + 1], uz = q[3 * b + 2] - q[3 * a + 2];
            int64_t vx = q[3 * c] - q[3 * b], vy = q[3 * c + 1] - q[3 * b + 1], vz = q[3 * c + 2] - q[3 * b + 2];
            if (uy * vz - uz * vy || uz * vx - ux * vz || ux * vy - uy * vx) corner[b] = 1;
        }
    /* one boundary edge per facet */
    for (int f = 0; f < nf; ++f) start[f] = -1;
    for (int f = 0; f < nf && !fail; ++f)
        for (int e = 0; e < 3; ++e)
            if (!(cop[f] & (1 << e))) {
                int L = ocn_find(uf, f);
                if (start[L] < 0) start[L] = 3 * f + e;
            }
    int nfac = 0, nt = 0;
    for (int L = 0; L < nf && !fail; ++L) {
        if (ocn_find(uf, L) != L || start[L] < 0) continue;
        int s0 = start[L], e = s0, m = 0, best = 0x7fffffff, eb = s0, guard = 0;
        do {
            int b = F[3 * (e / 3) + (e % 3 + 1) % 3];
            if (corner[b]) { ++m; if (b < best) { best = b; eb = e; } }
            e = succ[e];
            if (e < 0 || ++guard > 3 * nf) { fail = 1; break; }
        } while (e != s0);
        if (fail) break;
        int c1 = -1;
        for (e = succ[eb]; e != eb; e = succ[e]) {
            int b = F[3 * (e / 3) + (e % 3 + 1) % 3];
            if (corner[b]) { c1 = b; break; }
        }
        fl[nfac].c0 = best; fl[nfac].c1 = c1; fl[nfac].f = L; fl[nfac].e = eb; fl[nfac].m = m;
        ++nfac;
        if (m >= 3) nt += m - 2;
    }
    if (!fail) {
        qsort(fl, (size_t)nfac, sizeof(OFacet), ofacet_cmp);
        int t = 0;
        for (int k = 0; k < nfac; ++k) {
            int eb = fl[k].e, c0 = fl[k].c0, prev = -1;
            for (int e = succ[eb]; e != eb; e = succ[e]) {
                int b = F[3 * (e / 3) + (e % 3 + 1) % 3];
                if (!corner[b]) continue;
                if (prev >= 0) {
                    if (t < cap) { out[3 * t] = c0; out[3 * t + 1] = prev; out[3 * t + 2] = b; }
                    ++t;
                }
                prev = b;
            }
        }
        memcpy(is_vertex, corner, (size_t)n);
        nt = t;
    }
    free(F); free(em.k); free(em.v); free(uf); free(cop); free(succ); free(corner); free(fl); free(start);
    return fail ? -1 : nt;
}

/* 1 if the hull is in general position: no duplicate quantized points and no
 * input point other than a facet's three vertices lies on that facet's plane.
 * Then the vertex set and the facets are unique. */
int orc_general_position(const int* q, int n, const int* faces, int nf) {
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (q[3 * i] == q[3 * j] && q[3 * i + 1] == q[3 * j + 1] && q[3 * i + 2] == q[3 * j + 2]) return 0;
    for (int f = 0; f < nf; ++f) {
        const int* a = q + 3 * faces[3 * f];
        const int* b = q + 3 * faces[3 * f + 1];
        const int* c = q + 3 * faces[3 * f + 2];
        for (int p = 0; p < n; ++p) {
            if (p == faces[3 * f] || p == faces[3 * f + 1] || p == faces[3 * f + 2]) continue;
            if (orient64(a, b, c, q + 3 * p) == 0) return 0;
        }
    }
    return 1;
}

/* FP64 divergence-theorem volume of triangles over fp32 points (anchored at the
 * first vertex of the first face). */
double orc_mesh_volume(const float* pts, const int* tris, int nt) {
    if (nt <= 0) return 0.0;
    const float* o = pts + 3 * tris[0];
    double ax = o[0], ay = o[1], az = o[2], acc = 0.0;
    for (int t = 0; t < nt; ++t) {
        const float* p0 = pts + 3 * tris[3 * t];
        const float* p1 = pts + 3 * tris[3 * t + 1];
        const float* p2 = pts + 3 * tris[3 * t + 2];
        double x0 = p0[0] - ax, y0 = p0[1] - ay, z0 = p0[2] - az;
        double x1 = p1[0] - ax, y1 = p1[1] - ay, z1 = p1[2] - az;
        double x2 = p2[0] - ax, y2 = p2[1] - ay, z2 = p2[2] - az;
        acc += x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
    }
    return acc / 6.0;
}

/* Volume of a canonical hull mesh (orc_hull_canon's triangle order), summed
 * in the order the GPU library fixes: FP64 divergence terms anchored at the
 * first vertex of the first triangle, the triangle list cut into 64
 * contiguous chunks of ceil(nt / 64), each chunk summed in order, then the 64
 * partial sums in order. */
double orc_hull_volume(const float* pts, const int* tris, int nt) {
    if (nt <= 0) return 0.0;
    const float* o = pts + 3 * tris[0];
    double ax = o[0], ay = o[1], az = o[2], acc = 0.0;
    int chunk = (nt + 63) / 64;
    for (int l = 0; l < 64; ++l) {
        int b = l * chunk < nt ? l * chunk : nt, e = b + chunk < nt ? b + chunk : nt;
        double part = 0.0;
        for (int t = b; t < e; ++t) {
            const float* p0 = pts + 3 * tris[3 * t];
            const float* p1 = pts + 3 * tris[3 * t + 1];
            const float* p2 = pts + 3 * tris[3 * t + 2];
            double x0 = p0[0] - ax, y0 = p0[1] - ay, z0 = p0[2] - az;
            double x1 = p1[0] - ax, y1 = p1[1] - ay, z1 = p1[2] - az;
            double x2 = p2[0] - ax, y2 = p2[1] - ay, z2 = p2[2] - az;
            part += x0 * (y1 * z2 - z1 * y2) + y0 * (z1 * x2 - x1 * z2) + z0 * (x1 * y2 - y1 * x2);
        }
        acc += part;
    }
    return acc / 6.0;
}

/* Brute force (n <= ~60): vertex flags from every facet triple (i<j<k) whose
 * plane has all other points strictly on one side. Valid in general position. */
int orc_hull_bruteforce(const int* q, int n, unsigned char* is_vertex) {
    memset(is_vertex, 0, (size_t)n);
    int nfacets = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            for (int k = j + 1; k < n; ++k) {
                int pos = 0, neg = 0;
                for (int p = 0; p < n && !(pos && neg); ++p) {
                    if (p == i || p == j || p == k) continue;
                    int64_t o = orient64(q + 3 * i, q + 3 * j, q + 3 * k, q + 3 * p);
                    if (o > 0) pos = 1;
                    else if (o < 0) neg = 1;
                    else { pos = neg = 1; }
                }
                if (!(pos && neg)) {
                    is_vertex[i] = is_vertex[j] = is_vertex[k] = 1;
                    ++nfacets;
                }
            }
    return nfacets;
}

/* Batch: per cloud, vertex flags (total_pts), face count, volume (FP64),
 * general-position flag. OpenMP over clouds. Returns the number of clouds that
 * failed (allocation). */
int orc_hull_batch(const float* pts, const int* offsets, int n_clouds, unsigned char* is_vertex, int* n_faces,
                   double* volumes, int* general, int check_general) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : bad)
    for (int c = 0; c < n_clouds; ++c) {
        int s = offsets[c], n = offsets[c + 1] - s;
        int* q = (int*)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
        int cap = 2 * n + 8;
        int* faces = (int*)malloc(sizeof(int) * 3 * (size_t)cap);
        if (!q || !faces) { ++bad; free(q); free(faces); continue; }
        orc_quantize(pts + 3 * (size_t)s, n, q);
        int nf = orc_hull(q, n, is_vertex + s, faces, cap);
        if (nf < 0) { ++bad; nf = 0; }
        general[c] = check_general ? orc_general_position(q, n, faces, nf) : -1;
        nf = orc_hull_canon(q, n, is_vertex + s, faces, cap);
        if (nf < 0) { ++bad; nf = 0; }
        n_faces[c] = nf;
        volumes[c] = orc_hull_volume(pts + 3 * (size_t)s, faces, nf);
        free(q); free(faces);
    }
    return bad;
}

/* Single-thread variant for the 1-core CPU baseline. */
int orc_hull_batch_serial(const float* pts, const int* offsets, int n_clouds, unsigned char* is_vertex,
                          int* n_faces, double* volumes) {
    int bad = 0;
    for (int c = 0; c < n_clouds; ++c) {
        int s = offsets[c], n = offsets[c + 1] - s;
        int* q = (int*)malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
        int cap = 2 * n + 8;
        int* faces = (int*)malloc(sizeof(int) * 3 * (size_t)cap);
        if (!q || !faces) { ++bad; free(q); free(faces); continue; }
        orc_quantize(pts + 3 * (size_t)s, n, q);
        int nf = orc_hull_canon(q, n, is_vertex + s, faces, cap);
        if (nf < 0) { ++bad; nf = 0; }
        n_faces[c] = nf;
        volumes[c] = orc_hull_volume(pts + 3 * (size_t)s, faces, nf);
        free(q); free(faces);
    }
    return bad;
}

/* OpenMP threads of the batch functions (CPU baselines on one core and on
 * all of them). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }
