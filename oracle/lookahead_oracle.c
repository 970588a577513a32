/* lookahead_oracle.c - serial C pieces of the lookahead oracle.
 *
 * TEST INFRASTRUCTURE ONLY (tests/ and bench.py's CPU baseline leg).
 *
 * orc_score_cut: the leaf scoring of one cut - FP64 volumes of both sides as
 *   the divergence sum of the pieces the cut would produce (origin on the
 *   plane, so the caps contribute nothing), and the point sets whose hulls
 *   are the sides' hulls: the original vertices used by each side in index
 *   order, then the crossing points in (triangle, edge) order of first
 *   occurrence. Split rules as the cut (reference cuda/plane_cut.cuh:269-652).
 * orc_concave_edges: shared edges whose dihedral angle exceeds the threshold
 *   (reference cuda/la_concave.cu:133-233), one record per undirected edge with
 *   its two lowest faces, unit normals and midpoint, decided on the cosine with
 *   the caller's per-branch constants.
 * Pinned by tests/test_oracle_lookahead.py (closed forms and agreement with
 * the half-space volume oracle and the capped cut).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static float olo_side(float a, float b, float c, float d, const float* p) {
    return fmaf(a, p[0], fmaf(b, p[1], fmaf(c, p[2], d)));
}

typedef struct { uint64_t* k; int* v; uint32_t mask; } olo_map;
static uint32_t olo_h(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
    return (uint32_t)k;
}
static void olo_init(olo_map* m, int n) {
    uint32_t H = 16;
    while ((int)H < 2 * n + 2) H <<= 1;
    m->k = (uint64_t*)malloc(8ull * H);
    m->v = (int*)malloc(4ull * H);
    m->mask = H - 1;
    for (uint32_t i = 0; i < H; ++i) { m->k[i] = ~0ull; m->v[i] = -1; }
}
static int olo_put(olo_map* m, uint64_t key, int* fresh) {
    uint32_t h = olo_h(key) & m->mask;
    for (;;) {
        if (m->k[h] == ~0ull) { m->k[h] = key; *fresh = 1; return (int)h; }
        if (m->k[h] == key) { *fresh = 0; return (int)h; }
        h = (h + 1) & m->mask;
    }
}

static double olo_tet(const double* o, const float* a, const float* b, const float* c) {
    double ax = a[0] - o[0], ay = a[1] - o[1], az = a[2] - o[2];
    double bx = b[0] - o[0], by = b[1] - o[1], bz = b[2] - o[2];
    double cx = c[0] - o[0], cy = c[1] - o[1], cz = c[2] - o[2];
    return ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
}

static void olo_xpt(const float* V, int a, int b, float pa, float pb, float pc, float pd, float* o) {
    int lo = a < b ? a : b, hi = a < b ? b : a;
    const float* A = V + 3 * lo;
    const float* B = V + 3 * hi;
    float d0 = olo_side(pa, pb, pc, pd, A), d1 = olo_side(pa, pb, pc, pd, B);
    float t = d0 / (d0 - d1);
    o[0] = A[0] + t * (B[0] - A[0]);
    o[1] = A[1] + t * (B[1] - A[1]);
    o[2] = A[2] + t * (B[2] - A[2]);
}

/* out_vol[2] = (pos, neg); out_cnt[4] = (pos pieces, neg pieces, pos points, neg points).
 * pos_pts / neg_pts need room for nv + 3*nt points each. */
void orc_score_cut(const float* V, int nv, const int* T, int nt, float pa, float pb, float pc, float pd,
                   double* out_vol, int* out_cnt, float* pos_pts, float* neg_pts) {
    signed char* sg = (signed char*)malloc(nv > 0 ? nv : 1);
    int* mark = (int*)calloc(nv > 0 ? nv : 1, 4);
    for (int i = 0; i < nv; ++i) {
        float d = olo_side(pa, pb, pc, pd, V + 3 * i);
        sg[i] = d > 1e-6f ? 1 : (d < -1e-6f ? -1 : 0);
    }
    double n2 = (double)pa * pa + (double)pb * pb + (double)pc * pc;
    double O[3] = {-(double)pd * pa / n2, -(double)pd * pb / n2, -(double)pd * pc / n2};
    double vp = 0.0, vn = 0.0;
    int cp = 0, cn = 0;
    for (int k = 0; k < nt; ++k) {
        int i0 = T[3 * k], i1 = T[3 * k + 1], i2 = T[3 * k + 2];
        int s0 = sg[i0], s1 = sg[i1], s2 = sg[i2];
        const float *p0 = V + 3 * i0, *p1 = V + 3 * i1, *p2 = V + 3 * i2;
        int hp = (s0 > 0) | (s1 > 0) | (s2 > 0), hn = (s0 < 0) | (s1 < 0) | (s2 < 0);
        if (!hp || !hn) {
            int side;
            if (hn) side = 2;
            else if (hp) side = 1;
            else {
                float e1x = p1[0] - p0[0], e1y = p1[1] - p0[1], e1z = p1[2] - p0[2];
                float e2x = p2[0] - p0[0], e2y = p2[1] - p0[1], e2z = p2[2] - p0[2];
                float dt = pa * (e1y * e2z - e1z * e2y) + pb * (e1z * e2x - e1x * e2z) + pc * (e1x * e2y - e1y * e2x);
                side = dt > 0 ? 1 : 2;
            }
            double v = olo_tet(O, p0, p1, p2);
            if (side == 1) { vp += v; ++cp; } else { vn += v; ++cn; }
            mark[i0] |= side; mark[i1] |= side; mark[i2] |= side;
            continue;
        }
        int vi[3] = {i0, i1, i2}, si[3] = {s0, s1, s2};
        int on = -1;
        for (int q = 0; q < 3; ++q) if (si[q] == 0) { on = q; break; }
        if (on >= 0) {
            int ov = vi[on], a = vi[(on + 1) % 3], b = vi[(on + 2) % 3], sa = si[(on + 1) % 3];
            float x[3];
            olo_xpt(V, a, b, pa, pb, pc, pd, x);
            const float *Ov = V + 3 * ov, *A = V + 3 * a, *B = V + 3 * b;
            if (sa > 0) { vp += olo_tet(O, Ov, A, x); vn += olo_tet(O, Ov, x, B); }
            else { vn += olo_tet(O, Ov, A, x); vp += olo_tet(O, Ov, x, B); }
            ++cp; ++cn;
            mark[ov] |= 3;
            mark[a] |= sa > 0 ? 1 : 2;
            mark[b] |= sa > 0 ? 2 : 1;
        } else {
            int lone = 0;
            for (int q = 0; q < 3; ++q)
                if (si[q] != si[(q + 1) % 3] && si[q] != si[(q + 2) % 3]) { lone = q; break; }
            int lv = vi[lone], o1 = vi[(lone + 1) % 3], o2 = vi[(lone + 2) % 3];
            float x1[3], x2[3];
            olo_xpt(V, lv, o1, pa, pb, pc, pd, x1);
            olo_xpt(V, lv, o2, pa, pb, pc, pd, x2);
            const float *L = V + 3 * lv, *Q1 = V + 3 * o1, *Q2 = V + 3 * o2;
            double a1 = olo_tet(O, L, x1, x2);
            double a2 = olo_tet(O, Q1, x2, x1) + olo_tet(O, Q1, Q2, x2);
            int ls = si[lone] > 0 ? 1 : 2;
            if (ls == 1) { vp += a1; vn += a2; } else { vn += a1; vp += a2; }
            ++cp; ++cn;
            mark[lv] |= ls; mark[o1] |= 3 - ls; mark[o2] |= 3 - ls;
        }
    }
    out_vol[0] = vp / 6.0;
    out_vol[1] = vn / 6.0;
    int np_ = 0, nn_ = 0;
    for (int i = 0; i < nv; ++i) {
        if (mark[i] & 1) { memcpy(pos_pts + 3 * np_, V + 3 * i, 12); ++np_; }
        if (mark[i] & 2) { memcpy(neg_pts + 3 * nn_, V + 3 * i, 12); ++nn_; }
    }
    /* crossing points of off-plane edges, first occurrence in (triangle, edge) order */
    olo_map m;
    olo_init(&m, 3 * nt);
    for (int k = 0; k < nt; ++k)
        for (int e = 0; e < 3; ++e) {
            int a = T[3 * k + e], b = T[3 * k + (e == 2 ? 0 : e + 1)];
            if ((sg[a] >= 0) == (sg[b] >= 0)) continue;
            int lo = a < b ? a : b, hi = a < b ? b : a, fresh;
            olo_put(&m, ((uint64_t)lo << 32) | (uint32_t)hi, &fresh);
            if (!fresh || sg[a] == 0 || sg[b] == 0) continue;
            float x[3];
            olo_xpt(V, a, b, pa, pb, pc, pd, x);
            memcpy(pos_pts + 3 * np_, x, 12); ++np_;
            memcpy(neg_pts + 3 * nn_, x, 12); ++nn_;
        }
    free(m.k); free(m.v);
    out_cnt[0] = cp; out_cnt[1] = cn; out_cnt[2] = np_; out_cnt[3] = nn_;
    free(sg); free(mark);
}

static int olo_normal(const float* V, const int* T, int f, float* n) {
    int a = T[3 * f], b = T[3 * f + 1], c = T[3 * f + 2];
    float e1x = V[3 * b] - V[3 * a], e1y = V[3 * b + 1] - V[3 * a + 1], e1z = V[3 * b + 2] - V[3 * a + 2];
    float e2x = V[3 * c] - V[3 * a], e2y = V[3 * c + 1] - V[3 * a + 1], e2z = V[3 * c + 2] - V[3 * a + 2];
    float x = e1y * e2z - e1z * e2y, y = e1z * e2x - e1x * e2z, z = e1x * e2y - e1y * e2x;
    float len = sqrtf(x * x + y * y + z * z);
    if (len < 1e-12f) return 0;
    n[0] = x / len; n[1] = y / len; n[2] = z / len;
    return 1;
}

/* rec: 13 floats per record: lo, hi, f0, f1 (as float bits of int), n0[3], n1[3], mid[3].
 * Returns the record count (records beyond cap are not written). */
int orc_concave_edges(const float* V, int nv, const int* T, int nt, float cos_pos, float cos_neg, int pos_mode,
                      int neg_mode, int* rec_i, float* rec_f, int cap) {
    if (nv <= 0 || nt < 2) return 0;
    olo_map m;
    olo_init(&m, 3 * nt);
    int* f1 = (int*)malloc(4ull * (m.mask + 1));
    for (uint32_t i = 0; i <= m.mask; ++i) f1[i] = 0x7fffffff;
    for (int k = 0; k < nt; ++k)
        for (int e = 0; e < 3; ++e) {
            int a = T[3 * k + e], b = T[3 * k + (e == 2 ? 0 : e + 1)];
            int lo = a < b ? a : b, hi = a < b ? b : a, fresh;
            int s = olo_put(&m, ((uint64_t)lo << 32) | (uint32_t)hi, &fresh);
            if (fresh) m.v[s] = k;
            else if (m.v[s] != k && k < f1[s]) f1[s] = k;
        }
    int cnt = 0;
    for (int k = 0; k < nt; ++k)
        for (int e = 0; e < 3; ++e) {
            int a = T[3 * k + e], b = T[3 * k + (e == 2 ? 0 : e + 1)];
            int lo = a < b ? a : b, hi = a < b ? b : a, fresh;
            uint64_t key = ((uint64_t)lo << 32) | (uint32_t)hi;
            int first = 1;
            for (int e2 = 0; e2 < e; ++e2) {
                int a2 = T[3 * k + e2], b2 = T[3 * k + (e2 == 2 ? 0 : e2 + 1)];
                int lo2 = a2 < b2 ? a2 : b2, hi2 = a2 < b2 ? b2 : a2;
                if (lo2 == lo && hi2 == hi) first = 0;
            }
            if (!first) continue;
            int s = olo_put(&m, key, &fresh);
            if (m.v[s] != k || f1[s] == 0x7fffffff) continue;
            int g = f1[s];
            float n0[3], n1[3];
            if (!olo_normal(V, T, k, n0) || !olo_normal(V, T, g, n1)) continue;
            float ca = n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2];
            ca = fmaxf(-1.0f, fminf(1.0f, ca));
            int vo = -1;
            for (int q = 0; q < 3; ++q) {
                int vv = T[3 * k + q];
                if (vv != lo && vv != hi) { vo = vv; break; }
            }
            if (vo < 0) continue;
            float dx = V[3 * vo] - V[3 * lo], dy = V[3 * vo + 1] - V[3 * lo + 1], dz = V[3 * vo + 2] - V[3 * lo + 2];
            float d = dx * n1[0] + dy * n1[1] + dz * n1[2];
            int concave;
            if (d > 0.0f) concave = pos_mode == 1 || (pos_mode == 2 && ca < cos_pos);
            else concave = neg_mode == 1 || (neg_mode == 2 && ca > cos_neg);
            if (!concave) continue;
            if (cnt < cap) {
                rec_i[4 * cnt] = lo; rec_i[4 * cnt + 1] = hi; rec_i[4 * cnt + 2] = k; rec_i[4 * cnt + 3] = g;
                float* r = rec_f + 9 * cnt;
                for (int q = 0; q < 3; ++q) {
                    r[q] = n0[q];
                    r[3 + q] = n1[q];
                    r[6 + q] = (V[3 * lo + q] + V[3 * hi + q]) * 0.5f;
                }
            }
            ++cnt;
        }
    free(m.k); free(m.v); free(f1);
    return cnt;
}
