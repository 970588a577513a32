/* cut_oracle.c - CPU oracle for the volumes of a mesh cut by a plane.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, bench.py's CPU baseline leg).
 *
 * The volume of the part of a closed mesh on one side of a plane is computed
 * without building any cap: with the divergence-theorem origin O placed ON the
 * plane, the cap contributes nothing (its points x satisfy (x - O) . n = 0), so
 *   V(side) = sum over triangles clipped to the side of det(p0-O, p1-O, p2-O)/6.
 * Triangles are clipped exactly in FP64 (Sutherland-Hodgman against one
 * plane). This pins the GPU cut (reference cuda/plane_cut.cuh) independently of
 * its cap triangulation: a missing or mis-wound cap changes the GPU volume.
 * Pinned by closed forms in tests/test_oracle_cut.py (cube and box halves).
 */
#include <math.h>

static double tet(const double* o, const double* a, const double* b, const double* c) {
    double ax = a[0] - o[0], ay = a[1] - o[1], az = a[2] - o[2];
    double bx = b[0] - o[0], by = b[1] - o[1], bz = b[2] - o[2];
    double cx = c[0] - o[0], cy = c[1] - o[1], cz = c[2] - o[2];
    return ax * (by * cz - bz * cy) + ay * (bz * cx - bx * cz) + az * (bx * cy - by * cx);
}

/* out[0] = volume on the side n.x + d > 0, out[1] = the other side (signed sums / 6). */
void orc_halfspace_volumes(const float* verts, const int* tris, int nt, double pa, double pb, double pc,
                           double pd, double* out) {
    double n2 = pa * pa + pb * pb + pc * pc;
    double o[3] = {-pd * pa / n2, -pd * pb / n2, -pd * pc / n2};
    double vp = 0.0, vn = 0.0;
    for (int t = 0; t < nt; ++t) {
        double p[3][3], s[3];
        for (int k = 0; k < 3; ++k) {
            for (int j = 0; j < 3; ++j) p[k][j] = verts[3 * tris[3 * t + k] + j];
            s[k] = pa * p[k][0] + pb * p[k][1] + pc * p[k][2] + pd;
        }
        for (int side = 0; side < 2; ++side) {
            double sign = side == 0 ? 1.0 : -1.0;
            double poly[4][3];
            int m = 0;
            for (int k = 0; k < 3; ++k) {
                int k2 = (k + 1) % 3;
                double a = sign * s[k], b = sign * s[k2];
                if (a >= 0) {
                    for (int j = 0; j < 3; ++j) poly[m][j] = p[k][j];
                    ++m;
                }
                if ((a > 0 && b < 0) || (a < 0 && b > 0)) {
                    double u = a / (a - b);
                    for (int j = 0; j < 3; ++j) poly[m][j] = p[k][j] + u * (p[k2][j] - p[k][j]);
                    ++m;
                }
            }
            double v = 0.0;
            for (int k = 1; k + 1 < m; ++k) v += tet(o, poly[0], poly[k], poly[k + 1]);
            if (side == 0) vp += v;
            else vn += v;
        }
    }
    out[0] = vp / 6.0;
    out[1] = vn / 6.0;
}

/* ------------------------------------------------------------------------
 * orc_plane_cut: serial restatement of the plane cut with cap triangulation.
 *
 * Semantics follow the reference (cuda/plane_cut.cuh): sign eps 1e-6 with
 * s = fmaf(a, x, fmaf(b, y, fmaf(c, z, d))) (269-274); an edge crosses when
 * (s_a >= 0) != (s_b >= 0) and reuses an on-plane endpoint (301-318, 548-566);
 * split by lone vertex / on-plane vertex (582-652); coplanar triangles go to
 * the side their normal faces; boundary loops -> orientation by signed area ->
 * nesting by ray casting -> holes bridged from their rightmost vertex -> ear
 * clipping with collinear ears (1008-1308); cap failure returns the mesh whole
 * on the positive side (1339-1368).
 *
 * The canonical orders are the ones the GPU cut guarantees: new vertices in
 * (triangle, edge) order of first occurrence, side triangles in input order
 * followed by cap triangles, boundary edges in (positive triangle, edge) order,
 * the successor of an edge is the lowest-numbered boundary edge leaving its
 * end vertex, loops are labelled by their lowest edge. Given those orders the
 * output is a function of the input, so the GPU result must equal this one
 * exactly (vertices bit for bit, triangles index for index).
 * Pinned by tests/test_oracle_cut.py (closed forms, closedness, agreement of
 * the halves' volumes with orc_halfspace_volumes).
 * ---------------------------------------------------------------------- */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OPC_EPS 1e-6f
#define OPC_MAX_LOOPS 256

typedef struct { float u, v; } opc_p2;

static float opc_cross2(opc_p2 a, opc_p2 b, opc_p2 c) {
    return (b.u - a.u) * (c.v - a.v) - (b.v - a.v) * (c.u - a.u);
}
static uint32_t opc_fkey(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* open-addressing map u64 -> int */
typedef struct { uint64_t* k; int* v; uint32_t mask; } opc_map;
static uint32_t opc_h(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
    return (uint32_t)k;
}
static void opc_map_init(opc_map* m, int n) {
    uint32_t H = 16;
    while ((int)H < 2 * n + 2) H <<= 1;
    m->k = (uint64_t*)malloc(8ull * H);
    m->v = (int*)malloc(4ull * H);
    m->mask = H - 1;
    for (uint32_t i = 0; i < H; ++i) m->k[i] = ~0ull;
}
static void opc_map_free(opc_map* m) { free(m->k); free(m->v); }
/* returns slot; *fresh = 1 when inserted */
static int opc_map_put(opc_map* m, uint64_t key, int* fresh) {
    uint32_t h = opc_h(key) & m->mask;
    for (;;) {
        if (m->k[h] == ~0ull) { m->k[h] = key; *fresh = 1; return (int)h; }
        if (m->k[h] == key) { *fresh = 0; return (int)h; }
        h = (h + 1) & m->mask;
    }
}
static int opc_map_find(const opc_map* m, uint64_t key) {
    uint32_t h = opc_h(key) & m->mask;
    for (uint32_t p = 0; p <= m->mask; ++p) {
        if (m->k[h] == key) return (int)h;
        if (m->k[h] == ~0ull) return -1;
        h = (h + 1) & m->mask;
    }
    return -1;
}

/* ear clipping of one CCW polygon (ids poly[], coords uv[]); -1 on failure.
 * Reference semantics (cuda/plane_cut.cuh:1195-1290): a vertex with cross <
 * -1e-9 is not an ear; a collinear one (|cross| <= 1e-9) is clipped as a
 * zero-area ear; otherwise it is an ear when no other live vertex (farther
 * than 1e-6 from the ear's corners) lies in the closed triangle; after a clip
 * the walk continues at the next vertex. When a whole lap finds no ear, the
 * least reflex vertex is clipped if it is only reflex by rounding (cross above
 * -1e-6 of the squared extent): a sliver, not a failure. */
static int opc_pt_in_tri(opc_p2 X, opc_p2 A, opc_p2 B, opc_p2 C) {
    float d1 = opc_cross2(A, B, X), d2 = opc_cross2(B, C, X), d3 = opc_cross2(C, A, X);
    return !((d1 < -1e-10f || d2 < -1e-10f || d3 < -1e-10f) && (d1 > 1e-10f || d2 > 1e-10f || d3 > 1e-10f));
}
static int opc_earclip(const int* poly, int n, const opc_p2* uv, int* prv, int* nxt, unsigned char* unused,
                       int* out) {
    (void)unused;
    if (n < 3) return 0;
    float mnu = uv[0].u, mxu = uv[0].u, mnv = uv[0].v, mxv = uv[0].v;
    for (int i = 1; i < n; ++i) {
        mnu = fminf(mnu, uv[i].u); mxu = fmaxf(mxu, uv[i].u);
        mnv = fminf(mnv, uv[i].v); mxv = fmaxf(mxv, uv[i].v);
    }
    const float ext = fmaxf(mxu - mnu, mxv - mnv);
    const float sliver = -1e-6f * ext * ext;
    for (int i = 0; i < n; ++i) { prv[i] = i == 0 ? n - 1 : i - 1; nxt[i] = i == n - 1 ? 0 : i + 1; }
    int nt = 0, remaining = n, cur = 0, lap = 0;
    while (remaining > 3) {
        int p = prv[cur], q = nxt[cur];
        float cr = opc_cross2(uv[p], uv[cur], uv[q]);
        int ear;
        if (lap >= remaining) {
            /* a full lap without an ear: clip the least reflex vertex if it is a sliver */
            int best = cur, c = cur;
            float bc = cr;
            for (int k = 0; k < remaining; ++k) {
                float x = opc_cross2(uv[prv[c]], uv[c], uv[nxt[c]]);
                if (x > bc) { bc = x; best = c; }
                c = nxt[c];
            }
            if (bc < sliver) return -1;
            cur = best; p = prv[cur]; q = nxt[cur];
            ear = 1;
        } else if (cr < -1e-9f) {
            ear = 0;
        } else if (cr <= 1e-9f) {
            ear = 1;
        } else {
            opc_p2 A = uv[p], B = uv[cur], Cc = uv[q];
            ear = 1;
            for (int i = 0; i < n && ear; ++i) {
                if (i == p || i == cur || i == q || nxt[prv[i]] != i) continue;
                opc_p2 X = uv[i];
                if (fabsf(X.u - A.u) + fabsf(X.v - A.v) > 1e-6f && fabsf(X.u - B.u) + fabsf(X.v - B.v) > 1e-6f &&
                    fabsf(X.u - Cc.u) + fabsf(X.v - Cc.v) > 1e-6f && opc_pt_in_tri(X, A, B, Cc))
                    ear = 0;
            }
        }
        if (ear) {
            out[3 * nt] = poly[p]; out[3 * nt + 1] = poly[cur]; out[3 * nt + 2] = poly[q];
            nxt[p] = q; prv[q] = p;
            ++nt; --remaining;
            cur = q; lap = 0;
        } else {
            cur = q; ++lap;
        }
    }
    if (remaining == 3) {
        int p = prv[cur], q = nxt[cur];
        out[3 * nt] = poly[p]; out[3 * nt + 1] = poly[cur]; out[3 * nt + 2] = poly[q];
        ++nt;
    }
    return nt;
}

static int opc_emit_whole(const float* V, int nv, const int* T, int nt, float* ov, int* ot, int cap_v, int cap_t,
                          int* onv, int* ont) {
    if (nv > cap_v || nt > cap_t) return -1;
    memcpy(ov, V, 12ull * nv);
    memcpy(ot, T, 12ull * nt);
    *onv = nv; *ont = nt;
    return 0;
}

/* sizes[4] = pos nv, pos nt, neg nv, neg nt. Returns status 0/1/2, -1 capacity. */
int orc_plane_cut(const float* V, int nv, const int* T, int nt, float pa, float pb, float pc, float pd, float* opv,
                  int* opt, float* onv, int* ont, int cap_v, int cap_t, int* sizes) {
    sizes[0] = sizes[1] = sizes[2] = sizes[3] = 0;
    if (nv <= 0 || nt <= 0) return 2;
    signed char* sg = (signed char*)malloc(nv);
    for (int i = 0; i < nv; ++i) {
        float d = fmaf(pa, V[3 * i], fmaf(pb, V[3 * i + 1], fmaf(pc, V[3 * i + 2], pd)));
        sg[i] = d > OPC_EPS ? 1 : (d < -OPC_EPS ? -1 : 0);
    }
    int* tc = (int*)malloc(4ull * nt);
    int ncross_tot = 0, npt = 0, nnt = 0, any_pos = 0, any_neg = 0;
    for (int k = 0; k < nt; ++k) {
        int i0 = T[3 * k], i1 = T[3 * k + 1], i2 = T[3 * k + 2];
        int s0 = sg[i0], s1 = sg[i1], s2 = sg[i2];
        int e0 = s0 >= 0, e1 = s1 >= 0, e2 = s2 >= 0;
        int ncross = (e0 != e1) + (e1 != e2) + (e2 != e0);
        int hp = (s0 > 0) | (s1 > 0) | (s2 > 0), hn = (s0 < 0) | (s1 < 0) | (s2 < 0);
        int np_ = 0, nn_ = 0;
        if (!hp || !hn) {
            if (hn) nn_ = 1;
            else if (hp) np_ = 1;
            else {
                float e1x = V[3 * i1] - V[3 * i0], e1y = V[3 * i1 + 1] - V[3 * i0 + 1], e1z = V[3 * i1 + 2] - V[3 * i0 + 2];
                float e2x = V[3 * i2] - V[3 * i0], e2y = V[3 * i2 + 1] - V[3 * i0 + 1], e2z = V[3 * i2 + 2] - V[3 * i0 + 2];
                float dt = pa * (e1y * e2z - e1z * e2y) + pb * (e1z * e2x - e1x * e2z) + pc * (e1x * e2y - e1y * e2x);
                if (dt > 0) np_ = 1; else nn_ = 1;
            }
        } else if (s0 == 0 || s1 == 0 || s2 == 0) {
            np_ = 1; nn_ = 1;
        } else {
            int lone_pos = (s0 > 0) + (s1 > 0) + (s2 > 0) == 1;
            np_ = lone_pos ? 1 : 2;
            nn_ = lone_pos ? 2 : 1;
        }
        any_pos |= hp; any_neg |= hn;
        tc[k] = np_ | (nn_ << 8) | (ncross << 16);
        ncross_tot += ncross; npt += np_; nnt += nn_;
    }
    int status;
    if (ncross_tot == 0 && (!any_pos || !any_neg)) {
        int r;
        if (any_neg && !any_pos) r = opc_emit_whole(V, nv, T, nt, onv, ont, cap_v, cap_t, &sizes[2], &sizes[3]);
        else r = opc_emit_whole(V, nv, T, nt, opv, opt, cap_v, cap_t, &sizes[0], &sizes[1]);
        free(sg); free(tc);
        return r < 0 ? -1 : 2;
    }
    /* stage B: crossing edges, first occurrence owns, new vertices in order */
    opc_map em;
    opc_map_init(&em, ncross_tot);
    float* newv = (float*)malloc(12ull * (ncross_tot + 1));
    int n_new = 0;
    for (int k = 0; k < nt; ++k) {
        if (!((tc[k] >> 16) & 0xFF)) continue;
        for (int e = 0; e < 3; ++e) {
            int a = T[3 * k + e], b = T[3 * k + (e == 2 ? 0 : e + 1)];
            if ((sg[a] >= 0) == (sg[b] >= 0)) continue;
            int lo = a < b ? a : b, hi = a < b ? b : a, fresh;
            int slot = opc_map_put(&em, ((uint64_t)lo << 32) | (uint32_t)hi, &fresh);
            if (!fresh) continue;
            if (sg[lo] != 0 && sg[hi] != 0) {
                float ax = V[3 * lo], ay = V[3 * lo + 1], az = V[3 * lo + 2];
                float bx = V[3 * hi], by = V[3 * hi + 1], bz = V[3 * hi + 2];
                float d0 = fmaf(pa, ax, fmaf(pb, ay, fmaf(pc, az, pd)));
                float d1 = fmaf(pa, bx, fmaf(pb, by, fmaf(pc, bz, pd)));
                float t = d0 / (d0 - d1);
                newv[3 * n_new] = ax + t * (bx - ax);
                newv[3 * n_new + 1] = ay + t * (by - ay);
                newv[3 * n_new + 2] = az + t * (bz - az);
                em.v[slot] = nv + n_new++;
            } else {
                em.v[slot] = sg[lo] == 0 ? lo : hi;
            }
        }
    }
    const int nall = nv + n_new;
#define ISECT(a, b) em.v[opc_map_find(&em, (((uint64_t)((a) < (b) ? (a) : (b))) << 32) | (uint32_t)((a) < (b) ? (b) : (a)))]
#define PT(i) ((i) < nv ? V + 3 * (i) : newv + 3 * ((i) - nv))
    int* ptri = (int*)malloc(12ull * (npt + 1));
    int* ntri = (int*)malloc(12ull * (nnt + 1));
    int cp = 0, cn = 0;
    for (int k = 0; k < nt; ++k) {
        int i0 = T[3 * k], i1 = T[3 * k + 1], i2 = T[3 * k + 2];
        int s0 = sg[i0], s1 = sg[i1], s2 = sg[i2];
        int* P = ptri + 3 * cp;
        int* N = ntri + 3 * cn;
        int npk = tc[k] & 0xFF, nnk = (tc[k] >> 8) & 0xFF;
        int hp = (s0 > 0) | (s1 > 0) | (s2 > 0), hn = (s0 < 0) | (s1 < 0) | (s2 < 0);
        cp += npk; cn += nnk;
        if (!hp || !hn) {
            int* D = npk ? P : N;
            D[0] = i0; D[1] = i1; D[2] = i2;
            continue;
        }
        int vi[3] = {i0, i1, i2}, si[3] = {s0, s1, s2};
        int on = -1;
        for (int q = 0; q < 3; ++q) if (si[q] == 0) { on = q; break; }
        if (on >= 0) {
            int ov = vi[on], a = vi[(on + 1) % 3], b = vi[(on + 2) % 3], sa = si[(on + 1) % 3];
            int x = ISECT(a, b);
            if (sa > 0) { P[0] = ov; P[1] = a; P[2] = x; N[0] = ov; N[1] = x; N[2] = b; }
            else { N[0] = ov; N[1] = a; N[2] = x; P[0] = ov; P[1] = x; P[2] = b; }
        } else {
            int lone = 0;
            for (int q = 0; q < 3; ++q)
                if (si[q] != si[(q + 1) % 3] && si[q] != si[(q + 2) % 3]) { lone = q; break; }
            int lv = vi[lone], o1 = vi[(lone + 1) % 3], o2 = vi[(lone + 2) % 3];
            int x1 = ISECT(lv, o1), x2 = ISECT(lv, o2);
            int* L1 = si[lone] > 0 ? P : N;
            int* L2 = si[lone] > 0 ? N : P;
            L1[0] = lv; L1[1] = x1; L1[2] = x2;
            L2[0] = o1; L2[1] = x2; L2[2] = x1;
            L2[3] = o1; L2[4] = o2; L2[5] = x2;
        }
    }
    /* stage C: boundary of the positive side */
    int* mark = (int*)calloc(nall, 4);
    for (int k = 0; k < npt; ++k) for (int e = 0; e < 3; ++e) mark[ptri[3 * k + e]] |= 1;
    for (int k = 0; k < nnt; ++k) for (int e = 0; e < 3; ++e) mark[ntri[3 * k + e]] |= 2;
    int nbe = 0;
    for (int k = 0; k < npt; ++k)
        for (int e = 0; e < 3; ++e)
            nbe += mark[ptri[3 * k + e]] == 3 && mark[ptri[3 * k + (e == 2 ? 0 : e + 1)]] == 3;
    int* cap = NULL;
    int n_cap = 0, fail = 0;
    if (nbe > 0) {
        opc_map ek;
        opc_map_init(&ek, nbe);
        for (int k = 0; k < npt; ++k)
            for (int e = 0; e < 3; ++e) {
                int a = ptri[3 * k + e], b = ptri[3 * k + (e == 2 ? 0 : e + 1)], fr;
                if (mark[a] == 3 && mark[b] == 3) opc_map_put(&ek, ((uint64_t)a << 32) | (uint32_t)b, &fr);
            }
        int* bx = (int*)malloc(4ull * nbe);
        int* by = (int*)malloc(4ull * nbe);
        int nb = 0;
        for (int k = 0; k < npt; ++k)
            for (int e = 0; e < 3; ++e) {
                int a = ptri[3 * k + e], b = ptri[3 * k + (e == 2 ? 0 : e + 1)];
                if (mark[a] == 3 && mark[b] == 3 && opc_map_find(&ek, ((uint64_t)b << 32) | (uint32_t)a) < 0) {
                    bx[nb] = a; by[nb] = b; ++nb;
                }
            }
        opc_map_free(&ek);
        int* head = (int*)malloc(4ull * nall);
        for (int i = 0; i < nall; ++i) head[i] = 0x7fffffff;
        for (int j = nb - 1; j >= 0; --j) head[bx[j]] = j;
        int* succ = (int*)malloc(4ull * (nb + 1));
        int* indeg = (int*)calloc(nb + 1, 4);
        for (int j = 0; j < nb; ++j) succ[j] = head[by[j]] == 0x7fffffff ? -1 : head[by[j]];
        for (int j = 0; j < nb; ++j) if (succ[j] >= 0) indeg[succ[j]]++;
        for (int j = 0; j < nb && !fail; ++j) {
            if (succ[j] < 0) { fail = 1; break; }
            if (indeg[succ[j]] > 1) {
                int v = bx[succ[j]];
                for (int q = 0; q < nb; ++q)
                    if (bx[q] == v && indeg[q] == 0) { indeg[succ[j]]--; succ[j] = q; indeg[q] = 1; break; }
            }
        }
        for (int j = 0; j < nb && !fail; ++j) if (indeg[j] != 1) fail = 1;
        int* lab = (int*)malloc(4ull * (nb + 1));
        int* rank = (int*)malloc(4ull * (nb + 1));
        int nl = 0;
        int loop_start[OPC_MAX_LOOPS + 1], loop_parent[OPC_MAX_LOOPS], loop_hole[OPC_MAX_LOOPS];
        float loop_area[OPC_MAX_LOOPS];
        int* lid = (int*)malloc(4ull * (nb + 1));
        if (!fail) {
            for (int j = 0; j < nb; ++j) lab[j] = -1;
            for (int j = 0; j < nb; ++j) {
                if (lab[j] >= 0) continue;
                /* j is the lowest edge of an unvisited cycle */
                int r = 0, c = j;
                do { lab[c] = j; rank[c] = r++; c = succ[c]; } while (c != j && r <= nb);
                if (c != j) { fail = 1; break; }
                if (nl >= OPC_MAX_LOOPS) { fail = 1; break; }
                lid[j] = nl;
                loop_start[nl] = r; /* size for now */
                ++nl;
            }
        }
        int *poly = NULL, *wid = NULL, *prv = NULL, *nxt = NULL;
        opc_p2 *uv = NULL, *wuv = NULL;
        unsigned char* refl = NULL;
        if (!fail) {
            int off = 0;
            for (int l = 0; l < nl; ++l) { int sz = loop_start[l]; loop_start[l] = off; off += sz; }
            loop_start[nl] = off;
            poly = (int*)malloc(4ull * nb);
            for (int j = 0; j < nb; ++j) {
                int l = lid[lab[j]], sz = loop_start[l + 1] - loop_start[l];
                poly[loop_start[l] + (sz - 1 - rank[j])] = bx[j];
            }
            float anx = fabsf(pa), any_ = fabsf(pb), anz = fabsf(pc);
            int pu = (anx >= any_ && anx >= anz) ? 1 : 0;
            int pv = (anx >= any_ && anx >= anz) ? 2 : ((any_ >= anz) ? 2 : 1);
            uv = (opc_p2*)malloc(sizeof(opc_p2) * nb);
            for (int j = 0; j < nb; ++j) {
                const float* P3 = PT(poly[j]);
                uv[j].u = pu == 0 ? P3[0] : P3[1];
                uv[j].v = pv == 1 ? P3[1] : P3[2];
            }
            for (int l = 0; l < nl; ++l) {
                int b0 = loop_start[l], sz = loop_start[l + 1] - b0;
                double a = 0.0;
                for (int i = 0; i < sz; ++i) {
                    opc_p2 p = uv[b0 + i], q = uv[b0 + (i + 1 == sz ? 0 : i + 1)];
                    a += (double)p.u * (double)q.v - (double)q.u * (double)p.v;
                }
                loop_area[l] = (float)(a * 0.5);
            }
            for (int l = 0; l < nl; ++l) {
                opc_p2 q = uv[loop_start[l]];
                int par = -1;
                float best = 3.4e38f;
                for (int o = 0; o < nl; ++o) {
                    if (o == l) continue;
                    int b0 = loop_start[o], sz = loop_start[o + 1] - b0;
                    if (sz < 3) continue;
                    int cr = 0;
                    for (int i = 0; i < sz; ++i) {
                        opc_p2 a = uv[b0 + i], b = uv[b0 + (i + 1 == sz ? 0 : i + 1)];
                        if ((a.v <= q.v) == (b.v <= q.v)) continue;
                        float t = (q.v - a.v) / (b.v - a.v);
                        if (a.u + t * (b.u - a.u) > q.u) ++cr;
                    }
                    float ar = fabsf(loop_area[o]);
                    if ((cr & 1) && ar < best) { best = ar; par = o; }
                }
                loop_parent[l] = par;
            }
            for (int l = 0; l < nl && !fail; ++l) {
                int d = 0;
                for (int p = loop_parent[l]; p >= 0; p = loop_parent[p])
                    if (++d > nl) { fail = 1; break; }
                if (fail) break;
                loop_hole[l] = d & 1;
                if (!(d & 1) && d >= 2) loop_parent[l] = -1;
            }
        }
        if (!fail) {
            for (int l = 0; l < nl; ++l) {
                int b0 = loop_start[l], sz = loop_start[l + 1] - b0;
                int want_ccw = !loop_hole[l];
                if ((loop_area[l] > 0) != want_ccw)
                    for (int i = 0; i < sz / 2; ++i) {
                        int x = poly[b0 + i]; poly[b0 + i] = poly[b0 + sz - 1 - i]; poly[b0 + sz - 1 - i] = x;
                        opc_p2 y = uv[b0 + i]; uv[b0 + i] = uv[b0 + sz - 1 - i]; uv[b0 + sz - 1 - i] = y;
                    }
            }
            int W = nb + 2 * OPC_MAX_LOOPS + 4;
            wid = (int*)malloc(4ull * W);
            wuv = (opc_p2*)malloc(sizeof(opc_p2) * W);
            prv = (int*)malloc(4ull * W);
            nxt = (int*)malloc(4ull * W);
            refl = (unsigned char*)malloc(W);
            cap = (int*)malloc(12ull * W);
            for (int o = 0; o < nl && !fail; ++o) {
                if (loop_hole[o]) continue;
                int ob = loop_start[o], osz = loop_start[o + 1] - ob;
                if (osz < 3) continue;
                int n = osz;
                for (int i = 0; i < osz; ++i) { wid[i] = poly[ob + i]; wuv[i] = uv[ob + i]; }
                for (;;) {
                    int hbest = -1;
                    float hmax = -3.4e38f;
                    for (int h = 0; h < nl; ++h) {
                        if (!loop_hole[h] || loop_parent[h] != o) continue;
                        if (loop_area[h] == 3.4e38f) continue;
                        int hb = loop_start[h], hs = loop_start[h + 1] - hb;
                        float m = -3.4e38f;
                        for (int i = 0; i < hs; ++i) m = fmaxf(m, uv[hb + i].u);
                        if (m > hmax) { hmax = m; hbest = h; }
                    }
                    if (hbest < 0) break;
                    int hb = loop_start[hbest], hs = loop_start[hbest + 1] - hb;
                    uint64_t key = 0;
                    for (int i = 0; i < hs; ++i) {
                        uint64_t kk = ((uint64_t)opc_fkey(uv[hb + i].u) << 32) | (uint32_t)(0xFFFFFFFFu - (uint32_t)i);
                        if (kk > key) key = kk;
                    }
                    int mi = (int)(0xFFFFFFFFu - (uint32_t)key);
                    opc_p2 M = uv[hb + mi];
                    uint64_t ek2 = ~0ull;
                    for (int i = 0; i < n; ++i) {
                        opc_p2 a = wuv[i], b = wuv[i + 1 == n ? 0 : i + 1];
                        float dv = b.v - a.v;
                        if (fabsf(dv) < 1e-12f) continue;
                        float t = (M.v - a.v) / dv;
                        if (t < -1e-7f || t > 1.0f + 1e-7f) continue;
                        float x = a.u + t * (b.u - a.u) - M.u;
                        if (x <= 1e-12f) continue;
                        uint64_t kk = ((uint64_t)opc_fkey(x) << 32) | (uint32_t)i;
                        if (kk < ek2) ek2 = kk;
                    }
                    if (ek2 == ~0ull) { fail = 1; break; }
                    int be_ = (int)(uint32_t)ek2, bn = be_ + 1 == n ? 0 : be_ + 1;
                    int pp = wuv[be_].u >= wuv[bn].u ? be_ : bn;
                    {
                        opc_p2 a = wuv[be_], b = wuv[bn];
                        opc_p2 I;
                        I.u = a.u + (M.v - a.v) / (b.v - a.v) * (b.u - a.u);
                        I.v = M.v;
                        opc_p2 Pp = wuv[pp];
                        /* boundary points count as inside (a vertex on the bridge must be
                           bridged to instead), within a tolerance of the triangle's size */
                        float L = fmaxf(fmaxf(fabsf(Pp.u - M.u), fabsf(Pp.v - M.v)), I.u - M.u);
                        float tol = 1e-5f * L * L;
                        uint64_t vk = ~0ull, vt = ~0ull;
                        for (int i = 0; i < n; ++i) {
                            opc_p2 X = wuv[i];
                            if (X.u <= M.u) continue;  /* P itself competes: ties go to the nearest */
                            float d1 = opc_cross2(M, I, X), d2 = opc_cross2(I, Pp, X), d3 = opc_cross2(Pp, M, X);
                            int in = (d1 >= -tol && d2 >= -tol && d3 >= -tol) || (d1 <= tol && d2 <= tol && d3 <= tol);
                            if (!in) continue;
                            /* smallest angle to the ray, then nearest, then lowest index */
                            float tang = fabsf(X.v - M.v) / (X.u - M.u);
                            uint64_t k1 = (uint64_t)opc_fkey(tang);
                            uint64_t k2 = ((uint64_t)opc_fkey(X.u - M.u) << 32) | (uint32_t)i;
                            if (k1 < vt || (k1 == vt && k2 < vk)) { vt = k1; vk = k2; }
                        }
                        if (vk != ~0ull) pp = (int)(uint32_t)vk;
                    }
                    {
                        /* an earlier bridge may have duplicated the chosen vertex: take
                           the lowest copy whose interior wedge contains M */
                        opc_p2 Q = wuv[pp];
                        int best = -1;
                        for (int i = 0; i < n && best < 0; ++i) {
                            if (wuv[i].u != Q.u || wuv[i].v != Q.v) continue;
                            opc_p2 A = wuv[i == 0 ? n - 1 : i - 1], Bn = wuv[i + 1 == n ? 0 : i + 1];
                            int in;
                            if (opc_cross2(A, Q, Bn) > 0.0f) in = opc_cross2(Q, Bn, M) >= 0.0f && opc_cross2(Q, M, A) >= 0.0f;
                            else in = opc_cross2(Q, A, M) <= 0.0f || opc_cross2(Q, M, Bn) <= 0.0f;
                            if (in) best = i;
                        }
                        if (best >= 0) pp = best;
                    }
                    for (int i = n - 1; i > pp; --i) { wid[i + hs + 2] = wid[i]; wuv[i + hs + 2] = wuv[i]; }
                    for (int j = 0; j <= hs; ++j) {
                        int src = hb + (mi + j) % hs;
                        wid[pp + 1 + j] = poly[src];
                        wuv[pp + 1 + j] = uv[src];
                    }
                    wid[pp + hs + 2] = wid[pp];
                    wuv[pp + hs + 2] = wuv[pp];
                    loop_area[hbest] = 3.4e38f;
                    n += hs + 2;
                }
                if (fail) break;
                int got = opc_earclip(wid, n, wuv, prv, nxt, refl, cap + 3 * n_cap);
                if (got != n - 2) { fail = 1; break; }
                n_cap += got;
            }
        }
        free(bx); free(by); free(head); free(succ); free(indeg); free(lab); free(rank); free(lid);
        free(poly); free(uv); free(wid); free(wuv); free(prv); free(nxt); free(refl);
    }
    int rc = 0;
    if (fail) {
        status = 1;
        rc = opc_emit_whole(V, nv, T, nt, opv, opt, cap_v, cap_t, &sizes[0], &sizes[1]);
    } else {
        float anx = fabsf(pa), any_ = fabsf(pb), anz = fabsf(pc);
        float ccw_dot = (anx >= any_ && anx >= anz) ? pa : ((any_ >= anz) ? -pb : pc);
        int flip_pos = ccw_dot > 0.0f;
        int* remap = (int*)malloc(4ull * nall);
        for (int side = 0; side < 2 && rc == 0; ++side) {
            const int* st = side == 0 ? ptri : ntri;
            int sn = side == 0 ? npt : nnt;
            float* ov = side == 0 ? opv : onv;
            int* ot = side == 0 ? opt : ont;
            memset(remap, 0, 4ull * nall);
            for (int k = 0; k < sn; ++k) for (int e = 0; e < 3; ++e) remap[st[3 * k + e]] = 1;
            for (int k = 0; k < n_cap; ++k) for (int e = 0; e < 3; ++e) remap[cap[3 * k + e]] = 1;
            int c = 0;
            for (int i = 0; i < nall; ++i) { int u = remap[i]; remap[i] = u ? c : -1; c += u; }
            int ont_ = sn + n_cap;
            if (ont_ <= 0 || c <= 0) continue;
            if (c > cap_v || ont_ > cap_t) { rc = -1; break; }
            for (int i = 0; i < nall; ++i)
                if (remap[i] >= 0) { const float* P3 = PT(i); memcpy(ov + 3 * remap[i], P3, 12); }
            for (int k = 0; k < sn; ++k) for (int e = 0; e < 3; ++e) ot[3 * k + e] = remap[st[3 * k + e]];
            int flip = side == 0 ? flip_pos : !flip_pos;
            for (int k = 0; k < n_cap; ++k) {
                int a = cap[3 * k], b = cap[3 * k + 1], cc = cap[3 * k + 2];
                int* o = ot + 3 * (sn + k);
                o[0] = remap[a]; o[1] = remap[flip ? cc : b]; o[2] = remap[flip ? b : cc];
            }
            sizes[2 * side] = c;
            sizes[2 * side + 1] = ont_;
        }
        free(remap);
        status = (sizes[1] == 0 || sizes[3] == 0) ? 2 : 0;
    }
#undef ISECT
#undef PT
    free(cap); free(mark); free(ptri); free(ntri); free(newv); free(sg); free(tc);
    opc_map_free(&em);
    return rc < 0 ? -1 : status;
}
