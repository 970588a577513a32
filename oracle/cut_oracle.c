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
