/* hausdorff_oracle.c - CPU oracle for sampled directed Hausdorff distances.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py's CPU
 * baseline leg); never called by the product path.
 *
 * Restates the distance semantics of the reference's Hausdorff metric
 * (cuda/hausdorff.cuh:68-115 point-triangle distance, 1453 bidirectional max;
 * tests/ref_hausdorff.cpp:128-183 point-to-mesh distance) with an exact
 * nearest triangle by FP64 brute force instead of a BVH or a k-NN search.
 * The point-triangle distance is computed by projection onto the plane with an
 * inside test, falling back to the three edge segments (a different
 * construction from the GPU's Voronoi-region closest point).
 *
 * Pinned by closed forms in tests/test_oracle_hausdorff.py (points above a
 * face, beyond a vertex, beyond an edge; identical meshes -> 0; concentric
 * cubes -> half the size difference) and by dense brute-force sampling of the
 * triangle.
 */
#include <math.h>
#include <stdlib.h>

static double seg_d2(const double* p, const double* a, const double* b) {
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    double l2 = ab[0] * ab[0] + ab[1] * ab[1] + ab[2] * ab[2];
    double t = l2 > 0 ? (ap[0] * ab[0] + ap[1] * ab[1] + ap[2] * ab[2]) / l2 : 0.0;
    if (t < 0) t = 0;
    if (t > 1) t = 1;
    double d[3] = {ap[0] - t * ab[0], ap[1] - t * ab[1], ap[2] - t * ab[2]};
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

double orc_pt_tri_d2(const double* p, const double* a, const double* b, const double* c) {
    double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double n[3] = {ab[1] * ac[2] - ab[2] * ac[1], ab[2] * ac[0] - ab[0] * ac[2], ab[0] * ac[1] - ab[1] * ac[0]};
    double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    if (n2 > 0) {
        double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
        double h = (ap[0] * n[0] + ap[1] * n[1] + ap[2] * n[2]) / n2;
        double q[3] = {p[0] - h * n[0], p[1] - h * n[1], p[2] - h * n[2]};
        /* inside iff q is on the inner side of all three edges */
        const double* v[3] = {a, b, c};
        int inside = 1;
        for (int e = 0; e < 3 && inside; ++e) {
            const double* u = v[e];
            const double* w = v[(e + 1) % 3];
            double ew[3] = {w[0] - u[0], w[1] - u[1], w[2] - u[2]};
            double uq[3] = {q[0] - u[0], q[1] - u[1], q[2] - u[2]};
            double cr[3] = {ew[1] * uq[2] - ew[2] * uq[1], ew[2] * uq[0] - ew[0] * uq[2], ew[0] * uq[1] - ew[1] * uq[0]};
            if (cr[0] * n[0] + cr[1] * n[1] + cr[2] * n[2] < 0) inside = 0;
        }
        if (inside) return h * h * n2;
    }
    double d0 = seg_d2(p, a, b), d1 = seg_d2(p, b, c), d2 = seg_d2(p, c, a);
    double m = d0 < d1 ? d0 : d1;
    return m < d2 ? m : d2;
}

/* Squared distance from each point to the mesh (brute force), OpenMP. */
void orc_points_mesh_d2(const float* pts, int n, const float* verts, const int* tris, int nt, double* out_d2) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double p[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        double best = 1e300;
        for (int t = 0; t < nt; ++t) {
            double a[3], b[3], c[3];
            for (int k = 0; k < 3; ++k) {
                a[k] = verts[3 * tris[3 * t] + k];
                b[k] = verts[3 * tris[3 * t + 1] + k];
                c[k] = verts[3 * tris[3 * t + 2] + k];
            }
            double d = orc_pt_tri_d2(p, a, b, c);
            if (d < best) best = d;
        }
        out_d2[i] = best;
    }
}

/* Directed Hausdorff: max over points of the nearest distance (not squared). */
double orc_directed_hausdorff(const float* pts, int n, const float* verts, const int* tris, int nt) {
    if (n <= 0 || nt <= 0) return 0.0;
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    if (!d) return -1.0;
    orc_points_mesh_d2(pts, n, verts, tris, nt, d);
    double m = 0.0;
    for (int i = 0; i < n; ++i)
        if (d[i] > m) m = d[i];
    free(d);
    return sqrt(m);
}

/* Batch of directed distances for sample sets (n_pairs x n_samples x 3) to
 * meshes given by offsets; OpenMP over pairs, serial inside. */
void orc_directed_batch(const float* samples, int n_samples, const float* verts, const int* voff, const int* tris,
                        const int* toff, int n_pairs, double* out) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n_pairs; ++i) {
        const float* v = verts + 3 * (size_t)voff[i];
        const int* t = tris + 3 * (size_t)toff[i];
        int nt = toff[i + 1] - toff[i];
        const float* s = samples + 3 * (size_t)n_samples * i;
        double m = 0.0;
        for (int j = 0; j < n_samples; ++j) {
            double p[3] = {s[3 * j], s[3 * j + 1], s[3 * j + 2]};
            double best = 1e300;
            for (int k = 0; k < nt; ++k) {
                double a[3], b[3], c[3];
                for (int q = 0; q < 3; ++q) {
                    a[q] = v[3 * t[3 * k] + q];
                    b[q] = v[3 * t[3 * k + 1] + q];
                    c[q] = v[3 * t[3 * k + 2] + q];
                }
                double d = orc_pt_tri_d2(p, a, b, c);
                if (d < best) best = d;
            }
            if (best > m) m = best;
        }
        out[i] = sqrt(m);
    }
}
