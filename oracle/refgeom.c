/*
 * refgeom.c -- the reference's floating-point geometry, restated operation by operation.
 *
 * TEST INFRASTRUCTURE ONLY (oracle).  Compiled with -ffp-contract=off: Rust never fuses a
 * multiply and an add, so neither may this file.
 */
#include <math.h>
#include <stdint.h>
#include "oracle.h"

/* src/geometry.rs:2-22 -- perpendicular-bisector slope formula */
static void circumsphere_2d(const double v[3][2], double* c, double* r) {
    double x1 = v[0][0], y1 = v[0][1], x2 = v[1][0], y2 = v[1][1], x3 = v[2][0], y3 = v[2][1];
    double d0 = (x1 + x2) / 2.0, d1 = (y1 + y2) / 2.0;
    double e0 = (x2 + x3) / 2.0, e1 = (y2 + y3) / 2.0;
    double m_ab = (y2 - y1) / (x2 - x1);
    double m_bc = (y3 - y2) / (x3 - x2);
    double m_d = -1. / m_ab;
    double m_e = -1. / m_bc;
    double x = (m_d * d0 - m_e * e0 + e1 - d1) / (m_d - m_e);
    double y = m_d * (x - d0) + d1;
    c[0] = x;
    c[1] = y;
    *r = sqrt((x - x1) * (x - x1) + (y - y1) * (y - y1));
}

/* src/geometry.rs:24-56 -- rows D_i = v_{i+1} - v_0, rhs_i = sum_j D_ij * (v_{i+1,j} + v_0j)/2
 * (nalgebra column_sum: ((0 + m_i0) + m_i1) + m_i2), solved by nalgebra 0.32's LU:
 * partial pivoting on the first largest |a| (icamax, strict >), multipliers a * (1/pivot),
 * rank-1 updates y = fl(-(u_k) * l_j) + y (axpy, no fusion), then unit-lower forward and
 * column-oriented upper back substitution.  Returns 0 when a pivot is exactly zero (the
 * reference panics in unwrap()). */
static int circumsphere_3d(const double v[4][3], double* c, double* r) {
    double A[3][3], b[3];
    for (int i = 0; i < 3; ++i) {
        double m[3];
        for (int j = 0; j < 3; ++j) {
            A[i][j] = v[i + 1][j] - v[0][j];
            double mid = (v[i + 1][j] + v[0][j]) / 2.0;
            m[j] = A[i][j] * mid;
        }
        double s = 0.0;
        s = s + m[0];
        s = s + m[1];
        s = s + m[2];
        b[i] = s;
    }
    /* LU in place, transpositions recorded in order */
    int perm_i[3], perm_p[3], np = 0;
    for (int i = 0; i < 3; ++i) {
        int piv = i;
        double best = fabs(A[i][i]);
        for (int k = i + 1; k < 3; ++k)
            if (fabs(A[k][i]) > best) { best = fabs(A[k][i]); piv = k; }
        double diag = A[piv][i];
        if (diag == 0.0) continue;
        if (piv != i) {
            perm_i[np] = i; perm_p[np] = piv; np++;
            for (int j = 0; j < 3; ++j) { double t = A[i][j]; A[i][j] = A[piv][j]; A[piv][j] = t; }
        }
        double inv = 1.0 / diag;
        for (int k = i + 1; k < 3; ++k) A[k][i] = A[k][i] * inv;
        for (int j = i + 1; j < 3; ++j) {
            double a = -A[i][j];
            for (int k = i + 1; k < 3; ++k) {
                double ax = a * A[k][i];
                A[k][j] = ax + A[k][j];
            }
        }
    }
    for (int t = 0; t < np; ++t) { double s = b[perm_i[t]]; b[perm_i[t]] = b[perm_p[t]]; b[perm_p[t]] = s; }
    for (int i = 0; i < 2; ++i) {
        double coeff = b[i] / 1.0;
        double a = -coeff;
        for (int k = i + 1; k < 3; ++k) b[k] = a * A[k][i] + b[k];
    }
    for (int i = 2; i >= 0; --i) {
        double diag = A[i][i];
        if (diag == 0.0) return 0;
        double coeff = b[i] / diag;
        b[i] = coeff;
        double a = -coeff;
        for (int k = 0; k < i; ++k) b[k] = a * A[k][i] + b[k];
    }
    c[0] = b[0]; c[1] = b[1]; c[2] = b[2];
    *r = sqrt((v[0][0] - c[0]) * (v[0][0] - c[0]) + (v[0][1] - c[1]) * (v[0][1] - c[1]) +
              (v[0][2] - c[2]) * (v[0][2] - c[2]));
    return 1;
}

/* src/geometry.rs:58-88 */
void or_circumsphere(int dim, const double* verts, double* center, double* radius) {
    if (dim == 2) {
        double v[3][2];
        for (int i = 0; i < 3; ++i) { v[i][0] = verts[i * 2]; v[i][1] = verts[i * 2 + 1]; }
        circumsphere_2d(v, center, radius);
    } else {
        double v[4][3];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 3; ++j) v[i][j] = verts[i * 3 + j];
        if (!circumsphere_3d(v, center, radius)) {
            center[0] = center[1] = center[2] = NAN;
            *radius = NAN;
        }
    }
}

/* src/geometry.rs:91-97 -- strict test against a stored sphere */
int or_in_sphere(int dim, const double* p, const double* center, double radius) {
    double distance = 0.0;
    for (int i = 0; i < dim; ++i) distance += (center[i] - p[i]) * (center[i] - p[i]);
    return distance < radius * radius;
}

/* src/geometry.rs:99-141 -- per-axis min/max folds, center = (hi+lo)/2, radius = the larger
 * corner distance, then one pass in input order growing the radius by 1.5 for every point
 * the current sphere does not strictly contain. */
void or_bounding_sphere(int dim, const double* pts, int64_t n, double* center, double* radius) {
    double lo[3], hi[3];
    for (int i = 0; i < dim; ++i) {
        lo[i] = INFINITY;
        hi[i] = -INFINITY;
        for (int64_t k = 0; k < n; ++k) {
            double x = pts[k * dim + i];
            lo[i] = fmin(lo[i], x);
            hi[i] = fmax(hi[i], x);
        }
        center[i] = (hi[i] + lo[i]) / 2.0;
    }
    double su = 0.0, sl = 0.0;
    for (int i = 0; i < dim; ++i) su += (hi[i] - center[i]) * (hi[i] - center[i]);
    for (int i = 0; i < dim; ++i) sl += (lo[i] - center[i]) * (lo[i] - center[i]);
    double ud = sqrt(su), ld = sqrt(sl);
    double r = ud > ld ? ud : ld;
    for (int64_t k = 0; k < n; ++k)
        if (!or_in_sphere(dim, pts + k * dim, center, r)) r = r * 1.5;
    *radius = r;
}

void or_trig_constants(double* out4) {
    const double pi = 3.14159265358979323846;
    out4[0] = cos(2. * pi / 3.);
    out4[1] = sin(2. * pi / 3.);
    out4[2] = cos(4. * pi / 3.);
    out4[3] = sin(4. * pi / 3.);
}

/* DelaunayTree::new vertex coordinates: 3D src/delaunay_tree.rs:392-412 (super 0..3, ghosts
 * 4,5,6 = copies of 0, ghost 7 = copy of 1); 2D :547-566 (super 0..2, ghosts 3..5 = copies). */
void or_super_vertices(int dim, const double* pts, int64_t n, double* out) {
    double c[3], R, t[4];
    or_bounding_sphere(dim, pts, n, c, &R);
    R *= 10.0;
    or_trig_constants(t);
    if (dim == 3) {
        double v[4][3] = {
            {c[0], c[1], c[2] + R},
            {c[0] + R, c[1], c[2] - R},
            {c[0] + R * t[0], c[1] + R * t[1], c[2] - R},
            {c[0] + R * t[2], c[1] + R * t[3], c[2] - R},
        };
        double g[4][3] = {
            {0. + c[0], 0. + c[1], R + c[2]},
            {0. + c[0], 0. + c[1], R + c[2]},
            {0. + c[0], 0. + c[1], R + c[2]},
            {c[0] + R, c[1], c[2] - R},
        };
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 3; ++j) { out[i * 3 + j] = v[i][j]; out[(4 + i) * 3 + j] = g[i][j]; }
    } else {
        double v[3][2] = {
            {c[0] + R, c[1]},
            {c[0] + R * t[0], c[1] + R * t[1]},
            {c[0] + R * t[2], c[1] + R * t[3]},
        };
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 2; ++j) { out[i * 2 + j] = v[i][j]; out[(3 + i) * 2 + j] = v[i][j]; }
    }
}
