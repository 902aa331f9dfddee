// host_common.hpp -- host-side helpers shared by the one-set runtime and the multi-set engine:
// thread-local error text, the reference's bounding sphere and super simplex, the priority
// permutation and device allocation.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/voronoids.h"

namespace vo_host {

inline thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define VO_HIP(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return ::vo_host::fail(VO_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// bounding_sphere, src/geometry.rs:99-141: per-axis min/max folds, center (hi+lo)/2, radius
// the larger corner distance, then one pass in input order multiplying the radius by 1.5 for
// every point the current sphere does not strictly contain (in_sphere, :91-97).
inline void bounding_sphere(int D, const double* pts, int64_t n, double* c, double* r) {
    double lo[3], hi[3];
    for (int i = 0; i < D; ++i) {
        lo[i] = INFINITY;
        hi[i] = -INFINITY;
        for (int64_t k = 0; k < n; ++k) {
            lo[i] = std::fmin(lo[i], pts[k * D + i]);
            hi[i] = std::fmax(hi[i], pts[k * D + i]);
        }
        c[i] = (hi[i] + lo[i]) / 2.0;
    }
    double su = 0.0, sl = 0.0;
    for (int i = 0; i < D; ++i) su += (hi[i] - c[i]) * (hi[i] - c[i]);
    for (int i = 0; i < D; ++i) sl += (lo[i] - c[i]) * (lo[i] - c[i]);
    const double ud = std::sqrt(su), ld = std::sqrt(sl);
    double rad = ud > ld ? ud : ld;
    for (int64_t k = 0; k < n; ++k) {
        double dist = 0.0;
        for (int i = 0; i < D; ++i) dist += (c[i] - pts[k * D + i]) * (c[i] - pts[k * D + i]);
        if (!(dist < rad * rad)) rad = rad * 1.5;
    }
    *r = rad;
}

// circumsphere, src/geometry.rs:2-56, in the reference's evaluation order.  Returns false when
// the 3D solve meets a zero pivot (the reference unwraps a failed LU solve there, :49).
//  2D (:2-22): perpendicular bisectors through the edge midpoints by their slopes; an edge
//     parallel to an axis makes an infinite or zero slope, and the IEEE results of the formula
//     are kept (NaN when the first or second edge is horizontal).
//  3D (:24-56): rows D_i = v_i - v_0, Mid_i = (v_i + v_0) / 2, right-hand side
//     planes_i = ((0 + D_i0 Mid_i0) + D_i1 Mid_i1) + D_i2 Mid_i2 (nalgebra's column_sum starts
//     from zeros), center = the LU solve of D c = planes as nalgebra 0.32 computes it: for each
//     column the first row of largest |entry| is the pivot, the rows are swapped, the entries
//     below are scaled by 1 / pivot and the rest updated by (-u) * l + a (a product, then a
//     sum); the solve applies the swaps in order, then unit-lower forward and upper backward
//     substitution, each column subtracted as (-coeff) * entry + b.  The radius is measured
//     from v_0, summing x, y, z in order.
inline bool ref_circumsphere(int D, const double* v, double* center, double* radius) {
    if (D == 2) {
        const double x1 = v[0], y1 = v[1], x2 = v[2], y2 = v[3], x3 = v[4], y3 = v[5];
        const double d0 = (x1 + x2) / 2.0, d1 = (y1 + y2) / 2.0;
        const double e0 = (x2 + x3) / 2.0, e1 = (y2 + y3) / 2.0;
        const double m_ab = (y2 - y1) / (x2 - x1);
        const double m_bc = (y3 - y2) / (x3 - x2);
        const double m_d = -1. / m_ab;
        const double m_e = -1. / m_bc;
        const double x = (m_d * d0 - m_e * e0 + e1 - d1) / (m_d - m_e);
        const double y = m_d * (x - d0) + d1;
        center[0] = x;
        center[1] = y;
        *radius = std::sqrt((x - x1) * (x - x1) + (y - y1) * (y - y1));
        return true;
    }
    double a[3][3], b[3];
    for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) {
            a[i][k] = v[(i + 1) * 3 + k] - v[k];
            const double mid = (v[(i + 1) * 3 + k] + v[k]) / 2.0;
            s += a[i][k] * mid;
        }
        b[i] = s;
    }
    int swap_i[3], swap_p[3], n_swaps = 0;
    for (int i = 0; i < 3; ++i) {
        int piv = i;
        double best = std::fabs(a[i][i]);
        for (int r = i + 1; r < 3; ++r)
            if (std::fabs(a[r][i]) > best) {
                best = std::fabs(a[r][i]);
                piv = r;
            }
        const double diag = a[piv][i];
        if (diag == 0.0) continue;
        if (piv != i) {
            swap_i[n_swaps] = i;
            swap_p[n_swaps++] = piv;
            for (int k = 0; k < 3; ++k) std::swap(a[i][k], a[piv][k]);
        }
        const double inv = 1.0 / diag;
        for (int r = i + 1; r < 3; ++r) a[r][i] = a[r][i] * inv;
        for (int k = i + 1; k < 3; ++k)
            for (int r = i + 1; r < 3; ++r) a[r][k] = (-a[i][k]) * a[r][i] + a[r][k];
    }
    double x[3] = {b[0], b[1], b[2]};
    for (int s = 0; s < n_swaps; ++s) std::swap(x[swap_i[s]], x[swap_p[s]]);
    for (int i = 0; i < 2; ++i) {
        const double coeff = x[i] / 1.0;
        for (int r = i + 1; r < 3; ++r) x[r] = (-coeff) * a[r][i] + x[r];
    }
    for (int i = 2; i >= 0; --i) {
        if (a[i][i] == 0.0) {
            center[0] = center[1] = center[2] = *radius = NAN;
            return false;
        }
        const double coeff = x[i] / a[i][i];
        x[i] = coeff;
        for (int r = 0; r < i; ++r) x[r] = (-coeff) * a[r][i] + x[r];
    }
    for (int k = 0; k < 3; ++k) center[k] = x[k];
    *radius = std::sqrt((v[0] - x[0]) * (v[0] - x[0]) + (v[1] - x[1]) * (v[1] - x[1]) + (v[2] - x[2]) * (v[2] - x[2]));
    return true;
}

// in_sphere, src/geometry.rs:91-97: squared distance summed over the axes in order, strictly
// below radius * radius
inline bool ref_in_sphere(int D, const double* p, const double* c, double radius) {
    double distance = 0.0;
    for (int i = 0; i < D; ++i) distance += (c[i] - p[i]) * (c[i] - p[i]);
    return distance < radius * radius;
}

// center and radius the super simplex is built from: the reference's bounding sphere, except
// that a zero radius (one point, or all points equal), for which the reference has no
// super simplex (all its vertices coincide), is replaced by 1
inline void set_frame(int D, const double* pts, int64_t n, double* c, double* r) {
    c[0] = c[1] = c[2] = 0.0;
    *r = 1.0;
    if (n > 0) bounding_sphere(D, pts, n, c, r);
    if (!(*r > 0.0)) *r = 1.0;
}

// super vertices of DelaunayTree::new: 3D src/delaunay_tree.rs:393-406, 2D :548-558
inline void super_vertices(int D, const double* c, double R, double out[4][3]) {
    const double pi = 3.14159265358979323846;
    const double c1 = std::cos(2. * pi / 3.), s1 = std::sin(2. * pi / 3.);
    const double c2 = std::cos(4. * pi / 3.), s2 = std::sin(4. * pi / 3.);
    if (D == 3) {
        const double v[4][3] = {{c[0], c[1], c[2] + R},
                                {c[0] + R, c[1], c[2] - R},
                                {c[0] + R * c1, c[1] + R * s1, c[2] - R},
                                {c[0] + R * c2, c[1] + R * s2, c[2] - R}};
        std::memcpy(out, v, sizeof(v));
    } else {
        const double v[3][3] = {{c[0] + R, c[1], 0}, {c[0] + R * c1, c[1] + R * s1, 0}, {c[0] + R * c2, c[1] + R * s2, 0}};
        std::memcpy(out, v, sizeof(v));
    }
}

// vertex order of the first simplex giving it positive orientation (well conditioned: plain
// f64 decides it); tet0[3] = -1 in 2D
inline void first_simplex_order(int D, const double sv[4][3], int tet0[4]) {
    tet0[0] = 0; tet0[1] = 1; tet0[2] = 2; tet0[3] = 3;
    if (D == 3) {
        double u[3], v[3], w[3];
        for (int k = 0; k < 3; ++k) {
            u[k] = sv[1][k] - sv[0][k];
            v[k] = sv[2][k] - sv[0][k];
            w[k] = sv[3][k] - sv[0][k];
        }
        const double det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
                           u[2] * (v[0] * w[1] - v[1] * w[0]);
        if (det < 0) { tet0[1] = 2; tet0[2] = 1; }
    } else {
        const double det = (sv[1][0] - sv[0][0]) * (sv[2][1] - sv[0][1]) - (sv[1][1] - sv[0][1]) * (sv[2][0] - sv[0][0]);
        if (det < 0) { tet0[1] = 2; tet0[2] = 1; }
        tet0[3] = -1;
    }
}

// simplex record bytes (the device layout of kernels.hpp: Rec<3> / Rec<2>)
inline int record_bytes(int D) { return D == 3 ? 128 : 64; }

// the first simplex's record: vertex ids v0 + tet0[k], hull adjacency, fp32 coordinates of the
// super vertices translated by c, reservations at the largest key
inline void root_record(int D, const int tet0[4], int v0, const double sv[4][3], const double* c, unsigned char* out) {
    const int B = record_bytes(D), M = D + 1;
    // the device's record layout (Rec<D> in kernels.hpp): coordinates packed in vertex order
    const int V = 0, A = D == 3 ? 16 : 12, X = D == 3 ? 32 : 24, K = D == 3 ? 80 : 48;
    std::memset(out, 0, (size_t)B);
    int* v = reinterpret_cast<int*>(out + V);
    int* a = reinterpret_cast<int*>(out + A);
    float* x = reinterpret_cast<float*>(out + X);
    for (int k = 0; k < M; ++k) {
        v[k] = v0 + tet0[k];
        a[k] = -1;
        for (int j = 0; j < D; ++j) x[D * k + j] = (float)(sv[tet0[k]][j] - c[j]);
    }
    std::memset(out + K, 0xFF, 16);
}

inline uint64_t splitmix_next(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// insertion priority order: Fisher-Yates over splitmix64(seed); perm[i] = input index of the
// point with priority i
inline std::vector<int32_t> priority_order(int64_t n, uint64_t seed) {
    std::vector<int32_t> perm(n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    uint64_t sm = seed;
    for (int64_t i = n - 1; i > 0; --i) {
        const uint64_t j = (uint64_t)(((unsigned __int128)splitmix_next(sm) * (uint64_t)(i + 1)) >> 64);
        std::swap(perm[i], perm[j]);
    }
    return perm;
}

// 3D / 2D Morton code of a point in the box [c - r, c + r]^D (21 / 32 bits per axis)
inline uint64_t morton_code(int D, const double* x, const double* c, double r) {
    const int bits = D == 3 ? 21 : 32;
    const double scale = (double)((1ull << bits) - 1) / (2.0 * r);
    uint64_t code = 0;
    uint64_t q[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
        double t = (x[k] - (c[k] - r)) * scale;
        t = t < 0.0 ? 0.0 : (t > (double)((1ull << bits) - 1) ? (double)((1ull << bits) - 1) : t);
        q[k] = (uint64_t)t;
    }
    for (int b = bits - 1; b >= 0; --b)
        for (int k = D - 1; k >= 0; --k) code = (code << 1) | ((q[k] >> b) & 1ull);
    return code;
}

// Insertion priority with spatial coherence at fine scale: the seeded random order, cut into
// consecutive chunks of max(1, position / divisor) points, each chunk sorted along the Morton
// curve.  A round's window (about 1/20 of the points inserted so far) spans many whole chunks,
// so it is still a uniform random sample of the remaining points (same conflicts, same
// parallelism), while neighboring window slots -- the points of one wavefront -- lie close
// together, and so do the simplices they create and later revisit.
inline std::vector<int32_t> priority_order_local(int D, const double* pts, int64_t n, uint64_t seed, const double* c,
                                                 double r, int divisor = 256) {
    std::vector<int32_t> perm = priority_order(n, seed);
    std::vector<std::pair<uint64_t, int32_t>> key;
    for (int64_t s0 = 0; s0 < n;) {
        const int64_t len = std::min<int64_t>(n - s0, std::max<int64_t>(1, s0 / divisor));
        key.resize((size_t)len);
        for (int64_t i = 0; i < len; ++i) {
            const int32_t v = perm[s0 + i];
            key[i] = {morton_code(D, pts + (size_t)v * D, c, r), v};
        }
        std::sort(key.begin(), key.end());
        for (int64_t i = 0; i < len; ++i) perm[s0 + i] = key[i].second;
        s0 += len;
    }
    return perm;
}

// finest hint-grid level for n points: about 4 points a cell
inline int hint_levels(int D, int64_t n) {
    const double target = n / 4.0 > 1.0 ? n / 4.0 : 1.0;
    int L = (int)std::floor(std::log2(target) / D + 0.5);
    const int lmax = D == 3 ? 8 : 12;
    return L < 0 ? 0 : (L > lmax ? lmax : L);
}
inline int64_t hint_cells(int D, int L) {
    return D == 3 ? (((int64_t)1 << (3 * (L + 1))) - 1) / 7 : (((int64_t)1 << (2 * (L + 1))) - 1) / 3;
}

template <class T>
int dev_alloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
    if (e != hipSuccess) return fail(VO_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

// canonical output from live simplices in the reference id space (input point i = 2M + i):
// T = all-real simplices as sorted input indices, lexicographic; E = real-real pairs of any
// simplex, sorted, unique.  Arrays are malloc'd.
int canonical_from_simplices(int M, const int32_t* s, int64_t nl, int32_t** T, int64_t* n_T, int32_t** E,
                             int64_t* n_E);

}  // namespace vo_host
