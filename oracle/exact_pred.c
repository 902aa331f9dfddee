/*
 * exact_pred.c -- exact orientation / in-circle / in-sphere signs for the oracle.
 *
 * TEST INFRASTRUCTURE ONLY (oracle).  Method: every f64 coordinate is an integer times a
 * power of two; all coordinates of one predicate call are scaled to a common exponent and
 * the determinant is evaluated on arbitrary-precision integers (schoolbook, 32-bit limbs).
 * A double evaluation with a generous a-priori error bound answers first; the integer path
 * runs only when that bound cannot decide.  This is deliberately a different method from
 * the product's FP64 expansion tier, so a mistake in one does not hide in the other.
 *
 * The reference (src/geometry.rs:91-97) has no exact predicate; these signs define the
 * exact Delaunay triangulation DT(P u S) the reference computes whenever its floating-point
 * test agrees with the exact one.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "oracle.h"

#define BI_N 400 /* 12800 bits: degree-5 products of 2100-bit integers plus carries */

typedef struct {
    int n;          /* limbs in use, no leading zero limbs; 0 for zero */
    int s;          /* sign: -1, 0, +1 */
    uint32_t d[BI_N];
} bi;

static void bi_zero(bi* r) { r->n = 0; r->s = 0; }

static int mag_cmp(const bi* a, const bi* b) {
    if (a->n != b->n) return a->n < b->n ? -1 : 1;
    for (int i = a->n - 1; i >= 0; --i)
        if (a->d[i] != b->d[i]) return a->d[i] < b->d[i] ? -1 : 1;
    return 0;
}

static void mag_norm(bi* r) {
    while (r->n > 0 && r->d[r->n - 1] == 0) r->n--;
    if (r->n == 0) r->s = 0;
}

/* r = |a| + |b| (r distinct from a, b) */
static void mag_add(bi* r, const bi* a, const bi* b) {
    int n = a->n > b->n ? a->n : b->n;
    uint64_t c = 0;
    for (int i = 0; i < n; ++i) {
        uint64_t x = (i < a->n ? a->d[i] : 0) + (uint64_t)(i < b->n ? b->d[i] : 0) + c;
        r->d[i] = (uint32_t)x;
        c = x >> 32;
    }
    r->n = n;
    if (c) r->d[r->n++] = (uint32_t)c;
}

/* r = |a| - |b|, requires |a| >= |b| (r distinct from a, b) */
static void mag_sub(bi* r, const bi* a, const bi* b) {
    int64_t br = 0;
    for (int i = 0; i < a->n; ++i) {
        int64_t x = (int64_t)a->d[i] - (i < b->n ? b->d[i] : 0) - br;
        br = x < 0;
        r->d[i] = (uint32_t)(x + (br << 32));
    }
    r->n = a->n;
    r->s = 1;
    mag_norm(r);
}

/* r = a + sb * b (sb = +1 or -1), r distinct from a, b */
static void bi_addsub(bi* r, const bi* a, const bi* b, int sb) {
    int bs = b->s * sb;
    if (b->s == 0) { *r = *a; return; }
    if (a->s == 0) { *r = *b; r->s = bs; return; }
    if (a->s == bs) {
        mag_add(r, a, b);
        r->s = a->s;
        return;
    }
    int c = mag_cmp(a, b);
    if (c == 0) { bi_zero(r); return; }
    if (c > 0) { mag_sub(r, a, b); r->s = a->s; }
    else       { mag_sub(r, b, a); r->s = bs; }
}

static void bi_mul(bi* r, const bi* a, const bi* b) {
    if (a->s == 0 || b->s == 0) { bi_zero(r); return; }
    int n = a->n + b->n;
    memset(r->d, 0, sizeof(uint32_t) * (size_t)n);
    for (int i = 0; i < a->n; ++i) {
        uint64_t c = 0, ai = a->d[i];
        for (int j = 0; j < b->n; ++j) {
            uint64_t x = ai * b->d[j] + r->d[i + j] + c;
            r->d[i + j] = (uint32_t)x;
            c = x >> 32;
        }
        r->d[i + b->n] = (uint32_t)c;
    }
    r->n = n;
    r->s = a->s * b->s;
    mag_norm(r);
}

/* ---- conversion of doubles to integers at a common exponent ---- */
typedef struct { uint64_t m; int e; int neg; } split_t;   /* |x| = m * 2^e */

static split_t split_double(double x) {
    split_t s = {0, 0, 0};
    if (x == 0.0) return s;
    int E;
    double f = frexp(fabs(x), &E);                 /* |x| = f 2^E, f in [0.5, 1) */
    uint64_t m = (uint64_t)ldexp(f, 53);           /* exact: f has <= 53 significant bits */
    int e = E - 53;
    while ((m & 1u) == 0) { m >>= 1; e++; }
    s.m = m; s.e = e; s.neg = x < 0;
    return s;
}

static void bi_from_split(bi* r, split_t s, int emin) {
    if (s.m == 0) { bi_zero(r); return; }
    int sh = s.e - emin;                           /* >= 0 */
    int limb = sh / 32, bit = sh % 32;
    int n = limb + 3;
    memset(r->d, 0, sizeof(uint32_t) * (size_t)n);
    unsigned __int128 v = (unsigned __int128)s.m << bit;
    r->d[limb] = (uint32_t)v;
    r->d[limb + 1] = (uint32_t)(v >> 32);
    r->d[limb + 2] = (uint32_t)(v >> 64);
    r->n = n;
    r->s = s.neg ? -1 : 1;
    mag_norm(r);
}

/* k points of dimension dim -> integer coordinates at the common exponent */
static void to_ints(int k, int dim, const double* const* p, bi* out /* k*dim */) {
    split_t sp[5 * 3];
    int emin = 1 << 30;
    for (int i = 0; i < k; ++i)
        for (int j = 0; j < dim; ++j) {
            sp[i * dim + j] = split_double(p[i][j]);
            if (sp[i * dim + j].m && sp[i * dim + j].e < emin) emin = sp[i * dim + j].e;
        }
    for (int i = 0; i < k * dim; ++i) bi_from_split(&out[i], sp[i], emin);
}

/* det of 2x2 [[a b][c d]] */
static void det2(bi* r, const bi* a, const bi* b, const bi* c, const bi* d) {
    bi t1, t2;
    bi_mul(&t1, a, d);
    bi_mul(&t2, b, c);
    bi_addsub(r, &t1, &t2, -1);
}

/* det of 3x3 rows r0, r1, r2 (each 3 entries) */
static void det3(bi* r, const bi* r0, const bi* r1, const bi* r2) {
    bi m0, m1, m2, t0, t1, t2, s;
    det2(&m0, &r1[1], &r1[2], &r2[1], &r2[2]);
    det2(&m1, &r1[0], &r1[2], &r2[0], &r2[2]);
    det2(&m2, &r1[0], &r1[1], &r2[0], &r2[1]);
    bi_mul(&t0, &r0[0], &m0);
    bi_mul(&t1, &r0[1], &m1);
    bi_mul(&t2, &r0[2], &m2);
    bi_addsub(&s, &t0, &t1, -1);
    bi_addsub(r, &s, &t2, +1);
}

int or_orient2d_bigint(const double* a, const double* b, const double* c) {
    const double* p[3] = {a, b, c};
    bi X[6], u0, u1, v0, v1, r;
    to_ints(3, 2, p, X);
    bi_addsub(&u0, &X[2], &X[0], -1); bi_addsub(&u1, &X[3], &X[1], -1);
    bi_addsub(&v0, &X[4], &X[0], -1); bi_addsub(&v1, &X[5], &X[1], -1);
    det2(&r, &u0, &u1, &v0, &v1);
    return r.s;
}

int or_orient3d_bigint(const double* a, const double* b, const double* c, const double* d) {
    const double* p[4] = {a, b, c, d};
    bi X[12], R[9], r;
    to_ints(4, 3, p, X);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) bi_addsub(&R[i * 3 + j], &X[(i + 1) * 3 + j], &X[j], -1);
    det3(&r, &R[0], &R[3], &R[6]);
    return r.s;
}

int or_incircle_bigint(const double* a, const double* b, const double* c, const double* d) {
    const double* p[4] = {a, b, c, d};
    bi X[8], R[9], t0, t1, s, r;
    to_ints(4, 2, p, X);
    for (int i = 0; i < 3; ++i) {
        bi_addsub(&R[i * 3 + 0], &X[i * 2 + 0], &X[6], -1);
        bi_addsub(&R[i * 3 + 1], &X[i * 2 + 1], &X[7], -1);
        bi_mul(&t0, &R[i * 3 + 0], &R[i * 3 + 0]);
        bi_mul(&t1, &R[i * 3 + 1], &R[i * 3 + 1]);
        bi_addsub(&R[i * 3 + 2], &t0, &t1, +1);
    }
    det3(&r, &R[0], &R[3], &R[6]);
    (void)s;
    return r.s;
}

int or_insphere_bigint(const double* a, const double* b, const double* c, const double* d, const double* e) {
    const double* p[5] = {a, b, c, d, e};
    static __thread bi X[15], R[12], L[4];
    bi t0, t1, t2, D, T, acc, acc2;
    to_ints(5, 3, p, X);
    for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 3; ++j) bi_addsub(&R[i * 3 + j], &X[i * 3 + j], &X[12 + j], -1);
        bi_mul(&t0, &R[i * 3 + 0], &R[i * 3 + 0]);
        bi_mul(&t1, &R[i * 3 + 1], &R[i * 3 + 1]);
        bi_mul(&t2, &R[i * 3 + 2], &R[i * 3 + 2]);
        bi_addsub(&acc, &t0, &t1, +1);
        bi_addsub(&L[i], &acc, &t2, +1);
    }
    /* det4 [[R_i, L_i]] by cofactors of the lifted column:
     *   -L0 D(1,2,3) + L1 D(0,2,3) - L2 D(0,1,3) + L3 D(0,1,2)                      */
    static const int rows[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
    static const int sgn[4] = {-1, +1, -1, +1};
    bi_zero(&acc);
    for (int i = 0; i < 4; ++i) {
        det3(&D, &R[rows[i][0] * 3], &R[rows[i][1] * 3], &R[rows[i][2] * 3]);
        bi_mul(&T, &L[i], &D);
        bi_addsub(&acc2, &acc, &T, sgn[i]);
        acc = acc2;
    }
    /* in-sphere with this orientation convention is the negated lifted determinant */
    return -acc.s;
}

/* ---- double filters: evaluate, bound the rounding error by 1e-12 * permanent ----
 * The expressions below have rounding depth < 20, so their error is below 20 * 2^-53 times
 * the same expression evaluated on absolute values; 1e-12 exceeds that by > 400x.  Tiny or
 * non-finite magnitudes go straight to the integer path.                               */
#define FILT_REL 1e-12
#define FILT_MIN 1e-200

static int decide(double det, double perm) {
    if (!isfinite(det) || !isfinite(perm) || perm < FILT_MIN) return 2;
    double bound = FILT_REL * perm;
    if (det > bound) return 1;
    if (det < -bound) return -1;
    return 2;
}

int or_orient2d(const double* a, const double* b, const double* c) {
    double ux = b[0] - a[0], uy = b[1] - a[1], vx = c[0] - a[0], vy = c[1] - a[1];
    double det = ux * vy - uy * vx;
    double perm = fabs(ux * vy) + fabs(uy * vx);
    int s = decide(det, perm);
    return s != 2 ? s : or_orient2d_bigint(a, b, c);
}

int or_orient3d(const double* a, const double* b, const double* c, const double* d) {
    double u[3], v[3], w[3];
    for (int j = 0; j < 3; ++j) { u[j] = b[j] - a[j]; v[j] = c[j] - a[j]; w[j] = d[j] - a[j]; }
    double m0 = v[1] * w[2] - v[2] * w[1], m1 = v[0] * w[2] - v[2] * w[0], m2 = v[0] * w[1] - v[1] * w[0];
    double p0 = fabs(v[1] * w[2]) + fabs(v[2] * w[1]), p1 = fabs(v[0] * w[2]) + fabs(v[2] * w[0]),
           p2 = fabs(v[0] * w[1]) + fabs(v[1] * w[0]);
    double det = u[0] * m0 - u[1] * m1 + u[2] * m2;
    double perm = fabs(u[0]) * p0 + fabs(u[1]) * p1 + fabs(u[2]) * p2;
    int s = decide(det, perm);
    return s != 2 ? s : or_orient3d_bigint(a, b, c, d);
}

int or_incircle(const double* a, const double* b, const double* c, const double* d) {
    double r[3][3], q[3][3];
    const double* p[3] = {a, b, c};
    for (int i = 0; i < 3; ++i) {
        r[i][0] = p[i][0] - d[0];
        r[i][1] = p[i][1] - d[1];
        r[i][2] = r[i][0] * r[i][0] + r[i][1] * r[i][1];
        for (int j = 0; j < 3; ++j) q[i][j] = fabs(r[i][j]);
    }
    double det = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) - r[0][1] * (r[1][0] * r[2][2] - r[1][2] * r[2][0]) +
                 r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
    double perm = q[0][0] * (q[1][1] * q[2][2] + q[1][2] * q[2][1]) + q[0][1] * (q[1][0] * q[2][2] + q[1][2] * q[2][0]) +
                  q[0][2] * (q[1][0] * q[2][1] + q[1][1] * q[2][0]);
    int s = decide(det, perm);
    return s != 2 ? s : or_incircle_bigint(a, b, c, d);
}

static double d3(const double* r0, const double* r1, const double* r2) {
    return r0[0] * (r1[1] * r2[2] - r1[2] * r2[1]) - r0[1] * (r1[0] * r2[2] - r1[2] * r2[0]) +
           r0[2] * (r1[0] * r2[1] - r1[1] * r2[0]);
}
static double p3(const double* r0, const double* r1, const double* r2) {
    return r0[0] * (r1[1] * r2[2] + r1[2] * r2[1]) + r0[1] * (r1[0] * r2[2] + r1[2] * r2[0]) +
           r0[2] * (r1[0] * r2[1] + r1[1] * r2[0]);
}

int or_insphere(const double* a, const double* b, const double* c, const double* d, const double* e) {
    double r[4][3], q[4][3], l[4], ql[4];
    const double* p[4] = {a, b, c, d};
    for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 3; ++j) { r[i][j] = p[i][j] - e[j]; q[i][j] = fabs(r[i][j]); }
        l[i] = r[i][0] * r[i][0] + r[i][1] * r[i][1] + r[i][2] * r[i][2];
        ql[i] = l[i];
    }
    double det = -l[0] * d3(r[1], r[2], r[3]) + l[1] * d3(r[0], r[2], r[3]) - l[2] * d3(r[0], r[1], r[3]) +
                 l[3] * d3(r[0], r[1], r[2]);
    double perm = ql[0] * p3(q[1], q[2], q[3]) + ql[1] * p3(q[0], q[2], q[3]) + ql[2] * p3(q[0], q[1], q[3]) +
                  ql[3] * p3(q[0], q[1], q[2]);
    int s = decide(det, perm);
    return s != 2 ? -s : or_insphere_bigint(a, b, c, d, e);
}
