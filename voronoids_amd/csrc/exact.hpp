// exact.hpp -- tier 3: exact predicate signs by FP64 expansion arithmetic, and the cache
// through which the hot kernels read them.
//
// Method: Shewchuk (1997), "Adaptive Precision Floating-Point Arithmetic and Fast Robust
// Geometric Predicates": Two-Sum / Fast-Two-Sum / Two-Product (here through the hardware
// FMA), Fast-Expansion-Sum and Scale-Expansion with zero elimination.  Every difference of
// input coordinates is first made exact as a two-term expansion, then the determinant is
// expanded by cofactors entirely in expansion arithmetic.  The result is exact unless an
// intermediate underflows or overflows the f64 exponent range.
//
// The hot kernels never run this code: when tiers 1 and 2 cannot decide, they look the
// predicate up in a hash table keyed by (predicate, sorted vertex ids); on a miss they
// append it to a queue and defer.  k_exact (below) evaluates the queue into the table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "predicates.hpp"

#pragma clang fp contract(off)

namespace vo {

enum : uint32_t { PRED_ORIENT2 = 1, PRED_INCIRCLE = 2, PRED_ORIENT3 = 3, PRED_INSPHERE = 4 };
__host__ __device__ constexpr int pred_arity(uint32_t t) {
    return t == PRED_ORIENT2 ? 3 : (t == PRED_INSPHERE ? 5 : 4);
}

// ------------------------------------------------------------------ expansion arithmetic
struct Ws {                       // bump allocator over a per-thread global-memory slice
    double* base;
    int cap, top, ovf;
    __device__ double* take(int n) {
        if (top + n > cap) { ovf = 1; return nullptr; }
        double* r = base + top;
        top += n;
        return r;
    }
};
struct Ex { double* p; int n; };  // components in increasing magnitude, nonoverlapping

__device__ __forceinline__ void two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    double bv = x - a, av = x - bv, br = b - bv, ar = a - av;
    y = ar + br;
}
__device__ __forceinline__ void fast_two_sum(double a, double b, double& x, double& y) {
    x = a + b;
    double bv = x - a;
    y = b - bv;
}
__device__ __forceinline__ void two_product(double a, double b, double& x, double& y) {
    x = a * b;
    y = __builtin_fma(a, b, -x);
}

__device__ inline Ex ex_diff(Ws& w, double a, double b) {
    double* p = w.take(2);
    if (!p) return {nullptr, 0};
    double x = a - b;
    double bv = a - x, av = x + bv, br = bv - b, ar = a - av;
    double y = ar + br;
    if (y != 0.0) { p[0] = y; p[1] = x; return {p, 2}; }
    p[0] = x;
    return {p, 1};
}

// h = e + f (Fast-Expansion-Sum with zero elimination)
__device__ inline Ex ex_add(Ws& w, Ex e, Ex f) {
    if (!e.p || !f.p) return {nullptr, 0};
    double* h = w.take(e.n + f.n);
    if (!h) return {nullptr, 0};
    int ei = 0, fi = 0, hi = 0;
    double Q, Qn, hh;
    double en = e.p[0], fn = f.p[0];
    if ((fn > en) == (fn > -en)) { Q = en; ei++; } else { Q = fn; fi++; }
    if (ei < e.n && fi < f.n) {
        en = e.p[ei];
        fn = f.p[fi];
        if ((fn > en) == (fn > -en)) { fast_two_sum(en, Q, Qn, hh); ei++; }
        else { fast_two_sum(fn, Q, Qn, hh); fi++; }
        Q = Qn;
        if (hh != 0.0) h[hi++] = hh;
        while (ei < e.n && fi < f.n) {
            en = e.p[ei];
            fn = f.p[fi];
            if ((fn > en) == (fn > -en)) { two_sum(Q, en, Qn, hh); ei++; }
            else { two_sum(Q, fn, Qn, hh); fi++; }
            Q = Qn;
            if (hh != 0.0) h[hi++] = hh;
        }
    }
    while (ei < e.n) { two_sum(Q, e.p[ei++], Qn, hh); Q = Qn; if (hh != 0.0) h[hi++] = hh; }
    while (fi < f.n) { two_sum(Q, f.p[fi++], Qn, hh); Q = Qn; if (hh != 0.0) h[hi++] = hh; }
    if (Q != 0.0 || hi == 0) h[hi++] = Q;
    return {h, hi};
}

__device__ inline Ex ex_neg(Ws& w, Ex e) {
    if (!e.p) return e;
    double* h = w.take(e.n);
    if (!h) return {nullptr, 0};
    for (int i = 0; i < e.n; ++i) h[i] = -e.p[i];
    return {h, e.n};
}

__device__ inline Ex ex_sub(Ws& w, Ex e, Ex f) { return ex_add(w, e, ex_neg(w, f)); }

// h = e * b (Scale-Expansion with zero elimination)
__device__ inline Ex ex_scale(Ws& w, Ex e, double b) {
    if (!e.p) return e;
    double* h = w.take(2 * e.n);
    if (!h) return {nullptr, 0};
    int hi = 0;
    double Q, hh, p1, p0, s;
    two_product(e.p[0], b, Q, hh);
    if (hh != 0.0) h[hi++] = hh;
    for (int i = 1; i < e.n; ++i) {
        two_product(e.p[i], b, p1, p0);
        two_sum(Q, p0, s, hh);
        if (hh != 0.0) h[hi++] = hh;
        fast_two_sum(p1, s, Q, hh);
        if (hh != 0.0) h[hi++] = hh;
    }
    if (Q != 0.0 || hi == 0) h[hi++] = Q;
    return {h, hi};
}

// h = e * f, summing the scaled copies of e by each component of f
__device__ inline Ex ex_mul(Ws& w, Ex e, Ex f) {
    if (!e.p || !f.p) return {nullptr, 0};
    if (e.n < f.n) { Ex t = e; e = f; f = t; }
    Ex acc = ex_scale(w, e, f.p[0]);
    for (int j = 1; j < f.n; ++j) acc = ex_add(w, acc, ex_scale(w, e, f.p[j]));
    return acc;
}

__device__ inline int ex_sign(Ex e) {
    if (!e.p) return SIGN_UNDECIDED;
    double top = e.p[e.n - 1];
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

// exact signs, same conventions as the filters
__device__ inline int ex_orient2d(Ws& w, const double* a, const double* b, const double* c) {
    Ex ux = ex_diff(w, b[0], a[0]), uy = ex_diff(w, b[1], a[1]);
    Ex vx = ex_diff(w, c[0], a[0]), vy = ex_diff(w, c[1], a[1]);
    return ex_sign(ex_sub(w, ex_mul(w, ux, vy), ex_mul(w, uy, vx)));
}

__device__ inline int ex_orient3d(Ws& w, const double* a, const double* b, const double* c, const double* d) {
    Ex u[3], v[3], t[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = ex_diff(w, b[k], a[k]);
        v[k] = ex_diff(w, c[k], a[k]);
        t[k] = ex_diff(w, d[k], a[k]);
    }
    Ex m0 = ex_sub(w, ex_mul(w, v[1], t[2]), ex_mul(w, v[2], t[1]));
    Ex m1 = ex_sub(w, ex_mul(w, v[0], t[2]), ex_mul(w, v[2], t[0]));
    Ex m2 = ex_sub(w, ex_mul(w, v[0], t[1]), ex_mul(w, v[1], t[0]));
    Ex det = ex_add(w, ex_sub(w, ex_mul(w, u[0], m0), ex_mul(w, u[1], m1)), ex_mul(w, u[2], m2));
    return ex_sign(det);
}

__device__ inline int ex_incircle(Ws& w, const double* a, const double* b, const double* c, const double* d) {
    Ex x[3], y[3], l[3];
    const double* p[3] = {a, b, c};
    for (int i = 0; i < 3; ++i) {
        x[i] = ex_diff(w, p[i][0], d[0]);
        y[i] = ex_diff(w, p[i][1], d[1]);
        l[i] = ex_add(w, ex_mul(w, x[i], x[i]), ex_mul(w, y[i], y[i]));
    }
    Ex t0 = ex_mul(w, l[0], ex_sub(w, ex_mul(w, x[1], y[2]), ex_mul(w, x[2], y[1])));
    Ex t1 = ex_mul(w, l[1], ex_sub(w, ex_mul(w, x[2], y[0]), ex_mul(w, x[0], y[2])));
    Ex t2 = ex_mul(w, l[2], ex_sub(w, ex_mul(w, x[0], y[1]), ex_mul(w, x[1], y[0])));
    return ex_sign(ex_add(w, ex_add(w, t0, t1), t2));
}

// 3x3 determinant of rows (i, j, k) of the relative coordinates
__device__ inline Ex ex_det3(Ws& w, const Ex* x, const Ex* y, const Ex* z, int i, int j, int k) {
    Ex m0 = ex_sub(w, ex_mul(w, y[j], z[k]), ex_mul(w, z[j], y[k]));
    Ex m1 = ex_sub(w, ex_mul(w, x[j], z[k]), ex_mul(w, z[j], x[k]));
    Ex m2 = ex_sub(w, ex_mul(w, x[j], y[k]), ex_mul(w, y[j], x[k]));
    return ex_add(w, ex_sub(w, ex_mul(w, x[i], m0), ex_mul(w, y[i], m1)), ex_mul(w, z[i], m2));
}

__device__ inline int ex_insphere(Ws& w, const double* a, const double* b, const double* c, const double* d,
                                  const double* e) {
    Ex x[4], y[4], z[4], l[4];
    const double* p[4] = {a, b, c, d};
    for (int i = 0; i < 4; ++i) {
        x[i] = ex_diff(w, p[i][0], e[0]);
        y[i] = ex_diff(w, p[i][1], e[1]);
        z[i] = ex_diff(w, p[i][2], e[2]);
        l[i] = ex_add(w, ex_add(w, ex_mul(w, x[i], x[i]), ex_mul(w, y[i], y[i])), ex_mul(w, z[i], z[i]));
    }
    // lifted determinant = -L0 D(1,2,3) + L1 D(0,2,3) - L2 D(0,1,3) + L3 D(0,1,2)
    Ex acc = ex_mul(w, l[3], ex_det3(w, x, y, z, 0, 1, 2));
    acc = ex_sub(w, acc, ex_mul(w, l[2], ex_det3(w, x, y, z, 0, 1, 3)));
    acc = ex_add(w, acc, ex_mul(w, l[1], ex_det3(w, x, y, z, 0, 2, 3)));
    acc = ex_sub(w, acc, ex_mul(w, l[0], ex_det3(w, x, y, z, 1, 2, 3)));
    int s = ex_sign(acc);
    return s == SIGN_UNDECIDED ? s : -s;
}

// ------------------------------------------------------------------ queue + cache
struct QEntry { uint32_t type; uint32_t id[5]; };                  // ids in call order
struct CEntry { uint32_t k[5]; uint32_t tag; };                    // sorted ids, type | (sign+1)<<8
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// sort ids ascending, returning the parity (+1 even / -1 odd) of the permutation
__device__ __forceinline__ int sort_ids(uint32_t* k, int n) {
    int par = 1;
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && k[j - 1] > k[j]; --j) {
            uint32_t t = k[j]; k[j] = k[j - 1]; k[j - 1] = t;
            par = -par;
        }
    return par;
}

__device__ __forceinline__ uint32_t cache_hash(const uint32_t* k, int n, uint32_t type) {
    uint64_t h = 0x9E3779B97F4A7C15ull * (type + 1);
    for (int i = 0; i < n; ++i) {
        h ^= k[i] + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
        h *= 0xBF58476D1CE4E5B9ull;
    }
    return (uint32_t)(h ^ (h >> 31));
}

// sign of the predicate in the given id order, or SIGN_UNDECIDED when absent
__device__ inline int cache_lookup(const CEntry* tab, uint32_t mask, uint32_t type, const uint32_t* ids) {
    int n = pred_arity(type);
    uint32_t k[5];
    for (int i = 0; i < n; ++i) k[i] = ids[i];
    int par = sort_ids(k, n);
    uint32_t h = cache_hash(k, n, type);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const CEntry& e = tab[(h + probe) & mask];
        if (e.k[0] == kEmpty) return SIGN_UNDECIDED;
        bool same = (e.tag & 0xFF) == type;
        for (int i = 0; i < n && same; ++i) same = e.k[i] == k[i];
        if (same) return ((int)(e.tag >> 8) - 1) * par;
    }
    return SIGN_UNDECIDED;
}

__device__ inline void cache_insert(CEntry* tab, uint32_t mask, uint32_t type, const uint32_t* ids, int sign,
                                    uint32_t* full_flag) {
    int n = pred_arity(type);
    uint32_t k[5];
    for (int i = 0; i < n; ++i) k[i] = ids[i];
    int par = sort_ids(k, n);
    int s = sign * par;
    uint32_t h = cache_hash(k, n, type);
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        CEntry* e = &tab[(h + probe) & mask];
        uint32_t old = atomicCAS(&e->k[0], kEmpty, k[0]);
        if (old == kEmpty) {
            for (int i = 1; i < 5; ++i) e->k[i] = i < n ? k[i] : kEmpty;
            e->tag = type | ((uint32_t)(s + 1) << 8);
            return;
        }
    }
    atomicOr(full_flag, 1u);
}

}  // namespace vo
