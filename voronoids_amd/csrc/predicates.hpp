// predicates.hpp -- tiered exact orientation / in-sphere signs for gfx950.
//
// Replaces the reference's only conflict test, the stored-sphere distance comparison
// in_sphere (src/geometry.rs:91-97), and adds the orientation test the reference never had.
//
// Sign conventions (identical in every tier):
//   orient2d(a,b,c)     = det[b-a; c-a]                    > 0: counter-clockwise
//   orient3d(a,b,c,d)   = det[b-a; c-a; d-a]
//   incircle(a,b,c,d)   > 0  iff d inside circle(abc)  when orient2d(a,b,c) > 0
//   insphere(a,b,c,d,e) > 0  iff e inside sphere(abcd) when orient3d(a,b,c,d) > 0
//
// Tier 1 evaluates in FP32 on coordinates stored as fl32(x - c) (c = bounding-sphere
// center) and carries a running absolute error bound through every operation, starting
// from the storage error |s - (x - c)| <= 2^-23 |s| + 1e-36.  Tier 2 is the FP64 stage-A
// filter of Shewchuk (1997, "Adaptive Precision Floating-Point Arithmetic and Fast Robust
// Geometric Predicates"): evaluation on the raw f64 coordinates, bound = c * eps * permanent.
// Tier 3 is exact: FP64 expansion arithmetic (same paper, sections 2.2-2.6) on the exact
// differences, run by a separate kernel (see exact.hpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace vo {

enum : int { SIGN_UNDECIDED = 2 };

// ----------------------------------------------------------------------------- tier 1
struct ef { float v, e; };   // value and a bound on |value - exact|

constexpr float kU32 = 5.9604644775390625e-8f;    // 2^-24, unit roundoff of fp32
constexpr float kEta = 1.1920928955078125e-7f;    // 2^-23, relative storage error bound
constexpr float kFloor = 1e-36f;                  // covers underflow of any single op

__device__ __forceinline__ ef ef_diff(float a, float b) {
    float d = a - b;
    return {d, kEta * (fabsf(a) + fabsf(b)) + kU32 * fabsf(d) + kFloor};
}
__device__ __forceinline__ ef ef_mul(ef a, ef b) {
    float p = a.v * b.v;
    return {p, fabsf(a.v) * b.e + fabsf(b.v) * a.e + a.e * b.e + kU32 * fabsf(p) + kFloor};
}
__device__ __forceinline__ ef ef_add(ef a, ef b) {
    float s = a.v + b.v;
    return {s, a.e + b.e + kU32 * fabsf(s) + kFloor};
}
__device__ __forceinline__ ef ef_sub(ef a, ef b) {
    float s = a.v - b.v;
    return {s, a.e + b.e + kU32 * fabsf(s) + kFloor};
}
// The bound itself is computed with < 40 fp32 roundings, each relative <= 2^-24:
// inflating it by 1/1024 covers them.
__device__ __forceinline__ int ef_sign(ef r) {
    float b = r.e * 1.0009765625f;
    if (!(fabsf(r.v) < __builtin_inff()) || !(b < __builtin_inff())) return SIGN_UNDECIDED;
    if (r.v > b) return 1;
    if (r.v < -b) return -1;
    return SIGN_UNDECIDED;
}

__device__ __forceinline__ int t1_orient2d(float2 a, float2 b, float2 c) {
    ef ux = ef_diff(b.x, a.x), uy = ef_diff(b.y, a.y), vx = ef_diff(c.x, a.x), vy = ef_diff(c.y, a.y);
    return ef_sign(ef_sub(ef_mul(ux, vy), ef_mul(uy, vx)));
}

__device__ __forceinline__ int t1_orient3d(float4 a, float4 b, float4 c, float4 d) {
    ef ux = ef_diff(b.x, a.x), uy = ef_diff(b.y, a.y), uz = ef_diff(b.z, a.z);
    ef vx = ef_diff(c.x, a.x), vy = ef_diff(c.y, a.y), vz = ef_diff(c.z, a.z);
    ef wx = ef_diff(d.x, a.x), wy = ef_diff(d.y, a.y), wz = ef_diff(d.z, a.z);
    ef m0 = ef_sub(ef_mul(vy, wz), ef_mul(vz, wy));
    ef m1 = ef_sub(ef_mul(vx, wz), ef_mul(vz, wx));
    ef m2 = ef_sub(ef_mul(vx, wy), ef_mul(vy, wx));
    ef det = ef_add(ef_sub(ef_mul(ux, m0), ef_mul(uy, m1)), ef_mul(uz, m2));
    return ef_sign(det);
}

__device__ __forceinline__ int t1_incircle(float2 a, float2 b, float2 c, float2 d) {
    ef ax = ef_diff(a.x, d.x), ay = ef_diff(a.y, d.y);
    ef bx = ef_diff(b.x, d.x), by = ef_diff(b.y, d.y);
    ef cx = ef_diff(c.x, d.x), cy = ef_diff(c.y, d.y);
    ef al = ef_add(ef_mul(ax, ax), ef_mul(ay, ay));
    ef bl = ef_add(ef_mul(bx, bx), ef_mul(by, by));
    ef cl = ef_add(ef_mul(cx, cx), ef_mul(cy, cy));
    ef t0 = ef_mul(al, ef_sub(ef_mul(bx, cy), ef_mul(cx, by)));
    ef t1 = ef_mul(bl, ef_sub(ef_mul(cx, ay), ef_mul(ax, cy)));
    ef t2 = ef_mul(cl, ef_sub(ef_mul(ax, by), ef_mul(bx, ay)));
    return ef_sign(ef_add(ef_add(t0, t1), t2));
}

__device__ __forceinline__ int t1_insphere(float4 a, float4 b, float4 c, float4 d, float4 e) {
    ef ax = ef_diff(a.x, e.x), ay = ef_diff(a.y, e.y), az = ef_diff(a.z, e.z);
    ef bx = ef_diff(b.x, e.x), by = ef_diff(b.y, e.y), bz = ef_diff(b.z, e.z);
    ef cx = ef_diff(c.x, e.x), cy = ef_diff(c.y, e.y), cz = ef_diff(c.z, e.z);
    ef dx = ef_diff(d.x, e.x), dy = ef_diff(d.y, e.y), dz = ef_diff(d.z, e.z);
    ef ab = ef_sub(ef_mul(ax, by), ef_mul(bx, ay));
    ef bc = ef_sub(ef_mul(bx, cy), ef_mul(cx, by));
    ef cd = ef_sub(ef_mul(cx, dy), ef_mul(dx, cy));
    ef da = ef_sub(ef_mul(dx, ay), ef_mul(ax, dy));
    ef ac = ef_sub(ef_mul(ax, cy), ef_mul(cx, ay));
    ef bd = ef_sub(ef_mul(bx, dy), ef_mul(dx, by));
    ef abc = ef_add(ef_sub(ef_mul(az, bc), ef_mul(bz, ac)), ef_mul(cz, ab));
    ef bcd = ef_add(ef_sub(ef_mul(bz, cd), ef_mul(cz, bd)), ef_mul(dz, bc));
    ef cda = ef_add(ef_add(ef_mul(cz, da), ef_mul(dz, ac)), ef_mul(az, cd));
    ef dab = ef_add(ef_add(ef_mul(dz, ab), ef_mul(az, bd)), ef_mul(bz, da));
    ef al = ef_add(ef_add(ef_mul(ax, ax), ef_mul(ay, ay)), ef_mul(az, az));
    ef bl = ef_add(ef_add(ef_mul(bx, bx), ef_mul(by, by)), ef_mul(bz, bz));
    ef cl = ef_add(ef_add(ef_mul(cx, cx), ef_mul(cy, cy)), ef_mul(cz, cz));
    ef dl = ef_add(ef_add(ef_mul(dx, dx), ef_mul(dy, dy)), ef_mul(dz, dz));
    ef det = ef_add(ef_sub(ef_mul(dl, abc), ef_mul(cl, dab)), ef_sub(ef_mul(bl, cda), ef_mul(al, bcd)));
    int s = ef_sign(det);
    return s == SIGN_UNDECIDED ? s : -s;   // the lifted determinant is -insphere
}

// ----------------------------------------------------------------------------- tier 2
// Shewchuk's stage-A error bound coefficients, eps = 2^-53.
constexpr double kEps = 1.1102230246251565e-16;
constexpr double kCcwA = (3.0 + 16.0 * kEps) * kEps;
constexpr double kO3dA = (7.0 + 56.0 * kEps) * kEps;
constexpr double kIccA = (10.0 + 96.0 * kEps) * kEps;
constexpr double kIspA = (16.0 + 224.0 * kEps) * kEps;

__device__ __forceinline__ int sgn_or_undecided(double det, double bound) {
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return SIGN_UNDECIDED;     // also NaN
}

// det[a-c; b-c] (equal to det[b-a; c-a])
__device__ __forceinline__ int t2_orient2d(const double* a, const double* b, const double* c) {
    double detleft = (a[0] - c[0]) * (b[1] - c[1]);
    double detright = (a[1] - c[1]) * (b[0] - c[0]);
    double det = detleft - detright;
    double detsum;
    if (detleft > 0.0) {
        if (detright <= 0.0) return det > 0 ? 1 : (det < 0 ? -1 : 0);
        detsum = detleft + detright;
    } else if (detleft < 0.0) {
        if (detright >= 0.0) return det > 0 ? 1 : (det < 0 ? -1 : 0);
        detsum = -detleft - detright;
    } else {
        return det > 0 ? 1 : (det < 0 ? -1 : 0);
    }
    return sgn_or_undecided(det, kCcwA * detsum);
}

// orient3d as defined above is the negation of det[a-d; b-d; c-d]
__device__ __forceinline__ int t2_orient3d(const double* pa, const double* pb, const double* pc, const double* pd) {
    double adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    double ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    double adz = pa[2] - pd[2], bdz = pb[2] - pd[2], cdz = pc[2] - pd[2];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                  (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    int s = sgn_or_undecided(det, kO3dA * perm);
    return s == SIGN_UNDECIDED ? s : -s;
}

__device__ __forceinline__ int t2_incircle(const double* pa, const double* pb, const double* pc, const double* pd) {
    double adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    double ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double alift = adx * adx + ady * ady;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double blift = bdx * bdx + bdy * bdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double clift = cdx * cdx + cdy * cdy;
    double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * alift + (fabs(cdxady) + fabs(adxcdy)) * blift +
                  (fabs(adxbdy) + fabs(bdxady)) * clift;
    return sgn_or_undecided(det, kIccA * perm);
}

__device__ __forceinline__ int t2_insphere(const double* pa, const double* pb, const double* pc, const double* pd,
                                           const double* pe) {
    double aex = pa[0] - pe[0], bex = pb[0] - pe[0], cex = pc[0] - pe[0], dex = pd[0] - pe[0];
    double aey = pa[1] - pe[1], bey = pb[1] - pe[1], cey = pc[1] - pe[1], dey = pd[1] - pe[1];
    double aez = pa[2] - pe[2], bez = pb[2] - pe[2], cez = pc[2] - pe[2], dez = pd[2] - pe[2];
    double aexbey = aex * bey, bexaey = bex * aey, ab = aexbey - bexaey;
    double bexcey = bex * cey, cexbey = cex * bey, bc = bexcey - cexbey;
    double cexdey = cex * dey, dexcey = dex * cey, cd = cexdey - dexcey;
    double dexaey = dex * aey, aexdey = aex * dey, da = dexaey - aexdey;
    double aexcey = aex * cey, cexaey = cex * aey, ac = aexcey - cexaey;
    double bexdey = bex * dey, dexbey = dex * bey, bd = bexdey - dexbey;
    double abc = aez * bc - bez * ac + cez * ab;
    double bcd = bez * cd - cez * bd + dez * bc;
    double cda = cez * da + dez * ac + aez * cd;
    double dab = dez * ab + aez * bd + bez * da;
    double alift = aex * aex + aey * aey + aez * aez;
    double blift = bex * bex + bey * bey + bez * bez;
    double clift = cex * cex + cey * cey + cez * cez;
    double dlift = dex * dex + dey * dey + dez * dez;
    double det = (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
    double aezp = fabs(aez), bezp = fabs(bez), cezp = fabs(cez), dezp = fabs(dez);
    double aexbeyp = fabs(aexbey), bexaeyp = fabs(bexaey), bexceyp = fabs(bexcey), cexbeyp = fabs(cexbey);
    double cexdeyp = fabs(cexdey), dexceyp = fabs(dexcey), dexaeyp = fabs(dexaey), aexdeyp = fabs(aexdey);
    double aexceyp = fabs(aexcey), cexaeyp = fabs(cexaey), bexdeyp = fabs(bexdey), dexbeyp = fabs(dexbey);
    double perm = ((cexdeyp + dexceyp) * bezp + (dexbeyp + bexdeyp) * cezp + (bexceyp + cexbeyp) * dezp) * alift +
                  ((dexaeyp + aexdeyp) * cezp + (aexceyp + cexaeyp) * dezp + (cexdeyp + dexceyp) * aezp) * blift +
                  ((aexbeyp + bexaeyp) * dezp + (bexdeyp + dexbeyp) * aezp + (dexaeyp + aexdeyp) * bezp) * clift +
                  ((bexceyp + cexbeyp) * aezp + (cexaeyp + aexceyp) * bezp + (aexbeyp + bexaeyp) * cezp) * dlift;
    int s = sgn_or_undecided(det, kIspA * perm);
    return s == SIGN_UNDECIDED ? s : -s;
}

}  // namespace vo
