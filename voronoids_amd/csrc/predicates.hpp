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
// center): Shewchuk's stage-A expression plus a closed-form bound on the storage error
// |s - (x - c)| <= b |s|, b = 2^-24 (1 + 2^-22) (derivation below).  Tier 2 is the FP64 stage-A
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
// FP32 on stored coordinates s = fl32(fl64(x - c)).  The f64 difference is off by at most
// 2^-53 |x - c| and its rounding to fp32 by at most 2^-24 |s| (round to nearest; a subnormal s
// is off by at most 2^-150, inside the 2e-37 below), so |s - (x - c)| <= b |s| with
// b = 2^-24 (1 + 2^-22) > 2^-24 + 2^-52 (1 + 2^-23).  Each computed difference d~ = fl(s_a - s_b)
// then differs from the exact difference of the inputs by at most
//     eps = b (|s_a| + |s_b|) + 2^-24 |d~| + 2e-37  <=  2b m + 2^-24 delta + 2e-37,
// m = max |s|, delta = max |d~| over the call.  The determinant is then evaluated from the
// d~ with Shewchuk's stage-A expression (its bound c * u * permanent covers the evaluation,
// u = 2^-24), and the input error is bounded by multilinearity and Hadamard's inequality:
//   orient2d   2      ((delta+eps)^2 - delta^2)
//   orient3d   3 sqrt3((delta+eps)^3 - delta^3)
//   incircle  12      ((delta+eps)^4 - delta^4)
//   insphere  36 sqrt3((delta+eps)^5 - delta^5)
// (rows have norm <= sqrt(d) (delta+eps), lifts <= d (delta+eps)^2).  The bound's own fp32
// rounding is covered by a factor 1.01; magnitudes near the fp32 underflow range defer.
constexpr float kU32 = 5.9604644775390625e-8f;    // 2^-24
constexpr float kCcwA32 = (3.0f + 16.0f * kU32) * kU32;
constexpr float kO3dA32 = (7.0f + 56.0f * kU32) * kU32;
constexpr float kIccA32 = (10.0f + 96.0f * kU32) * kU32;
constexpr float kIspA32 = (16.0f + 224.0f * kU32) * kU32;
constexpr float kTiny32 = 1e-30f;

// 2b m + 2^-24 delta (2b = 2^-23 + 2^-45, exact in fp32), plus 2e-37 for coordinates that
// underflowed in the fp32 store
__device__ __forceinline__ float in_err(float m, float dl) { return m * 1.1920931797249068e-7f + dl * kU32 + 2e-37f; }

__device__ __forceinline__ int t1_decide(float det, float perm, float pin) {
    if (!(perm > kTiny32) || !(perm < 1e30f)) return SIGN_UNDECIDED;   // also NaN
    const float bound = (perm + pin) * 1.01f;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return SIGN_UNDECIDED;
}

__device__ __forceinline__ float amax3(float a, float b, float c) { return fmaxf(fabsf(a), fmaxf(fabsf(b), fabsf(c))); }

__device__ __forceinline__ int t1_orient2d(float2 a, float2 b, float2 c) {
    const float acx = a.x - c.x, bcx = b.x - c.x, acy = a.y - c.y, bcy = b.y - c.y;
    const float l = acx * bcy, r = acy * bcx;
    const float det = l - r;
    const float m = fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fmaxf(fabsf(b.x), fabsf(b.y)), fmaxf(fabsf(c.x), fabsf(c.y))));
    const float dl = fmaxf(fmaxf(fabsf(acx), fabsf(bcx)), fmaxf(fabsf(acy), fabsf(bcy)));
    const float e = in_err(m, dl);
    const float pin = 2.0f * e * (2.0f * dl + e);
    return t1_decide(det, kCcwA32 * (fabsf(l) + fabsf(r)), pin);
}

// orient3d(a,b,c,d) = -det[a-d; b-d; c-d]
__device__ __forceinline__ int t1_orient3d(float4 pa, float4 pb, float4 pc, float4 pd) {
    const float adx = pa.x - pd.x, bdx = pb.x - pd.x, cdx = pc.x - pd.x;
    const float ady = pa.y - pd.y, bdy = pb.y - pd.y, cdy = pc.y - pd.y;
    const float adz = pa.z - pd.z, bdz = pb.z - pd.z, cdz = pc.z - pd.z;
    const float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const float cdxady = cdx * ady, adxcdy = adx * cdy;
    const float adxbdy = adx * bdy, bdxady = bdx * ady;
    const float det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    const float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * fabsf(adz) + (fabsf(cdxady) + fabsf(adxcdy)) * fabsf(bdz) +
                       (fabsf(adxbdy) + fabsf(bdxady)) * fabsf(cdz);
    const float m = fmaxf(fmaxf(amax3(pa.x, pa.y, pa.z), amax3(pb.x, pb.y, pb.z)),
                          fmaxf(amax3(pc.x, pc.y, pc.z), amax3(pd.x, pd.y, pd.z)));
    const float dl = fmaxf(fmaxf(amax3(adx, ady, adz), amax3(bdx, bdy, bdz)), amax3(cdx, cdy, cdz));
    const float e = in_err(m, dl);
    const float pin = 5.196152422706632f * e * (3.0f * dl * dl + e * (3.0f * dl + e));
    const int s = t1_decide(det, kO3dA32 * perm, pin);
    return s == SIGN_UNDECIDED ? s : -s;
}

__device__ __forceinline__ int t1_incircle(float2 pa, float2 pb, float2 pc, float2 pd) {
    const float adx = pa.x - pd.x, bdx = pb.x - pd.x, cdx = pc.x - pd.x;
    const float ady = pa.y - pd.y, bdy = pb.y - pd.y, cdy = pc.y - pd.y;
    const float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const float alift = adx * adx + ady * ady;
    const float cdxady = cdx * ady, adxcdy = adx * cdy;
    const float blift = bdx * bdx + bdy * bdy;
    const float adxbdy = adx * bdy, bdxady = bdx * ady;
    const float clift = cdx * cdx + cdy * cdy;
    const float det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) + clift * (adxbdy - bdxady);
    const float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * alift + (fabsf(cdxady) + fabsf(adxcdy)) * blift +
                       (fabsf(adxbdy) + fabsf(bdxady)) * clift;
    const float m = fmaxf(fmaxf(fmaxf(fabsf(pa.x), fabsf(pa.y)), fmaxf(fabsf(pb.x), fabsf(pb.y))),
                          fmaxf(fmaxf(fabsf(pc.x), fabsf(pc.y)), fmaxf(fabsf(pd.x), fabsf(pd.y))));
    const float dl = fmaxf(fmaxf(fmaxf(fabsf(adx), fabsf(ady)), fmaxf(fabsf(bdx), fabsf(bdy))),
                           fmaxf(fabsf(cdx), fabsf(cdy)));
    const float e = in_err(m, dl);
    const float pin = 12.0f * e * (4.0f * dl * dl * dl + e * (6.0f * dl * dl + e * (4.0f * dl + e)));
    return t1_decide(det, kIccA32 * perm, pin);
}

__device__ __forceinline__ int t1_insphere(float4 pa, float4 pb, float4 pc, float4 pd, float4 pe) {
    const float aex = pa.x - pe.x, bex = pb.x - pe.x, cex = pc.x - pe.x, dex = pd.x - pe.x;
    const float aey = pa.y - pe.y, bey = pb.y - pe.y, cey = pc.y - pe.y, dey = pd.y - pe.y;
    const float aez = pa.z - pe.z, bez = pb.z - pe.z, cez = pc.z - pe.z, dez = pd.z - pe.z;
    const float aexbey = aex * bey, bexaey = bex * aey, ab = aexbey - bexaey;
    const float bexcey = bex * cey, cexbey = cex * bey, bc = bexcey - cexbey;
    const float cexdey = cex * dey, dexcey = dex * cey, cd = cexdey - dexcey;
    const float dexaey = dex * aey, aexdey = aex * dey, da = dexaey - aexdey;
    const float aexcey = aex * cey, cexaey = cex * aey, ac = aexcey - cexaey;
    const float bexdey = bex * dey, dexbey = dex * bey, bd = bexdey - dexbey;
    const float abc = aez * bc - bez * ac + cez * ab;
    const float bcd = bez * cd - cez * bd + dez * bc;
    const float cda = cez * da + dez * ac + aez * cd;
    const float dab = dez * ab + aez * bd + bez * da;
    const float alift = aex * aex + aey * aey + aez * aez;
    const float blift = bex * bex + bey * bey + bez * bez;
    const float clift = cex * cex + cey * cey + cez * cez;
    const float dlift = dex * dex + dey * dey + dez * dez;
    const float det = (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
    const float aezp = fabsf(aez), bezp = fabsf(bez), cezp = fabsf(cez), dezp = fabsf(dez);
    const float aexbeyp = fabsf(aexbey), bexaeyp = fabsf(bexaey), bexceyp = fabsf(bexcey), cexbeyp = fabsf(cexbey);
    const float cexdeyp = fabsf(cexdey), dexceyp = fabsf(dexcey), dexaeyp = fabsf(dexaey), aexdeyp = fabsf(aexdey);
    const float aexceyp = fabsf(aexcey), cexaeyp = fabsf(cexaey), bexdeyp = fabsf(bexdey), dexbeyp = fabsf(dexbey);
    const float perm =
        ((cexdeyp + dexceyp) * bezp + (dexbeyp + bexdeyp) * cezp + (bexceyp + cexbeyp) * dezp) * alift +
        ((dexaeyp + aexdeyp) * cezp + (aexceyp + cexaeyp) * dezp + (cexdeyp + dexceyp) * aezp) * blift +
        ((aexbeyp + bexaeyp) * dezp + (bexdeyp + dexbeyp) * aezp + (dexaeyp + aexdeyp) * bezp) * clift +
        ((bexceyp + cexbeyp) * aezp + (cexaeyp + aexceyp) * bezp + (aexbeyp + bexaeyp) * cezp) * dlift;
    const float m = fmaxf(fmaxf(fmaxf(amax3(pa.x, pa.y, pa.z), amax3(pb.x, pb.y, pb.z)),
                                fmaxf(amax3(pc.x, pc.y, pc.z), amax3(pd.x, pd.y, pd.z))),
                          amax3(pe.x, pe.y, pe.z));
    const float dl = fmaxf(fmaxf(amax3(aex, aey, aez), amax3(bex, bey, bez)),
                           fmaxf(amax3(cex, cey, cez), amax3(dex, dey, dez)));
    const float e = in_err(m, dl);
    const float d2 = dl * dl;
    const float pin = 62.35382907247958f * e *
                      (5.0f * d2 * d2 + e * (10.0f * d2 * dl + e * (10.0f * d2 + e * (5.0f * dl + e))));
    const int s = t1_decide(det, kIspA32 * perm, pin);
    return s == SIGN_UNDECIDED ? s : -s;    // the lifted determinant is -insphere
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
