// Conservative FP32 ray-triangle filter (mixed precision occlusion).
//
// Decides, from FP32 data, whether the reference's FP64 Moller-Trumbore test
// (_ray_triangle with t_min = 0 and t < L, BVH.jl:181-201) can report a hit
// for this ray and triangle.  0 means it certainly cannot; 1 means it may, and
// the exact FP64 test decides.  Division-free form: with h = d x e2,
// a = e1.h, s = o - v0, U = s.h, q = s x e1, V = d.q, T = e2.q, the FP64 test
// hits iff |a| >= 1e-10, 0 <= U/a <= 1, V/a >= 0, (U+V)/a <= 1 and
// 0 < T/a < L.  Each FP32 quantity is compared with a margin that bounds its
// distance from the exact value of the FP64 inputs:
//   FP32 inputs:  |o' - o| <= 3u Sm, |v0' - v0| <= u Sm, |e' - e| <= u|e|
//                 (componentwise), |d' - d| <= dd (componentwise, per ray);
//   magnitudes:   |d|_inf <= 1, |o|, |v0| <= Sm, |s|_inf <= 2 Sm,
//                 E1 = |e1|_1, E2 = |e2|_1;
//   first-order bounds (u = 2^-24, FP32 rounding of each product and sum):
//     |a' - a| <= E1 E2 (dd + 7u)        |U' - U| <= Sm E2 (6 dd + 54u)
//     |V' - V| <= Sm E1 (6 dd + 32u)     |T' - T| <= 20u Sm E1 E2
//     |L' - L| <= L (2 dd + 4u).
// The per-ray margins below are at least twice these bounds; the FP64 test's
// own rounding (2^-53 relative) is far inside them.  Second-order terms (the
// products of two first-order relative errors) stay below 1e-3 of the
// first-order bounds because rays with dd > 1e-3 are always left undecided.
#pragma once

#ifndef RVF_HD
#define RVF_HD __host__ __device__
#endif

struct RvfFRay {
    float ox, oy, oz;   // origin
    float dx, dy, dz;   // unit direction
    float L;            // segment length
    float A, B, C;      // margins: A = 4 (dd + 8u), B = 2 Sm (12 dd + 120u), C = 2 Sm (12 dd + 64u)
};

// Per-ray margins from the direction error bound dd and the scene magnitude Sm.
RVF_HD inline void rvf_fray_margins(RvfFRay& R, float dd, float Sm) {
    const float u = 5.9604645e-08f;
    R.A = 4.0f * (dd + 8.0f * u);
    R.B = 2.0f * Sm * (12.0f * dd + 120.0f * u);
    R.C = 2.0f * Sm * (12.0f * dd + 64.0f * u);
}

// Classifies a ray against a triangle from FP32 data:
//   0  the FP64 test certainly misses,
//   1  undecided (the exact FP64 test decides),
//   2  the FP64 test certainly hits.
// v0, e1, e2: the triangle as FP32; E1, E2: |e1|_1, |e2|_1 of the FP64 edges
// rounded up; K = 80u Sm.  A miss needs one comparison to fail by more than
// its margin.  A hit needs every comparison to hold by twice its margin: the
// exact value then holds by at least 1.5 margins, which covers the FP64
// test's own rounding (each margin is at least 16u relative to the scale of
// what it bounds, that rounding below 2^-50) and the FP32 rounding of the
// comparison itself (a few u relative).
RVF_HD inline int rvf_tri_classify(const float* v0, const float* e1, const float* e2, float E1, float E2,
                                   const RvfFRay& R, float K) {
    const float hx = R.dy * e2[2] - R.dz * e2[1];
    const float hy = R.dz * e2[0] - R.dx * e2[2];
    const float hz = R.dx * e2[1] - R.dy * e2[0];
    const float a = e1[0] * hx + e1[1] * hy + e1[2] * hz;
    const float sx = R.ox - v0[0], sy = R.oy - v0[1], sz = R.oz - v0[2];
    const float U = sx * hx + sy * hy + sz * hz;
    const float qx = sy * e1[2] - sz * e1[1];
    const float qy = sz * e1[0] - sx * e1[2];
    const float qz = sx * e1[1] - sy * e1[0];
    const float V = R.dx * qx + R.dy * qy + R.dz * qz;
    const float T = e2[0] * qx + e2[1] * qy + e2[2] * qz;
    const float E12 = E1 * E2;
    const float da = R.A * E12, dU = R.B * E2, dV = R.C * E1, dT = K * E12;
    const bool neg = a < 0.0f;
    const float aa = neg ? -a : a;
    const float ua = neg ? -U : U, va = neg ? -V : V, ta = neg ? -T : T;
    const float top = aa + da;
    // rays whose direction error bound dd exceeds 1e-3 (A > 4e-3) are left to
    // the exact test: below it the second-order terms of the error expansion
    // stay far inside the margins' slack over the first-order bounds
    const bool undecided = aa <= da + 1e-10f || R.A > 4.0e-3f;
    const bool miss = (ua < -dU) | (ua > top + dU) | (va < -dV) | (ua + va > top + dU + dV) | (ta < -dT) |
                      (ta > R.L * top * (1.0f + R.A) + dT);
    if (!undecided && miss) return 0;
    const float low = aa - 2.0f * da;  // exact |a| exceeds this
    const bool sure = (low > 2e-10f) & (ua >= 2.0f * dU) & (va >= 2.0f * dV) &
                      (ua + va + 2.0f * (dU + dV) <= low) & (ta >= 2.0f * dT) &
                      (ta + 2.0f * dT <= R.L * low * (1.0f - 2.0f * R.A));
    return sure ? 2 : 1;
}

// 0 when the FP64 test certainly misses, 1 when it may hit.
RVF_HD inline int rvf_tri_filter(const float* v0, const float* e1, const float* e2, float E1, float E2,
                                 const RvfFRay& R, float K) {
    return rvf_tri_classify(v0, e1, e2, E1, E2, R, K) != 0;
}
