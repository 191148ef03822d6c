// CPU check of the conservative FP32 ray-triangle classifier (rvf_filter.h):
// over random and adversarial rays, one whose FP64 Moller-Trumbore test (the
// reference's _ray_triangle semantics, BVH.jl:181-201) hits is never
// classified a certain miss, and one it misses is never classified a certain
// hit.  Built and driven by tests/test_filter_cpu.py.
#define RVF_HD
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>

#include "../../radiativeviewfactor.jl_amd/csrc/rvf_filter.h"

static bool mt64(const double* v0, const double* e1, const double* e2, const double* o, const double* d, double L) {
    double hx = d[1] * e2[2] - d[2] * e2[1], hy = d[2] * e2[0] - d[0] * e2[2], hz = d[0] * e2[1] - d[1] * e2[0];
    double a = (e1[0] * hx + e1[1] * hy) + e1[2] * hz;
    if (std::fabs(a) < 1e-10) return false;
    double f = 1.0 / a;
    double sx = o[0] - v0[0], sy = o[1] - v0[1], sz = o[2] - v0[2];
    double u = f * ((sx * hx + sy * hy) + sz * hz);
    if (u < 0.0 || u > 1.0) return false;
    double qx = sy * e1[2] - sz * e1[1], qy = sz * e1[0] - sx * e1[2], qz = sx * e1[1] - sy * e1[0];
    double v = f * ((d[0] * qx + d[1] * qy) + d[2] * qz);
    if (v < 0.0 || u + v > 1.0) return false;
    double t = f * ((e2[0] * qx + e2[1] * qy) + e2[2] * qz);
    return t > 0.0 && t < L;
}

static float up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

// mode 0: the classifier as built; mode 1: margins forced to zero (must fail).
// Returns the FP64 hits classified as certain misses; *sure_out counts the
// certain hits, *false_sure_out those among them the FP64 test misses.
extern "C" int64_t filter_check(uint64_t seed, int64_t n, int mode, int64_t* hits_out, int64_t* sure_out,
                                int64_t* false_sure_out) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    const float uu = 5.9604645e-08f;
    int64_t bad = 0, hits = 0, sure = 0, false_sure = 0;
    for (int64_t k = 0; k < n; k++) {
        const double Sm = 1.0 + 3.0 * U01(g);
        double v[3][3];
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) v[a][c] = (2.0 * U01(g) - 1.0) * Sm * 0.9;
        if (g() % 3 == 0) {  // small triangles (occluder-sized)
            for (int a = 1; a < 3; a++)
                for (int c = 0; c < 3; c++) v[a][c] = v[0][c] + (2.0 * U01(g) - 1.0) * 0.02;
        }
        double e1[3], e2[3];
        for (int c = 0; c < 3; c++) { e1[c] = v[1][c] - v[0][c]; e2[c] = v[2][c] - v[0][c]; }
        // target point on or near the triangle: barycentrics at the edges/vertices often
        double bu = U01(g), bv = U01(g);
        const int kind = (int)(g() % 8);  // 6, 7: generic rays (no boundary nearby)
        const double tiny = std::ldexp(1.0, -(int)(20 + g() % 30));
        if (kind == 0) bu = (g() & 1) ? tiny : -tiny;
        if (kind == 1) bv = (g() & 1) ? tiny : -tiny;
        if (kind == 2) { bv = 1.0 - bu + ((g() & 1) ? tiny : -tiny); }
        if (bu + bv > 1.0 && kind > 2) { bu = 1.0 - bu; bv = 1.0 - bv; }
        double P[3];
        for (int c = 0; c < 3; c++) P[c] = v[0][c] + bu * e1[c] + bv * e2[c];
        // origin: random, or grazing (in the triangle's plane), or at the plane (t ~ 0)
        double o[3];
        for (int c = 0; c < 3; c++) o[c] = (2.0 * U01(g) - 1.0) * Sm * 0.9;
        if (kind == 3) {  // nearly in-plane ray
            double w = U01(g);
            for (int c = 0; c < 3; c++) o[c] = P[c] + (w - 0.5) * (e1[c] - e2[c]) * 3.0 + tiny * (U01(g) - 0.5);
        }
        double dv[3] = {P[0] - o[0], P[1] - o[1], P[2] - o[2]};
        double L0 = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        if (!(L0 > 1e-9)) continue;
        // segment end: beyond P, or stopping just before / after it (t ~ L)
        double scale = (kind == 4) ? 1.0 + ((g() & 1) ? tiny : -tiny) : 1.0 + 2.0 * U01(g);
        if (kind == 5) scale = tiny * 4;  // origin almost on the triangle, t ~ 0
        double L = L0 * scale;
        double d[3] = {dv[0] / L0, dv[1] / L0, dv[2] / L0};
        if (kind == 5) for (int c = 0; c < 3; c++) o[c] = P[c] - d[c] * L0 * tiny;
        const bool hit = mt64(v[0], e1, e2, o, d, L);
        hits += hit;
        // FP32 data as the kernel has it: rounded inputs, direction within dd
        const double ddd = 1e-7 * std::pow(10.0, 2.0 * U01(g));
        RvfFRay R;
        R.ox = (float)o[0]; R.oy = (float)o[1]; R.oz = (float)o[2];
        float dd = (float)ddd;
        R.dx = (float)(d[0] + ddd * (2.0 * U01(g) - 1.0) * 0.5);
        R.dy = (float)(d[1] + ddd * (2.0 * U01(g) - 1.0) * 0.5);
        R.dz = (float)(d[2] + ddd * (2.0 * U01(g) - 1.0) * 0.5);
        R.L = (float)L;
        rvf_fray_margins(R, dd, (float)Sm);
        float K = 80.0f * uu * (float)Sm;
        if (mode == 1) { R.A = R.B = R.C = 0.0f; K = 0.0f; }
        float fv0[3], fe1[3], fe2[3];
        double E1 = 0, E2 = 0;
        for (int c = 0; c < 3; c++) {
            fv0[c] = (float)v[0][c]; fe1[c] = (float)e1[c]; fe2[c] = (float)e2[c];
            E1 += std::fabs(e1[c]); E2 += std::fabs(e2[c]);
        }
        const int cls = rvf_tri_classify(fv0, fe1, fe2, up(E1 * (1 + 1e-6)), up(E2 * (1 + 1e-6)), R, K);
        if (hit && cls == 0) bad++;
        sure += cls == 2;
        false_sure += cls == 2 && !hit;
    }
    if (hits_out) *hits_out = hits;
    if (sure_out) *sure_out = sure;
    if (false_sure_out) *false_sure_out = false_sure;
    return bad;
}

// ---------------------------------------------------------------------------
// Adversarial search at the decision boundaries.  The FP64 ray (o, unit d, L)
// and triangle are built so that one of the reference test's quantities sits
// a tiny signed distance from its threshold: U = 0, V = 0, U + V = a, T = 0,
// T = L a (each relative to a), or |a| just above 1e-10 (grazing); triangle
// sizes down to 1e-4 Sm.  The FP32 inputs are then moved to the full extent
// of the classifier's stated input bounds, in the direction that pushes the
// boundary quantity across its threshold: |o' - o| = 3u Sm, |d' - d| = dd,
// |L' - L| = L (2 dd + 4u) (each up to its final FP32 rounding), with v0 and
// the edges rounded to nearest.  `scale` multiplies every margin (1 = as
// built; 0.5 = margins equal to the first-order bounds).  Returns FP64 hits
// classified certain misses; *false_sure_out: FP64 misses classified certain
// hits; *near_out: cases whose FP64 decision is within 1e-6 of a threshold.
static void cross3(const double* a, const double* b, double* c) {
    c[0] = a[1] * b[2] - a[2] * b[1]; c[1] = a[2] * b[0] - a[0] * b[2]; c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// the boundary quantity Q (Q > 0 on the hit side) and its gradients in o and d
static double boundary_q(int kind, const double* o, const double* d, double L, const double* v0, const double* e1,
                         const double* e2) {
    double h[3], s[3], q[3];
    cross3(d, e2, h);
    const double a = dot3(e1, h);
    for (int c = 0; c < 3; c++) s[c] = o[c] - v0[c];
    const double U = dot3(s, h);
    cross3(s, e1, q);
    const double V = dot3(d, q), T = dot3(e2, q);
    const double sg = a < 0 ? -1.0 : 1.0;
    switch (kind) {
        case 0: return sg * U;                 // U / a >= 0
        case 1: return sg * V;                 // V / a >= 0
        case 2: return sg * (a - U - V);       // (U + V) / a <= 1
        case 3: return sg * T;                 // T / a > 0
        case 4: return sg * (L * a - T);       // T / a < L
        default: return std::fabs(a) - 1e-10;  // |a| >= 1e-10
    }
}

extern "C" int64_t filter_check_adv(uint64_t seed, int64_t n, double scale, int64_t* false_sure_out, int64_t* near_out,
                                    int64_t* hits_out) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    const double u = 5.9604644775390625e-08;
    int64_t false_miss = 0, false_sure = 0, nearc = 0, hits = 0;
    for (int64_t k = 0; k < n; k++) {
        const double Sm = 1.0 + 3.0 * U01(g);
        const double size = Sm * std::pow(10.0, -4.0 * U01(g));
        double v0[3], e1[3], e2[3];
        for (int c = 0; c < 3; c++) v0[c] = (2.0 * U01(g) - 1.0) * Sm * 0.8;
        for (int c = 0; c < 3; c++) { e1[c] = (2.0 * U01(g) - 1.0) * size; e2[c] = (2.0 * U01(g) - 1.0) * size; }
        const int kind = (int)(g() % 6);
        // a point of the triangle's plane, with barycentrics at the boundary of
        // the kind (or inside), and the segment through it
        double bu = U01(g), bv = U01(g);
        if (bu + bv > 1.0) { bu = 1.0 - bu; bv = 1.0 - bv; }
        if (kind == 0) bu = 0.0;
        if (kind == 1) bv = 0.0;
        if (kind == 2) bv = 1.0 - bu;
        double P[3];
        for (int c = 0; c < 3; c++) P[c] = v0[c] + bu * e1[c] + bv * e2[c];
        double nrm[3];
        cross3(e1, e2, nrm);
        const double nn = std::sqrt(dot3(nrm, nrm));
        if (!(nn > 0)) continue;
        double dir[3];
        for (int c = 0; c < 3; c++) dir[c] = 2.0 * U01(g) - 1.0;
        if (kind == 5) {  // grazing: the direction nearly in the plane, |a| ~ 1e-10
            const double dn = dot3(dir, nrm) / nn;
            for (int c = 0; c < 3; c++) dir[c] -= dn * nrm[c] / nn;
        }
        double dl = std::sqrt(dot3(dir, dir));
        if (!(dl > 1e-6)) continue;
        for (int c = 0; c < 3; c++) dir[c] /= dl;
        if (kind == 5) {  // tilt so that |a| = (1 .. 8) x 1e-10
            double h[3];
            cross3(dir, e2, h);
            const double want = 1e-10 * (1.0 + 7.0 * U01(g));
            for (int c = 0; c < 3; c++) dir[c] += want / nn * nrm[c] / nn;
            dl = std::sqrt(dot3(dir, dir));
            for (int c = 0; c < 3; c++) dir[c] /= dl;
        }
        const double tP = Sm * (0.05 + 1.5 * U01(g));
        double o[3];
        for (int c = 0; c < 3; c++) o[c] = P[c] - tP * dir[c];
        double L = tP * (1.0 + 2.0 * U01(g));
        if (kind == 3) {  // segment starts at the plane
            for (int c = 0; c < 3; c++) o[c] = P[c];
            L = Sm * (0.05 + U01(g));
        }
        if (kind == 4) L = tP;  // segment ends at the plane
        // nudge the boundary quantity to a tiny signed distance from zero along o
        // (o moves along the gradient of Q until Q = target)
        const double sgn = (g() & 1) ? 1.0 : -1.0;
        const double target = sgn * std::ldexp(Sm * size * size, -(int)(18 + g() % 26));
        for (int it = 0; it < 3 && kind != 5; it++) {
            const double Q = boundary_q(kind, o, dir, L, v0, e1, e2);
            double go[3];
            for (int c = 0; c < 3; c++) {
                double o2[3] = {o[0], o[1], o[2]};
                const double step = 1e-6 * Sm;
                o2[c] += step;
                go[c] = (boundary_q(kind, o2, dir, L, v0, e1, e2) - Q) / step;
            }
            const double gg = dot3(go, go);
            if (!(gg > 0)) break;
            for (int c = 0; c < 3; c++) o[c] += (target - Q) * go[c] / gg;
        }
        const bool hit = mt64(v0, e1, e2, o, dir, L);
        hits += hit;
        const double Qf = boundary_q(kind, o, dir, L, v0, e1, e2);
        nearc += std::fabs(Qf) < 1e-6 * Sm * size * size;
        // adversarial FP32 inputs: push Q toward the other side of zero
        const double push = hit ? -1.0 : 1.0;
        const double ddd = 4.0 * u + (1e-4 - 4.0 * u) * std::pow(U01(g), 3.0);
        double go[3], gd[3];
        {
            const double Q = boundary_q(kind, o, dir, L, v0, e1, e2);
            for (int c = 0; c < 3; c++) {
                double o2[3] = {o[0], o[1], o[2]}, d2[3] = {dir[0], dir[1], dir[2]};
                o2[c] += 1e-6 * Sm; d2[c] += 1e-6;
                go[c] = boundary_q(kind, o2, dir, L, v0, e1, e2) - Q;
                gd[c] = boundary_q(kind, o, d2, L, v0, e1, e2) - Q;
            }
        }
        double dL;
        {
            const double Q = boundary_q(kind, o, dir, L, v0, e1, e2);
            dL = boundary_q(kind, o, dir, L * (1 + 1e-6), v0, e1, e2) - Q;
        }
        RvfFRay R;
        float of[3], df[3];
        for (int c = 0; c < 3; c++) {
            of[c] = (float)(o[c] + push * (go[c] > 0 ? 1.0 : -1.0) * 2.0 * u * Sm);
            df[c] = (float)(dir[c] + push * (gd[c] > 0 ? 1.0 : -1.0) * (ddd - u));
        }
        R.ox = of[0]; R.oy = of[1]; R.oz = of[2];
        R.dx = df[0]; R.dy = df[1]; R.dz = df[2];
        R.L = (float)(L * (1.0 + push * (dL > 0 ? 1.0 : -1.0) * (2.0 * ddd + 3.0 * u)));
        rvf_fray_margins(R, (float)ddd, (float)Sm);
        R.A *= (float)scale; R.B *= (float)scale; R.C *= (float)scale;
        const float K = (float)(80.0 * u * Sm * scale);
        float fv0[3], fe1[3], fe2[3];
        double E1 = 0, E2 = 0;
        for (int c = 0; c < 3; c++) {
            fv0[c] = (float)v0[c]; fe1[c] = (float)e1[c]; fe2[c] = (float)e2[c];
            E1 += std::fabs(e1[c]); E2 += std::fabs(e2[c]);
        }
        const int cls = rvf_tri_classify(fv0, fe1, fe2, up(E1 * (1 + 1e-6)), up(E2 * (1 + 1e-6)), R, K);
        false_miss += hit && cls == 0;
        false_sure += !hit && cls == 2;
    }
    if (false_sure_out) *false_sure_out = false_sure;
    if (near_out) *near_out = nearc;
    if (hits_out) *hits_out = hits;
    return false_miss;
}

// ---------------------------------------------------------------------------
// Rays formed as the mixed kernel forms them (sray): FP64 anchors a_i, a_j and
// points x = a + offset; the kernel holds the offsets and a_j - a_i as FP32,
// the origin as FP32(a_i) + offset, r = ((a_j - a_i) - off_i) + off_j, r2 by
// fused multiply-adds, 1/|r| by the reciprocal square root (emulated here at
// its worst case, 1 ulp off either way), d = r / |r|, L = r2 / |r|, and dd from
// the kernel's formula.  The reference ray is the FP64 one from x_i to x_j
// (is_visible, RayCast.jl:27-30).  Triangles are placed across the segment
// with barycentrics near edges and vertices.  Returns FP64 hits classified
// certain misses; *false_sure_out: FP64 misses classified certain hits.
extern "C" int64_t filter_check_sray(uint64_t seed, int64_t n, int64_t* false_sure_out, int64_t* sure_out,
                                     int64_t* hits_out) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    const float uu = 5.9604645e-08f;
    int64_t false_miss = 0, false_sure = 0, sure = 0, hits = 0;
    for (int64_t k = 0; k < n; k++) {
        const double Sm0 = 1.0 + 3.0 * U01(g);
        double ai[3], aj[3], xi[3], xj[3];
        const double esz = 0.001 + 0.05 * U01(g);
        for (int c = 0; c < 3; c++) {
            ai[c] = (2.0 * U01(g) - 1.0) * Sm0 * 0.8;
            aj[c] = (g() % 4 == 0) ? ai[c] + (2.0 * U01(g) - 1.0) * 4 * esz : (2.0 * U01(g) - 1.0) * Sm0 * 0.8;
            xi[c] = ai[c] + (2.0 * U01(g) - 1.0) * esz;
            xj[c] = aj[c] + (2.0 * U01(g) - 1.0) * esz;
        }
        const bool rev = g() & 1;
        // the kernel's FP32 ray (sray)
        float dxi[3], qx[3], dax[3], fa[3], fj[3];
        for (int c = 0; c < 3; c++) {
            dxi[c] = (float)(xi[c] - ai[c]); qx[c] = (float)(xj[c] - aj[c]);
            dax[c] = (float)(aj[c] - ai[c]); fa[c] = (float)ai[c]; fj[c] = (float)aj[c];
        }
        float b[3], r[3];
        for (int c = 0; c < 3; c++) { b[c] = dax[c] - dxi[c]; r[c] = b[c] + qx[c]; }
        const float r2 = std::fmaf(r[0], r[0], std::fmaf(r[1], r[1], r[2] * r[2]));
        if (!(r2 > 1e-20f)) continue;
        float rs = (float)(1.0 / std::sqrt((double)r2));
        const int ulp = (int)(g() % 3) - 1;  // rsq off by -1, 0 or +1 ulp
        if (ulp > 0) rs = std::nextafter(rs, INFINITY);
        if (ulp < 0) rs = std::nextafter(rs, 0.0f);
        const float sgn = rev ? -rs : rs;
        RvfFRay R;
        R.ox = rev ? fj[0] + qx[0] : fa[0] + dxi[0];
        R.oy = rev ? fj[1] + qx[1] : fa[1] + dxi[1];
        R.oz = rev ? fj[2] + qx[2] : fa[2] + dxi[2];
        R.dx = r[0] * sgn; R.dy = r[1] * sgn; R.dz = r[2] * sgn; R.L = r2 * rs;
        const float dd = 3.0f * uu * (std::fabs(dax[0]) + std::fabs(dax[1]) + std::fabs(dax[2]) + std::fabs(b[0]) +
                                      std::fabs(b[1]) + std::fabs(b[2]) + std::fabs(r[0]) + std::fabs(r[1]) +
                                      std::fabs(r[2])) * rs + 4.0f * uu;
        // the reference ray
        const double o[3] = {rev ? xj[0] : xi[0], rev ? xj[1] : xi[1], rev ? xj[2] : xi[2]};
        const double dv[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
        const double L = std::sqrt((dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2]);
        double d[3] = {dv[0] / L, dv[1] / L, dv[2] / L};
        if (rev) for (int c = 0; c < 3; c++) d[c] = -d[c];
        // a triangle across the segment, at a parameter near 0, L or inside
        const double tt = (g() % 3 == 0) ? L * std::ldexp(1.0, -(int)(10 + g() % 30)) * ((g() & 1) ? 1 : -1)
                          : (g() % 2 == 0) ? L * (1.0 + std::ldexp(1.0, -(int)(10 + g() % 30)) * ((g() & 1) ? 1 : -1))
                                           : L * U01(g);
        double P[3];
        for (int c = 0; c < 3; c++) P[c] = o[c] + tt * d[c];
        const double size = esz * std::pow(10.0, -2.0 * U01(g));
        double e1[3], e2[3];
        for (int c = 0; c < 3; c++) { e1[c] = (2.0 * U01(g) - 1.0) * size; e2[c] = (2.0 * U01(g) - 1.0) * size; }
        double bu = U01(g), bv = U01(g);
        if (bu + bv > 1.0) { bu = 1.0 - bu; bv = 1.0 - bv; }
        const int kind = (int)(g() % 4);
        const double tiny = std::ldexp(1.0, -(int)(20 + g() % 30)) * ((g() & 1) ? 1 : -1);
        if (kind == 0) bu = tiny;
        if (kind == 1) bv = tiny;
        if (kind == 2) bv = 1.0 - bu + tiny;
        double v0[3];
        for (int c = 0; c < 3; c++) v0[c] = P[c] - bu * e1[c] - bv * e2[c];
        double Sm = Sm0;
        for (int c = 0; c < 3; c++) {
            Sm = std::max(Sm, std::fabs(v0[c]) + 2 * size);
            Sm = std::max(Sm, std::max(std::fabs(xi[c]), std::fabs(xj[c])));
        }
        Sm *= 1.000001;
        const bool hit = mt64(v0, e1, e2, o, d, L);
        hits += hit;
        rvf_fray_margins(R, dd, (float)Sm);
        const float K = 80.0f * uu * (float)Sm;
        float fv0[3], fe1[3], fe2[3];
        double E1 = 0, E2 = 0;
        for (int c = 0; c < 3; c++) {
            fv0[c] = (float)v0[c]; fe1[c] = (float)e1[c]; fe2[c] = (float)e2[c];
            E1 += std::fabs(e1[c]); E2 += std::fabs(e2[c]);
        }
        const int cls = rvf_tri_classify(fv0, fe1, fe2, up(E1 * (1 + 1e-6)), up(E2 * (1 + 1e-6)), R, K);
        false_miss += hit && cls == 0;
        false_sure += !hit && cls == 2;
        sure += cls == 2;
    }
    if (false_sure_out) *false_sure_out = false_sure;
    if (sure_out) *sure_out = sure;
    if (hits_out) *hits_out = hits;
    return false_miss;
}
