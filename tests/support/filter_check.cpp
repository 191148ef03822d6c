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
