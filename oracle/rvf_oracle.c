/*
 * rvf_oracle.c — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, Float64 restatement of RadiativeViewFactor.jl's CPU view-factor
 * path, used as the parity reference for the HIP library and as the CPU baseline that
 * bench.py times beside the GPU.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path never links,
 * loads or calls this file.
 *
 * Arithmetic follows the reference's Julia expressions operation by operation
 * (left-associative, no FMA contraction: build with -ffp-contract=off).
 * Every function cites the reference file:line it restates.
 *
 * Parity status (what pins each part; see tests/test_oracle_*.py):
 *   gl1d / quad rule        closed forms + test/runtests.jl:16-48 assertions
 *   tri rule (Dunavant)     verbatim constants + test/runtests.jl:258-275
 *   shape functions         partition of unity / nodal delta (runtests.jl:51-67,227-255)
 *   element geometry        runtests.jl:81-99 (x(0,0), n, dA, area)
 *   vf_kernel               runtests.jl:137-153 (1/pi, back-face, coincident)
 *   pair integrator         runtests.jl:160-198, 278-300, 420-436 + Howell closed
 *                           forms for parallel / perpendicular squares
 *   BVH + ray tests         runtests.jl:102-134 + brute-force any-hit over all
 *                           triangles (tree-independent) on random soups
 *   MC stream               closed-form integer recurrences (splitmix64,
 *                           xorshift64) + 3-sigma agreement with quadrature
 *                           (coaxial plates) and with orc_mc_expect, the
 *                           paired-cell expectation of the stream's estimate
 *   Duffy                   no reference test covers it; pinned by the
 *                           invariants non-adjacent == standard and non-Quad8 ==
 *                           standard quadrature, by closed-form moments of the
 *                           common-vertex rule (a constant integrates to 2, the
 *                           Q2 Jacobian), by the common-edge rule's vanishing-
 *                           Jacobian hyperplanes, in-square weight sums and
 *                           square rotations, and bit for bit by vectors of the
 *                           reference's own region-map expressions.
 *   Golub-Welsch (nquad>5)  runtests.jl:34-48, numpy's leggauss to 1e-14, and
 *                           the reference's eigenvalue route through LAPACK
 *                           dsyevr to 8e-15 (plate view factor to 1e-13); the
 *                           reference's own last bits depend on the machine's
 *                           BLAS kernels, so no bit-level target exists.
 *   curve meshes (Line2/3)  runtests.jl:241-247 (line2 shape) + Hottel's crossed
 *                           strings (exact 2-D view factors between straight
 *                           strips, with and without a blocking strip; arcs of a
 *                           circle, self view factor 1 - chord/arc) + brute-force
 *                           any-hit over all segments.
 *   self_vf                 pairs unchanged, diagonal = pair(i, i), flat = 0,
 *                           sphere closed form with O(h^2) convergence.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXPTS 64
#define LEAF_MAX 8

/* ------------------------------------------------------------------------- */
/* Julia float min/max: NaN-propagating, -0.0 < +0.0 (Base.min/max for      */
/* IEEE floats).  Used by BVH.jl:89-97,165-166.                               */
/* ------------------------------------------------------------------------- */
static inline double jl_min(double x, double y) {
    double diff = x - y;
    double am = signbit(diff) ? x : y;
    return (isnan(x) || isnan(y)) ? diff : am;
}
static inline double jl_max(double x, double y) {
    double diff = x - y;
    double am = signbit(diff) ? y : x;
    return (isnan(x) || isnan(y)) ? diff : am;
}
/* Julia isless for Float64 (sort! default lt). */
static inline int jl_isless(double a, double b) {
    if (isnan(a)) return 0;
    if (isnan(b)) return 1;
    if (a < b) return 1;
    if (a > b) return 0;
    return signbit(a) && !signbit(b);
}

/* ------------------------------------------------------------------------- */
/* Quadrature — src/Quadrature.jl:35-61 (_GL_TABLE), :69-84, :96-109           */
/* ------------------------------------------------------------------------- */
static void legendre_newton(int n, double* p, double* w) {
    /* Golub-Welsch replacement for n > 5 (reference: eigen of the Jacobi
     * matrix, Quadrature.jl:75-84).  Roots by Newton on P_n in long double,
     * weights 2/((1-x^2) P_n'(x)^2); ascending order like sortperm(vals). */
    for (int k = 0; k < n; k++) {
        long double x = cosl(3.14159265358979323846264338327950288L * (k + 0.75L) / (n + 0.5L));
        long double dp = 0;
        for (int it = 0; it < 100; it++) {
            long double p0 = 1, p1 = x;
            for (int m = 2; m <= n; m++) {
                long double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
                p0 = p1; p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1);
            long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-30L) break;
        }
        {
            long double p0 = 1, p1 = x;
            for (int m = 2; m <= n; m++) {
                long double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
                p0 = p1; p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1);
        }
        /* cos ordering is descending; store ascending */
        p[n - 1 - k] = (double)x;
        w[n - 1 - k] = (double)(2.0L / ((1 - x * x) * dp * dp));
    }
}

int orc_gl1d(int n, double* p, double* w) {
    switch (n) {
    case 1: p[0] = 0.0; w[0] = 2.0; return 1;
    case 2: p[0] = -1.0 / sqrt(3.0); p[1] = 1.0 / sqrt(3.0); w[0] = 1.0; w[1] = 1.0; return 2;
    case 3: {
        double a = sqrt(3.0 / 5.0);
        p[0] = -a; p[1] = 0.0; p[2] = a;
        w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0; return 3;
    }
    case 4: {
        double a = sqrt((3.0 + 2.0 * sqrt(6.0 / 5.0)) / 7.0);
        double b = sqrt((3.0 - 2.0 * sqrt(6.0 / 5.0)) / 7.0);
        double wa = (18.0 - sqrt(30.0)) / 36.0;
        double wb = (18.0 + sqrt(30.0)) / 36.0;
        p[0] = -a; p[1] = -b; p[2] = b; p[3] = a;
        w[0] = wa; w[1] = wb; w[2] = wb; w[3] = wa; return 4;
    }
    case 5: {
        double a = sqrt(5.0 + 2.0 * sqrt(10.0 / 7.0)) / 3.0;
        double b = sqrt(5.0 - 2.0 * sqrt(10.0 / 7.0)) / 3.0;
        double wa = (322.0 - 13.0 * sqrt(70.0)) / 900.0;
        double wb = (322.0 + 13.0 * sqrt(70.0)) / 900.0;
        double wc = 128.0 / 225.0;
        p[0] = -a; p[1] = -b; p[2] = 0.0; p[3] = b; p[4] = a;
        w[0] = wa; w[1] = wb; w[2] = wc; w[3] = wb; w[4] = wa; return 5;
    }
    default:
        if (n < 1 || n > 8) return -1;
        legendre_newton(n, p, w);
        return n;
    }
}

/* gauss_legendre_2d (Quadrature.jl:96-109): k = i + (j-1) n, xi fastest. */
int orc_quad_rule(int nquad, double* pts, double* wts) {
    double p[8], w[8];
    int n = orc_gl1d(nquad, p, w);
    if (n < 0) return -1;
    int k = 0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            pts[2 * k] = p[i]; pts[2 * k + 1] = p[j];
            wts[k] = w[i] * w[j];
            k++;
        }
    return k;
}

/* tri_quad_rule (ViewFactorKernel.jl:42-71), constants verbatim (Q3 kept). */
int orc_tri_rule(int nquad, double* pts, double* wts) {
    if (nquad <= 1) {
        pts[0] = 1.0 / 3.0; pts[1] = 1.0 / 3.0; wts[0] = 0.5; return 1;
    } else if (nquad == 2) {
        double x[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0};
        double y[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
        for (int k = 0; k < 3; k++) { pts[2 * k] = x[k]; pts[2 * k + 1] = y[k]; wts[k] = 0.5 / 3.0; }
        return 3;
    } else if (nquad == 3) {
        double a1 = 0.101286507323456, b1 = 1.0 - 2.0 * a1;
        double a2 = 0.470142064105115, b2 = 1.0 - 2.0 * a2;
        double x[7] = {a1, b1, a1, a2, b2, a2, 1.0 / 3.0};
        double y[7] = {a1, a1, b1, a2, a2, b2, 1.0 / 3.0};
        double w1 = 0.125939180544827 / 2.0, w2 = 0.132394440720100 / 2.0, w3 = 0.225 / 2.0;
        double w[7] = {w1, w1, w1, w2, w2, w2, w3};
        for (int k = 0; k < 7; k++) { pts[2 * k] = x[k]; pts[2 * k + 1] = y[k]; wts[k] = w[k]; }
        return 7;
    } else {
        double cen = 1.0 / 3.0;
        double a1 = 0.260345966079038, b1 = 1.0 - 2.0 * a1;
        double a2 = 0.065130102902216, b2 = 1.0 - 2.0 * a2;
        double a3 = 0.048690315425316, b3 = 0.312865496004875, c3 = (1.0 - a3) - b3;
        double x[13] = {a1, b1, a1, a2, b2, a2, a3, b3, a3, c3, b3, c3, cen};
        double y[13] = {a1, a1, b1, a2, a2, b2, b3, a3, c3, a3, c3, b3, cen};
        double wc = -0.149570044467670 / 2.0, wA = 0.175615257433204 / 2.0;
        double wB = 0.053347235608839 / 2.0, wC = 0.077113760890257 / 2.0;
        double w[13] = {wA, wA, wA, wB, wB, wB, wC, wC, wC, wC, wC, wC, wc};
        for (int k = 0; k < 13; k++) { pts[2 * k] = x[k]; pts[2 * k + 1] = y[k]; wts[k] = w[k]; }
        return 13;
    }
}

/* ------------------------------------------------------------------------- */
/* Shape functions.  Family codes as GPUKernels.jl:344-375:                  */
/* 0 = Quad8 (:quad), 1 = Tri6 (:tri), 2 = Quad4, 3 = Tri3; curve meshes     */
/* (mesh_dim 1, CPU only in the reference) 4 = Line3, 5 = Line2.             */
/* ------------------------------------------------------------------------- */

/* quad8_shape — Geometry.jl:36-71 */
static void quad8_shape(double x, double e, double* N, double* dx, double* de) {
    N[0] = ((0.25 * (1 - x)) * (1 - e)) * ((-x - e) - 1);
    N[1] = ((0.25 * (1 + x)) * (1 - e)) * ((x - e) - 1);
    N[2] = ((0.25 * (1 + x)) * (1 + e)) * ((x + e) - 1);
    N[3] = ((0.25 * (1 - x)) * (1 + e)) * ((-x + e) - 1);
    N[4] = (0.5 * (1 - x * x)) * (1 - e);
    N[5] = (0.5 * (1 + x)) * (1 - e * e);
    N[6] = (0.5 * (1 - x * x)) * (1 + e);
    N[7] = (0.5 * (1 - x)) * (1 - e * e);
    dx[0] = (0.25 * (1 - e)) * (2 * x + e);
    dx[1] = (0.25 * (1 - e)) * (2 * x - e);
    dx[2] = (0.25 * (1 + e)) * (2 * x + e);
    dx[3] = (0.25 * (1 + e)) * (2 * x - e);
    dx[4] = (-x) * (1 - e);
    dx[5] = 0.5 * (1 - e * e);
    dx[6] = (-x) * (1 + e);
    dx[7] = (-0.5) * (1 - e * e);
    de[0] = (0.25 * (1 - x)) * (x + 2 * e);
    de[1] = (0.25 * (1 + x)) * (-x + 2 * e);
    de[2] = (0.25 * (1 + x)) * (x + 2 * e);
    de[3] = (0.25 * (1 - x)) * (-x + 2 * e);
    de[4] = (-0.5) * (1 - x * x);
    de[5] = (-(1 + x)) * e;
    de[6] = 0.5 * (1 - x * x);
    de[7] = (-(1 - x)) * e;
}
/* quad4_shape — Geometry.jl:124-134 */
static void quad4_shape(double x, double e, double* N, double* dx, double* de) {
    N[0] = (0.25 * (1 - x)) * (1 - e);
    N[1] = (0.25 * (1 + x)) * (1 - e);
    N[2] = (0.25 * (1 + x)) * (1 + e);
    N[3] = (0.25 * (1 - x)) * (1 + e);
    dx[0] = (-0.25) * (1 - e); dx[1] = 0.25 * (1 - e); dx[2] = 0.25 * (1 + e); dx[3] = (-0.25) * (1 + e);
    de[0] = (-0.25) * (1 - x); de[1] = (-0.25) * (1 + x); de[2] = 0.25 * (1 + x); de[3] = 0.25 * (1 - x);
}
/* tri6_shape — ViewFactorKernel.jl:77-83 */
static void tri6_shape(double x, double e, double* N, double* dx, double* de) {
    double L1 = (1.0 - x) - e, L2 = x, L3 = e;
    N[0] = L1 * (2 * L1 - 1); N[1] = L2 * (2 * L2 - 1); N[2] = L3 * (2 * L3 - 1);
    N[3] = (4 * L1) * L2; N[4] = (4 * L2) * L3; N[5] = (4 * L1) * L3;
    dx[0] = (4 * L1 - 1) * (-1.0); dx[1] = 4 * L2 - 1; dx[2] = 0.0;
    dx[3] = 4 * (L2 * (-1.0) + L1); dx[4] = 4 * L3; dx[5] = (4 * L3) * (-1.0);
    de[0] = (4 * L1 - 1) * (-1.0); de[1] = 0.0; de[2] = 4 * L3 - 1;
    de[3] = (4 * L2) * (-1.0); de[4] = 4 * L2; de[5] = 4 * (L3 * (-1.0) + L1);
}
/* tri3_shape — ViewFactorKernel.jl:116-121 */
static void tri3_shape(double x, double e, double* N, double* dx, double* de) {
    N[0] = (1.0 - x) - e; N[1] = x; N[2] = e;
    dx[0] = -1.0; dx[1] = 1.0; dx[2] = 0.0;
    de[0] = -1.0; de[1] = 0.0; de[2] = 1.0;
}
/* line3_shape — Geometry.jl:226-235 (nodes: left, right, midpoint) */
static void line3_shape(double x, double* N, double* dx) {
    N[0] = ((-0.5) * x) * (1.0 - x);
    N[1] = (0.5 * x) * (1.0 + x);
    N[2] = 1.0 - x * x;
    dx[0] = -0.5 + x; dx[1] = 0.5 + x; dx[2] = (-2.0) * x;
}
/* line2_shape — Geometry.jl:292-294 */
static void line2_shape(double x, double* N, double* dx) {
    N[0] = 0.5 * (1.0 - x); N[1] = 0.5 * (1.0 + x);
    dx[0] = -0.5; dx[1] = 0.5;
}
int orc_shape(int fam, double x, double e, double* N, double* dx, double* de) {
    switch (fam) {
    case 0: quad8_shape(x, e, N, dx, de); return 8;
    case 1: tri6_shape(x, e, N, dx, de); return 6;
    case 2: quad4_shape(x, e, N, dx, de); return 4;
    case 3: tri3_shape(x, e, N, dx, de); return 3;
    case 4: line3_shape(x, N, dx); return 3;
    case 5: line2_shape(x, N, dx); return 2;
    }
    return -1;
}

/* norm of a 3-vector (quirk Q13: the reference calls StaticArrays' norm,
 * Geometry.jl:108,171, ViewFactorKernel.jl:108,146, RayCast.jl:28, and
 * Project.toml pins no version).  Mode 0 (default): the plain form
 * sqrt((x^2 + y^2) + z^2), which the current StaticArrays returns whenever
 * the result is finite and nonzero.  Mode 1: the scaled form of the versions
 * that always scale, m = max |x_i|, m * sqrt(((x/m)^2 + (y/m)^2) + (z/m)^2),
 * 0 when m = 0.  Set by orc_set_norm before a computation. */
static int g_norm_mode = 0;
void orc_set_norm(int mode) { g_norm_mode = mode; }
static inline double norm3(double x, double y, double z) {
    if (g_norm_mode == 0) return sqrt((x * x + y * y) + z * z);
    const double m = fmax(fmax(fabs(x), fabs(y)), fabs(z));
    if (m == 0.0) return 0.0;
    const double a = x / m, b = y / m, c = z / m;
    return m * sqrt((a * a + b * b) + c * c);
}

/* *_physical_point + *_normal_and_area_element (Geometry.jl:78-110,142-173;
 * ViewFactorKernel.jl:85-110,123-148): x = sum N_a x_a, c = dxdxi x dxdeta,
 * dA = norm(c) (norm3: plain sqrt of the left-folded sum of squares, or the
 * scaled form, Q13), n = c/dA. */
void orc_eval_point(int fam, const double* coords, const int32_t* nodes,
                    double xi, double eta, double* x, double* n, double* dA) {
    double N[8], dx[8], de[8];
    int nn = orc_shape(fam, xi, eta, N, dx, de);
    if (fam >= 4) {
        /* line*_physical_point + line*_normal_and_length_element (Geometry.jl:242-282,
         * 302-332): x = sum N_a x_a, dL = norm(dx/dxi), t = (dx/dxi)/dL, n = (-t_y, t_x, 0) */
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int a = 0; a < nn; a++) {
            const double* xa = coords + 3 * (int64_t)(nodes[a] - 1);
            p0 = p0 + N[a] * xa[0]; p1 = p1 + N[a] * xa[1]; p2 = p2 + N[a] * xa[2];
        }
        for (int a = 0; a < nn; a++) {
            const double* xa = coords + 3 * (int64_t)(nodes[a] - 1);
            a0 = a0 + dx[a] * xa[0]; a1 = a1 + dx[a] * xa[1]; a2 = a2 + dx[a] * xa[2];
        }
        double d = norm3(a0, a1, a2);
        x[0] = p0; x[1] = p1; x[2] = p2;
        n[0] = -(a1 / d); n[1] = a0 / d; n[2] = 0.0;
        *dA = d;
        return;
    }
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    for (int a = 0; a < nn; a++) {
        const double* xa = coords + 3 * (int64_t)(nodes[a] - 1);
        p0 = p0 + N[a] * xa[0]; p1 = p1 + N[a] * xa[1]; p2 = p2 + N[a] * xa[2];
    }
    for (int a = 0; a < nn; a++) {
        const double* xa = coords + 3 * (int64_t)(nodes[a] - 1);
        a0 = a0 + dx[a] * xa[0]; a1 = a1 + dx[a] * xa[1]; a2 = a2 + dx[a] * xa[2];
        b0 = b0 + de[a] * xa[0]; b1 = b1 + de[a] * xa[1]; b2 = b2 + de[a] * xa[2];
    }
    double c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
    double d = norm3(c0, c1, c2);
    x[0] = p0; x[1] = p1; x[2] = p2;
    n[0] = c0 / d; n[1] = c1 / d; n[2] = c2 / d;
    *dA = d;
}

/* vf_kernel — ViewFactorKernel.jl:155-166 */
double orc_vf_kernel(const double* xi, const double* ni, const double* xj, const double* nj) {
    double r0 = xj[0] - xi[0], r1 = xj[1] - xi[1], r2 = xj[2] - xi[2];
    double rr = (r0 * r0 + r1 * r1) + r2 * r2;
    if (rr < 1e-30) return 0.0;
    double r = sqrt(rr);
    double h0 = r0 / r, h1 = r1 / r, h2 = r2 / r;
    double ci = (ni[0] * h0 + ni[1] * h1) + ni[2] * h2;
    double cj = (nj[0] * (-h0) + nj[1] * (-h1)) + nj[2] * (-h2);
    if (ci <= 0.0 || cj <= 0.0) return 0.0;
    return (ci * cj) / (3.141592653589793 * rr);
}

/* vf_kernel_2d — ViewFactorKernel.jl:169-180 (per unit depth) */
double orc_vf_kernel_2d(const double* xi, const double* ni, const double* xj, const double* nj) {
    double r0 = xj[0] - xi[0], r1 = xj[1] - xi[1], r2 = xj[2] - xi[2];
    double rr = (r0 * r0 + r1 * r1) + r2 * r2;
    if (rr < 1e-30) return 0.0;
    double r = sqrt(rr);
    double h0 = r0 / r, h1 = r1 / r, h2 = r2 / r;
    double ci = (ni[0] * h0 + ni[1] * h1) + ni[2] * h2;
    double cj = (nj[0] * (-h0) + nj[1] * (-h1)) + nj[2] * (-h2);
    if (ci <= 0.0 || cj <= 0.0) return 0.0;
    return (ci * cj) / (2.0 * r);
}

/* ------------------------------------------------------------------------- */
/* BVH — src/BVH.jl:71-250 (build + stack traversal), GPUBVH.jl:69-82        */
/* (miss links, used only for the canonical stackless test counts).          */
/* ------------------------------------------------------------------------- */
typedef struct {
    double lo[3], hi[3];
    int left, right, tri_start, tri_count; /* 1-based like the reference; 0 = leaf */
} OrcNode;

typedef struct {
    OrcNode* nodes; int nn, cap;
    int* tri_idx;          /* 1-based permutation, length T */
    const double* soup;    /* 3*nv*T, Julia (3,3,T) or segment (3,2,T) column-major */
    int64_t T;
    int nv;                /* vertices per item: 3 (triangles) or 2 (segments, [xyz,endpoint]) */
    int aabb_v, aabb_c;    /* strides for AABB/centroid read: coord c of vertex v at c*aabb_c + v*aabb_v */
    int mt_v, mt_c;        /* strides for Moller-Trumbore read */
    int* miss;             /* 1-based miss links (0 = end) */
} OrcBVH;

/* soup_read modes: 0 = reference (AABB reads [xyz,vertex], MT reads [vertex,xyz], Q1),
 * 1 = geometric, soup stored [xyz,vertex,tri], 2 = geometric, soup stored [vertex,xyz,tri]. */
static void set_read(OrcBVH* b, int mode) {
    if (mode == 0) { b->aabb_c = 1; b->aabb_v = 3; b->mt_v = 1; b->mt_c = 3; }
    else if (mode == 1) { b->aabb_c = 1; b->aabb_v = 3; b->mt_c = 1; b->mt_v = 3; }
    else { b->aabb_v = 1; b->aabb_c = 3; b->mt_v = 1; b->mt_c = 3; }
}
static inline double sa(const OrcBVH* b, int64_t t, int v, int c) {
    return b->soup[3 * b->nv * t + v * b->aabb_v + c * b->aabb_c];
}
static inline double sm(const OrcBVH* b, int64_t t, int v, int c) {
    return b->soup[9 * t + v * b->mt_v + c * b->mt_c];
}

/* _triangle_aabb — BVH.jl:81-94 (also segment soups: nv endpoints) */
static void tri_aabb(const OrcBVH* b, int64_t t, double* lo, double* hi) {
    for (int c = 0; c < 3; c++) { lo[c] = sa(b, t, 0, c); hi[c] = lo[c]; }
    for (int v = 1; v < b->nv; v++)
        for (int c = 0; c < 3; c++) {
            double x = sa(b, t, v, c);
            lo[c] = jl_min(lo[c], x); hi[c] = jl_max(hi[c], x);
        }
}
/* _centroid — BVH.jl:100-109: ((0 + x1) + x2) + x3, then / nv */
static double centroid(const OrcBVH* b, int64_t t, int axis) {
    double s = 0.0;
    for (int v = 0; v < b->nv; v++) s += sa(b, t, v, axis);
    return s / (double)b->nv;
}

static int push_node(OrcBVH* b, OrcNode nd) {
    if (b->nn == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 64;
        b->nodes = (OrcNode*)realloc(b->nodes, sizeof(OrcNode) * b->cap);
    }
    b->nodes[b->nn++] = nd;
    return b->nn; /* 1-based index */
}

/* stable merge sort of idx[0..n) by key (Julia isless) — sort! is stable (Julia >= 1.9) */
static void stable_sort(int* idx, double* key, int n, int* tmpi, double* tmpk) {
    if (n <= 1) return;
    int h = n / 2;
    stable_sort(idx, key, h, tmpi, tmpk);
    stable_sort(idx + h, key + h, n - h, tmpi, tmpk);
    int i = 0, j = h, k = 0;
    while (i < h && j < n) {
        if (jl_isless(key[j], key[i])) { tmpi[k] = idx[j]; tmpk[k] = key[j]; j++; }
        else { tmpi[k] = idx[i]; tmpk[k] = key[i]; i++; }
        k++;
    }
    while (i < h) { tmpi[k] = idx[i]; tmpk[k] = key[i]; i++; k++; }
    while (j < n) { tmpi[k] = idx[j]; tmpk[k] = key[j]; j++; k++; }
    memcpy(idx, tmpi, sizeof(int) * n);
    memcpy(key, tmpk, sizeof(double) * n);
}

/* _build_recursive! — BVH.jl:112-148 (start/stop 1-based) */
static int build_rec(OrcBVH* b, int start, int stop, int* tmpi, double* tmpk, double* keys) {
    double lo[3], hi[3];
    tri_aabb(b, b->tri_idx[start - 1] - 1, lo, hi);
    for (int i = start + 1; i <= stop; i++) {
        double l2[3], h2[3];
        tri_aabb(b, b->tri_idx[i - 1] - 1, l2, h2);
        for (int c = 0; c < 3; c++) { lo[c] = jl_min(lo[c], l2[c]); hi[c] = jl_max(hi[c], h2[c]); }
    }
    int count = stop - start + 1;
    OrcNode nd;
    memcpy(nd.lo, lo, sizeof lo); memcpy(nd.hi, hi, sizeof hi);
    if (count <= LEAF_MAX) {
        nd.left = 0; nd.right = 0; nd.tri_start = start; nd.tri_count = count;
        return push_node(b, nd);
    }
    double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    int axis = 0; /* argmax: first maximum */
    for (int c = 1; c < 3; c++) if (jl_isless(ext[axis], ext[c])) axis = c;
    int mid = (start + stop) / 2;
    for (int i = start; i <= stop; i++) keys[i - start] = centroid(b, b->tri_idx[i - 1] - 1, axis);
    stable_sort(b->tri_idx + start - 1, keys, count, tmpi, tmpk);
    nd.left = 0; nd.right = 0; nd.tri_start = 0; nd.tri_count = 0;
    int me = push_node(b, nd);
    int l = build_rec(b, start, mid, tmpi, tmpk, keys);
    int r = build_rec(b, mid + 1, stop, tmpi, tmpk, keys);
    b->nodes[me - 1].left = l; b->nodes[me - 1].right = r;
    return me;
}

/* _assign_miss_links! — GPUBVH.jl:69-82 */
static void assign_miss(OrcBVH* b, int node, int next) {
    b->miss[node - 1] = next;
    OrcNode* nd = &b->nodes[node - 1];
    if (nd->left != 0) {
        int l = nd->left, r = nd->right;
        assign_miss(b, l, r);
        assign_miss(b, r, next);
    }
}

/* nv = 2: segment soup (3,2,T), read [xyz, endpoint] everywhere (MeshIO.jl:641-645, BVH.jl:307) */
static OrcBVH* bvh_build(const double* soup, int64_t T, int mode, int nv) {
    OrcBVH* b = (OrcBVH*)calloc(1, sizeof(OrcBVH));
    b->soup = soup; b->T = T; b->nv = nv; set_read(b, nv == 2 ? 1 : mode);
    b->tri_idx = (int*)malloc(sizeof(int) * (T ? T : 1));
    for (int64_t t = 0; t < T; t++) b->tri_idx[t] = (int)(t + 1);
    int* tmpi = (int*)malloc(sizeof(int) * (T ? T : 1));
    double* tmpk = (double*)malloc(sizeof(double) * (T ? T : 1));
    double* keys = (double*)malloc(sizeof(double) * (T ? T : 1));
    if (T > 0) build_rec(b, 1, (int)T, tmpi, tmpk, keys);
    free(tmpi); free(tmpk); free(keys);
    b->miss = (int*)calloc(b->nn ? b->nn : 1, sizeof(int));
    if (b->nn) assign_miss(b, 1, 0);
    return b;
}
static void bvh_free(OrcBVH* b) {
    if (!b) return;
    free(b->nodes); free(b->tri_idx); free(b->miss); free(b);
}

/* _ray_aabb — BVH.jl:155-169 */
static inline int ray_aabb(const double* o, const double* inv, const OrcNode* nd, double tmax) {
    double t1x = (nd->lo[0] - o[0]) * inv[0], t2x = (nd->hi[0] - o[0]) * inv[0];
    double t1y = (nd->lo[1] - o[1]) * inv[1], t2y = (nd->hi[1] - o[1]) * inv[1];
    double t1z = (nd->lo[2] - o[2]) * inv[2], t2z = (nd->hi[2] - o[2]) * inv[2];
    double tmin = jl_max(jl_max(jl_max(jl_min(t1x, t2x), jl_min(t1y, t2y)), jl_min(t1z, t2z)), 0.0);
    double tmx = jl_min(jl_min(jl_min(jl_max(t1x, t2x), jl_max(t1y, t2y)), jl_max(t1z, t2z)), tmax);
    return tmin <= tmx;
}

/* _ray_triangle — BVH.jl:181-201 with t_min = 0.0; returns t or +Inf */
static inline double ray_tri(const double* o, const double* d, const double* v0, const double* v1,
                             const double* v2) {
    double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
    double h[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double a = (e1[0] * h[0] + e1[1] * h[1]) + e1[2] * h[2];
    if (fabs(a) < 1e-10) return INFINITY;
    double f = 1.0 / a;
    double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    double u = f * ((s[0] * h[0] + s[1] * h[1]) + s[2] * h[2]);
    if (u < 0.0 || u > 1.0) return INFINITY;
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    double v = f * ((d[0] * q[0] + d[1] * q[1]) + d[2] * q[2]);
    if (v < 0.0 || u + v > 1.0) return INFINITY;
    double t = f * ((e2[0] * q[0] + e2[1] * q[1]) + e2[2] * q[2]);
    return t > 0.0 ? t : INFINITY;
}

static inline void tri_mt(const OrcBVH* b, int64_t t, double* v0, double* v1, double* v2) {
    for (int c = 0; c < 3; c++) { v0[c] = sm(b, t, 0, c); v1[c] = sm(b, t, 1, c); v2[c] = sm(b, t, 2, c); }
}

typedef struct { int64_t rays, aabb, tri; } OrcCounters;

/* intersect_ray_bvh — BVH.jl:213-250 (stack: push left then right) */
static int intersect_ray_bvh(const OrcBVH* b, const double* o, const double* d, double tmax) {
    if (b->nn == 0) return 0;
    double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
    int stack[128]; int sp = 0;
    stack[sp++] = 1;
    while (sp > 0) {
        const OrcNode* nd = &b->nodes[stack[--sp] - 1];
        if (!ray_aabb(o, inv, nd, tmax)) continue;
        if (nd->left == 0) {
            for (int k = nd->tri_start; k <= nd->tri_start + nd->tri_count - 1; k++) {
                int64_t t = b->tri_idx[k - 1] - 1;
                double v0[3], v1[3], v2[3];
                tri_mt(b, t, v0, v1, v2);
                double th = ray_tri(o, d, v0, v1, v2);
                if (th < tmax) return 1;
            }
        } else {
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return 0;
}

/* Canonical stackless order (GPUBVH.jl:202-285 control flow, CPU decision
 * arithmetic) — used only to count AABB / triangle tests for the flop model. */
static int intersect_stackless_count(const OrcBVH* b, const double* o, const double* d, double tmax,
                                     OrcCounters* cnt) {
    if (b->nn == 0) return 0;
    double inv[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
    int nidx = 1;
    while (nidx > 0) {
        const OrcNode* nd = &b->nodes[nidx - 1];
        cnt->aabb++;
        if (ray_aabb(o, inv, nd, tmax)) {
            if (nd->tri_count > 0) {
                for (int k = nd->tri_start; k <= nd->tri_start + nd->tri_count - 1; k++) {
                    int64_t t = b->tri_idx[k - 1] - 1;
                    double v0[3], v1[3], v2[3];
                    tri_mt(b, t, v0, v1, v2);
                    cnt->tri++;
                    if (ray_tri(o, d, v0, v1, v2) < tmax) return 1;
                }
                nidx = b->miss[nidx - 1];
            } else {
                nidx = nd->left;
            }
        } else {
            nidx = b->miss[nidx - 1];
        }
    }
    return 0;
}

static int intersect_seg_bvh(const OrcBVH* b, const double* xi, const double* xj, OrcCounters* cnt);

/* is_visible — RayCast.jl:20-32 */
static int is_visible(const OrcBVH* b, const double* xi, const double* xj, OrcCounters* cnt) {
    if (b->nv == 2) { /* mesh_dim = 1 */
        if (cnt) cnt->rays++;
        return !intersect_seg_bvh(b, xi, xj, cnt);
    }
    double d[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
    double L = norm3(d[0], d[1], d[2]);
    if (L < DBL_EPSILON) return 1;
    double dir[3] = {d[0] / L, d[1] / L, d[2] / L};
    if (cnt) {
        cnt->rays++;
        return !intersect_stackless_count(b, xi, dir, L, cnt);
    }
    return !intersect_ray_bvh(b, xi, dir, L);
}

/* _seg_blocks_ray — BVH.jl:264-277 (tol = 1e-8; xy only) */
static inline int seg_blocks_ray(const double* xi, const double* xj, const double* v0, const double* v1) {
    const double tol = 1e-8;
    double dx = xj[0] - xi[0], dy = xj[1] - xi[1];
    double ex = v1[0] - v0[0], ey = v1[1] - v0[1];
    double denom = dx * ey - dy * ex;
    if (fabs(denom) < tol) return 0;
    double fx = v0[0] - xi[0], fy = v0[1] - xi[1];
    double s = (fx * ey - fy * ex) / denom;
    double t = (fx * dy - fy * dx) / denom;
    return (tol < s && s < 1.0 - tol) && (tol < t && t < 1.0 - tol);
}

/* intersect_seg_bvh — BVH.jl:286-320 (stack: push left then right; 2-D box
 * overlap with the ray's bounding box) */
static int intersect_seg_bvh(const OrcBVH* b, const double* xi, const double* xj, OrcCounters* cnt) {
    if (b->nn == 0) return 0;
    double xlo = jl_min(xi[0], xj[0]), xhi = jl_max(xi[0], xj[0]);
    double ylo = jl_min(xi[1], xj[1]), yhi = jl_max(xi[1], xj[1]);
    int stack[128]; int sp = 0;
    stack[sp++] = 1;
    while (sp > 0) {
        const OrcNode* nd = &b->nodes[stack[--sp] - 1];
        if (cnt) cnt->aabb++;
        if (nd->lo[0] > xhi || nd->hi[0] < xlo || nd->lo[1] > yhi || nd->hi[1] < ylo) continue;
        if (nd->left == 0) {
            for (int k = nd->tri_start; k <= nd->tri_start + nd->tri_count - 1; k++) {
                int64_t t = b->tri_idx[k - 1] - 1;
                const double* s = b->soup + 6 * t;
                double v0[3] = {s[0], s[1], 0.0}, v1[3] = {s[3], s[4], 0.0};
                if (cnt) cnt->tri++;
                if (seg_blocks_ray(xi, xj, v0, v1)) return 1;
            }
        } else {
            stack[sp++] = nd->left;
            stack[sp++] = nd->right;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Mesh + pair integrator                                                     */
/* ------------------------------------------------------------------------- */
typedef struct {
    const double* coords;   /* 3*Nn */
    int64_t N;
    const int8_t* fam;      /* N */
    const int32_t* nodes;   /* 8*N, 1-based node ids, padded */
    const int32_t* group;   /* N */
    int nquad;
    int nq_quad, nq_tri, nq_line;
    double qp[2 * ORC_MAXPTS], qw[ORC_MAXPTS];
    double tp[2 * 13], tw[13];
    double lp[8], lw[8];
    int dim2;               /* curve mesh (mesh_dim 1): Line3/Line2 elements, segment occluders */
} OrcMesh;

static void mesh_rules(OrcMesh* m, int nquad) {
    m->nquad = nquad;
    m->nq_quad = orc_quad_rule(nquad, m->qp, m->qw);
    m->nq_tri = orc_tri_rule(nquad, m->tp, m->tw);
    m->nq_line = orc_gl1d(nquad, m->lp, m->lw);
}

static int quad_points(const OrcMesh* m, int64_t e, double* xs, double* ns, double* dA, double* w) {
    int f = m->fam[e];
    if (f >= 4) { /* _quad_points, curve branch (ViewFactorKernel.jl:284-301): 1-D Gauss-Legendre */
        for (int k = 0; k < m->nq_line; k++) {
            orc_eval_point(f, m->coords, m->nodes + 8 * e, m->lp[k], 0.0, xs + 3 * k, ns + 3 * k, dA + k);
            w[k] = m->lw[k];
        }
        return m->nq_line;
    }
    int isq = (f == 0 || f == 2);
    int np = isq ? m->nq_quad : m->nq_tri;
    const double* pts = isq ? m->qp : m->tp;
    const double* wts = isq ? m->qw : m->tw;
    for (int k = 0; k < np; k++) {
        orc_eval_point(f, m->coords, m->nodes + 8 * e, pts[2 * k], pts[2 * k + 1], xs + 3 * k, ns + 3 * k, dA + k);
        w[k] = wts[k];
    }
    return np;
}

static inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

/* element_pair_view_factor — ViewFactorKernel.jl:199-241 */
static double pair_vf(const OrcMesh* m, int64_t i, int64_t j, const OrcBVH* bvh, double* Li,
                      uint32_t* bl_count, uint32_t* bl_hash, OrcCounters* cnt) {
    double xi[3 * ORC_MAXPTS], ni[3 * ORC_MAXPTS], ai[ORC_MAXPTS], wi[ORC_MAXPTS];
    double xj[3 * ORC_MAXPTS], nj[3 * ORC_MAXPTS], aj[ORC_MAXPTS], wj[ORC_MAXPTS];
    int npi = quad_points(m, i, xi, ni, ai, wi);
    int npj = quad_points(m, j, xj, nj, aj, wj);
    double F = 0.0, L = 0.0;
    uint32_t bc = 0, bh = 0;
    for (int p = 0; p < npi; p++) {
        L += wi[p] * ai[p];
        double inner = 0.0;
        for (int q = 0; q < npj; q++) {
            double K = m->dim2 ? orc_vf_kernel_2d(xi + 3 * p, ni + 3 * p, xj + 3 * q, nj + 3 * q)
                               : orc_vf_kernel(xi + 3 * p, ni + 3 * p, xj + 3 * q, nj + 3 * q);
            if (K == 0.0) continue;
            if (bvh && !is_visible(bvh, xi + 3 * p, xj + 3 * q, cnt)) {
                bc++; bh += mix32((uint32_t)(p * npj + q) + 0x9e3779b9U);
                continue;
            }
            inner += (wj[q] * K) * aj[q];
        }
        F += (wi[p] * inner) * ai[p];
    }
    if (Li) *Li = L;
    if (bl_count) *bl_count = bc;
    if (bl_hash) *bl_hash = bh;
    return F;
}

/* ------------------------------------------------------------------------- */
/* Duffy / Sauter-Schwab — src/DuffyKernel.jl (restated verbatim, Q2 kept)   */
/* ------------------------------------------------------------------------- */
static const double QUAD_CORNER_UV[4][2] = {{0.0, 0.0}, {1.0, 0.0}, {1.0, 1.0}, {0.0, 1.0}};

/* singularity_type — DuffyKernel.jl:95-115; returns 0 none, 1 vertex, 2 edge */
int orc_singularity(const int32_t* ni, int fi, const int32_t* nj, int fj, int* ci, int* cj) {
    if (fi != 0 || fj != 0) return 0;
    int n = 0;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++)
            if (ni[a] == nj[b]) {
                if (n < 2) { ci[n] = a + 1; cj[n] = b + 1; }
                n++;
            }
    if (n == 0) return 0;
    if (n == 1) return 1;
    return 2;
}

/* _eval_quad — DuffyKernel.jl:134-139 */
static void eval_quad01(const OrcMesh* m, int64_t e, double u, double v, double* x, double* n, double* dA) {
    double xi = 2 * u - 1.0, eta = 2 * v - 1.0;
    orc_eval_point(0, m->coords, m->nodes + 8 * e, xi, eta, x, n, dA);
    *dA = *dA * 4.0;
}

/* _vertex_map — DuffyKernel.jl:288-314 */
static void vertex_map(int region, double r, double e1, double e2, double e3, double u0i, double v0i,
                       double u0j, double v0j, double* u, double* v, double* s, double* t) {
    double du, dv, ds, dt;
    switch (region) {
    case 1: du = r; dv = r * e1; ds = r * e2; dt = (r * e2) * e3; break;
    case 2: du = r; dv = r * e1; ds = (r * e2) * e3; dt = r * e2; break;
    case 3: du = r * e1; dv = r; ds = r * e2; dt = (r * e2) * e3; break;
    case 4: du = r * e1; dv = r; ds = (r * e2) * e3; dt = r * e2; break;
    case 5: du = r * e2; dv = (r * e2) * e3; ds = r; dt = r * e1; break;
    case 6: du = (r * e2) * e3; dv = r * e2; ds = r; dt = r * e1; break;
    case 7: du = r * e2; dv = (r * e2) * e3; ds = r * e1; dt = r; break;
    default: du = (r * e2) * e3; dv = r * e2; ds = r * e1; dt = r; break;
    }
    *u = u0i + (u0i == 0.0 ? du : -du);
    *v = v0i + (v0i == 0.0 ? dv : -dv);
    *s = u0j + (u0j == 0.0 ? ds : -ds);
    *t = v0j + (v0j == 0.0 ? dt : -dt);
}

/* _edge_map — DuffyKernel.jl:375-408 (region 4 uses its final "clean" form) */
static void edge_map(int region, double r, double e1, double e2, double e3, double* u, double* v, double* s,
                     double* t) {
    switch (region) {
    case 1:
        *u = r * (1 - e1 * (1 - e3));
        *v = ((r * e1) * e2) * e3;
        *s = (r * (1 - e1 * (1 - e3)) - ((r * e1) * (1 - e2)) * e3) + ((r * e1) * e2) * e3;
        *t = ((r * e1) * (1 - e2)) * e3;
        break;
    case 2:
        *u = r * (1 - (e1 * e2) * e3);
        *v = ((r * e1) * e3) * (1 - e2);
        *s = (r * (1 - (e1 * e2) * e3) + ((r * e1) * e2) * e3) - (r * e1) * e3;
        *t = ((r * e1) * e2) * e3;
        break;
    case 3:
        *u = r * (1 - e1 * e2);
        *v = ((r * e1) * e2) * e3;
        *s = r * (1 - e1 * e2) - ((r * e1) * e2) * (1 - e3);
        *t = ((r * e1) * e2) * (1 - e3);
        break;
    case 4:
        *u = r * (e2 * e3 + e1 * (1 - e2 * e3));
        *v = ((r * (1 - e1)) * e2) * e3;
        *s = (r * e2) * e3;
        *t = ((r * (1 - e1)) * e2) * e3;
        break;
    default:
        *u = (r * e1) * e3;
        *v = ((r * (1 - e1)) * e2) * e3;
        *s = ((r * e1) * e3) * (1 - e2) + ((r * e1) * e2) * e3;
        *t = ((r * (1 - e1)) * e2) * e3;
        break;
    }
}

/* _corners_to_edge — DuffyKernel.jl:411-419 */
static int corners_to_edge(int c1, int c2) {
    int a = c1 < c2 ? c1 : c2, b = c1 < c2 ? c2 : c1;
    if (a == 1 && b == 2) return 1;
    if (a == 2 && b == 3) return 2;
    if (a == 3 && b == 4) return 3;
    if (a == 1 && b == 4) return 4;
    return 1;
}
/* _rotate_to_edge — DuffyKernel.jl:423-429 */
static void rotate_to_edge(double* u, double* v, int e) {
    double a = *u, b = *v;
    if (e == 2) { *u = 1.0 - b; *v = a; }
    else if (e == 3) { *u = 1.0 - a; *v = 1.0 - b; }
    else if (e == 4) { *u = b; *v = 1.0 - a; }
}

/* One region map at one point (test entry): kind 1 the common-vertex map's
 * corner-relative offsets (du, dv, ds, dt), kind 2 the common-edge map in the
 * bottom-edge frame (u, v, s, t). */
void orc_duffy_map(int kind, int region, double r, double e1, double e2, double e3, double* out) {
    if (kind == 1) vertex_map(region, r, e1, e2, e3, 0.0, 0.0, 0.0, 0.0, out, out + 1, out + 2, out + 3);
    else edge_map(region, r, e1, e2, e3, out, out + 1, out + 2, out + 3);
}

/* The point rule of _vertex_integral / _edge_integral (DuffyKernel.jl:222-283,
 * 326-371) for one pair of shared corners: every (region, rho, eta1, eta2, eta3)
 * in the reference's loop order whose (u, v, s, t) falls inside [0,1]^4, with its
 * weight ((((w_rho w_eta1) w_eta2) w_eta3) rho^3).  kind 1: common vertex
 * (ci[0], cj[0]); kind 2: common edge (ci[0..1], cj[0..1]).  uvst / w / region
 * may be NULL (count only); returns the number of points. */
int64_t orc_duffy_rule(int kind, int nquad, const int* ci, const int* cj, double* uvst, double* w, int* region_out) {
    double p1[8], w1[8], p01[8], w01[8];
    int nq = orc_gl1d(nquad, p1, w1);
    for (int k = 0; k < nq; k++) { p01[k] = (p1[k] + 1.0) / 2.0; w01[k] = w1[k] / 2.0; }
    double u0i = 0, v0i = 0, u0j = 0, v0j = 0;
    int edge_i = 1, edge_j = 1;
    if (kind == 1) {
        u0i = QUAD_CORNER_UV[ci[0] - 1][0]; v0i = QUAD_CORNER_UV[ci[0] - 1][1];
        u0j = QUAD_CORNER_UV[cj[0] - 1][0]; v0j = QUAD_CORNER_UV[cj[0] - 1][1];
    } else {
        edge_i = corners_to_edge(ci[0], ci[1]);
        edge_j = corners_to_edge(cj[0], cj[1]);
    }
    int nreg = kind == 1 ? 8 : 5;
    int64_t n = 0;
    for (int region = 1; region <= nreg; region++)
        for (int a = 0; a < nq; a++)
            for (int b = 0; b < nq; b++)
                for (int c = 0; c < nq; c++)
                    for (int d = 0; d < nq; d++) {
                        double r = p01[a], wr = w01[a], e1 = p01[b], we1 = w01[b];
                        double e2 = p01[c], we2 = w01[c], e3 = p01[d], we3 = w01[d];
                        double u, v, s, t;
                        if (kind == 1) vertex_map(region, r, e1, e2, e3, u0i, v0i, u0j, v0j, &u, &v, &s, &t);
                        else {
                            edge_map(region, r, e1, e2, e3, &u, &v, &s, &t);
                            rotate_to_edge(&u, &v, edge_i);
                            rotate_to_edge(&s, &t, edge_j);
                        }
                        if (u < 0 || u > 1 || v < 0 || v > 1 || s < 0 || s > 1 || t < 0 || t > 1) continue;
                        if (uvst) { uvst[4 * n] = u; uvst[4 * n + 1] = v; uvst[4 * n + 2] = s; uvst[4 * n + 3] = t; }
                        if (w) w[n] = (((wr * we1) * we2) * we3) * ((r * r) * r);
                        if (region_out) region_out[n] = region;
                        n++;
                    }
    return n;
}

/* _vertex_integral / _edge_integral — DuffyKernel.jl:222-283, 326-371: the
 * kernel summed over the point rule above, term by term in the same order */
static double duffy_integral(const OrcMesh* m, int64_t i, int64_t j, int kind, const int* ci, const int* cj,
                             const OrcBVH* bvh) {
    int64_t n = orc_duffy_rule(kind, m->nquad, ci, cj, NULL, NULL, NULL);
    double* uvst = (double*)malloc(sizeof(double) * 4 * (size_t)(n ? n : 1));
    double* w = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
    orc_duffy_rule(kind, m->nquad, ci, cj, uvst, w, NULL);
    double F = 0.0;
    for (int64_t k = 0; k < n; k++) {
        double xi[3], ni[3], ai, xj[3], nj[3], aj;
        eval_quad01(m, i, uvst[4 * k], uvst[4 * k + 1], xi, ni, &ai);
        eval_quad01(m, j, uvst[4 * k + 2], uvst[4 * k + 3], xj, nj, &aj);
        double K = orc_vf_kernel(xi, ni, xj, nj);
        if (K == 0.0) continue;
        if (bvh && !is_visible(bvh, xi, xj, 0)) continue;
        F += ((w[k] * K) * ai) * aj;
    }
    free(uvst);
    free(w);
    return F;
}

/* element_pair_view_factor_duffy — DuffyKernel.jl:452-487 (raw only) */
static double pair_duffy(const OrcMesh* m, int64_t i, int64_t j, const OrcBVH* bvh) {
    if (m->fam[i] != 0 || m->fam[j] != 0) return pair_vf(m, i, j, bvh, 0, 0, 0, 0);
    int ci[2], cj[2];
    int kind = orc_singularity(m->nodes + 8 * i, 0, m->nodes + 8 * j, 0, ci, cj);
    if (kind == 0) return pair_vf(m, i, j, bvh, 0, 0, 0, 0); /* _standard_integral == pair_vf bitwise */
    return duffy_integral(m, i, j, kind, ci, cj, bvh);
}

/* ------------------------------------------------------------------------- */
/* Monte Carlo: the GPU stream definition (GPUMCKernels.jl:37-50,146-262),   */
/* restated on the CPU, with the CPU visibility semantics (Q4, Q5).          */
/* ------------------------------------------------------------------------- */
static inline double xorshift64(uint64_t* s) {
    uint64_t x = *s;
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    *s = x;
    return (double)(x >> 11) / (double)0x001FFFFFFFFFFFFFULL;
}
uint64_t orc_init_rng(uint64_t seed, int64_t tid) {
    uint64_t z = seed + (uint64_t)tid * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
/* _sample_on_element — GPUMCKernels.jl:266-281 */
static void mc_sample(const OrcMesh* m, int64_t e, double u1, double u2, double* x, double* n, double* dA) {
    int f = m->fam[e];
    double xi, eta;
    if (f >= 4) { xi = u1; eta = 0.0; } /* curve elements: u1 is the point on [-1, 1] */
    else if (f == 0 || f == 2) { xi = 2.0 * u1 - 1.0; eta = 2.0 * u2 - 1.0; }
    else { xi = u1; eta = u2; if (xi + eta > 1.0) { xi = 1.0 - xi; eta = 1.0 - eta; } }
    orc_eval_point(f, m->coords, m->nodes + 8 * e, xi, eta, x, n, dA);
}
/* per pair raw (and the element-i area estimate), pair (i,j) 0-based, tid = i*N + j + 1 */
/* Curve meshes (no reference GPU stream exists for them): the same per-pair
 * stream; sample k of both elements lies in stratum k of n equal strata of
 * [-1, 1], as _stratified_line_points (MCKernel.jl:107-115) places them:
 * xi = -1 + ((k + u) * 2) / n, with u1 for element i and u3 for element j. */
static double mc_pair_line(const OrcMesh* m, int64_t i, int64_t j, int nsamp, uint64_t seed, const OrcBVH* bvh,
                           double* Ai_out) {
    uint64_t st = orc_init_rng(seed, i * m->N + j + 1);
    double Ai = 0.0, Ks = 0.0;
    for (int k = 0; k < nsamp; k++) {
        double u1 = xorshift64(&st), u2 = xorshift64(&st);
        (void)u2;
        double xi[3], ni[3], ai, xj[3], nj[3], aj;
        mc_sample(m, i, -1.0 + ((k + u1) * 2.0) / nsamp, 0.0, xi, ni, &ai);
        Ai += ai;
        double u3 = xorshift64(&st), u4 = xorshift64(&st);
        (void)u4;
        mc_sample(m, j, -1.0 + ((k + u3) * 2.0) / nsamp, 0.0, xj, nj, &aj);
        double K = orc_vf_kernel_2d(xi, ni, xj, nj);
        if (K > 0.0 && bvh && !is_visible(bvh, xi, xj, 0)) K = 0.0;
        Ks += (K * ai) * aj;
    }
    if (Ai_out) *Ai_out = (Ai * 2.0) / nsamp;
    return ((Ks * 2.0) * 2.0) / nsamp;
}

static double mc_pair(const OrcMesh* m, int64_t i, int64_t j, int nsamp, uint64_t seed, const OrcBVH* bvh,
                      double* Ai_out) {
    if (m->dim2) return mc_pair_line(m, i, j, nsamp, seed, bvh, Ai_out);
    uint64_t st = orc_init_rng(seed, i * m->N + j + 1);
    int s = (int)floor(sqrt((double)nsamp));
    double Ai = 0.0, Aj = 0.0, Ks = 0.0;
    int k = 0;
    for (int si = 0; si < s; si++)
        for (int sj = 0; sj < s; sj++) {
            k++;
            double u1 = xorshift64(&st), u2 = xorshift64(&st);
            double xi[3], ni[3], ai, xj[3], nj[3], aj;
            mc_sample(m, i, (si + u1) / s, (sj + u2) / s, xi, ni, &ai);
            Ai += ai;
            double u3 = xorshift64(&st), u4 = xorshift64(&st);
            mc_sample(m, j, (si + u3) / s, (sj + u4) / s, xj, nj, &aj);
            Aj += aj;
            double K = orc_vf_kernel(xi, ni, xj, nj);
            if (K > 0.0 && bvh && !is_visible(bvh, xi, xj, 0)) K = 0.0;
            Ks += (K * ai) * aj;
        }
    for (; k < nsamp; k++) {
        double u1 = xorshift64(&st), u2 = xorshift64(&st);
        double xi[3], ni[3], ai, xj[3], nj[3], aj;
        mc_sample(m, i, u1, u2, xi, ni, &ai);
        double u3 = xorshift64(&st), u4 = xorshift64(&st);
        mc_sample(m, j, u3, u4, xj, nj, &aj);
        Ai += ai; Aj += aj;
        double K = orc_vf_kernel(xi, ni, xj, nj);
        if (K > 0.0 && bvh && !is_visible(bvh, xi, xj, 0)) K = 0.0;
        Ks += (K * ai) * aj;
    }
    double ri = (m->fam[i] == 0 || m->fam[i] == 2) ? 4.0 : 0.5;
    double rj = (m->fam[j] == 0 || m->fam[j] == 2) ? 4.0 : 0.5;
    if (Ai_out) *Ai_out = (Ai * ri) / nsamp;
    return ((Ks * ri) * rj) / nsamp;
}

/* ------------------------------------------------------------------------- */
/* _compute_cpu — src/Assembly.jl:102-225                                    */
/* ------------------------------------------------------------------------- */
typedef struct { int n; int* groups; OrcBVH* bvh; double* merged; } ActiveSet;

static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/*
 * method: 0 quadrature, 1 Monte Carlo (GPU stream), 2 Duffy.
 * schedule: 0 = contiguous equal row chunks (Threads.@threads), 1 = dynamic rows.
 * Rows [row_begin, row_end) of the upper triangle are computed (full run: 0..N).
 * raw: N*N column-major (or NULL); A: N.  audit_*: N*N or NULL.
 * counters (3): rays, aabb tests, triangle tests in the canonical stackless order
 * (only if count != 0; then visibility uses the stackless order, same decisions).
 */
int orc_compute(const double* coords, int64_t Nn, int64_t N, const int8_t* fam, const int32_t* nodes8,
                const int32_t* group, int nquad, const double* soup, const int32_t* soup_group, int64_t T,
                const int32_t* obs, int n_obs, int soup_mode, int method, int n_samples, uint64_t seed,
                int nthreads, int schedule, int64_t row_begin, int64_t row_end, double* raw, double* A,
                uint32_t* audit_count, uint32_t* audit_hash, int count, int64_t* counters, int self_vf) {
    (void)Nn;
    OrcMesh m;
    memset(&m, 0, sizeof m);
    m.coords = coords; m.N = N; m.fam = fam; m.nodes = nodes8; m.group = group;
    mesh_rules(&m, nquad);
    if (m.nq_quad < 0) return -1;
    /* mesh_dim 1: every element a curve element; occluders are segment soups (3,2,T) */
    m.dim2 = N > 0 && fam[0] >= 4;
    for (int64_t e = 0; e < N; e++) if ((fam[e] >= 4) != m.dim2) return -2;
    const int nv = m.dim2 ? 2 : 3;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    /* areas (serial pass, Assembly.jl:186-190 / 164-169) */
    if (A) {
        for (int64_t i = 0; i < N; i++) {
            if (method == 1) {
                double Ai;
                mc_pair(&m, i, i, n_samples, seed, 0, &Ai);
                A[i] = Ai;
            } else {
                double Li;
                pair_vf(&m, i, i, 0, &Li, 0, 0, 0);
                A[i] = Li;
            }
        }
    }
    /* BVH cache per active set, prebuilt before the threaded loop (Q8) */
    int ngrp_max = 0;
    int* gtags = NULL;
    int G = 0;
    ActiveSet* sets = NULL; int nsets = 0;
    int* set_of = NULL; /* G*G */
    if (n_obs > 0 && T > 0) {
        /* distinct radiating group tags */
        gtags = (int*)malloc(sizeof(int) * (N ? N : 1));
        for (int64_t i = 0; i < N; i++) {
            int g = group[i], found = 0;
            for (int k = 0; k < G; k++) if (gtags[k] == g) { found = 1; break; }
            if (!found) gtags[G++] = g;
        }
        qsort(gtags, G, sizeof(int), cmp_int);
        set_of = (int*)malloc(sizeof(int) * G * G);
        sets = (ActiveSet*)calloc(G * G, sizeof(ActiveSet));
        int* act = (int*)malloc(sizeof(int) * (n_obs ? n_obs : 1));
        for (int a = 0; a < G; a++)
            for (int b = 0; b < G; b++) {
                int na = 0;
                for (int k = 0; k < n_obs; k++)
                    if (obs[k] != gtags[a] && obs[k] != gtags[b]) act[na++] = obs[k];
                qsort(act, na, sizeof(int), cmp_int);
                int found = -1;
                for (int s = 0; s < nsets; s++)
                    if (sets[s].n == na && memcmp(sets[s].groups, act, sizeof(int) * na) == 0) { found = s; break; }
                if (found < 0) {
                    ActiveSet* as = &sets[nsets];
                    as->n = na;
                    as->groups = (int*)malloc(sizeof(int) * (na ? na : 1));
                    memcpy(as->groups, act, sizeof(int) * na);
                    /* merged soup: groups in sorted active order, triangles in given order */
                    int64_t tot = 0;
                    for (int g = 0; g < na; g++)
                        for (int64_t t = 0; t < T; t++) if (soup_group[t] == act[g]) tot++;
                    as->merged = NULL; as->bvh = NULL;
                    if (tot > 0) {
                        as->merged = (double*)malloc(sizeof(double) * 3 * nv * tot);
                        int64_t w = 0;
                        for (int g = 0; g < na; g++)
                            for (int64_t t = 0; t < T; t++)
                                if (soup_group[t] == act[g]) {
                                    memcpy(as->merged + 3 * nv * w, soup + 3 * nv * t, 3 * nv * sizeof(double));
                                    w++;
                                }
                        as->bvh = bvh_build(as->merged, tot, soup_mode, nv);
                    }
                    found = nsets++;
                }
                set_of[a * G + b] = found;
            }
        free(act);
        ngrp_max = G;
    }
    (void)ngrp_max;
    int64_t nrays = 0, naabb = 0, ntri = 0;
    int64_t rb = row_begin < 0 ? 0 : row_begin, re = row_end > N ? N : row_end;
    int64_t next_row = 0;
#pragma omp parallel reduction(+ : nrays, naabb, ntri)
    {
        OrcCounters cnt = {0, 0, 0};
        int64_t nrow = re - rb;
        int nth = 1, tid = 0;
#ifdef _OPENMP
        nth = omp_get_num_threads(); tid = omp_get_thread_num();
#endif
        int64_t c0 = rb + (nrow * tid) / nth, c1 = rb + (nrow * (tid + 1)) / nth;
        int64_t i = schedule == 0 ? c0 : 0;
        for (;;) {
            if (schedule == 0) {
                if (i >= c1) break;
            } else {
                int64_t got;
#pragma omp atomic capture
                got = next_row++;
                i = rb + got;
                if (i >= re) break;
            }
            int a = -1;
            if (sets) {
                int g = group[i];
                for (int k = 0; k < G; k++) if (gtags[k] == g) { a = k; break; }
            }
            /* self_vf: the pair (i, i) too (Assembly.jl:173,193) */
            for (int64_t j = self_vf ? i : i + 1; j < N; j++) {
                const OrcBVH* bvh = NULL;
                if (sets) {
                    int g = group[j], b = 0;
                    for (int k = 0; k < G; k++) if (gtags[k] == g) { b = k; break; }
                    bvh = sets[set_of[a * G + b]].bvh;
                }
                double v;
                uint32_t bc = 0, bh = 0;
                if (method == 1) v = mc_pair(&m, i, j, n_samples, seed, bvh, 0);
                else if (method == 2) v = pair_duffy(&m, i, j, bvh);
                else v = pair_vf(&m, i, j, bvh, 0, &bc, &bh, count ? &cnt : 0);
                if (raw) { raw[i + N * j] = v; raw[j + N * i] = v; }
                if (audit_count) { audit_count[i + N * j] = bc; audit_count[j + N * i] = bc; }
                if (audit_hash) { audit_hash[i + N * j] = bh; audit_hash[j + N * i] = bh; }
            }
            if (schedule == 0) i++;
        }
        nrays += cnt.rays; naabb += cnt.aabb; ntri += cnt.tri;
    }
    if (counters) { counters[0] = nrays; counters[1] = naabb; counters[2] = ntri; }
    if (sets) {
        for (int s = 0; s < nsets; s++) { bvh_free(sets[s].bvh); free(sets[s].merged); free(sets[s].groups); }
        free(sets); free(set_of); free(gtags);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Small entry points used by the unit tests                                  */
/* ------------------------------------------------------------------------- */
double orc_pair(const double* coords, const int8_t* fam, const int32_t* nodes8, int64_t i, int64_t j, int nquad,
                double* Li) {
    OrcMesh m;
    memset(&m, 0, sizeof m);
    m.coords = coords; m.fam = fam; m.nodes = nodes8; m.N = (i > j ? i : j) + 1;
    mesh_rules(&m, nquad);
    m.dim2 = fam[i] >= 4;
    return pair_vf(&m, i, j, 0, Li, 0, 0, 0);
}

/* Test entries take mode 0..2 (triangle soup readings) or 3 (segment soup (3,2,T)). */
static int mode_nv(int mode) { return mode == 3 ? 2 : 3; }

/* Build a BVH and export it: nodes (lo3, hi3 as 6 doubles; left,right,start,count,miss as 5 ints),
 * tri_idx (1-based).  Returns node count, or -(needed) if cap is too small. */
int orc_bvh_export(const double* soup, int64_t T, int mode, int cap, double* box, int32_t* meta, int32_t* tri_idx) {
    OrcBVH* b = bvh_build(soup, T, mode, mode_nv(mode));
    int nn = b->nn;
    if (nn > cap) { bvh_free(b); return -nn; }
    for (int k = 0; k < nn; k++) {
        for (int c = 0; c < 3; c++) { box[6 * k + c] = b->nodes[k].lo[c]; box[6 * k + 3 + c] = b->nodes[k].hi[c]; }
        meta[5 * k + 0] = b->nodes[k].left; meta[5 * k + 1] = b->nodes[k].right;
        meta[5 * k + 2] = b->nodes[k].tri_start; meta[5 * k + 3] = b->nodes[k].tri_count;
        meta[5 * k + 4] = b->miss[k];
    }
    for (int64_t t = 0; t < T; t++) tri_idx[t] = b->tri_idx[t];
    bvh_free(b);
    return nn;
}

/* intersect_ray_bvh on a freshly built tree (test entry).  mode as above. */
int orc_ray_hits(const double* soup, int64_t T, int mode, const double* o, const double* d, double tmax) {
    OrcBVH* b = bvh_build(soup, T, mode, 3);
    int r = intersect_ray_bvh(b, o, d, tmax);
    bvh_free(b);
    return r;
}

/* brute force any-hit over all triangles (tree-independent check) */
int orc_ray_hits_brute(const double* soup, int64_t T, int mode, const double* o, const double* d, double tmax) {
    OrcBVH b;
    memset(&b, 0, sizeof b);
    b.soup = soup; b.T = T; set_read(&b, mode);
    for (int64_t t = 0; t < T; t++) {
        double v0[3], v1[3], v2[3];
        tri_mt(&b, t, v0, v1, v2);
        if (ray_tri(o, d, v0, v1, v2) < tmax) return 1;
    }
    return 0;
}

/* is_visible on a freshly built tree, batch of rays (test entry) */
int orc_visible_batch(const double* soup, int64_t T, int mode, const double* xi, const double* xj, int64_t n,
                      uint8_t* out) {
    OrcBVH* b = bvh_build(soup, T, mode, mode_nv(mode));
    for (int64_t k = 0; k < n; k++) out[k] = (uint8_t)is_visible(b, xi + 3 * k, xj + 3 * k, 0);
    bvh_free(b);
    return 0;
}

/* brute force: does any segment of a (3,2,T) soup block the 2-D path xi -> xj (test entry) */
int orc_seg_hits_brute(const double* soup, int64_t T, const double* xi, const double* xj) {
    for (int64_t t = 0; t < T; t++) {
        const double* s = soup + 6 * t;
        double v0[3] = {s[0], s[1], 0.0}, v1[3] = {s[3], s[4], 0.0};
        if (seg_blocks_ray(xi, xj, v0, v1)) return 1;
    }
    return 0;
}

/* Duffy pair (test entry) */
double orc_pair_duffy(const double* coords, const int8_t* fam, const int32_t* nodes8, int64_t i, int64_t j,
                      int nquad) {
    OrcMesh m;
    memset(&m, 0, sizeof m);
    m.coords = coords; m.fam = fam; m.nodes = nodes8; m.N = (i > j ? i : j) + 1;
    mesh_rules(&m, nquad);
    m.dim2 = fam[i] >= 4;
    return pair_duffy(&m, i, j, 0);
}

/* Expectation of the Monte Carlo pair estimate of two quadrilaterals (mc_pair,
 * GPUMCKernels.jl:146-262) with n_samples = s^2: sample (si, sj) puts element
 * i's point and element j's point in the SAME parameter cell (si, sj), so
 *   E[raw] = (ri rj / n) sum_cells E_{u in cell x cell}[K a_i a_j],
 * evaluated here by a g x g Gauss-Legendre rule on [0,1]^2 inside each cell of
 * each element (g^4 points per cell pair).  No visibility (unoccluded pairs).
 * This pins the stream's stratum mapping statistically; it is not the pair's
 * view-factor integral, which sums over all cell pairs (si, sj) x (si', sj'). */
double orc_mc_expect(const double* coords, const int8_t* fam, const int32_t* nodes8, int64_t i, int64_t j,
                     int nsamp, int g) {
    OrcMesh m;
    memset(&m, 0, sizeof m);
    m.coords = coords; m.fam = fam; m.nodes = nodes8; m.N = (i > j ? i : j) + 1;
    if (!((fam[i] == 0 || fam[i] == 2) && (fam[j] == 0 || fam[j] == 2))) return NAN;
    int s = (int)floor(sqrt((double)nsamp));
    if (s * s != nsamp || g < 1 || g > 8) return NAN;
    double p[8], w[8];
    orc_gl1d(g, p, w);
    for (int k = 0; k < g; k++) { p[k] = (p[k] + 1.0) / 2.0; w[k] = w[k] / 2.0; }  /* [0,1] */
    double tot = 0.0;
    for (int si = 0; si < s; si++)
        for (int sj = 0; sj < s; sj++) {
            double cell = 0.0;
            for (int a1 = 0; a1 < g; a1++)
                for (int a2 = 0; a2 < g; a2++) {
                    double xi[3], ni[3], ai;
                    mc_sample(&m, i, (si + p[a1]) / s, (sj + p[a2]) / s, xi, ni, &ai);
                    for (int b1 = 0; b1 < g; b1++)
                        for (int b2 = 0; b2 < g; b2++) {
                            double xj[3], nj[3], aj;
                            mc_sample(&m, j, (si + p[b1]) / s, (sj + p[b2]) / s, xj, nj, &aj);
                            double K = orc_vf_kernel(xi, ni, xj, nj);
                            cell += (((w[a1] * w[a2]) * (w[b1] * w[b2])) * K) * (ai * aj);
                        }
                }
            tot += cell;
        }
    return (tot * 16.0) / nsamp;
}

/* xorshift stream (test entry) */
void orc_xorshift_stream(uint64_t seed, int64_t tid, int n, double* out) {
    uint64_t s = orc_init_rng(seed, tid);
    for (int k = 0; k < n; k++) out[k] = xorshift64(&s);
}
