// HIP kernels for the element-pair view-factor path on MI355X (gfx950).
//
// Layout and mapping
//   * Element e's quadrature points are precomputed once (k_precompute) into
//     two layouts: point-major SoA P[(p*8+c)*NS + e] read by lanes (coalesced
//     over e) and element-major Q[(e*maxpts+p)*8 + c] read with wave-uniform
//     (scalar) loads.
//   * A pair tile is one wavefront: lane = row element i of a 64-element strip,
//     the column element j (a run of <= 16 elements of one group) is uniform
//     across the wave.  Each lane owns its pair sum, so the FP64 mode sums in
//     exactly the reference's order (ViewFactorKernel.jl:215-238).
//   * Row sums and group sums come from per-tile partials written without
//     atomics (deterministic), reduced by k_reduce_rows.
//
// FP64 mode: reference arithmetic verbatim (vf_kernel, ViewFactorKernel.jl:
// 155-166; is_visible / _ray_aabb / _ray_triangle, RayCast.jl:20-32,
// BVH.jl:155-201), compiled with -ffp-contract=off.
// Mixed mode: FP32 anchored coordinates for the kernel, FP64 outer sums, FP32
// box traversal on padded bounds with an FP32 ray-triangle filter whose
// undecided cases are re-tested in FP64 with the reference arithmetic.
#include "rvf_common.h"
#include "rvf_filter.h"

#include <algorithm>
#include <type_traits>

#define PI_D 3.141592653589793

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void neumaier(double& s, double& c, double x) {
    double t = s + x;
    c += (fabs(s) >= fabs(x)) ? ((s - t) + x) : ((x - t) + s);
    s = t;
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

// ---------------------------------------------------------------------------
// Shape functions (Geometry.jl:36-71,124-134; ViewFactorKernel.jl:77-83,
// 116-121), same expression trees as the reference, no contraction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int shape_fn(int fam, double x, double e, double* N, double* dx, double* de) {
    if (fam == 0) {
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
        return 8;
    } else if (fam == 1) {
        double L1 = (1.0 - x) - e, L2 = x, L3 = e;
        N[0] = L1 * (2 * L1 - 1); N[1] = L2 * (2 * L2 - 1); N[2] = L3 * (2 * L3 - 1);
        N[3] = (4 * L1) * L2; N[4] = (4 * L2) * L3; N[5] = (4 * L1) * L3;
        dx[0] = (4 * L1 - 1) * (-1.0); dx[1] = 4 * L2 - 1; dx[2] = 0.0;
        dx[3] = 4 * (L2 * (-1.0) + L1); dx[4] = 4 * L3; dx[5] = (4 * L3) * (-1.0);
        de[0] = (4 * L1 - 1) * (-1.0); de[1] = 0.0; de[2] = 4 * L3 - 1;
        de[3] = (4 * L2) * (-1.0); de[4] = 4 * L2; de[5] = 4 * (L3 * (-1.0) + L1);
        return 6;
    } else if (fam == 2) {
        N[0] = (0.25 * (1 - x)) * (1 - e);
        N[1] = (0.25 * (1 + x)) * (1 - e);
        N[2] = (0.25 * (1 + x)) * (1 + e);
        N[3] = (0.25 * (1 - x)) * (1 + e);
        dx[0] = (-0.25) * (1 - e); dx[1] = 0.25 * (1 - e); dx[2] = 0.25 * (1 + e); dx[3] = (-0.25) * (1 + e);
        de[0] = (-0.25) * (1 - x); de[1] = (-0.25) * (1 + x); de[2] = 0.25 * (1 + x); de[3] = 0.25 * (1 - x);
        return 4;
    } else if (fam == 3) {
        N[0] = (1.0 - x) - e; N[1] = x; N[2] = e;
        dx[0] = -1.0; dx[1] = 1.0; dx[2] = 0.0;
        de[0] = -1.0; de[1] = 0.0; de[2] = 1.0;
        return 3;
    } else if (fam == 4) {  // Line3 [left, right, midpoint] (Geometry.jl:226-235)
        N[0] = ((-0.5) * x) * (1.0 - x); N[1] = (0.5 * x) * (1.0 + x); N[2] = 1.0 - x * x;
        dx[0] = -0.5 + x; dx[1] = 0.5 + x; dx[2] = (-2.0) * x;
        return 3;
    } else {  // Line2 (Geometry.jl:292-294)
        N[0] = 0.5 * (1.0 - x); N[1] = 0.5 * (1.0 + x);
        dx[0] = -0.5; dx[1] = 0.5;
        return 2;
    }
}

// x, n = c/|c|, dA = |c| at (xi, eta) (Geometry.jl:78-110 and twins)
// (coords, nodes: plain pointers, or constant-address-space ones where the
// element is the same for the whole wave and the loads can be scalar)
template <class PD, class PI>
__device__ __forceinline__ void eval_point(int fam, PD coords, PI nodes,
                                           double xi, double eta, double* x, double* n, double* dA) {
    double N[8], dx[8], de[8];
    int nn = shape_fn(fam, xi, eta, N, dx, de);
    if (fam >= 4) {
        // curve element: dL = |dx/dxi|, n = tangent turned 90 degrees counter-
        // clockwise in xy (Geometry.jl:265-282, 321-332)
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
        for (int k = 0; k < nn; k++) {
            const auto xa = coords + 3 * (int64_t)(nodes[k] - 1);
            double X = xa[0], Y = xa[1], Z = xa[2];
            p0 = p0 + N[k] * X; p1 = p1 + N[k] * Y; p2 = p2 + N[k] * Z;
            a0 = a0 + dx[k] * X; a1 = a1 + dx[k] * Y; a2 = a2 + dx[k] * Z;
        }
        double d = sqrt((a0 * a0 + a1 * a1) + a2 * a2);
        x[0] = p0; x[1] = p1; x[2] = p2;
        n[0] = -(a1 / d); n[1] = a0 / d; n[2] = 0.0;
        *dA = d;
        return;
    }
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    for (int k = 0; k < nn; k++) {
        const auto xa = coords + 3 * (int64_t)(nodes[k] - 1);
        double X = xa[0], Y = xa[1], Z = xa[2];
        p0 = p0 + N[k] * X; p1 = p1 + N[k] * Y; p2 = p2 + N[k] * Z;
        a0 = a0 + dx[k] * X; a1 = a1 + dx[k] * Y; a2 = a2 + dx[k] * Z;
        b0 = b0 + de[k] * X; b1 = b1 + de[k] * Y; b2 = b2 + de[k] * Z;
    }
    double c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
    double d = sqrt((c0 * c0 + c1 * c1) + c2 * c2);
    x[0] = p0; x[1] = p1; x[2] = p2;
    n[0] = c0 / d; n[1] = c1 / d; n[2] = c2 / d;
    *dA = d;
}

// eval_point with the element's node coordinates already in registers
// (cx/cy/cz[k] for node k < the family's node count): the same arithmetic in
// the same order, so the same bits.
__device__ __forceinline__ void eval_point_r(int fam, const double* cx, const double* cy, const double* cz, double xi,
                                             double eta, double* x, double* n, double* dA) {
    double N[8], dx[8], de[8];
    int nn = shape_fn(fam, xi, eta, N, dx, de);
    if (fam >= 4) {
        double p0 = 0.0, p1 = 0.0, p2 = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (k < nn) {
                p0 = p0 + N[k] * cx[k]; p1 = p1 + N[k] * cy[k]; p2 = p2 + N[k] * cz[k];
                a0 = a0 + dx[k] * cx[k]; a1 = a1 + dx[k] * cy[k]; a2 = a2 + dx[k] * cz[k];
            }
        }
        double d = sqrt((a0 * a0 + a1 * a1) + a2 * a2);
        x[0] = p0; x[1] = p1; x[2] = p2;
        n[0] = -(a1 / d); n[1] = a0 / d; n[2] = 0.0;
        *dA = d;
        return;
    }
    double p0 = 0.0, p1 = 0.0, p2 = 0.0, a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (k < nn) {
            p0 = p0 + N[k] * cx[k]; p1 = p1 + N[k] * cy[k]; p2 = p2 + N[k] * cz[k];
            a0 = a0 + dx[k] * cx[k]; a1 = a1 + dx[k] * cy[k]; a2 = a2 + dx[k] * cz[k];
            b0 = b0 + de[k] * cx[k]; b1 = b1 + de[k] * cy[k]; b2 = b2 + de[k] * cz[k];
        }
    }
    double c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
    double d = sqrt((c0 * c0 + c1 * c1) + c2 * c2);
    x[0] = p0; x[1] = p1; x[2] = p2;
    n[0] = c0 / d; n[1] = c1 / d; n[2] = c2 / d;
    *dA = d;
}

// ---------------------------------------------------------------------------
// Precompute: one thread per element, points in rule order.  Writes FP64 SoA
// and AoS, the area A_e = sum_p w_p dA_p in p order (Assembly.jl:186-190 via
// ViewFactorKernel.jl:221), and the FP32 anchored data for mixed mode.
// ---------------------------------------------------------------------------


__global__ __launch_bounds__(64) void k_precompute(PrecompArgs a) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // slot
    if (e >= a.N) return;
    const int64_t o = a.slot_orig[e];                             // original element
    int f = a.fam[o];
    int isq = (f == 0 || f == 2);
    const bool line = f >= 4;  // 1-D Gauss-Legendre (ViewFactorKernel.jl:284-301), points (xi, 0)
    int np = line ? a.nq_line : isq ? a.nq_quad : a.nq_tri;
    const double* pts = line ? a.rule + RVF_RULE_LINE : isq ? a.rule : a.rule + 3 * 64;
    const double* wts = line ? a.rule + RVF_RULE_LINE + 16 : isq ? a.rule + 2 * 64 : a.rule + 3 * 64 + 26;
    const int32_t* nd = a.nodes8 + 8 * o;
    double A = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int p = 0; p < np; p++) {
        double x[3], n[3], dA;
        eval_point(f, a.coords, nd, pts[2 * p], pts[2 * p + 1], x, n, &dA);
        double w = wts[p];
        A += w * dA;
        if (!(dA > 0.0) || !isfinite(dA)) atomicExch(a.bad, 1);
        double v[8] = {x[0], x[1], x[2], n[0], n[1], n[2], w, dA};
        for (int c = 0; c < 8; c++) {
            a.P64[(int64_t)(p * 8 + c) * a.NS + e] = v[c];
            a.Q64[((int64_t)e * a.maxpts + p) * 8 + c] = v[c];
        }
        cx += x[0]; cy += x[1]; cz += x[2];
    }
    a.area[e] = A;
    a.area_orig[o] = A;
    a.elem_np[e] = np;
    double ax = cx / np, ay = cy / np, az = cz / np;
    // anchor[3]: radius of the element's points about the anchor (occlusion
    // culling of whole pairs, tile_occ_queue), rounded up
    double r2 = 0.0;
    for (int p = 0; p < np; p++) {
        const double dx = a.P64[(int64_t)(p * 8 + 0) * a.NS + e] - ax;
        const double dy = a.P64[(int64_t)(p * 8 + 1) * a.NS + e] - ay;
        const double dz = a.P64[(int64_t)(p * 8 + 2) * a.NS + e] - az;
        r2 = fmax(r2, dx * dx + dy * dy + dz * dz);
    }
    a.anchor[e * 4 + 0] = ax; a.anchor[e * 4 + 1] = ay; a.anchor[e * 4 + 2] = az;
    a.anchor[e * 4 + 3] = sqrt(r2) * (1.0 + 1e-9);
    for (int p = 0; p < np; p++) {
        const double* P = a.P64;
        int64_t NS = a.NS;
        float v[8];
        v[0] = (float)(P[(int64_t)(p * 8 + 0) * NS + e] - ax);
        v[1] = (float)(P[(int64_t)(p * 8 + 1) * NS + e] - ay);
        v[2] = (float)(P[(int64_t)(p * 8 + 2) * NS + e] - az);
        v[3] = (float)P[(int64_t)(p * 8 + 3) * NS + e];
        v[4] = (float)P[(int64_t)(p * 8 + 4) * NS + e];
        v[5] = (float)P[(int64_t)(p * 8 + 5) * NS + e];
        v[6] = (float)(P[(int64_t)(p * 8 + 6) * NS + e] * P[(int64_t)(p * 8 + 7) * NS + e]);
        v[7] = 0.0f;
        for (int c = 0; c < 8; c++) {
            a.P32[(int64_t)(p * 8 + c) * NS + e] = v[c];
            a.Q32[((int64_t)e * a.maxpts + p) * 8 + c] = v[c];
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 reference kernel and visibility (exact reference arithmetic)
// ---------------------------------------------------------------------------
struct Ray64 { double ox, oy, oz, dx, dy, dz, L; };

// vf_kernel (ViewFactorKernel.jl:155-166).  Also returns the normalised
// direction and |r|, which equal is_visible's d/L and L (RayCast.jl:27-30).
__device__ __forceinline__ double vf_exact(double xi, double yi, double zi, double nxi, double nyi, double nzi,
                                           double xj, double yj, double zj, double nxj, double nyj, double nzj,
                                           Ray64& ray) {
    double r0 = xj - xi, r1 = yj - yi, r2 = zj - zi;
    double rr = (r0 * r0 + r1 * r1) + r2 * r2;
    double r = sqrt(rr);
    double h0 = r0 / r, h1 = r1 / r, h2 = r2 / r;
    double ci = (nxi * h0 + nyi * h1) + nzi * h2;
    double cj = (nxj * (-h0) + nyj * (-h1)) + nzj * (-h2);
    ray.ox = xi; ray.oy = yi; ray.oz = zi; ray.dx = h0; ray.dy = h1; ray.dz = h2; ray.L = r;
    double K = (ci * cj) / (PI_D * rr);
    return (rr < 1e-30 || ci <= 0.0 || cj <= 0.0) ? 0.0 : K;
}

// vf_kernel_2d (ViewFactorKernel.jl:169-180), per unit depth: cos cos / (2 r)
__device__ __forceinline__ double vf_exact2d(double xi, double yi, double zi, double nxi, double nyi, double nzi,
                                             double xj, double yj, double zj, double nxj, double nyj, double nzj) {
    double r0 = xj - xi, r1 = yj - yi, r2 = zj - zi;
    double rr = (r0 * r0 + r1 * r1) + r2 * r2;
    double r = sqrt(rr);
    double h0 = r0 / r, h1 = r1 / r, h2 = r2 / r;
    double ci = (nxi * h0 + nyi * h1) + nzi * h2;
    double cj = (nxj * (-h0) + nyj * (-h1)) + nzj * (-h2);
    double K = (ci * cj) / (2.0 * r);
    return (rr < 1e-30 || ci <= 0.0 || cj <= 0.0) ? 0.0 : K;
}

// _seg_blocks_ray (BVH.jl:264-277): the 2-D path xi -> xj against the segment
// v0 -> v0 + e (e = v1 - v0, the same double the reference forms), tol 1e-8.
__device__ __forceinline__ bool seg_blocks64(double xi, double yi, double xj, double yj, const Tri64& S) {
    const double tol = 1e-8;
    double dx = xj - xi, dy = yj - yi;
    double ex = S.e1[0], ey = S.e1[1];
    double denom = dx * ey - dy * ex;
    if (fabs(denom) < tol) return false;
    double fx = S.v0[0] - xi, fy = S.v0[1] - yi;
    double sp = (fx * ey - fy * ex) / denom;
    double tp = (fx * dy - fy * dx) / denom;
    return (tol < sp && sp < 1.0 - tol) && (tol < tp && tp < 1.0 - tol);
}

// intersect_seg_bvh (BVH.jl:286-320) over the reference's tree in stackless
// preorder: a node is entered iff its box overlaps the path's xy bounding box,
// as in the reference, so every segment the reference tests is tested and the
// any-hit decision is the reference's whatever the visiting order.
__device__ __forceinline__ bool blocked2d(const BvhNode64* __restrict__ nodes, const Tri64* __restrict__ segs, int root,
                                          double xi, double yi, double xj, double yj, unsigned long long& naabb,
                                          unsigned long long& ntri) {
    // finite points: the order of equal values and the sign of zero do not change the comparisons below
    const double xlo = xi < xj ? xi : xj, xhi = xi < xj ? xj : xi;
    const double ylo = yi < yj ? yi : yj, yhi = yi < yj ? yj : yi;
    int n = root;
    while (n >= 0) {
        const BvhNode64& nd = nodes[n];
        naabb++;
        if (nd.lo[0] > xhi || nd.hi[0] < xlo || nd.lo[1] > yhi || nd.hi[1] < ylo) { n = nd.miss; continue; }
        if (nd.count > 0) {
            for (int k = nd.start; k < nd.start + nd.count; k++) {
                ntri++;
                if (seg_blocks64(xi, yi, xj, yj, segs[k])) return true;
            }
            n = nd.miss;
        } else {
            n = n + 1;
        }
    }
    return false;
}

// _ray_triangle with t_min = 0 and the caller's t < t_max (BVH.jl:181-201,241-242)
__device__ __forceinline__ bool mt_hit64(const Tri64& T, const Ray64& R) {
    double e1x = T.e1[0], e1y = T.e1[1], e1z = T.e1[2];
    double e2x = T.e2[0], e2y = T.e2[1], e2z = T.e2[2];
    double hx = R.dy * e2z - R.dz * e2y;
    double hy = R.dz * e2x - R.dx * e2z;
    double hz = R.dx * e2y - R.dy * e2x;
    double av = (e1x * hx + e1y * hy) + e1z * hz;
    if (fabs(av) < 1e-10) return false;
    double f = 1.0 / av;
    double sx = R.ox - T.v0[0], sy = R.oy - T.v0[1], sz = R.oz - T.v0[2];
    double u = f * ((sx * hx + sy * hy) + sz * hz);
    if (u < 0.0 || u > 1.0) return false;
    double qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
    double v = f * ((R.dx * qx + R.dy * qy) + R.dz * qz);
    if (v < 0.0 || u + v > 1.0) return false;
    double t = f * ((e2x * qx + e2y * qy) + e2z * qz);
    return t > 0.0 && t < R.L;
}

// intersect_ray_bvh (BVH.jl:213-250) in stackless order; the any-hit result
// does not depend on visiting order.  _ray_aabb (BVH.jl:155-169) keeps Julia's
// NaN-propagating min/max semantics: any NaN slab value makes the box a miss.
__device__ bool blocked64(const BvhNode64* __restrict__ nodes, const Tri64* __restrict__ tris, int root,
                          const Ray64& R, unsigned long long& naabb, unsigned long long& ntri) {
    double ix = 1.0 / R.dx, iy = 1.0 / R.dy, iz = 1.0 / R.dz;
    int n = root;
    while (n >= 0) {
        const BvhNode64 nd = nodes[n];
        naabb++;
        double t1x = (nd.lo[0] - R.ox) * ix, t2x = (nd.hi[0] - R.ox) * ix;
        double t1y = (nd.lo[1] - R.oy) * iy, t2y = (nd.hi[1] - R.oy) * iy;
        double t1z = (nd.lo[2] - R.oz) * iz, t2z = (nd.hi[2] - R.oz) * iz;
        bool nan = (t1x != t1x) | (t2x != t2x) | (t1y != t1y) | (t2y != t2y) | (t1z != t1z) | (t2z != t2z);
        double tmin = fmax(fmax(fmax(fmin(t1x, t2x), fmin(t1y, t2y)), fmin(t1z, t2z)), 0.0);
        double tmax = fmin(fmin(fmin(fmax(t1x, t2x), fmax(t1y, t2y)), fmax(t1z, t2z)), R.L);
        if (!nan && tmin <= tmax) {
            if (nd.count > 0) {
                for (int k = nd.start; k < nd.start + nd.count; k++) {
                    ntri++;
                    if (mt_hit64(tris[k], R)) return true;
                }
                n = nd.miss;
            } else {
                n = n + 1;
            }
        } else {
            n = nd.miss;
        }
    }
    return false;
}

// One exact slab test (BVH.jl:155-169 semantics; a NaN slab value is a miss).
__device__ __forceinline__ bool slab64(const BvhNode64& nd, const Ray64& R, double ix, double iy, double iz) {
    double t1x = (nd.lo[0] - R.ox) * ix, t2x = (nd.hi[0] - R.ox) * ix;
    double t1y = (nd.lo[1] - R.oy) * iy, t2y = (nd.hi[1] - R.oy) * iy;
    double t1z = (nd.lo[2] - R.oz) * iz, t2z = (nd.hi[2] - R.oz) * iz;
    bool nan = (t1x != t1x) | (t2x != t2x) | (t1y != t1y) | (t2y != t2y) | (t1z != t1z) | (t2z != t2z);
    double tmin = fmax(fmax(fmax(fmin(t1x, t2x), fmin(t1y, t2y)), fmin(t1z, t2z)), 0.0);
    double tmax = fmin(fmin(fmin(fmax(t1x, t2x), fmax(t1y, t2y)), fmax(t1z, t2z)), R.L);
    return !nan && tmin <= tmax;
}

// Per-lane any-hit for every lane with need set; returns the lane mask of
// blocked rays.  Each lane walks its own ray (independent dependent-load
// chains keep many L2 requests in flight).
__device__ __forceinline__ uint64_t occlusion64(const PairArgs& a, bool need, int root, const Ray64& R,
                                                unsigned long long& naabb, unsigned long long& ntri) {
    bool b = need && blocked64(a.nodes64, a.tris64, root, R, naabb, ntri);
    return __ballot(b);
}

// ---------------------------------------------------------------------------
// Pair tile prologue shared by all pair kernels
// ---------------------------------------------------------------------------
struct TileCtx {
    int tile, s, chunk, lane, jb, je;
    int64_t i;   // slot of this lane's element
    int64_t oi;  // its original index
    int64_t jlo; // columns below jlo have no pair with any lane (j > i; j >= i with self_vf)
    bool vi;
};

__device__ __forceinline__ TileCtx tile_ctx(const PairArgs& a, int tile) {
    TileCtx t;
    // a tile belongs to one wave: its index and fields are read back as
    // wave-uniform values, so what derives from them stays in scalar
    // registers (fewer spills in the occluded kernels: 162.3 -> 161.9 ms)
    tile = __builtin_amdgcn_readfirstlane(tile);
    int2 tl = a.tiles[tile];
    t.tile = tile;
    t.s = __builtin_amdgcn_readfirstlane(tl.x); t.chunk = __builtin_amdgcn_readfirstlane(tl.y);
    t.lane = threadIdx.x;
    t.i = (int64_t)t.s * RVF_STRIP + t.lane;
    t.vi = t.i < a.N;
    t.oi = t.vi ? a.slot_orig[t.i] : 0;
    t.jlo = (int64_t)t.s * RVF_STRIP + (a.self_vf ? 0 : 1);
    Chunk c = a.chunks[t.chunk];
    t.jb = __builtin_amdgcn_readfirstlane(c.begin); t.je = __builtin_amdgcn_readfirstlane(c.end);
    return t;
}

// Quad8 pairs sharing a corner node are integrated by k_duffy instead
// (singularity_type, DuffyKernel.jl:95-115).
struct Corners { int32_t c[4]; int fam; };

__device__ __forceinline__ Corners corners_of(const PairArgs& a, int64_t e, bool valid) {
    Corners k;
    k.fam = valid ? a.fam[e] : -1;
    for (int q = 0; q < 4; q++) k.c[q] = (valid && k.fam == 0) ? a.nodes8[8 * e + q] : -1 - q;
    return k;
}

__device__ __forceinline__ bool shares_corner(const Corners& x, const Corners& y) {
    if (x.fam != 0 || y.fam != 0) return false;
    bool s = false;
#pragma unroll
    for (int p = 0; p < 4; p++)
#pragma unroll
        for (int q = 0; q < 4; q++) s |= (x.c[p] == y.c[q]);
    return s;
}

// Per-j epilogue: raw output, row partial (F_ij = raw/A_i) and the
// transposed partial for row j (sum over lanes of raw/A_j).
__device__ __forceinline__ void pair_out(const PairArgs& a, const TileCtx& t, int64_t j, int64_t oj, bool act,
                                         double raw, double Ai, double& ls, double& lc) {
    const bool self = t.i == j;  // only active with self_vf: kept out of both partials
    if (act) {
        if (a.F) a.F[min(t.oi, oj) + a.N * max(t.oi, oj)] = raw;
        if (self) a.diag[j] = raw;
        else neumaier(ls, lc, raw / Ai);
    }
    double Aj = a.area[j];
    double v = act && !self ? raw / Aj : 0.0;
    v = wave_sum_d(v);
    if (t.lane == 0) a.partT[(int64_t)t.s * a.NS + j] = v;
}

// ---------------------------------------------------------------------------
// FP64, unoccluded: i-points in registers, PC per chunk of the p loop.  D2:
// curve meshes (2-D kernel per unit depth).
// ---------------------------------------------------------------------------
template <int PC, bool D2 = false>
__device__ __forceinline__ void tile_fp64(const PairArgs& a, int tile) {
    const TileCtx t = tile_ctx(a, tile);
    const Corners cwi = corners_of(a, t.oi, t.vi && a.duffy);
    const int64_t NS = a.NS;
    const int npi = t.vi ? a.elem_np[t.i] : 0;
    int npmax = npi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) npmax = max(npmax, __shfl_xor(npmax, o, 64));
    npmax = __builtin_amdgcn_readfirstlane(npmax);
    const double Ai = t.vi ? a.area[t.i] : 1.0;
    const double* __restrict__ P = a.P64;
    double ls = 0.0, lc = 0.0;
    for (int64_t j = t.jb; j < t.je; j++) {
        if (j < t.jlo) {  // no lane pairs with column j
            if (t.lane == 0) a.partT[(int64_t)t.s * a.NS + j] = 0.0;
            continue;
        }
        bool act = t.vi && t.i < j + a.self_vf;
        const int64_t oj = a.slot_orig[j];
        if (a.duffy && shares_corner(cwi, corners_of(a, oj, true))) act = false;
        const int npj = a.elem_np[j];
        const double* __restrict__ qj = a.Q64 + j * a.maxpts * 8;
        double Fij = 0.0;
        for (int c0 = 0; c0 < npmax; c0 += PC) {
            double X[PC], Y[PC], Z[PC], NX[PC], NY[PC], NZ[PC], inner[PC];
#pragma unroll
            for (int p = 0; p < PC; p++) {
                int pp = c0 + p;
                inner[p] = 0.0;
                if (pp < npmax) {
                    X[p] = P[(int64_t)(pp * 8 + 0) * NS + t.i];
                    Y[p] = P[(int64_t)(pp * 8 + 1) * NS + t.i];
                    Z[p] = P[(int64_t)(pp * 8 + 2) * NS + t.i];
                    NX[p] = P[(int64_t)(pp * 8 + 3) * NS + t.i];
                    NY[p] = P[(int64_t)(pp * 8 + 4) * NS + t.i];
                    NZ[p] = P[(int64_t)(pp * 8 + 5) * NS + t.i];
                } else {
                    X[p] = Y[p] = Z[p] = NX[p] = NY[p] = NZ[p] = 0.0;
                }
            }
            for (int q = 0; q < npj; q++) {
                const double xj = qj[q * 8 + 0], yj = qj[q * 8 + 1], zj = qj[q * 8 + 2];
                const double nxj = qj[q * 8 + 3], nyj = qj[q * 8 + 4], nzj = qj[q * 8 + 5];
                const double wj = qj[q * 8 + 6], dAj = qj[q * 8 + 7];
#pragma unroll
                for (int p = 0; p < PC; p++) {
                    if (c0 + p < npmax) {
                        Ray64 R;
                        double K = D2 ? vf_exact2d(X[p], Y[p], Z[p], NX[p], NY[p], NZ[p], xj, yj, zj, nxj, nyj, nzj)
                                      : vf_exact(X[p], Y[p], Z[p], NX[p], NY[p], NZ[p], xj, yj, zj, nxj, nyj, nzj, R);
                        inner[p] += (wj * K) * dAj;
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < PC; p++) {
                int pp = c0 + p;
                if (pp < npmax && pp < npi) {
                    double wi = P[(int64_t)(pp * 8 + 6) * NS + t.i], dAi = P[(int64_t)(pp * 8 + 7) * NS + t.i];
                    Fij += (wi * inner[p]) * dAi;
                }
            }
        }
        pair_out(a, t, j, oj, act, Fij, Ai, ls, lc);
    }
    a.part[(int64_t)t.tile * RVF_STRIP + t.lane] = ls + lc;
}

// ---------------------------------------------------------------------------
// FP64, occluded.  The pair's outer element is the one with the lower
// original index (the reference loops j > i, Assembly.jl:191-205): its points
// are the ray origins and the outer sum runs over them.  When the lane's
// element is the inner one ("rev", possible under the spatial slot layout),
// the ray starts at the column element's point and the per-outer-point sums
// are accumulated in LDS (accq[q][lane]) so the summation order is still the
// reference's.  K is symmetric bit for bit, so one evaluation serves both.
// ---------------------------------------------------------------------------
template <bool D2 = false>
__device__ __forceinline__ void tile_fp64_occ(const PairArgs& a, int tile, double* accq /* [maxpts][64] LDS */) {
    const TileCtx t = tile_ctx(a, tile);
    const Corners cwi = corners_of(a, t.oi, t.vi && a.duffy);
    const int64_t NS = a.NS;
    const int npi = t.vi ? a.elem_np[t.i] : 0;
    int npmax = npi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) npmax = max(npmax, __shfl_xor(npmax, o, 64));
    npmax = __builtin_amdgcn_readfirstlane(npmax);
    const double Ai = t.vi ? a.area[t.i] : 1.0;
    const int gi = t.vi ? a.elem_grp[t.i] : 0;
    const double* __restrict__ P = a.P64;
    double ls = 0.0, lc = 0.0;
    unsigned long long nrays = 0, naabb = 0, ntri = 0;
    for (int64_t j = t.jb; j < t.je; j++) {
        if (j < t.jlo) {
            if (t.lane == 0) a.partT[(int64_t)t.s * a.NS + j] = 0.0;
            continue;
        }
        bool act = t.vi && t.i < j + a.self_vf;
        const int64_t oj = a.slot_orig[j];
        if (a.duffy && shares_corner(cwi, corners_of(a, oj, true))) act = false;
        const bool rev = t.oi > oj;
        const int npj = a.elem_np[j];
        const int gj = a.elem_grp[j];
        const int root = a.tree_root[gi * a.G + gj];
        const double* __restrict__ qj = a.Q64 + j * a.maxpts * 8;
        for (int q = 0; q < npj; q++) accq[q * 64 + t.lane] = 0.0;
        double Fij = 0.0;
        uint32_t bc = 0, bh = 0;
        for (int p = 0; p < npmax; p++) {
            const bool pv = act && p < npi;
            const double xi = P[(int64_t)(p * 8 + 0) * NS + t.i], yi = P[(int64_t)(p * 8 + 1) * NS + t.i];
            const double zi = P[(int64_t)(p * 8 + 2) * NS + t.i];
            const double nxi = P[(int64_t)(p * 8 + 3) * NS + t.i], nyi = P[(int64_t)(p * 8 + 4) * NS + t.i];
            const double nzi = P[(int64_t)(p * 8 + 5) * NS + t.i];
            const double wi = P[(int64_t)(p * 8 + 6) * NS + t.i], dAi = P[(int64_t)(p * 8 + 7) * NS + t.i];
            double inner = 0.0;
            for (int q = 0; q < npj; q++) {
                const double xj = qj[q * 8 + 0], yj = qj[q * 8 + 1], zj = qj[q * 8 + 2];
                const double nxj = qj[q * 8 + 3], nyj = qj[q * 8 + 4], nzj = qj[q * 8 + 5];
                const double wj = qj[q * 8 + 6], dAj = qj[q * 8 + 7];
                Ray64 R;
                double K = D2 ? vf_exact2d(xi, yi, zi, nxi, nyi, nzi, xj, yj, zj, nxj, nyj, nzj)
                              : vf_exact(xi, yi, zi, nxi, nyi, nzi, xj, yj, zj, nxj, nyj, nzj, R);
                if (!D2 && rev) {  // is_visible(x_outer = x_j, x_inner = x_i): d = x_i - x_j = -(x_j - x_i) exactly
                    R.ox = xj; R.oy = yj; R.oz = zj;
                    R.dx = -R.dx; R.dy = -R.dy; R.dz = -R.dz;
                }
                const bool need = pv && K != 0.0 && root >= 0;
                nrays += need;
                bool blk;
                if (D2)  // is_visible(bvh, x_outer, x_inner; mesh_dim = 1)
                    blk = need && (rev ? blocked2d(a.nodes64, a.tris64, root, xj, yj, xi, yi, naabb, ntri)
                                       : blocked2d(a.nodes64, a.tris64, root, xi, yi, xj, yj, naabb, ntri));
                else
                    blk = (occlusion64(a, need, root, R, naabb, ntri) >> t.lane) & 1;
                if (blk) {
                    K = 0.0;
                    bc++;
                    bh += mix32((uint32_t)(rev ? q * npi + p : p * npj + q) + 0x9e3779b9U);
                }
                inner += (wj * K) * dAj;
                if (rev && p < npi) accq[q * 64 + t.lane] += (wi * K) * dAi;
            }
            if (p < npi) Fij += (wi * inner) * dAi;
        }
        if (rev) {
            Fij = 0.0;
            for (int q = 0; q < npj; q++) Fij += (qj[q * 8 + 6] * accq[q * 64 + t.lane]) * qj[q * 8 + 7];
        }
        if (act && a.audit_count) {
            a.audit_count[min(t.oi, oj) + a.N * max(t.oi, oj)] = bc;
            a.audit_hash[min(t.oi, oj) + a.N * max(t.oi, oj)] = bh;
        }
        pair_out(a, t, j, oj, act, Fij, Ai, ls, lc);
    }
    a.part[(int64_t)t.tile * RVF_STRIP + t.lane] = ls + lc;
    if (a.counters) {
        unsigned long long c1 = nrays, c2 = naabb, c3 = ntri;
        for (int o = 32; o > 0; o >>= 1) {
            c1 += __shfl_xor(c1, o, 64); c2 += __shfl_xor(c2, o, 64); c3 += __shfl_xor(c3, o, 64);
        }
        if (t.lane == 0) { atomicAdd(&a.counters[1], c1); atomicAdd(&a.counters[2], c2); atomicAdd(&a.counters[3], c3); }
    }
}

// ---------------------------------------------------------------------------
// Mixed precision kernel evaluation.
//   r = (a_j - a_i) + (d_j - d_i) with FP64 anchors a and FP32 offsets d;
//   K' = max(n_i.r, 0) max(-n_j.r, 0) / r^4  (1/pi applied once per pair)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float vf_mixed(float bx, float by, float bz, float nxi, float nyi, float nzi,
                                          float qx, float qy, float qz, float nxj, float nyj, float nzj,
                                          float& rx, float& ry, float& rz, float& r2) {
#pragma clang fp contract(fast)
    rx = bx + qx; ry = by + qy; rz = bz + qz;
    r2 = rx * rx + ry * ry + rz * rz;
    float di = nxi * rx + nyi * ry + nzi * rz;
    float dj = nxj * rx + nyj * ry + nzj * rz;
    float rr = __builtin_amdgcn_rcpf(r2);
    float k = (fmaxf(di, 0.0f) * rr) * (fmaxf(-dj, 0.0f) * rr);
    return (r2 >= 1e-30f) ? k : 0.0f;
}

// Curve meshes: K'' = max(n_i.r, 0) max(-n_j.r, 0) / |r|^3 (1/2 applied once
// per pair: cos cos / (2 r), ViewFactorKernel.jl:169-180).
__device__ __forceinline__ float vf_mixed2d(float bx, float by, float bz, float nxi, float nyi, float nzi,
                                            float qx, float qy, float qz, float nxj, float nyj, float nzj) {
#pragma clang fp contract(fast)
    const float rx = bx + qx, ry = by + qy, rz = bz + qz;
    const float r2 = rx * rx + ry * ry + rz * rz;
    const float di = nxi * rx + nyi * ry + nzi * rz;
    const float dj = nxj * rx + nyj * ry + nzj * rz;
    const float rs = __builtin_amdgcn_rsqf(r2);
    const float k = ((fmaxf(di, 0.0f) * rs) * (fmaxf(-dj, 0.0f) * rs)) * rs;
    return (r2 >= 1e-30f) ? k : 0.0f;
}

template <int PC, bool SINGLE, bool D2 = false>
__device__ __forceinline__ void tile_mixed(const PairArgs& a, int tile) {
    // SINGLE: every element has <= PC points, so the strip's points stay in
    // registers for the whole tile; otherwise chunks of PC points are reloaded
    // per column element (L1/L2 hits).
    const TileCtx t = tile_ctx(a, tile);
    const Corners cwi = corners_of(a, t.oi, t.vi && a.duffy);
    const int64_t NS = a.NS;
    const int npi = t.vi ? a.elem_np[t.i] : 0;
    int npmax = npi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) npmax = max(npmax, __shfl_xor(npmax, o, 64));
    npmax = __builtin_amdgcn_readfirstlane(npmax);
    const double Ai = t.vi ? a.area[t.i] : 1.0;
    const float* __restrict__ P = a.P32;
    const double* __restrict__ P64 = a.P64;
    const double aix = a.anchor[t.i * 4 + 0], aiy = a.anchor[t.i * 4 + 1], aiz = a.anchor[t.i * 4 + 2];
    float DX[PC], DY[PC], DZ[PC], NX[PC], NY[PC], NZ[PC];
    double WD[PC];
    auto load_chunk = [&](int c0) {
#pragma unroll
        for (int p = 0; p < PC; p++) {
            int pp = c0 + p;
            if (pp < npmax) {
                DX[p] = P[(int64_t)(pp * 8 + 0) * NS + t.i];
                DY[p] = P[(int64_t)(pp * 8 + 1) * NS + t.i];
                DZ[p] = P[(int64_t)(pp * 8 + 2) * NS + t.i];
                NX[p] = P[(int64_t)(pp * 8 + 3) * NS + t.i];
                NY[p] = P[(int64_t)(pp * 8 + 4) * NS + t.i];
                NZ[p] = P[(int64_t)(pp * 8 + 5) * NS + t.i];
                WD[p] = pp < npi ? P64[(int64_t)(pp * 8 + 6) * NS + t.i] * P64[(int64_t)(pp * 8 + 7) * NS + t.i] : 0.0;
            } else {
                DX[p] = DY[p] = DZ[p] = NX[p] = NY[p] = NZ[p] = 0.0f;
                WD[p] = 0.0;
            }
        }
    };
    if (SINGLE) load_chunk(0);
    double ls = 0.0, lc = 0.0;
    for (int64_t j = t.jb; j < t.je; j++) {
        if (j < t.jlo) {
            if (t.lane == 0) a.partT[(int64_t)t.s * a.NS + j] = 0.0;
            continue;
        }
        bool act = t.vi && t.i < j + a.self_vf;
        const int64_t oj = a.slot_orig[j];
        if (a.duffy && shares_corner(cwi, corners_of(a, oj, true))) act = false;
        const int npj = a.elem_np[j];
        const float* __restrict__ qj = a.Q32 + j * a.maxpts * 8;
        const float dax = (float)(a.anchor[j * 4 + 0] - aix);
        const float day = (float)(a.anchor[j * 4 + 1] - aiy);
        const float daz = (float)(a.anchor[j * 4 + 2] - aiz);
        double Fs = 0.0, Fc = 0.0;
        for (int c0 = 0; c0 < (SINGLE ? 1 : npmax); c0 += PC) {
            if (!SINGLE) load_chunk(c0);
            float BX[PC], BY[PC], BZ[PC], inner[PC];
#pragma unroll
            for (int p = 0; p < PC; p++) {
                BX[p] = dax - DX[p]; BY[p] = day - DY[p]; BZ[p] = daz - DZ[p];
                inner[p] = 0.0f;
            }
            for (int q = 0; q < npj; q++) {
                const float qx = qj[q * 8 + 0], qy = qj[q * 8 + 1], qz = qj[q * 8 + 2];
                const float nxj = qj[q * 8 + 3], nyj = qj[q * 8 + 4], nzj = qj[q * 8 + 5];
                const float wdj = qj[q * 8 + 6];
#pragma unroll
                for (int p = 0; p < PC; p++) {
                    if (c0 + p < npmax) {
                        float rx, ry, rz, r2;
                        float k = D2 ? vf_mixed2d(BX[p], BY[p], BZ[p], NX[p], NY[p], NZ[p], qx, qy, qz, nxj, nyj, nzj)
                                     : vf_mixed(BX[p], BY[p], BZ[p], NX[p], NY[p], NZ[p], qx, qy, qz, nxj, nyj, nzj,
                                                rx, ry, rz, r2);
                        inner[p] = fmaf(wdj, k, inner[p]);
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < PC; p++)
                if (c0 + p < npmax) neumaier(Fs, Fc, WD[p] * (double)inner[p]);
        }
        pair_out(a, t, j, oj, act, (Fs + Fc) * (D2 ? 0.5 : 1.0 / PI_D), Ai, ls, lc);
    }
    a.part[(int64_t)t.tile * RVF_STRIP + t.lane] = ls + lc;
}

// ---------------------------------------------------------------------------
// Mixed precision occlusion: FP32 traversal of conservatively padded boxes,
// FP32 Moller-Trumbore filter with rigorous error bounds; triangles the
// filter cannot reject are decided by mt_hit64 on the exact FP64 ray.
// ---------------------------------------------------------------------------
struct Ray32 { float ox, oy, oz, dx, dy, dz, L; };

// Returns 0 = definite miss, 1 = undecided (FP64 decides).  Error bounds
// (u = 2^-24, safety factor 2): |d origin| <= 3u Sm, |d v0| <= u Sm,
// |d e| <= u |e|, |d dir| <= dd (per ray, from the anchored difference), Sm =
// scene coordinate magnitude, E1/E2 = |e1|_1/|e2|_1, sa = E1 E2.
__device__ __forceinline__ int mt_filter32(const Tri32& T, const Ray32& R, float Sm, float dd) {
#pragma clang fp contract(fast)
    const float u = 5.9604645e-08f;
    float e1x = T.e1[0], e1y = T.e1[1], e1z = T.e1[2];
    float e2x = T.e2[0], e2y = T.e2[1], e2z = T.e2[2];
    float hx = R.dy * e2z - R.dz * e2y;
    float hy = R.dz * e2x - R.dx * e2z;
    float hz = R.dx * e2y - R.dy * e2x;
    float av = e1x * hx + e1y * hy + e1z * hz;
    float E1 = fabsf(e1x) + fabsf(e1y) + fabsf(e1z);
    float E2 = fabsf(e2x) + fabsf(e2y) + fabsf(e2z);
    float KA = 2.0f * T.sa * (dd + 9.0f * u);
    if (fabsf(av) <= KA + 1e-10f) return 1;  // grazing or near the 1e-10 cut
    float f = __builtin_amdgcn_rcpf(av);
    float af = fabsf(f) * 1.0000005f;
    float sx = R.ox - T.v0[0], sy = R.oy - T.v0[1], sz = R.oz - T.v0[2];
    float ub = f * (sx * hx + sy * hy + sz * hz);
    float eu = (2.0f * E2 * Sm * (64.0f * u + 6.0f * dd) + 2.0f * KA) * af + 12.0f * u;
    if (ub < -eu || ub > 1.0f + eu) return 0;
    float qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
    float vb = f * (R.dx * qx + R.dy * qy + R.dz * qz);
    float ev = (2.0f * E1 * Sm * (12.0f * dd + 68.0f * u) + 2.0f * KA) * af + 12.0f * u;
    if (vb < -ev || ub + vb > 1.0f + eu + ev) return 0;
    float tb = f * (e2x * qx + e2y * qy + e2z * qz);
    float et = (132.0f * u * T.sa * Sm + fabsf(tb) * KA) * af + 6.0f * u * fabsf(tb);
    if (tb < -et || tb > R.L * (1.0f + 2.0f * (dd + 3.0f * u)) + et) return 0;
    return 1;
}

struct Exact64Ctx {
    // lazily built exact FP64 ray between (slot i, point p) and (slot j, point q)
    const double* Pi;  // P64 base
    int64_t NS, i;
    int p;
    const double* qjq; // Q64 row of (j, q)
    bool rev;          // the column element is the outer one: the ray starts at (j, q)
};

// is_visible's ray (RayCast.jl:27-30) from the outer point to the inner one
__device__ __forceinline__ void exact_ray(const Exact64Ctx& c, Ray64& R) {
    double xi = c.Pi[(int64_t)(c.p * 8 + 0) * c.NS + c.i];
    double yi = c.Pi[(int64_t)(c.p * 8 + 1) * c.NS + c.i];
    double zi = c.Pi[(int64_t)(c.p * 8 + 2) * c.NS + c.i];
    double ox = c.rev ? c.qjq[0] : xi, oy = c.rev ? c.qjq[1] : yi, oz = c.rev ? c.qjq[2] : zi;
    double tx = c.rev ? xi : c.qjq[0], ty = c.rev ? yi : c.qjq[1], tz = c.rev ? zi : c.qjq[2];
    double d0 = tx - ox, d1 = ty - oy, d2 = tz - oz;
    double L = sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    R.ox = ox; R.oy = oy; R.oz = oz;
    R.dx = d0 / L; R.dy = d1 / L; R.dz = d2 / L; R.L = L;
}

// The exact FP64 traversal reaches leaf n iff the FP64 slab test passes at n
// and at every ancestor; checked by walking the parent links.
__device__ __forceinline__ bool reaches64(const BvhNode64* __restrict__ nodes, const int32_t* __restrict__ parent,
                                          int n, const Ray64& R, double ix, double iy, double iz) {
    for (; n >= 0; n = parent[n])
        if (!slab64(nodes[n], R, ix, iy, iz)) return false;
    return true;
}

// Mixed-precision traversal (per lane): FP32 slab tests on padded bounds (a
// superset of the FP64 traversal's visits), the FP32 triangle filter, and for
// every triangle the filter keeps the exact FP64 ray and mt_hit64.  With the
// reference soup reading (triangles tested need not lie in their boxes) a hit
// is also confirmed by the leaf's exact FP64 slab chain (reaches64); with the
// geometric reading a hit triangle lies inside every box on its path, so the
// chain is only re-tested when `verify` is set.
__device__ bool blocked_mixed(const PairArgs& a, int root, const Ray32& R, float Sm, float dd, const Exact64Ctx& ec,
                              bool verify, unsigned long long& nretest) {
    float ix = __builtin_amdgcn_rcpf(R.dx), iy = __builtin_amdgcn_rcpf(R.dy), iz = __builtin_amdgcn_rcpf(R.dz);
    bool have64 = false;
    Ray64 R64;
    int n = root;
    while (n >= 0) {
        const BvhNode32 nd = a.nodes32[n];
        float t1x = (nd.lo[0] - R.ox) * ix, t2x = (nd.hi[0] - R.ox) * ix;
        float t1y = (nd.lo[1] - R.oy) * iy, t2y = (nd.hi[1] - R.oy) * iy;
        float t1z = (nd.lo[2] - R.oz) * iz, t2z = (nd.hi[2] - R.oz) * iz;
        float tmin = fmaxf(fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z)), 0.0f);
        float tmax = fminf(fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z)), R.L * 1.0001f);
        if (tmin <= tmax) {
            const int cnt = nd.start_count & 15;
            if (cnt > 0) {
                const int st = nd.start_count >> 4;
                for (int k = st; k < st + cnt; k++) {
                    if (mt_filter32(a.tris32[k], R, Sm, dd)) {
                        if (!have64) { exact_ray(ec, R64); have64 = true; }
                        nretest++;
                        if (mt_hit64(a.tris64[k], R64) &&
                            (!verify || reaches64(a.nodes64, a.parent, n, R64, 1.0 / R64.dx, 1.0 / R64.dy, 1.0 / R64.dz)))
                            return true;
                    }
                }
                n = nd.miss;
            } else {
                n = n + 1;
            }
        } else {
            n = nd.miss;
        }
    }
    return false;
}

__device__ __forceinline__ void tile_mixed_occ(const PairArgs& a, int tile) {
    const float Sm = a.Sm;
    const TileCtx t = tile_ctx(a, tile);
    const Corners cwi = corners_of(a, t.oi, t.vi && a.duffy);
    const int64_t NS = a.NS;
    const int npi = t.vi ? a.elem_np[t.i] : 0;
    int npmax = npi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) npmax = max(npmax, __shfl_xor(npmax, o, 64));
    npmax = __builtin_amdgcn_readfirstlane(npmax);
    const double Ai = t.vi ? a.area[t.i] : 1.0;
    const int gi = t.vi ? a.elem_grp[t.i] : 0;
    const float* __restrict__ P = a.P32;
    const double* __restrict__ P64 = a.P64;
    const double aix = a.anchor[t.i * 4 + 0], aiy = a.anchor[t.i * 4 + 1], aiz = a.anchor[t.i * 4 + 2];
    const float fax = (float)aix, fay = (float)aiy, faz = (float)aiz;
    double ls = 0.0, lc = 0.0;
    unsigned long long nret = 0, nrays = 0;
    for (int64_t j = t.jb; j < t.je; j++) {
        if (j < t.jlo) {
            if (t.lane == 0) a.partT[(int64_t)t.s * a.NS + j] = 0.0;
            continue;
        }
        bool act = t.vi && t.i < j + a.self_vf;
        const int64_t oj = a.slot_orig[j];
        if (a.duffy && shares_corner(cwi, corners_of(a, oj, true))) act = false;
        const int npj = a.elem_np[j];
        const int gj = a.elem_grp[j];
        const int root = a.tree_root[gi * a.G + gj];
        const float* __restrict__ qj = a.Q32 + j * a.maxpts * 8;
        const double* __restrict__ qj64 = a.Q64 + j * a.maxpts * 8;
        const bool rev = t.oi > oj;
        const float fjx = (float)a.anchor[j * 4 + 0], fjy = (float)a.anchor[j * 4 + 1], fjz = (float)a.anchor[j * 4 + 2];
        const float dax = (float)(a.anchor[j * 4 + 0] - aix);
        const float day = (float)(a.anchor[j * 4 + 1] - aiy);
        const float daz = (float)(a.anchor[j * 4 + 2] - aiz);
        double Fs = 0.0, Fc = 0.0;
        uint32_t bc = 0, bh = 0;
        for (int p = 0; p < npmax; p++) {
            const bool pv = act && p < npi;
            const float dxi = P[(int64_t)(p * 8 + 0) * NS + t.i], dyi = P[(int64_t)(p * 8 + 1) * NS + t.i];
            const float dzi = P[(int64_t)(p * 8 + 2) * NS + t.i];
            const float nxi = P[(int64_t)(p * 8 + 3) * NS + t.i], nyi = P[(int64_t)(p * 8 + 4) * NS + t.i];
            const float nzi = P[(int64_t)(p * 8 + 5) * NS + t.i];
            const float bx = dax - dxi, by = day - dyi, bz = daz - dzi;
            float inner = 0.0f;
            for (int q = 0; q < npj; q++) {
                const float qx = qj[q * 8 + 0], qy = qj[q * 8 + 1], qz = qj[q * 8 + 2];
                const float nxj = qj[q * 8 + 3], nyj = qj[q * 8 + 4], nzj = qj[q * 8 + 5];
                const float wdj = qj[q * 8 + 6];
                float rx, ry, rz, r2;
                float k = vf_mixed(bx, by, bz, nxi, nyi, nzi, qx, qy, qz, nxj, nyj, nzj, rx, ry, rz, r2);
                const bool need = pv && k != 0.0f && root >= 0;
                uint64_t pending = __ballot(need);
                if (pending) {
                    nrays += need;
                    float rs = __builtin_amdgcn_rsqf(r2);
                    Ray32 R;
                    const float sg = rev ? -rs : rs;
                    R.ox = rev ? fjx + qx : fax + dxi; R.oy = rev ? fjy + qy : fay + dyi;
                    R.oz = rev ? fjz + qz : faz + dzi;
                    R.dx = rx * sg; R.dy = ry * sg; R.dz = rz * sg; R.L = r2 * rs;
                    const float uu = 5.9604645e-08f;
                    float dd = 3.0f * uu * (fabsf(dax) + fabsf(day) + fabsf(daz) + fabsf(bx) + fabsf(by) + fabsf(bz) +
                                            fabsf(rx) + fabsf(ry) + fabsf(rz)) * rs + 4.0f * uu;
                    Exact64Ctx ec{P64, NS, t.i, p, qj64 + q * 8, rev};
                    const bool blk = need && blocked_mixed(a, root, R, Sm, dd, ec, a.verify != 0, nret);
                    if (blk) {
                        k = 0.0f;
                        bc++;
                        bh += mix32((uint32_t)(rev ? q * npi + p : p * npj + q) + 0x9e3779b9U);
                    }
                }
                inner = fmaf(wdj, k, inner);
            }
            if (p < npi) {
                double wd = P64[(int64_t)(p * 8 + 6) * NS + t.i] * P64[(int64_t)(p * 8 + 7) * NS + t.i];
                neumaier(Fs, Fc, wd * (double)inner);
            }
        }
        if (act && a.audit_count) {
            a.audit_count[min(t.oi, oj) + a.N * max(t.oi, oj)] = bc;
            a.audit_hash[min(t.oi, oj) + a.N * max(t.oi, oj)] = bh;
        }
        pair_out(a, t, j, oj, act, (Fs + Fc) * (1.0 / PI_D), Ai, ls, lc);
    }
    a.part[(int64_t)t.tile * RVF_STRIP + t.lane] = ls + lc;
    if (a.counters) {
        unsigned long long r0 = nret, r1 = nrays;
        for (int o = 32; o > 0; o >>= 1) { r0 += __shfl_xor(r0, o, 64); r1 += __shfl_xor(r1, o, 64); }
        if (t.lane == 0) { atomicAdd(&a.counters[0], r0); atomicAdd(&a.counters[1], r1); }
    }
}
// ---------------------------------------------------------------------------
// Occluded pair tiles, wave-queue form.
//
// Per column element j the (p, q) point combinations are handled in batches
// of at most RVF_QCOMBO.  Phase A evaluates the kernel for every lane and
// combination and stores the wave's need mask per combination in LDS.  Phase
// B drains those rays as one queue shared by the whole wavefront: every lane
// keeps S rays in flight (independent fetch chains) and refills a finished
// slot with the next pending ray in queue order, whichever lane's element it
// starts from, so lanes stay busy until the batch is empty and neighbouring
// lanes trace neighbouring rays.  A traversal step fetches one item (node or
// triangle) and applies the slab and triangle tests branch-free.  Blocked
// rays set a bit per (combination, lane) in LDS.  Phase C re-evaluates the
// kernel and sums the unblocked terms in each lane's own order, so FP64 sums
// stay in the reference's order and visibility decisions are per ray as in
// is_visible.
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_rank(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// Point data of a column element, which no pair kernel writes, read at
// wave-uniform addresses: through a constant-address-space pointer the loads
// go to the scalar cache and land in SGPRs.
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ cptr<T> as_const(const T* p) { return (cptr<T>)p; }

__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

struct WaveQueue {
    int c, nc;      // current combination, number of combinations (uniform)
    uint64_t m;     // pending source lanes of combination c (uniform)
};

// Hands pending rays, in queue order, to the lanes with `want` set, in lane
// order.  Must be called with all 64 lanes active.  Returns true in lanes
// that received a ray, with its combination and source lane.
__device__ __forceinline__ bool queue_take(uint64_t req, WaveQueue& Q, const uint64_t* need, int lane, int& c_out,
                                           int& b_out) {
    bool got = false;
    while (req != 0 && Q.c < Q.nc) {
        if (Q.m == 0) {
            Q.c++;
            Q.m = Q.c < Q.nc ? uniform64(need[Q.c]) : 0;
            continue;
        }
        const int take = min(__popcll(Q.m), __popcll(req));
        const bool holder = (Q.m >> lane) & 1;
        const int hr = lane_rank(Q.m);
        // forward permute: the holder of rank r sends its lane index to lane r
        // (non-holders write lane 63, which is read only when every lane holds)
        const int src = __builtin_amdgcn_ds_permute((holder ? hr : 63) << 2, lane);
        const bool isreq = (req >> lane) & 1;
        const int rr = lane_rank(req);
        const int b = __builtin_amdgcn_ds_bpermute(rr << 2, src);
        if (isreq && rr < take) { got = true; c_out = Q.c; b_out = b; }
        Q.m &= ~__ballot(holder && hr < take);
        req &= ~__ballot(isreq && rr < take);
    }
    return got;
}

// FP32 slab test on a padded box (a superset of the FP64 test's passes).
__device__ __forceinline__ bool slab32(const float* v, float ox, float oy, float oz, float ix, float iy, float iz,
                                       float L) {
    float t1x = (v[0] - ox) * ix, t2x = (v[3] - ox) * ix;
    float t1y = (v[1] - oy) * iy, t2y = (v[4] - oy) * iy;
    float t1z = (v[2] - oz) * iz, t2z = (v[5] - oz) * iz;
    float tmin = fmaxf(fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z)), 0.0f);
    float tmax = fminf(fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z)), L * 1.0001f);
    return tmin <= tmax;
}

// mt_filter32 on an item record (same bounds; sa computed from the FP32 edges
// with a relative margin covering its own rounding).
__device__ __forceinline__ int mt_filter32_item(const float* v, float ox, float oy, float oz, float dx, float dy,
                                                float dz, float L, float Sm, float dd) {
#pragma clang fp contract(fast)
    const float u = 5.9604645e-08f;
    float e1x = v[3], e1y = v[4], e1z = v[5];
    float e2x = v[6], e2y = v[7], e2z = v[8];
    float hx = dy * e2z - dz * e2y;
    float hy = dz * e2x - dx * e2z;
    float hz = dx * e2y - dy * e2x;
    float av = e1x * hx + e1y * hy + e1z * hz;
    float E1 = fabsf(e1x) + fabsf(e1y) + fabsf(e1z);
    float E2 = fabsf(e2x) + fabsf(e2y) + fabsf(e2z);
    float sa = E1 * E2 * 1.000002f;
    float KA = 2.0f * sa * (dd + 9.0f * u);
    float f = __builtin_amdgcn_rcpf(av);
    float af = fabsf(f) * 1.0000005f;
    float sx = ox - v[0], sy = oy - v[1], sz = oz - v[2];
    float ub = f * (sx * hx + sy * hy + sz * hz);
    float eu = (2.0f * E2 * Sm * (64.0f * u + 6.0f * dd) + 2.0f * KA) * af + 12.0f * u;
    float qx = sy * e1z - sz * e1y, qy = sz * e1x - sx * e1z, qz = sx * e1y - sy * e1x;
    float vb = f * (dx * qx + dy * qy + dz * qz);
    float ev = (2.0f * E1 * Sm * (12.0f * dd + 68.0f * u) + 2.0f * KA) * af + 12.0f * u;
    float tb = f * (e2x * qx + e2y * qy + e2z * qz);
    float et = (132.0f * u * sa * Sm + fabsf(tb) * KA) * af + 6.0f * u * fabsf(tb);
    const bool band = fabsf(av) <= KA + 1e-10f;  // grazing or near the 1e-10 cut: undecided
    const bool miss = (ub < -eu) | (ub > 1.0f + eu) | (vb < -ev) | (ub + vb > 1.0f + eu + ev) | (tb < -et) |
                      (tb > L * (1.0f + 2.0f * (dd + 3.0f * u)) + et);
    return band || !miss;
}

// Tile-level data one queue batch needs.
struct QCtx {
    int64_t sbase;       // first slot of the strip
    int64_t j, oj;       // column element (slot, original index)
    int npj, gj, p0, pch;
    uint32_t magic;      // floor(c / npj) = (c * magic) >> 16 for c < RVF_QCOMBO
    uint64_t revmask;    // lanes whose element is the pair's inner element
    const int* lroot;    // LDS: per lane, tree root of (lane's group, column group)
    int* lhint;          // LDS (mixed, split form): per lane, a triangle that certainly blocked a ray of its pair, or -1
    const int* lfr;      // LDS (mixed, split form): per lane, the frontier subtrees its pair's rays walk, [k * 64 + lane]
    const float* stage;  // LDS (mixed, split form): per lane [dax day daz fax fay faz] x 64
    float fjx, fjy, fjz; // column element's anchor as FP32
    // the strip's row points and the column element's points (FP32 and
    // FP64), indexed with 32-bit offsets: every plan has fewer than 2^31 pair
    // tiles of at most 16 columns, so NS < 2.1M and 8 maxpts NS < 2^30
    // (maxpts <= RVF_MAXPTS = 64); the host also refuses the mixed split
    // variants past 2^32.  Against 64-bit index arithmetic: mixed 164.5 ->
    // 162.7 ms (FP32 points), then 163.4 -> 162.6 ms and FP64 669.1 -> 665.8
    // ms (FP64 points) on the occluded cube.
    const float* p32;    // a.P32 + sbase
    const float* q32;    // a.Q32 + j maxpts 8
    const double* p64;   // a.P64 + sbase
    const double* q64;   // a.Q64 + j maxpts 8
    uint32_t ns32;       // a.NS
};

// Per-wave statistics of the wave-queue kernels, kept in LDS (u64 each) so
// that no counter stays in a register across a trace: rays cast, c0, c1,
// loop iterations, slot steps, task batches, exact batches, culled rays,
// skipped columns.
enum { QS_RAYS, QS_C0, QS_C1, QS_ITERS, QS_STEPS, QS_TBATCH, QS_EBATCH, QS_CULLED, QS_SKIPPED };
static_assert(QS_SKIPPED < RVF_QSTATS, "per-wave statistics block too small");

// Traversal counters of one drain of the wave-queue kernels: per lane, or
// wave-uniform where the whole wave counts one event.
struct QCount {
    uint32_t c0 = 0, c1 = 0;     // FP64: box tests, triangle tests; mixed: exact re-tests, certain hits
    uint32_t iters = 0;          // traversal loop iterations (uniform)
    uint32_t steps = 0;          // slot steps that advanced a ray
    uint32_t wsteps = 0;         // the same as a wave total (uniform; split kernels: a ballot's popcount, no VALU)
    uint32_t tbatch = 0;         // triangle-task batches (uniform)
    uint32_t ebatch = 0;         // exact-test batches (uniform)
    // adds the drain's wave totals to the per-wave block (every lane active)
    __device__ __forceinline__ void flush(uint64_t* w, int lane) const {
        uint32_t v[3] = {c0, c1, steps};
#pragma unroll
        for (int o = 32; o > 0; o >>= 1)
#pragma unroll
            for (int e = 0; e < 3; e++) v[e] += __shfl_xor(v[e], o, 64);
        if (lane == 0) {
            w[QS_C0] += v[0]; w[QS_C1] += v[1]; w[QS_STEPS] += v[2] + wsteps;
            w[QS_ITERS] += iters; w[QS_TBATCH] += tbatch; w[QS_EBATCH] += ebatch;
        }
    }
};

template <bool MIX>
struct QRay;

template <>
struct QRay<false> {  // FP64 ray: origin, direction, inverse direction
    double ox, oy, oz, dx, dy, dz, ix, iy, iz, L;
};
template <>
struct QRay<true> {   // FP32 ray with its direction error bound
    float ox, oy, oz, dx, dy, dz, ix, iy, iz, L, dd;
};

// Combination c of a queue batch: p = p0 + c / npj and, serpentine, q runs
// up for even p and down for odd p, so consecutive rays of a pair stay close
// (the blocker hint of the ray before serves more often; measured 1.9 % fewer
// node steps, 178.0 -> 176.9 ms on the occluded cube).  combo_index is the
// inverse; the sums still run in the reference's (p, q) order.
__device__ __forceinline__ void combo_pq(const QCtx& X, int c, int& p, int& q) {
    const int pl = (int)(((uint32_t)c * X.magic) >> 16);
    p = X.p0 + pl;
    q = c - pl * X.npj;
    q = (p & 1) ? X.npj - 1 - q : q;
}
__device__ __forceinline__ int combo_index(int p, int q, int p0, int npj) {
    return (p - p0) * npj + ((p & 1) ? npj - 1 - q : q);
}

// FP64 ray of combination c from lane b's element (is_visible, RayCast.jl:27-30,
// direction as vf_exact computes it; reversed when the column element is the
// pair's outer element).  Returns the tree root, -1 if none.
__device__ __forceinline__ int qray_setup(const PairArgs& a, const QCtx& X, int c, int b, QRay<false>& R) {
    int p, q;
    combo_pq(X, c, p, q);
    const uint32_t o = (uint32_t)(p * 8) * X.ns32 + (uint32_t)b;
    const double xi = X.p64[o], yi = X.p64[o + X.ns32];
    const double zi = X.p64[o + 2 * X.ns32];
    const double* qj = X.q64 + q * 8;
    const double xj = qj[0], yj = qj[1], zj = qj[2];
    const bool rev = (X.revmask >> b) & 1;
    const double r0 = xj - xi, r1 = yj - yi, r2 = zj - zi;
    const double rr = (r0 * r0 + r1 * r1) + r2 * r2;
    const double r = sqrt(rr);
    double h0 = r0 / r, h1 = r1 / r, h2 = r2 / r;
    if (rev) { h0 = -h0; h1 = -h1; h2 = -h2; }
    R.ox = rev ? xj : xi; R.oy = rev ? yj : yi; R.oz = rev ? zj : zi;
    R.dx = h0; R.dy = h1; R.dz = h2; R.L = r;
    R.ix = 1.0 / h0; R.iy = 1.0 / h1; R.iz = 1.0 / h2;
    return X.lroot[b];
}

// FP32 ray of combination c from lane b's element, from the anchored FP32
// coordinates exactly as the mixed kernel evaluation forms r.
__device__ __forceinline__ int qray_setup(const PairArgs& a, const QCtx& X, int c, int b, QRay<true>& R) {
    int p, q;
    combo_pq(X, c, p, q);
    const int64_t ib = X.sbase + b;
    const int64_t NS = a.NS;
    const float dxi = a.P32[(int64_t)(p * 8 + 0) * NS + ib], dyi = a.P32[(int64_t)(p * 8 + 1) * NS + ib];
    const float dzi = a.P32[(int64_t)(p * 8 + 2) * NS + ib];
    const float* qj = a.Q32 + (X.j * a.maxpts + q) * 8;
    const float qx = qj[0], qy = qj[1], qz = qj[2];
    const double aix = a.anchor[ib * 4 + 0], aiy = a.anchor[ib * 4 + 1], aiz = a.anchor[ib * 4 + 2];
    const double ajx = a.anchor[X.j * 4 + 0], ajy = a.anchor[X.j * 4 + 1], ajz = a.anchor[X.j * 4 + 2];
    const bool rev = (X.revmask >> b) & 1;
    const float dax = (float)(ajx - aix), day = (float)(ajy - aiy), daz = (float)(ajz - aiz);
    const float bx = dax - dxi, by = day - dyi, bz = daz - dzi;
    const float rx = bx + qx, ry = by + qy, rz = bz + qz;
    float r2;
    {
#pragma clang fp contract(fast)
        r2 = rx * rx + ry * ry + rz * rz;
    }
    const float rs = __builtin_amdgcn_rsqf(r2);
    const float sg = rev ? -rs : rs;
    R.ox = rev ? (float)ajx + qx : (float)aix + dxi;
    R.oy = rev ? (float)ajy + qy : (float)aiy + dyi;
    R.oz = rev ? (float)ajz + qz : (float)aiz + dzi;
    R.dx = rx * sg; R.dy = ry * sg; R.dz = rz * sg; R.L = r2 * rs;
    const float uu = 5.9604645e-08f;
    R.dd = 3.0f * uu * (fabsf(dax) + fabsf(day) + fabsf(daz) + fabsf(bx) + fabsf(by) + fabsf(bz) + fabsf(rx) +
                        fabsf(ry) + fabsf(rz)) * rs + 4.0f * uu;
    R.ix = __builtin_amdgcn_rcpf(R.dx); R.iy = __builtin_amdgcn_rcpf(R.dy); R.iz = __builtin_amdgcn_rcpf(R.dz);
    return X.lroot[b];
}

// Exact FP64 decision for a triangle the FP32 filter kept (mixed mode): the
// reference's ray (is_visible, RayCast.jl:27-30) and _ray_triangle; with the
// reference soup reading also the leaf's exact slab chain.
__device__ __forceinline__ bool qexact_hit(const PairArgs& a, const QCtx& X, int c, int b, int tri, int leaf) {
    int p, q;
    combo_pq(X, c, p, q);
    const uint32_t o = (uint32_t)(p * 8) * X.ns32 + (uint32_t)b;
    const double xi = X.p64[o], yi = X.p64[o + X.ns32];
    const double zi = X.p64[o + 2 * X.ns32];
    const double* qj = X.q64 + q * 8;
    const double xj = qj[0], yj = qj[1], zj = qj[2];
    const bool rev = (X.revmask >> b) & 1;
    const double r0 = xj - xi, r1 = yj - yi, r2 = zj - zi;
    const double r = sqrt((r0 * r0 + r1 * r1) + r2 * r2);
    Ray64 R;
    // one division at a time: this path runs while every slot's ray is live,
    // so it must not widen the register peak
    R.dx = r0 / r;
    __builtin_amdgcn_sched_barrier(0);
    R.dy = r1 / r;
    __builtin_amdgcn_sched_barrier(0);
    R.dz = r2 / r;
    __builtin_amdgcn_sched_barrier(0);
    R.L = r;
    if (rev) { R.dx = -R.dx; R.dy = -R.dy; R.dz = -R.dz; }
    R.ox = rev ? xj : xi; R.oy = rev ? yj : yi; R.oz = rev ? zj : zi;
    if (!mt_hit64(a.tris64[tri], R)) return false;
    return !a.verify || reaches64(a.nodes64, a.parent, leaf, R, 1.0 / R.dx, 1.0 / R.dy, 1.0 / R.dz);
}

// One traversal step of an FP64 ray: returns the next item, -1 when the ray
// is done; sets hit when a triangle blocks it.
__device__ __forceinline__ int qstep(const PairArgs& a, const QCtx&, const Item64& it, int cur, bool act, int,
                                     const QRay<false>& R, bool& hit, uint32_t& naabb, uint32_t& ntri) {
    const bool is_node = cur < a.n_nodes;
    // slab (BVH.jl:155-169, NaN slab values make the box a miss)
    double t1x = (it.v[0] - R.ox) * R.ix, t2x = (it.v[3] - R.ox) * R.ix;
    double t1y = (it.v[1] - R.oy) * R.iy, t2y = (it.v[4] - R.oy) * R.iy;
    double t1z = (it.v[2] - R.oz) * R.iz, t2z = (it.v[5] - R.oz) * R.iz;
    bool nan = (t1x != t1x) | (t2x != t2x) | (t1y != t1y) | (t2y != t2y) | (t1z != t1z) | (t2z != t2z);
    double tmin = fmax(fmax(fmax(fmin(t1x, t2x), fmin(t1y, t2y)), fmin(t1z, t2z)), 0.0);
    double tmax = fmin(fmin(fmin(fmax(t1x, t2x), fmax(t1y, t2y)), fmax(t1z, t2z)), R.L);
    const bool pass = !nan && tmin <= tmax;
    // Moller-Trumbore (BVH.jl:181-201) on the same record read as a triangle
    Tri64 T;
    T.v0[0] = it.v[0]; T.v0[1] = it.v[1]; T.v0[2] = it.v[2];
    T.e1[0] = it.v[3]; T.e1[1] = it.v[4]; T.e1[2] = it.v[5];
    T.e2[0] = it.v[6]; T.e2[1] = it.v[7]; T.e2[2] = it.v[8];
    Ray64 r64{R.ox, R.oy, R.oz, R.dx, R.dy, R.dz, R.L};
    const bool th = mt_hit64(T, r64);
    naabb += act && is_node; ntri += act && !is_node;
    hit = act && !is_node && th;
    return (is_node && !pass) ? it.link_y : it.link_x;
}

// One traversal step of an FP32 ray (mixed mode).  A triangle the FP32
// filter keeps is reported as a candidate (cand) for the exact FP64 test,
// which the caller runs after all slots have stepped.
__device__ __forceinline__ int qstep(const PairArgs& a, const QCtx&, const Item32& it, int cur, bool act, int,
                                     const QRay<true>& R, bool& cand, uint32_t& nret, uint32_t&) {
    const bool is_node = cur < a.n_nodes;
    const bool pass = slab32(it.v, R.ox, R.oy, R.oz, R.ix, R.iy, R.iz, R.L);
    const int f = mt_filter32_item(it.v, R.ox, R.oy, R.oz, R.dx, R.dy, R.dz, R.L, a.Sm, R.dd);
    cand = act && !is_node && f;
    nret += cand;
    return (is_node && !pass) ? it.link_y : it.link_x;
}

template <bool MIX>
using QItem = typename std::conditional<MIX, Item32, Item64>::type;

template <bool MIX>
__device__ __forceinline__ const QItem<MIX>* qitems(const PairArgs& a);
template <>
__device__ __forceinline__ const Item32* qitems<true>(const PairArgs& a) { return a.items32; }
template <>
__device__ __forceinline__ const Item64* qitems<false>(const PairArgs& a) { return a.items64; }

// Phase B: drain the batch's rays; blocked rays set blk bits.  A slot is
// refilled once at least refill_min lanes have it empty (or the wave has no
// ray in flight), so the refill's cross-lane exchange and ray set-up are
// amortised over several traversal steps.  The item fetches of all slots are
// issued before any is used.
template <bool MIX, int S>
__device__ __forceinline__ void queue_drain(const PairArgs& a, const QCtx& X, int nc, const uint64_t* need,
                                            uint32_t* blk, int lane, QCount& K) {
    const QItem<MIX>* __restrict__ items = qitems<MIX>(a);
    WaveQueue Q;
    Q.c = 0; Q.nc = nc; Q.m = uniform64(need[0]);
    int cur[S], tag[S], ctri[S], cleaf[S];
    QRay<MIX> R[S];
#pragma unroll
    for (int s = 0; s < S; s++) { cur[s] = -1; tag[s] = 0; ctri[s] = 0; cleaf[s] = 0; }
    for (;;) {
        if (Q.c < Q.nc) {
            bool any = false;
#pragma unroll
            for (int s = 0; s < S; s++) any |= cur[s] >= 0;
            const bool idle = __ballot(any) == 0;
#pragma unroll
            for (int s = 0; s < S; s++) {
                const uint64_t req = __ballot(cur[s] < 0);
                if (Q.c < Q.nc && req != 0 && (idle || __popcll(req) >= a.refill_min)) {
                    int c = 0, b = 0;
                    if (queue_take(req, Q, need, lane, c, b)) {
                        tag[s] = (c << 6) | b;
                        cur[s] = qray_setup(a, X, c, b, R[s]);
                    }
                }
            }
        }
        bool any = false;
#pragma unroll
        for (int s = 0; s < S; s++) any |= cur[s] >= 0;
        if (__ballot(any) == 0) break;
        K.iters++;
        bool hit[S];
        {
            QItem<MIX> it[S];
#pragma unroll
            for (int s = 0; s < S; s++) it[s] = items[max(cur[s], 0)];
#pragma unroll
            for (int s = 0; s < S; s++) {
                const bool act = cur[s] >= 0;
                const int tri = cur[s] - a.n_nodes;
                const int nx = qstep(a, X, it[s], max(cur[s], 0), act, tag[s], R[s], hit[s], K.c0, K.c1);
                K.steps += act;
                if (MIX) { ctri[s] = tri; cleaf[s] = it[s].link_y; }
                cur[s] = act ? nx : -1;
            }
        }
        if (MIX) {  // exact FP64 decisions for the FP32 candidates
#pragma unroll
            for (int s = 0; s < S; s++)
                if (hit[s]) hit[s] = qexact_hit(a, X, tag[s] >> 6, tag[s] & 63, ctri[s], cleaf[s]);
        }
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (hit[s]) {
                const int c = tag[s] >> 6, b = tag[s] & 63;
                atomicOr(&blk[c * 2 + (b >> 5)], 1u << (b & 31));
                cur[s] = -1;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Split form of the wave-queue traversal: the S ray slots of a lane only walk
// nodes (a slot holds origin, inverse direction and length); each triangle of
// a leaf a ray reaches becomes a task in an LDS ring, and the whole wavefront
// runs those tasks 64 at a time, so the triangle test runs at full width
// instead of riding along with every node step.  A ray keeps walking past a
// leaf whose tasks are pending; a task that finds a hit sets the ray's blocked
// bit, and the slot stops at the next iteration's check.  The blocked
// decision is the OR over every triangle of every leaf the traversal reaches,
// which is the stackless any-hit result (BVH.jl:213-250).  The FP64 variants
// walk the reference trees; the mixed ones their own traversal trees (small
// leaves), whose decisions are the same because every triangle an exact test
// could hit lies inside the padded boxes of its leaf and their ancestors.
// ---------------------------------------------------------------------------
// RVF_TQCAP, RVF_TQCAP_MIX (powers of two): one iteration of S <= 3 slots
// pushes at most S x 64 lanes x the leaf size (8 in the reference trees,
// RVF_TLEAF_MAX in the traversal trees); RVF_CQCAP >= 64 + 63.
// (RVF_TQ_TRI_BITS in rvf_common.h: a task packs (tag << 19) | triangle)
static_assert(RVF_QCOMBO * 64 <= (1 << (32 - RVF_TQ_TRI_BITS)), "task tag does not fit");
static_assert(RVF_FRONTIER <= 16, "frontier row index and remaining count take 5 and 4 bits of a slot tag");
static_assert(RVF_QCOMBO <= 128, "a mixed slot tag holds the combination in bits 16-22");

template <bool MIX>
struct NRay;
template <>
struct NRay<false> { double ox, oy, oz, ix, iy, iz, L; };
template <>
struct NRay<true> {  // slab form t = fma(bound, i, n): n = -o * i; L with the slab test's slack
    float nx, ny, nz, ix, iy, iz, L;
};

__device__ __forceinline__ int nray_setup(const PairArgs& a, const QCtx& X, int c, int b, NRay<false>& N) {
    QRay<false> R;
    const int root = qray_setup(a, X, c, b, R);
    N.ox = R.ox; N.oy = R.oy; N.oz = R.oz; N.ix = R.ix; N.iy = R.iy; N.iz = R.iz; N.L = R.L;
    return root;
}
struct SRay { float ox, oy, oz, dx, dy, dz, L, dd; };  // FP32 ray and its direction error bound
__device__ __forceinline__ SRay sray(const PairArgs& a, const QCtx& X, int c, int b);
__device__ __forceinline__ int nray_from(const SRay& S, int root, NRay<true>& N);
__device__ __forceinline__ int nray_setup(const PairArgs& a, const QCtx& X, int c, int b, NRay<true>& N) {
    return nray_from(sray(a, X, c, b), X.lroot[b], N);  // the same ray as the triangle tasks'
}
__device__ __forceinline__ int nray_from(const SRay& S, int root, NRay<true>& N) {
    struct { float ox, oy, oz, ix, iy, iz, L; } R;
    R.ox = S.ox; R.oy = S.oy; R.oz = S.oz; R.L = S.L;
    R.ix = __builtin_amdgcn_rcpf(S.dx); R.iy = __builtin_amdgcn_rcpf(S.dy); R.iz = __builtin_amdgcn_rcpf(S.dz);
    // A zero direction component gives an infinite inverse; clamped to 1e30 the
    // slab of that axis is still (-huge, +huge) inside the box's extent and
    // empty outside, and fma(bound, i, -o*i) stays finite (|bound|, |o| <= Sm).
    // The FMA form's rounding, about u |o i|, is far inside the boxes' padding
    // of 128 u Sm (128 u Sm |i| in t).
    N.ix = fminf(fmaxf(R.ix, -1e30f), 1e30f);
    N.iy = fminf(fmaxf(R.iy, -1e30f), 1e30f);
    N.iz = fminf(fmaxf(R.iz, -1e30f), 1e30f);
    N.nx = -R.ox * N.ix; N.ny = -R.oy * N.iy; N.nz = -R.oz * N.iz;
    N.L = R.L * 1.0001f;
    return root;
}

// Node step: returns the next node (-1 = done); a passed leaf reports its
// first triangle and count.
__device__ __forceinline__ int nstep(const BvhNode64& nd, int cur, const NRay<false>& N, int& first, int& cnt) {
    double t1x = (nd.lo[0] - N.ox) * N.ix, t2x = (nd.hi[0] - N.ox) * N.ix;
    double t1y = (nd.lo[1] - N.oy) * N.iy, t2y = (nd.hi[1] - N.oy) * N.iy;
    double t1z = (nd.lo[2] - N.oz) * N.iz, t2z = (nd.hi[2] - N.oz) * N.iz;
    bool nan = (t1x != t1x) | (t2x != t2x) | (t1y != t1y) | (t2y != t2y) | (t1z != t1z) | (t2z != t2z);
    double tmin = fmax(fmax(fmax(fmin(t1x, t2x), fmin(t1y, t2y)), fmin(t1z, t2z)), 0.0);
    double tmax = fmin(fmin(fmin(fmax(t1x, t2x), fmax(t1y, t2y)), fmax(t1z, t2z)), N.L);
    const bool pass = !nan && tmin <= tmax;
    const bool leaf = nd.count > 0;
    first = nd.start;
    cnt = (pass && leaf) ? nd.count : 0;
    return (pass && !leaf) ? cur + 1 : nd.miss;
}
// Mixed traversal tree: single-triangle leaves, exits instead of miss links
// (the next node is never negative; a walk past the tree lands at INT32_MAX).
__device__ __forceinline__ int nstep(const TNode32& nd, int cur, const NRay<true>& N, int& first, int& cnt) {
    const float t1x = fmaf(nd.lo[0], N.ix, N.nx), t2x = fmaf(nd.hi[0], N.ix, N.nx);
    const float t1y = fmaf(nd.lo[1], N.iy, N.ny), t2y = fmaf(nd.hi[1], N.iy, N.ny);
    const float t1z = fmaf(nd.lo[2], N.iz, N.nz), t2z = fmaf(nd.hi[2], N.iz, N.nz);
    const float tmin = fmaxf(fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z)), 0.0f);
    const float tmax = fminf(fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z)), N.L);
    const bool pass = tmin <= tmax;
    const bool leaf = nd.link < 0;
    first = ~nd.link;
    cnt = (pass && leaf) ? 1 : 0;
    return (pass && !leaf) ? cur + 1 : nd.exit;
}

template <bool MIX>
__device__ __forceinline__ const typename std::conditional<MIX, TNode32, BvhNode64>::type* nnodes(const PairArgs& a);
template <>
__device__ __forceinline__ const TNode32* nnodes<true>(const PairArgs& a) { return a.tnodes32; }  // traversal tree
template <>
__device__ __forceinline__ const BvhNode64* nnodes<false>(const PairArgs& a) { return a.nodes64; }

// One triangle task: the ray of (combination, lane) against one triangle,
// with is_visible's origin and direction (RayCast.jl:27-30).
__device__ __forceinline__ bool ttest(const PairArgs& a, const QCtx& X, int c, int b, int tri, uint32_t&) {
    int p, q;
    combo_pq(X, c, p, q);
    const uint32_t o = (uint32_t)(p * 8) * X.ns32 + (uint32_t)b;
    const double xi = X.p64[o], yi = X.p64[o + X.ns32];
    const double zi = X.p64[o + 2 * X.ns32];
    const double* qj = X.q64 + q * 8;
    const double xj = qj[0], yj = qj[1], zj = qj[2];
    const bool rev = (X.revmask >> b) & 1;
    const double r0 = xj - xi, r1 = yj - yi, r2 = zj - zi;
    const double r = sqrt((r0 * r0 + r1 * r1) + r2 * r2);
    Ray64 R64;
    R64.dx = r0 / r; R64.dy = r1 / r; R64.dz = r2 / r; R64.L = r;
    if (rev) { R64.dx = -R64.dx; R64.dy = -R64.dy; R64.dz = -R64.dz; }
    R64.ox = rev ? xj : xi; R64.oy = rev ? yj : yi; R64.oz = rev ? zj : zi;
    return mt_hit64(a.tris64[tri], R64);
}

// Mixed split traversal: the FP32 ray of combination c and lane b from the
// staged anchored offsets, as the kernel evaluation forms r.  dd bounds the
// direction's error (the classifier's margins).
__device__ __forceinline__ SRay sray(const PairArgs&, const QCtx& X, int c, int b) {
    int p, q;
    combo_pq(X, c, p, q);
    const uint32_t o = (uint32_t)(p * 8) * X.ns32 + (uint32_t)b;
    const float dxi = X.p32[o], dyi = X.p32[o + X.ns32];
    const float dzi = X.p32[o + 2 * X.ns32];
    const float* qj = X.q32 + q * 8;
    const float qx = qj[0], qy = qj[1], qz = qj[2];
    const float dax = X.stage[0 * 64 + b], day = X.stage[1 * 64 + b], daz = X.stage[2 * 64 + b];
    const float fax = X.stage[3 * 64 + b], fay = X.stage[4 * 64 + b], faz = X.stage[5 * 64 + b];
    const bool rev = (X.revmask >> b) & 1;
    const float bx = dax - dxi, by = day - dyi, bz = daz - dzi;
    const float rx = bx + qx, ry = by + qy, rz = bz + qz;
    float r2;
    {
#pragma clang fp contract(fast)
        r2 = rx * rx + ry * ry + rz * rz;
    }
    const float rs = __builtin_amdgcn_rsqf(r2);
    const float sg = rev ? -rs : rs;
    SRay R;
    R.ox = rev ? X.fjx + qx : fax + dxi;
    R.oy = rev ? X.fjy + qy : fay + dyi;
    R.oz = rev ? X.fjz + qz : faz + dzi;
    R.dx = rx * sg; R.dy = ry * sg; R.dz = rz * sg; R.L = r2 * rs;
    const float uu = 5.9604645e-08f;
    R.dd = 3.0f * uu * (fabsf(dax) + fabsf(day) + fabsf(daz) + fabsf(bx) + fabsf(by) + fabsf(bz) +
                        fabsf(rx) + fabsf(ry) + fabsf(rz)) * rs + 4.0f * uu;
    return R;
}

// Mixed mode triangle task: the conservative classifier on that ray:
// 0 certain miss, 1 candidate for the exact FP64 test, 2 certain hit.
__device__ __forceinline__ int tclass(const PairArgs& a, const SRay& S, int tri);
__device__ __forceinline__ int ttest_mixed(const PairArgs& a, const QCtx& X, int c, int b, int tri) {
    return tclass(a, sray(a, X, c, b), tri);
}
__device__ __forceinline__ int tclass(const PairArgs& a, const SRay& S, int tri) {
    const Tri32 T = a.ttris32[tri];  // traversal-tree order
    RvfFRay R;
    R.ox = S.ox; R.oy = S.oy; R.oz = S.oz; R.dx = S.dx; R.dy = S.dy; R.dz = S.dz; R.L = S.L;
    rvf_fray_margins(R, S.dd, a.Sm);
    const float uu = 5.9604645e-08f;
    return rvf_tri_classify(T.v0, T.e1, T.e2, T.E1, T.E2, R, 80.0f * uu * a.Sm);
}

// Mixed split traversal, after a node step of a slot (cur = the step's next
// node, never negative): the first step in a subtree reads its root, whose
// exit ends the subtree's preorder range; a walk that leaves the range (or
// the tree, INT32_MAX) moves on to the pair's next frontier subtree, or ends
// (cur = -1) after the last.  A mixed slot's tag packs the pair's lane
// (bits 0-5), the frontier row of the next subtree (bits 6-10, 1..16), the
// number of subtrees after the current one (bits 11-15) and the combination
// (bits 16-22), so bits 0-9 index the next entry's LDS word directly (row 16
// wraps to row 0, read but never taken) and a switch is one add.
// Branch-free: the next entry is read from LDS whether or not it is taken
// (13 % fewer VALU instructions per node step than decoding an index and a
// count: 165.2 -> 164.5 ms on the occluded cube).
__device__ __forceinline__ int mix_tag(int c, int b, int nf) { return b | (1 << 6) | ((nf - 1) << 11) | (c << 16); }
__device__ __forceinline__ uint32_t mix_tag_task(int tag) {  // combination << 6 | lane, as the tasks pack it
    return ((uint32_t)tag & 63u) | (((uint32_t)tag >> 10) & ~63u);
}
__device__ __forceinline__ void frontier_next(const QCtx& X, int exit, int& cur, int& tag, int& end) {
    end = end == -2 ? exit : end;
    const bool leave = cur >= end;
    const int nxt = X.lfr[tag & 0x3ff];
    const bool more = leave && (tag & 0xf800) != 0;
    cur = more ? nxt : (leave ? -1 : cur);
    end = more ? -2 : end;
    tag = more ? tag + (64 - 2048) : tag;  // next row, one subtree fewer
}

template <bool MIX, int S>
__device__ __forceinline__ void split_drain(const PairArgs& a, const QCtx& X, int nc, const uint64_t* need,
                                            uint32_t* blk, int lane, uint32_t* tq, QCount& K) {
    // ring capacity: one iteration pushes at most S x 64 leaves' triangles on
    // top of fewer than 64 pending tasks
    constexpr uint32_t TQC = MIX ? RVF_TQCAP_MIX : RVF_TQCAP;
    static_assert(S * 64 * (MIX ? RVF_TLEAF_MAX : 8) + 64 <= TQC, "task ring too small for S slots");
    const auto* __restrict__ nodes = nnodes<MIX>(a);
    WaveQueue Q;
    Q.c = 0; Q.nc = nc; Q.m = uniform64(need[0]);
    uint32_t head = 0, tail = 0;   // uniform ring cursors
    uint32_t chead = 0, ctail = 0; // candidate ring (mixed)
    bool bdirty = false;           // a batch may have set blocked bits (uniform)
    uint32_t* cq = tq + TQC;
    // a slot: current node, tag (FP64: combination << 6 | lane; mixed:
    // mix_tag) and (mixed) the end of the current subtree's preorder range
    // (-2 until its root has been read)
    int cur[S], tag[S], end[S];
    NRay<MIX> R[S];
#pragma unroll
    for (int s = 0; s < S; s++) { cur[s] = -1; tag[s] = 0; end[s] = -2; }
    for (;;) {
        // A ray whose blocked bit a task has set stops here, and its slot is
        // refilled below.  Blocked bits change only in task and exact batches,
        // so they are read after one ran.
        if (bdirty) {
#pragma unroll
            for (int s = 0; s < S; s++) {
                const int c = MIX ? tag[s] >> 16 : (tag[s] & 0xffff) >> 6, b = tag[s] & 63;
                if (cur[s] >= 0 && ((blk[c * 2 + (b >> 5)] >> (b & 31)) & 1u)) cur[s] = -1;
            }
            bdirty = false;
        }
        if (Q.c < Q.nc) {
            bool any = false;
#pragma unroll
            for (int s = 0; s < S; s++) any |= cur[s] >= 0;
            const bool idle = __ballot(any) == 0;
            if constexpr (MIX) {
                // mixed: a pass gives each lane that has an empty slot one ray
                // for its first empty slot, once enough lanes wait (however
                // their empty slots are spread); a further pass only while
                // enough lanes still wait.  A refilled ray is formed once for
                // the blocker-hint test and the walk: a triangle that
                // certainly blocked another ray of the pair often blocks this
                // one too, and a certain hit on it settles the ray at once.
#pragma unroll
                for (int pass = 0; pass < S; pass++) {
                    int es = -1;
#pragma unroll
                    for (int s = S - 1; s >= 0; s--) es = cur[s] < 0 ? s : es;
                    const uint64_t req = __ballot(es >= 0);
                    if (!(Q.c < Q.nc && req != 0 && (idle || __popcll(req) >= a.refill_min))) break;
                    int c = 0, b = 0;
                    if (queue_take(req, Q, need, lane, c, b)) {
                        const SRay SR = sray(a, X, c, b);
                        bool settled = false;
                        if (a.hint) {
                            const int h = X.lhint[b];
                            settled = h >= 0 && tclass(a, SR, h) == 2;
                        }
                        if (settled) {
                            K.c1++;
                            atomicOr(&blk[c * 2 + (b >> 5)], 1u << (b & 31));
                        } else {
                            // the walk starts at the pair's first frontier subtree
                            NRay<MIX> nr;
                            const int root = nray_from(SR, X.lfr[b], nr);
                            const int nf = X.lroot[b];
#pragma unroll
                            for (int s = 0; s < S; s++)
                                if (s == es) { tag[s] = mix_tag(c, b, nf); cur[s] = root; end[s] = -2; R[s] = nr; }
                        }
                    }
                }
            } else {
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const uint64_t req = __ballot(cur[s] < 0);
                    if (Q.c < Q.nc && req != 0 && (idle || __popcll(req) >= a.refill_min)) {
                        int c = 0, b = 0;
                        if (queue_take(req, Q, need, lane, c, b)) {
                            tag[s] = (c << 6) | b;
                            cur[s] = nray_setup(a, X, c, b, R[s]);
                        }
                    }
                }
            }
        }
        // node steps of every slot; passed leaves queue their triangles
        K.iters++;
        int first[S], cnt[S];
        uint32_t total = 0;
        {
            typename std::remove_cv<typename std::remove_reference<decltype(nodes[0])>::type>::type nd[S];
#pragma unroll
            for (int s = 0; s < S; s++)
                nd[s] = *(decltype(&nodes[0]))((const char*)nodes + (uint32_t)max(cur[s], 0) * (uint32_t)sizeof(nd[0]));
#pragma unroll
            for (int s = 0; s < S; s++) {
                const bool act = cur[s] >= 0;
                const int nx = nstep(nd[s], max(cur[s], 0), R[s], first[s], cnt[s]);
                if (!MIX) K.c0 += act;
                K.wsteps += (uint32_t)__popcll(__ballot(act));
                cnt[s] = act ? cnt[s] : 0;
                if constexpr (MIX) {
                    int c2 = nx, t2 = tag[s], e2 = end[s];
                    frontier_next(X, nd[s].exit, c2, t2, e2);
                    cur[s] = act ? c2 : -1;
                    tag[s] = act ? t2 : tag[s];
                    end[s] = act ? e2 : end[s];
                } else {
                    cur[s] = act ? nx : -1;
                }
            }
            // mixed: a slot whose step queued no leaf takes one more step, so
            // the iteration's fixed cost (refill, ring offsets, batches) is
            // spread over more steps while each slot still queues at most
            // one leaf per iteration (the ring bound).  Measured: 2 steps
            // -5 % kernel time, 3 and 4 no better.
            if constexpr (MIX) {
#pragma unroll
                for (int s = 0; s < S; s++)
                    nd[s] = *(decltype(&nodes[0]))((const char*)nodes + (uint32_t)max(cur[s], 0) * (uint32_t)sizeof(nd[0]));
#pragma unroll
                for (int s = 0; s < S; s++) {
                    const bool act = cur[s] >= 0 && cnt[s] == 0;
                    int f2 = 0, c2 = 0;
                    const int nx = nstep(nd[s], max(cur[s], 0), R[s], f2, c2);
                    K.wsteps += (uint32_t)__popcll(__ballot(act));
                    int cx = nx, tx = tag[s], ex = end[s];
                    frontier_next(X, nd[s].exit, cx, tx, ex);
                    first[s] = act ? f2 : first[s];
                    cnt[s] = act ? c2 : cnt[s];
                    cur[s] = act ? cx : cur[s];
                    tag[s] = act ? tx : tag[s];
                    end[s] = act ? ex : end[s];
                }
            }
#pragma unroll
            for (int s = 0; s < S; s++) total += (uint32_t)cnt[s];
        }
        // ring offsets: exclusive prefix of the lane totals over the wave (FP64:
        // <= 8 S < 32; mixed, single-triangle leaves: <= S <= 3)
        static_assert(!MIX || S * RVF_TLEAF_MAX < 4, "mixed lane totals need more bits");
        constexpr int TBITS = MIX ? 2 : 5;
        uint32_t off = 0, tot = 0;
#pragma unroll
        for (int bit = 0; bit < TBITS; bit++) {
            const uint64_t bm = __ballot((total >> bit) & 1u);
            off += (uint32_t)lane_rank(bm) << bit;
            tot += (uint32_t)__popcll(bm) << bit;
        }
        if (tot) {
#pragma unroll
            for (int s = 0; s < S; s++) {
                for (int k = 0; k < cnt[s]; k++)
                    tq[(tail + off + k) & (TQC - 1)] =
                        ((MIX ? mix_tag_task(tag[s]) : (uint32_t)(tag[s] & 0xffff)) << RVF_TQ_TRI_BITS) |
                        (uint32_t)(first[s] + k);
                off += (uint32_t)cnt[s];
            }
            tail += tot;
        }
        __builtin_amdgcn_wave_barrier();
        bool any = false;
#pragma unroll
        for (int s = 0; s < S; s++) any |= cur[s] >= 0;
        const bool wany = __ballot(any) != 0;
        // triangle tasks, 64 at a time; everything once no ray is in flight.
        // Mixed mode: the filter's candidates go to a second ring whose exact
        // FP64 tests also run 64 (or, at the end, all remaining) at a time.
        // (mixed: from 48 pending tasks, so blocked rays stop sooner; measured
        // 181.3 -> 179.2 ms on the occluded cube, 40: 180.1, 56: 179.7)
        constexpr uint32_t TBATCH = MIX ? 48 : 64;
        while (tail - head >= TBATCH || (!wany && (tail != head || ctail != chead))) {
            const uint32_t n = min(64u, tail - head);
            bool keep = false;
            uint32_t e = 0;
            if ((uint32_t)lane < n) {
                e = tq[(head + lane) & (TQC - 1)];
                const int tg = (int)(e >> RVF_TQ_TRI_BITS), tri = (int)(e & ((1u << RVF_TQ_TRI_BITS) - 1));
                const int c = tg >> 6, b = tg & 63;
                if (MIX) {
                    // a certain hit blocks the ray at once; with the reference
                    // soup reading the exact test also checks the leaf's box
                    // chain, so every kept triangle goes to it
                    const int cls = ttest_mixed(a, X, c, b, tri);
                    keep = cls == 1 || (cls == 2 && a.verify);
                    if (cls == 2 && !a.verify) {
                        K.c1++;
                        atomicOr(&blk[c * 2 + (b >> 5)], 1u << (b & 31));
                        X.lhint[b] = tri;  // any lane's certain hit will do
                    }
                } else {
                    K.c1++;
                    if (ttest(a, X, c, b, tri, K.c1)) atomicOr(&blk[c * 2 + (b >> 5)], 1u << (b & 31));
                }
            }
            K.tbatch += n > 0;
            head += n;
            bdirty = true;
            if (MIX) {
                const uint64_t km = __ballot(keep);
                if (keep) cq[(ctail + (uint32_t)lane_rank(km)) & (RVF_CQCAP - 1)] = e;
                ctail += (uint32_t)__popcll(km);
            }
            __builtin_amdgcn_wave_barrier();
            if (MIX) {
                while (ctail - chead >= 64 || (!wany && tail == head && ctail != chead)) {
                    const uint32_t m = min(64u, ctail - chead);
                    if ((uint32_t)lane < m) {
                        const uint32_t ce = cq[(chead + lane) & (RVF_CQCAP - 1)];
                        const int tg = (int)(ce >> RVF_TQ_TRI_BITS), tri = (int)(ce & ((1u << RVF_TQ_TRI_BITS) - 1));
                        const int c = tg >> 6, b = tg & 63;
                        K.c0++;
                        const int rt = a.tmap[tri];  // the triangle in the reference trees' order
                        if (qexact_hit(a, X, c, b, rt, a.items32[a.n_nodes + rt].link_y))
                            atomicOr(&blk[c * 2 + (b >> 5)], 1u << (b & 31));
                    }
                    K.ebatch++;
                    chead += m;
                    __builtin_amdgcn_wave_barrier();
                }
            }
        }
        if (!wany && Q.c >= Q.nc) break;
    }
}

// The subtrees a pair's rays walk (mixed split traversal): a cut of the
// traversal tree of at most RVF_FRONTIER nodes whose padded boxes the pair's
// capsule meets, nearest the pair's axis first, written to out[k * 64] (LDS, one column per
// lane; RVF_FRONTIER <= 16: a slot tag holds a row index and a count of 5 and 4 bits).  Every
// triangle a ray of the pair can reach lies in a leaf box the capsule meets,
// and every such leaf lies in one of these subtrees, so walking each subtree
// from its root to its root's exit visits every leaf the ray's walk from the
// root would pass.  Returns their number (0: the capsule meets no leaf box,
// nothing can block).  After `limit` box tests the walk stops and the root
// alone is returned.
__device__ __forceinline__ int capsule_frontier(const double* ai, const double* aj, double R,
                                                const TNode32* __restrict__ nodes, int root, int limit, int* out,
                                                float Sm) {
    // FP32 slab test of the segment a_i -> a_j (t in [0, 1]) against the box
    // grown by R, in the traversal's FMA form.  The boxes are padded outward
    // by 128 u Sm; the anchors' rounding to FP32 (u Sm), the FMA form's
    // rounding (a few u |o i| in t, i.e. a few u Sm in space) and the radius
    // (rounded up, plus 4 u Sm) stay inside that padding, so a capsule that
    // meets the exact box is never reported missing it.
    const float uu = 5.9604645e-08f;
    const float Rf = (float)R * (1.0f + 4.0f * uu) + 4.0f * uu * Sm;
    float o[3], iv[3], nn[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        o[c] = (float)ai[c];
        const float d = (float)aj[c] - o[c];
        iv[c] = fminf(fmaxf(__builtin_amdgcn_rcpf(d), -1e30f), 1e30f);
        nn[c] = -o[c] * iv[c];
    }
    auto meets = [&](const TNode32& nd) {
        float tmin = 0.0f, tmax = 1.0001f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float t1 = fmaf(nd.lo[c] - Rf, iv[c], nn[c]), t2 = fmaf(nd.hi[c] + Rf, iv[c], nn[c]);
            tmin = fmaxf(tmin, fminf(t1, t2));
            tmax = fminf(tmax, fmaxf(t1, t2));
        }
        return tmin <= tmax;
    };
    // Breadth first: each round replaces every open entry by the children
    // its capsule meets, as long as the list keeps at most RVF_FRONTIER
    // entries.  A round runs from the last entry to the first and writes the
    // new list from the top of the column down, so an entry not yet read is
    // never overwritten (entry u has the slots u .. w-1 to itself and the
    // entries before it); the list stays in preorder.  Leaves, and entries
    // that found no room (room only shrinks), are closed: stored as ~node.
    int tests = 1;
    if (!meets(nodes[root])) return 0;
    out[0] = root;
    int k = 1;
    for (bool open = true; open;) {
        open = false;
        int w = RVF_FRONTIER;
        for (int u = k - 1; u >= 0; u--) {
            const int e = out[u * 64];
            if (e >= 0) {
                const int lk = nodes[e].link;
                if (lk >= 0) {  // interior: children e + 1 and lk
                    if (tests + 2 > limit) { out[0] = root; return 1; }
                    tests += 2;
                    const int l = e + 1, r = lk;
                    const bool ml = meets(nodes[l]), mr = meets(nodes[r]);
                    if ((int)ml + (int)mr <= w - u) {
                        if (mr) out[--w * 64] = r;
                        if (ml) out[--w * 64] = l;
                        open = true;
                        continue;
                    }
                }
            }
            out[--w * 64] = e >= 0 ? ~e : e;  // closed (a leaf, or no room)
        }
        k = RVF_FRONTIER - w;
        for (int v = 0; v < k; v++) out[v * 64] = out[(w + v) * 64];
    }
    // the last round found every entry closed
    for (int v = 0; v < k; v++) out[v * 64] = ~out[v * 64];
    // Order: entries whose boxes lie nearest the pair's axis first (a stable
    // sort, preorder among equal keys).  Each entry is walked whole whatever
    // the order, so visibility is unchanged; a blocked ray stops at its first
    // blocker, and those lie mostly where the axis crosses the occluder.
    // key = distance from the box centre to the anchors' line in units of
    // (R + box half-diagonal) / 8, capped at 31 (5 bits above the node's 27).
    // Measured on the occluded cube: node steps -15.8 %, 173.7 -> 165.3 ms;
    // finer keys (16 per unit) 173.8 ms, distance in units of R alone 190.7,
    // 4 and 6 per unit 165.9 and 165.3.
    {
        const float ax = (float)ai[0], ay = (float)ai[1], az = (float)ai[2];
        float ux = (float)aj[0] - ax, uy = (float)aj[1] - ay, uz = (float)aj[2] - az;
        const float un = __builtin_amdgcn_rsqf(ux * ux + uy * uy + uz * uz);
        ux *= un; uy *= un; uz *= un;
        for (int v = 0; v < k; v++) {
            const int e = out[v * 64];
            const TNode32 nd = nodes[e];
            const float cx = 0.5f * (nd.lo[0] + nd.hi[0]) - ax, cy = 0.5f * (nd.lo[1] + nd.hi[1]) - ay;
            const float cz = 0.5f * (nd.lo[2] + nd.hi[2]) - az;
            const float px = cy * uz - cz * uy, py = cz * ux - cx * uz, pz = cx * uy - cy * ux;
            const float hx = nd.hi[0] - nd.lo[0], hy = nd.hi[1] - nd.lo[1], hz = nd.hi[2] - nd.lo[2];
            // hardware square root and reciprocal: the key orders the walk and
            // nothing else (160.7 against 161.2 ms with IEEE sqrtf and division)
            const float dist = __builtin_amdgcn_sqrtf(px * px + py * py + pz * pz);
            const float scale = (float)R + 0.5f * __builtin_amdgcn_sqrtf(hx * hx + hy * hy + hz * hz);
            const int q = (int)fminf(31.0f, fmaxf(0.0f, 8.0f * dist * __builtin_amdgcn_rcpf(scale)));
            out[v * 64] = (q << 27) | e;  // e < 2^27 (tree size limit)
        }
        for (int v = 1; v < k; v++) {
            const int x = out[v * 64];
            int w = v - 1;
            while (w >= 0 && out[w * 64] > x) { out[(w + 1) * 64] = out[w * 64]; w--; }
            out[(w + 1) * 64] = x;
        }
        for (int v = 0; v < k; v++) out[v * 64] &= (1 << 27) - 1;
    }
    return k;
}

// Can a ray between a point of element i and a point of element j reach the
// box?  Such rays lie in the convex hull of the two elements' points, inside
// the capsule of radius max(R_i, R_j) around the anchors' segment, hence
// inside that segment swept by the cube of that half-width: the test is the
// segment against the box grown by the radius.  The box is a tree's padded
// FP32 root box (128 u Sm beyond the triangles), so this FP64 test's own
// rounding cannot drop a ray that could touch a triangle, nor one whose exact
// FP64 slab test at the root (BVH.jl:155-169) could pass.
__device__ __forceinline__ bool capsule_meets_box(const double* ai, const double* aj, double R, const BvhNode32& b) {
    double tmin = 0.0, tmax = 1.0;
    for (int c = 0; c < 3; c++) {
        const double lo = (double)b.lo[c] - R, hi = (double)b.hi[c] + R;
        const double o = ai[c], d = aj[c] - ai[c];
        if (d == 0.0) {
            if (o < lo || o > hi) return false;
            continue;
        }
        double t1 = (lo - o) / d, t2 = (hi - o) / d;
        if (t1 > t2) { const double t = t1; t1 = t2; t2 = t; }
        tmin = fmax(tmin, t1);
        tmax = fmin(tmax, t2);
    }
    return tmin <= tmax;
}

// ---------------------------------------------------------------------------
// Mixed, occluded curve meshes: FP32 kernel, each cast path decided by the
// exact FP64 segment test from the FP64 points (so no decision can differ
// from the FP64 solve), from the outer element's point as the reference does.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void tile_mixed_occ2d(const PairArgs& a, int tile) {
    const TileCtx t = tile_ctx(a, tile);
    const int64_t NS = a.NS;
    const int npi = t.vi ? a.elem_np[t.i] : 0;
    int npmax = npi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) npmax = max(npmax, __shfl_xor(npmax, o, 64));
    npmax = __builtin_amdgcn_readfirstlane(npmax);
    const double Ai = t.vi ? a.area[t.i] : 1.0;
    const int gi = t.vi ? a.elem_grp[t.i] : 0;
    const float* __restrict__ P = a.P32;
    const double* __restrict__ P64 = a.P64;
    const double aix = a.anchor[t.i * 4 + 0], aiy = a.anchor[t.i * 4 + 1], aiz = a.anchor[t.i * 4 + 2];
    double ls = 0.0, lc = 0.0;
    unsigned long long nrays = 0, naabb = 0, ntri = 0;
    for (int64_t j = t.jb; j < t.je; j++) {
        if (j < t.jlo) {
            if (t.lane == 0) a.partT[(int64_t)t.s * a.NS + j] = 0.0;
            continue;
        }
        const bool act = t.vi && t.i < j + a.self_vf;
        const int64_t oj = a.slot_orig[j];
        const int npj = a.elem_np[j];
        const int root = a.tree_root[gi * a.G + a.elem_grp[j]];
        const float* __restrict__ qj = a.Q32 + j * a.maxpts * 8;
        const double* __restrict__ qj64 = a.Q64 + j * a.maxpts * 8;
        const bool rev = t.oi > oj;
        const float dax = (float)(a.anchor[j * 4 + 0] - aix);
        const float day = (float)(a.anchor[j * 4 + 1] - aiy);
        const float daz = (float)(a.anchor[j * 4 + 2] - aiz);
        double Fs = 0.0, Fc = 0.0;
        uint32_t bc = 0, bh = 0;
        for (int p = 0; p < npmax; p++) {
            const bool pv = act && p < npi;
            const float bx = dax - P[(int64_t)(p * 8 + 0) * NS + t.i], by = day - P[(int64_t)(p * 8 + 1) * NS + t.i];
            const float bz = daz - P[(int64_t)(p * 8 + 2) * NS + t.i];
            const float nxi = P[(int64_t)(p * 8 + 3) * NS + t.i], nyi = P[(int64_t)(p * 8 + 4) * NS + t.i];
            const float nzi = P[(int64_t)(p * 8 + 5) * NS + t.i];
            const double xi = P64[(int64_t)(p * 8 + 0) * NS + t.i], yi = P64[(int64_t)(p * 8 + 1) * NS + t.i];
            float inner = 0.0f;
            for (int q = 0; q < npj; q++) {
                float k = vf_mixed2d(bx, by, bz, nxi, nyi, nzi, qj[q * 8 + 0], qj[q * 8 + 1], qj[q * 8 + 2],
                                     qj[q * 8 + 3], qj[q * 8 + 4], qj[q * 8 + 5]);
                const bool need = pv && k != 0.0f && root >= 0;
                nrays += need;
                if (need) {
                    const double xj = qj64[q * 8 + 0], yj = qj64[q * 8 + 1];
                    const bool blk = rev ? blocked2d(a.nodes64, a.tris64, root, xj, yj, xi, yi, naabb, ntri)
                                         : blocked2d(a.nodes64, a.tris64, root, xi, yi, xj, yj, naabb, ntri);
                    if (blk) {
                        k = 0.0f;
                        bc++;
                        bh += mix32((uint32_t)(rev ? q * npi + p : p * npj + q) + 0x9e3779b9U);
                    }
                }
                inner = fmaf(qj[q * 8 + 6], k, inner);
            }
            if (p < npi) {
                double wd = P64[(int64_t)(p * 8 + 6) * NS + t.i] * P64[(int64_t)(p * 8 + 7) * NS + t.i];
                neumaier(Fs, Fc, wd * (double)inner);
            }
        }
        if (act && a.audit_count) {
            a.audit_count[min(t.oi, oj) + a.N * max(t.oi, oj)] = bc;
            a.audit_hash[min(t.oi, oj) + a.N * max(t.oi, oj)] = bh;
        }
        pair_out(a, t, j, oj, act, (Fs + Fc) * 0.5, Ai, ls, lc);
    }
    a.part[(int64_t)t.tile * RVF_STRIP + t.lane] = ls + lc;
    if (a.counters) {
        unsigned long long c1 = nrays, c2 = naabb, c3 = ntri;
        for (int o = 32; o > 0; o >>= 1) {
            c1 += __shfl_xor(c1, o, 64); c2 += __shfl_xor(c2, o, 64); c3 += __shfl_xor(c3, o, 64);
        }
        if (t.lane == 0) { atomicAdd(&a.counters[1], c1); atomicAdd(&a.counters[2], c2); atomicAdd(&a.counters[3], c3); }
    }
}

// The same capsule walked down a tree of padded FP32 boxes (preorder, miss
// links, interior start_count = right child << 4).  Returns -1 when it reaches
// no leaf's box: then no ray of the pair can reach a triangle, and none is
// blocked.  Otherwise returns the node the pair's rays may start from: the
// descent from the root follows the only child the capsule meets while there
// is one, so every node a ray of the pair could pass is in that node's
// subtree or is a sibling the capsule misses (which a ray leaving the subtree
// then tests and fails).  Box tests use the capsule's inverse direction; their
// rounding is far inside the boxes' padding.  After `limit` box tests the walk
// stops without culling.
__device__ __forceinline__ int capsule_entry(const double* ai, const double* aj, double R,
                                             const BvhNode32* __restrict__ nodes, int root, int limit) {
    double inv[3];
    bool flat[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const double d = aj[c] - ai[c];
        flat[c] = d == 0.0;
        inv[c] = flat[c] ? 0.0 : 1.0 / d;
    }
    auto meets = [&](const BvhNode32& nd) {
        double tmin = 0.0, tmax = 1.0;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double lo = (double)nd.lo[c] - R, hi = (double)nd.hi[c] + R;
            if (flat[c]) {
                if (ai[c] < lo || ai[c] > hi) tmax = -1.0;
            } else {
                double t1 = (lo - ai[c]) * inv[c], t2 = (hi - ai[c]) * inv[c];
                if (t1 > t2) { const double t = t1; t1 = t2; t2 = t; }
                tmin = fmax(tmin, t1);
                tmax = fmin(tmax, t2);
            }
        }
        return tmin <= tmax;
    };
    if (limit < 1) return root;
    int n = root, k = 1;
    BvhNode32 nd = nodes[n];
    if (!meets(nd)) return -1;
    while ((nd.start_count & 15) == 0) {  // descend while one child alone is met
        if (k + 2 > limit) return n;
        const int l = n + 1, r = nd.start_count >> 4;
        const BvhNode32 L = nodes[l], Rt = nodes[r];
        const bool ml = meets(L), mr = meets(Rt);
        k += 2;
        if (!ml && !mr) return -1;
        if (ml && mr) break;
        n = ml ? l : r;
        nd = ml ? L : Rt;
    }
    if (nd.start_count & 15) return n;  // a leaf the capsule meets
    // both children met: is any leaf below n met?
    for (int m = n + 1; m != nd.miss;) {
        if (k++ >= limit) return n;
        const BvhNode32 x = nodes[m];
        if (meets(x)) {
            if (x.start_count & 15) return n;
            m = m + 1;
        } else {
            m = x.miss;
        }
    }
    return -1;
}

// Per-wave statistics block of the wave-queue tiles in LDS (its place in the
// tile's LDS layout below), and its flush into the plan's counters.
template <bool MIX, bool SPLIT>
__device__ __forceinline__ uint64_t* occ_wst(const PairArgs& a, double* lds) {
    const char* p = (const char*)(lds + (MIX ? 0 : a.maxpts * 64));
    p += RVF_QCOMBO * (sizeof(uint64_t) + 2 * sizeof(uint32_t)) + 2 * 64 * sizeof(int);
    p += (SPLIT ? (MIX ? RVF_TQCAP_MIX + RVF_CQCAP : RVF_TQCAP) : 0) * sizeof(uint32_t);
    p += (MIX && SPLIT ? 6 * 64 * sizeof(float) + RVF_FRONTIER * 64 * sizeof(int) : 0);
    return (uint64_t*)p;
}
template <bool MIX>
__device__ __forceinline__ void occ_flush(const PairArgs& a, const uint64_t* wst) {
    if (a.counters && threadIdx.x == 0) {
        unsigned long long v[9];
#pragma unroll
        for (int e = 0; e < 9; e++) v[e] = wst[e];
        if (!MIX) v[QS_C0] += v[QS_CULLED];  // FP64 box tests: a culled ray's root test
        atomicAdd(&a.counters[1], v[QS_RAYS]);
        if (MIX) { atomicAdd(&a.counters[0], v[QS_C0]); atomicAdd(&a.counters[8], v[QS_C1]); }
        else { atomicAdd(&a.counters[2], v[QS_C0]); atomicAdd(&a.counters[3], v[QS_C1]); }
        for (int e = QS_ITERS; e <= QS_EBATCH; e++) atomicAdd(&a.counters[e + 1], v[e]);
        atomicAdd(&a.counters[9], v[QS_CULLED]);
        atomicAdd(&a.counters[10], v[QS_SKIPPED]);
    }
}

// WS: the per-wave statistics block is zeroed and flushed by the caller once
// per wave (persistent launches; per tile, every wave's eleven counter
// atomics contended with every other's: 26 ms of a tile loop that did no
// other work), not per tile.
template <bool MIX, int S, bool SPLIT, bool WS = false>
__device__ __forceinline__ void tile_occ_queue(const PairArgs& a, int tile, double* lds) {
    const TileCtx t = tile_ctx(a, tile);
    const int64_t NS = a.NS;
    const int lane = t.lane;
    const int npi = t.vi ? a.elem_np[t.i] : 0;
    int npmax = npi;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) npmax = max(npmax, __shfl_xor(npmax, o, 64));
    npmax = __builtin_amdgcn_readfirstlane(npmax);
    const double Ai = t.vi ? a.area[t.i] : 1.0;
    const int gi = t.vi ? a.elem_grp[t.i] : 0;
    // LDS: [accq: maxpts x 64 doubles (FP64 only)][need: RVF_QCOMBO u64][blk: 2 RVF_QCOMBO u32]
    double* accq = lds;
    uint64_t* need = (uint64_t*)(lds + (MIX ? 0 : a.maxpts * 64));
    uint32_t* blk = (uint32_t*)(need + RVF_QCOMBO);
    int* lroot = (int*)(blk + 2 * RVF_QCOMBO);
    int* lhint = lroot + 64;
    uint32_t* tq = (uint32_t*)(lhint + 64);
    float* stage = (float*)(tq + (SPLIT ? (MIX ? RVF_TQCAP_MIX + RVF_CQCAP : RVF_TQCAP) : 0));
    int* lfr = (int*)(stage + (MIX && SPLIT ? 6 * 64 : 0));  // (mixed split) RVF_FRONTIER x 64 frontier subtrees
    // per-wave statistics (QS_*): rays cast, and of those the rays of pairs
    // culled before any traversal, columns whose trace the wave skipped (all
    // its pairs culled), the drains' counters
    uint64_t* wst = (uint64_t*)(lfr + (MIX && SPLIT ? RVF_FRONTIER * 64 : 0));
    if (!WS) {
        if (lane < RVF_QSTATS) wst[lane] = 0;
        __syncthreads();
    }
    const double* __restrict__ P64 = a.P64;
    const float* __restrict__ P = a.P32;
    double ls = 0.0, lc = 0.0;
    QCtx X;
    X.sbase = (int64_t)t.s * RVF_STRIP;
    X.p32 = a.P32 + X.sbase;
    X.p64 = a.P64 + X.sbase;
    X.ns32 = (uint32_t)a.NS;
    X.lroot = lroot;
    X.lhint = lhint;
    X.stage = stage;
    X.lfr = lfr;
    for (int64_t j = t.jb; j < t.je; j++) {
        if (j < t.jlo) {
            if (lane == 0) a.partT[(int64_t)t.s * a.NS + j] = 0.0;
            continue;
        }
        bool act = t.vi && t.i < j + a.self_vf;
        const int64_t oj = a.slot_orig[j];
        // (the row element's corners are read again here rather than held
        // in registers across the column loop)
        if (a.duffy && shares_corner(corners_of(a, t.oi, t.vi), corners_of(a, oj, true))) act = false;
        const bool rev = t.oi > oj;
        const int npj = a.elem_np[j];
        const int gj = a.elem_grp[j];
        // mixed split traversal walks its own tree over the same triangles
        int root = (MIX && SPLIT ? a.ttree_root : a.tree_root)[gi * a.G + gj];
        // pairs whose rays cannot reach the occluders trace nothing (their
        // rays count as cast; FP64: with one box test, as the reference's)
        bool cap = false;
        if (act && root >= 0) {
            const double ai[3] = {a.anchor[t.i * 4 + 0], a.anchor[t.i * 4 + 1], a.anchor[t.i * 4 + 2]};
            const double aj[3] = {a.anchor[j * 4 + 0], a.anchor[j * 4 + 1], a.anchor[j * 4 + 2]};
            const double R = fmax(a.anchor[t.i * 4 + 3], a.anchor[j * 4 + 3]);
            // mixed: down the tree the rays walk, which also gives their start
            // node; FP64 keeps the root box alone, so its box-test counts stay
            // the reference's
            if (MIX && SPLIT) {
                // the subtrees the pair's rays walk; lroot keeps their number
                const int nf = capsule_frontier(ai, aj, R, a.tnodes32, root, 2 * a.cull_limit, lfr + lane, a.Sm);
                cap = nf > 0;
                if (cap) root = nf;
            } else if (MIX) {
                const int e = capsule_entry(ai, aj, R, a.nodes32, root, a.cull_limit);
                cap = e >= 0;
                if (cap) root = e;
            } else {
                cap = capsule_meets_box(ai, aj, R, a.nodes32[a.tree_root[gi * a.G + gj]]);
            }
        }
        X.j = j; X.oj = oj; X.npj = npj; X.gj = gj;
        X.q32 = a.Q32 + j * a.maxpts * 8;
        X.q64 = a.Q64 + j * a.maxpts * 8;
        X.revmask = __ballot(rev);
        lroot[lane] = root;
        lhint[lane] = -1;
        X.pch = max(1, RVF_QCOMBO / npj);
        X.magic = (65536u + (uint32_t)npj - 1u) / (uint32_t)npj;
        const cptr<double> qj64 = as_const(a.Q64 + j * a.maxpts * 8);
        const cptr<float> qj32 = as_const(a.Q32 + j * a.maxpts * 8);
        if (!MIX)
            for (int q = 0; q < npj; q++) accq[q * 64 + lane] = 0.0;
        // mixed mode: anchored offsets of the column element
        const double aix = a.anchor[t.i * 4 + 0], aiy = a.anchor[t.i * 4 + 1], aiz = a.anchor[t.i * 4 + 2];
        const float dax = (float)(a.anchor[j * 4 + 0] - aix);
        const float day = (float)(a.anchor[j * 4 + 1] - aiy);
        const float daz = (float)(a.anchor[j * 4 + 2] - aiz);
        if (MIX && SPLIT) {
            stage[0 * 64 + lane] = dax; stage[1 * 64 + lane] = day; stage[2 * 64 + lane] = daz;
            stage[3 * 64 + lane] = (float)aix; stage[4 * 64 + lane] = (float)aiy; stage[5 * 64 + lane] = (float)aiz;
            X.fjx = (float)a.anchor[j * 4 + 0]; X.fjy = (float)a.anchor[j * 4 + 1]; X.fjz = (float)a.anchor[j * 4 + 2];
        }
        double Fij = 0.0, Fs = 0.0, Fc = 0.0;
        uint32_t bc = 0, bh = 0;  // audit: blocked rays of the pair and their order-free hash
        const bool aud = a.audit_count != nullptr;
        // when no pair of the wave can reach an occluder (all culled), the need
        // masks and the trace are skipped; the sums below count the rays
        // instead, from the same kernel values
        const bool trace = __ballot(cap) != 0;
        const bool anyp = __ballot(act && root >= 0) != 0;
        if (lane == 0 && !trace && anyp) wst[QS_SKIPPED]++;
        for (int p0 = 0; p0 < npmax; p0 += X.pch) {
            const int p1 = min(npmax, p0 + X.pch);
            const int nc = (p1 - p0) * npj;
            X.p0 = p0;
            for (int w = lane; w < 2 * nc; w += 64) blk[w] = 0u;
            // Phase A: need masks (ray counts as wave totals): one ballot per
            // combination, the culled pairs' lanes masked out with capm
            const uint64_t capm = __ballot(cap);
            uint32_t wrays = 0, wculled = 0;
            for (int p = p0; p < p1 && trace; p++) {
                const bool pv = act && p < npi && root >= 0;
                if (MIX) {
                    const float dxi = X.p32[(uint32_t)(p * 8 + 0) * X.ns32 + (uint32_t)lane], dyi = X.p32[(uint32_t)(p * 8 + 1) * X.ns32 + (uint32_t)lane];
                    const float dzi = X.p32[(uint32_t)(p * 8 + 2) * X.ns32 + (uint32_t)lane];
                    const float nxi = X.p32[(uint32_t)(p * 8 + 3) * X.ns32 + (uint32_t)lane], nyi = X.p32[(uint32_t)(p * 8 + 4) * X.ns32 + (uint32_t)lane];
                    const float nzi = X.p32[(uint32_t)(p * 8 + 5) * X.ns32 + (uint32_t)lane];
                    const float bx = dax - dxi, by = day - dyi, bz = daz - dzi;
                    for (int q = 0; q < npj; q++) {
                        float rx, ry, rz, r2;
                        float k = vf_mixed(bx, by, bz, nxi, nyi, nzi, qj32[q * 8 + 0], qj32[q * 8 + 1], qj32[q * 8 + 2],
                                           qj32[q * 8 + 3], qj32[q * 8 + 4], qj32[q * 8 + 5], rx, ry, rz, r2);
                        const uint64_t ndm = __ballot(pv && k != 0.0f);
                        wrays += __popcll(ndm);
                        wculled += __popcll(ndm & ~capm);
                        const int c = combo_index(p, q, p0, npj);
                        if (lane == 0) need[c] = ndm & capm;
                    }
                } else {
                    const double xi = X.p64[(uint32_t)(p * 8 + 0) * X.ns32 + (uint32_t)lane], yi = X.p64[(uint32_t)(p * 8 + 1) * X.ns32 + (uint32_t)lane];
                    const double zi = X.p64[(uint32_t)(p * 8 + 2) * X.ns32 + (uint32_t)lane];
                    const double nxi = X.p64[(uint32_t)(p * 8 + 3) * X.ns32 + (uint32_t)lane];
                    const double nyi = X.p64[(uint32_t)(p * 8 + 4) * X.ns32 + (uint32_t)lane];
                    const double nzi = X.p64[(uint32_t)(p * 8 + 5) * X.ns32 + (uint32_t)lane];
                    for (int q = 0; q < npj; q++) {
                        Ray64 Rr;
                        const double K = vf_exact(xi, yi, zi, nxi, nyi, nzi, qj64[q * 8 + 0], qj64[q * 8 + 1],
                                                  qj64[q * 8 + 2], qj64[q * 8 + 3], qj64[q * 8 + 4], qj64[q * 8 + 5], Rr);
                        const uint64_t ndm = __ballot(pv && K != 0.0);
                        wrays += __popcll(ndm);
                        wculled += __popcll(ndm & ~capm);
                        const int c = combo_index(p, q, p0, npj);
                        if (lane == 0) need[c] = ndm & capm;
                    }
                }
            }
            if (lane == 0) { wst[QS_RAYS] += wrays; wst[QS_CULLED] += wculled; }
            __syncthreads();
            // Phase B: trace
            if (trace) {
                QCount K;
                if constexpr (SPLIT) split_drain<MIX, S>(a, X, nc, need, blk, lane, tq, K);
                else queue_drain<MIX, S>(a, X, nc, need, blk, lane, K);
                K.flush(wst, lane);
            }
            __syncthreads();
            // Phase C: sums in each lane's own order
            uint32_t wcl = 0;  // rays of a skipped trace (all culled)
            for (int p = p0; p < p1; p++) {
                if (MIX) {
                    const float dxi = X.p32[(uint32_t)(p * 8 + 0) * X.ns32 + (uint32_t)lane], dyi = X.p32[(uint32_t)(p * 8 + 1) * X.ns32 + (uint32_t)lane];
                    const float dzi = X.p32[(uint32_t)(p * 8 + 2) * X.ns32 + (uint32_t)lane];
                    const float nxi = X.p32[(uint32_t)(p * 8 + 3) * X.ns32 + (uint32_t)lane], nyi = X.p32[(uint32_t)(p * 8 + 4) * X.ns32 + (uint32_t)lane];
                    const float nzi = X.p32[(uint32_t)(p * 8 + 5) * X.ns32 + (uint32_t)lane];
                    const float bx = dax - dxi, by = day - dyi, bz = daz - dzi;
                    float inner = 0.0f;
                    if (trace) {
                        for (int q = 0; q < npj; q++) {
                            float rx, ry, rz, r2;
                            float k = vf_mixed(bx, by, bz, nxi, nyi, nzi, qj32[q * 8 + 0], qj32[q * 8 + 1],
                                               qj32[q * 8 + 2], qj32[q * 8 + 3], qj32[q * 8 + 4], qj32[q * 8 + 5], rx, ry,
                                               rz, r2);
                            const int c = combo_index(p, q, p0, npj);
                            const bool blocked = (blk[2 * c + (lane >> 5)] >> (lane & 31)) & 1u;
                            k = blocked ? 0.0f : k;
                            if (aud && blocked) {
                                bc++;
                                bh += mix32((uint32_t)(rev ? q * npi + p : p * npj + q) + 0x9e3779b9U);
                            }
                            inner = fmaf(qj32[q * 8 + 6], k, inner);
                        }
                    } else {
                        // no ray of the wave was traced (every blocked bit is
                        // 0); the rays are counted as the need masks would
                        // count them, every one culled
                        const bool pv = act && p < npi && root >= 0;
                        for (int q = 0; q < npj; q++) {
                            float rx, ry, rz, r2;
                            const float k = vf_mixed(bx, by, bz, nxi, nyi, nzi, qj32[q * 8 + 0], qj32[q * 8 + 1],
                                                     qj32[q * 8 + 2], qj32[q * 8 + 3], qj32[q * 8 + 4], qj32[q * 8 + 5],
                                                     rx, ry, rz, r2);
                            wcl += __popcll(__ballot(pv && k != 0.0f));
                            inner = fmaf(qj32[q * 8 + 6], k, inner);
                        }
                    }
                    if (p < npi) {
                        double wd = X.p64[(uint32_t)(p * 8 + 6) * X.ns32 + (uint32_t)lane] * X.p64[(uint32_t)(p * 8 + 7) * X.ns32 + (uint32_t)lane];
                        neumaier(Fs, Fc, wd * (double)inner);
                    }
                } else {
                    const double xi = X.p64[(uint32_t)(p * 8 + 0) * X.ns32 + (uint32_t)lane], yi = X.p64[(uint32_t)(p * 8 + 1) * X.ns32 + (uint32_t)lane];
                    const double zi = X.p64[(uint32_t)(p * 8 + 2) * X.ns32 + (uint32_t)lane];
                    const double nxi = X.p64[(uint32_t)(p * 8 + 3) * X.ns32 + (uint32_t)lane];
                    const double nyi = X.p64[(uint32_t)(p * 8 + 4) * X.ns32 + (uint32_t)lane];
                    const double nzi = X.p64[(uint32_t)(p * 8 + 5) * X.ns32 + (uint32_t)lane];
                    const double wi = X.p64[(uint32_t)(p * 8 + 6) * X.ns32 + (uint32_t)lane], dAi = X.p64[(uint32_t)(p * 8 + 7) * X.ns32 + (uint32_t)lane];
                    double inner = 0.0;
                    for (int q = 0; q < npj; q++) {
                        Ray64 Rr;
                        double K = vf_exact(xi, yi, zi, nxi, nyi, nzi, qj64[q * 8 + 0], qj64[q * 8 + 1], qj64[q * 8 + 2],
                                            qj64[q * 8 + 3], qj64[q * 8 + 4], qj64[q * 8 + 5], Rr);
                        const int c = combo_index(p, q, p0, npj);
                        if (!trace) {  // as the need masks count them: every such ray is culled
                            const bool nd = act && p < npi && root >= 0 && K != 0.0;
                            wcl += __popcll(__ballot(nd));
                        }
                        const bool blocked = (blk[2 * c + (lane >> 5)] >> (lane & 31)) & 1u;
                        K = blocked ? 0.0 : K;
                        if (aud && blocked) {
                            bc++;
                            bh += mix32((uint32_t)(rev ? q * npi + p : p * npj + q) + 0x9e3779b9U);
                        }
                        inner += (qj64[q * 8 + 6] * K) * qj64[q * 8 + 7];
                        if (rev && p < npi) accq[q * 64 + lane] += (wi * K) * dAi;
                    }
                    if (p < npi) Fij += (wi * inner) * dAi;
                }
            }
            if (lane == 0 && !trace) { wst[QS_RAYS] += wcl; wst[QS_CULLED] += wcl; }
            __syncthreads();
        }
        double raw;
        if (MIX) {
            raw = (Fs + Fc) * (1.0 / PI_D);
        } else {
            if (rev) {
                Fij = 0.0;
                for (int q = 0; q < npj; q++) Fij += (qj64[q * 8 + 6] * accq[q * 64 + lane]) * qj64[q * 8 + 7];
            }
            raw = Fij;
        }
        if (act && a.audit_count) {
            a.audit_count[min(t.oi, oj) + a.N * max(t.oi, oj)] = bc;
            a.audit_hash[min(t.oi, oj) + a.N * max(t.oi, oj)] = bh;
        }
        pair_out(a, t, j, oj, act, raw, Ai, ls, lc);
    }
    a.part[(int64_t)t.tile * RVF_STRIP + t.lane] = ls + lc;
    if (!WS) occ_flush<MIX>(a, wst);
}

// ---------------------------------------------------------------------------
// Monte Carlo (GPUMCKernels.jl:37-50,146-281): the reference's per-pair
// stream (splitmix64-seeded xorshift64, thread id (i-1)N + j), stratum k of
// element i paired with stratum k of element j, FP64 throughout; visibility
// with the CPU path's semantics (per-active-set trees, t > 0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t init_rng(uint64_t seed, int64_t tid) {
    uint64_t z = seed + (uint64_t)tid * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double xorshift64(uint64_t& s) {
    uint64_t x = s;
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    s = x;
    // (x >> 11) / (2^53 - 1), correctly rounded, without the division: for
    // 1 <= m < 2^53, m / (2^53 - 1) = y (1 + 2^-53 / (1 - 2^-53)) with
    // y = m 2^-53 exact, and the excess over y lies strictly between half an
    // ulp of y and y's ulp times (1 + 2^-52), so the rounded quotient is the
    // double after y (0 for m = 0).  Checked against the division for every m
    // below 5,000 and above 2^53 - 5,000, at m = 2^e + {-2..2} and on 35 M
    // random m; the Monte Carlo streams stay bit-identical to the oracle's.
    const uint64_t m = x >> 11;
    const double y = (double)m * 0x1p-53;
    return m ? __longlong_as_double(__double_as_longlong(y) + 1) : 0.0;
}
// _sample_on_element (GPUMCKernels.jl:266-281)
__device__ __forceinline__ void mc_param(int f, double u1, double u2, double& xi, double& eta) {
    if (f >= 4) { xi = u1; eta = 0.0; }  // curve element: u1 is the point on [-1, 1]
    else if (f == 0 || f == 2) { xi = 2.0 * u1 - 1.0; eta = 2.0 * u2 - 1.0; }
    else { xi = u1; eta = u2; if (xi + eta > 1.0) { xi = 1.0 - xi; eta = 1.0 - eta; } }
}
__device__ __forceinline__ void mc_sample(const PairArgs& a, int f, int64_t e, double u1, double u2, double* x,
                                          double* n, double* dA) {
    double xi, eta;
    mc_param(f, u1, u2, xi, eta);
    // read-only data: where e is wave-uniform (the pair kernel's column
    // element) the node list and coordinates come through the scalar cache
    // (653 -> 619 ms on the cylinder); elsewhere these are ordinary loads
    eval_point(f, as_const(a.coords), as_const(a.nodes8 + 8 * e), xi, eta, x, n, dA);
}
// the same with the element's node coordinates in registers
__device__ __forceinline__ void mc_sample_r(int f, const double* cx, const double* cy, const double* cz, double u1,
                                            double u2, double* x, double* n, double* dA) {
    double xi, eta;
    mc_param(f, u1, u2, xi, eta);
    eval_point_r(f, cx, cy, cz, xi, eta, x, n, dA);
}

// Curve meshes (the reference's GPU Monte Carlo has no curve form): the same
// per-pair stream, sample k of both elements in stratum k of n equal strata of
// [-1, 1] as _stratified_line_points places them (MCKernel.jl:107-115):
// xi = -1 + ((k + u) * 2) / n, u1 for element i, u3 for element j.
__device__ __forceinline__ double mc_line_point(int k, double u, int n) { return -1.0 + ((k + u) * 2.0) / n; }

// MC area of element e from the stream of the pair (e, e), mirroring the CPU
// path's area pass element_pair_view_factor_mc(i, i) (Assembly.jl:164-169).
__global__ __launch_bounds__(64) void k_mc_area(PairArgs a, double* area, double* area_orig) {
    int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // slot == element (identity layout)
    if (e >= a.N) return;
    int f = a.fam[e];
    uint64_t st = init_rng(a.mc_seed, e * a.N + e + 1);
    int s = a.mc_s;
    double Ai = 0.0;
    int k = 0;
    double x[3], n[3], dA;
    if (a.dim2) {
        for (int kk = 0; kk < a.mc_nsamp; kk++) {
            double u1 = xorshift64(st);
            xorshift64(st);
            mc_sample(a, f, e, mc_line_point(kk, u1, a.mc_nsamp), 0.0, x, n, &dA);
            Ai += dA;
            xorshift64(st); xorshift64(st);
        }
        area[e] = (Ai * 2.0) / a.mc_nsamp;
        area_orig[a.slot_orig[e]] = area[e];
        return;
    }
    for (int si = 0; si < s; si++)
        for (int sj = 0; sj < s; sj++) {
            k++;
            double u1 = xorshift64(st), u2 = xorshift64(st);
            mc_sample(a, f, e, (si + u1) / s, (sj + u2) / s, x, n, &dA);
            Ai += dA;
            xorshift64(st); xorshift64(st);
        }
    for (; k < a.mc_nsamp; k++) {
        double u1 = xorshift64(st), u2 = xorshift64(st);
        mc_sample(a, f, e, u1, u2, x, n, &dA);
        Ai += dA;
        xorshift64(st); xorshift64(st);
    }
    double ri = (f == 0 || f == 2) ? 4.0 : 0.5;
    area[e] = (Ai * ri) / a.mc_nsamp;
    area_orig[a.slot_orig[e]] = area[e];
}

__device__ __forceinline__ void tile_mc(const PairArgs& a, int tile) {
    const TileCtx t = tile_ctx(a, tile);
    const Corners cwi = corners_of(a, t.oi, false);
    (void)cwi;
    const int64_t NS = a.NS;
    // Monte Carlo solves use the identity slot layout (slot == element), so
    // i < j is the reference's thread ordering (GPUMCKernels.jl:167).
    // lanes past the last element sample element 0 with its own family (its
    // node list; another family's shape would read node slots it leaves 0)
    const int fi = a.fam[t.vi ? t.oi : 0];
    const double Ai = t.vi ? a.area[t.i] : 1.0;
    const int gi = t.vi ? a.elem_grp[t.i] : 0;
    const int64_t ii = t.vi ? t.oi : 0;
    const int s = a.mc_s;
    // the row element's node coordinates, the same for every column and
    // sample: loaded once (gathered per sample, 64 elements' nodes per wave
    // made the L1 the bound: 1,400 cache accesses per sample and wavefront)
    double cx[8], cy[8], cz[8];
    {
        const int32_t* nd = a.nodes8 + 8 * ii;
        const int nni = fi >= 4 ? (fi == 4 ? 3 : 2) : (fi == 0 ? 8 : fi == 1 ? 6 : fi == 2 ? 4 : 3);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const double* xa = a.coords + 3 * (int64_t)(nd[k < nni ? k : 0] - 1);
            cx[k] = xa[0]; cy[k] = xa[1]; cz[k] = xa[2];
        }
    }
    double ls = 0.0, lc = 0.0;
    unsigned long long nrays = 0, naabb = 0, ntri = 0;
    for (int64_t j = t.jb; j < t.je; j++) {
        if (j < t.jlo) {
            if (t.lane == 0) a.partT[(int64_t)t.s * NS + j] = 0.0;
            continue;
        }
        const bool act = t.vi && t.i < j + a.self_vf;
        const int64_t oj = __builtin_amdgcn_readfirstlane(a.slot_orig[j]);  // wave-uniform
        const int fj = a.fam[oj];
        const int root = a.tree_root ? a.tree_root[gi * a.G + a.elem_grp[j]] : -1;
        uint64_t st = init_rng(a.mc_seed, ii * a.N + oj + 1);
        double Ks = 0.0;
        int k = 0;
        double xi[3], ni[3], ai, xj[3], nj[3], aj;
        for (int kk = 0; kk < a.mc_nsamp; kk++) {
            double u1 = xorshift64(st), u2 = xorshift64(st);
            double si = (double)(kk / s), sj = (double)(kk % s);
            bool strat = kk < s * s;
            if (a.dim2) mc_sample_r(fi, cx, cy, cz, mc_line_point(kk, u1, a.mc_nsamp), 0.0, xi, ni, &ai);
            else mc_sample_r(fi, cx, cy, cz, strat ? (si + u1) / s : u1, strat ? (sj + u2) / s : u2, xi, ni, &ai);
            double u3 = xorshift64(st), u4 = xorshift64(st);
            if (a.dim2) mc_sample(a, fj, oj, mc_line_point(kk, u3, a.mc_nsamp), 0.0, xj, nj, &aj);
            else mc_sample(a, fj, oj, strat ? (si + u3) / s : u3, strat ? (sj + u4) / s : u4, xj, nj, &aj);
            Ray64 R;
            double K = a.dim2 ? vf_exact2d(xi[0], xi[1], xi[2], ni[0], ni[1], ni[2], xj[0], xj[1], xj[2], nj[0], nj[1],
                                           nj[2])
                              : vf_exact(xi[0], xi[1], xi[2], ni[0], ni[1], ni[2], xj[0], xj[1], xj[2], nj[0], nj[1],
                                         nj[2], R);
            const bool need = act && K > 0.0 && root >= 0;
            nrays += need;
            bool blk;
            if (a.dim2) blk = need && blocked2d(a.nodes64, a.tris64, root, xi[0], xi[1], xj[0], xj[1], naabb, ntri);
            else blk = (occlusion64(a, need, root, R, naabb, ntri) >> t.lane) & 1;
            if (blk) K = 0.0;
            Ks += (K * ai) * aj;
            k++;
        }
        (void)k;
        double ri = fi >= 4 ? 2.0 : (fi == 0 || fi == 2) ? 4.0 : 0.5;
        double rj = fj >= 4 ? 2.0 : (fj == 0 || fj == 2) ? 4.0 : 0.5;
        double raw = ((Ks * ri) * rj) / a.mc_nsamp;
        pair_out(a, t, j, oj, act, raw, Ai, ls, lc);
    }
    a.part[(int64_t)t.tile * RVF_STRIP + t.lane] = ls + lc;
    if (a.counters) {
        unsigned long long c1 = nrays;
        for (int o = 32; o > 0; o >>= 1) c1 += __shfl_xor(c1, o, 64);
        if (t.lane == 0) atomicAdd(&a.counters[1], c1);
    }
}

// ---------------------------------------------------------------------------
// Duffy / Sauter-Schwab for Quad8 pairs sharing corners (DuffyKernel.jl:
// 222-487, maps restated verbatim including the reference's Jacobians).  One
// wavefront per singular pair; lanes split the regions x nq^4 points and the
// non-negative terms are summed (per-lane sequential, then a wave tree).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void vertex_map(int region, double r, double e1, double e2, double e3, double u0i,
                                           double v0i, double u0j, double v0j, double& u, double& v, double& s,
                                           double& t) {
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
    u = u0i + (u0i == 0.0 ? du : -du);
    v = v0i + (v0i == 0.0 ? dv : -dv);
    s = u0j + (u0j == 0.0 ? ds : -ds);
    t = v0j + (v0j == 0.0 ? dt : -dt);
}

__device__ __forceinline__ void edge_map(int region, double r, double e1, double e2, double e3, double& u, double& v,
                                         double& s, double& t) {
    switch (region) {
    case 1:
        u = r * (1 - e1 * (1 - e3)); v = ((r * e1) * e2) * e3;
        s = (r * (1 - e1 * (1 - e3)) - ((r * e1) * (1 - e2)) * e3) + ((r * e1) * e2) * e3;
        t = ((r * e1) * (1 - e2)) * e3;
        break;
    case 2:
        u = r * (1 - (e1 * e2) * e3); v = ((r * e1) * e3) * (1 - e2);
        s = (r * (1 - (e1 * e2) * e3) + ((r * e1) * e2) * e3) - (r * e1) * e3;
        t = ((r * e1) * e2) * e3;
        break;
    case 3:
        u = r * (1 - e1 * e2); v = ((r * e1) * e2) * e3;
        s = r * (1 - e1 * e2) - ((r * e1) * e2) * (1 - e3);
        t = ((r * e1) * e2) * (1 - e3);
        break;
    case 4:
        u = r * (e2 * e3 + e1 * (1 - e2 * e3)); v = ((r * (1 - e1)) * e2) * e3;
        s = (r * e2) * e3; t = ((r * (1 - e1)) * e2) * e3;
        break;
    default:
        u = (r * e1) * e3; v = ((r * (1 - e1)) * e2) * e3;
        s = ((r * e1) * e3) * (1 - e2) + ((r * e1) * e2) * e3;
        t = ((r * (1 - e1)) * e2) * e3;
        break;
    }
}

__device__ __forceinline__ int corners_to_edge(int c1, int c2) {
    int a = c1 < c2 ? c1 : c2, b = c1 < c2 ? c2 : c1;
    if (a == 1 && b == 2) return 1;
    if (a == 2 && b == 3) return 2;
    if (a == 3 && b == 4) return 3;
    if (a == 1 && b == 4) return 4;
    return 1;
}
__device__ __forceinline__ void rotate_to_edge(double& u, double& v, int e) {
    double a = u, b = v;
    if (e == 2) { u = 1.0 - b; v = a; }
    else if (e == 3) { u = 1.0 - a; v = 1.0 - b; }
    else if (e == 4) { u = b; v = 1.0 - a; }
}

__global__ __launch_bounds__(64) void k_duffy(PairArgs a, const DuffyPair* __restrict__ pairs, int npairs,
                                              const double* __restrict__ gl01, int nq, double* __restrict__ out) {
    int pidx = blockIdx.x;
    if (pidx >= npairs) return;
    const DuffyPair P = pairs[pidx];
    const double* p01 = gl01;
    const double* w01 = gl01 + 8;
    const double CUV[4][2] = {{0.0, 0.0}, {1.0, 0.0}, {1.0, 1.0}, {0.0, 1.0}};
    double u0i = 0, v0i = 0, u0j = 0, v0j = 0;
    int edge_i = 1, edge_j = 1;
    if (P.kind == 1) {
        u0i = CUV[P.ci0 - 1][0]; v0i = CUV[P.ci0 - 1][1];
        u0j = CUV[P.cj0 - 1][0]; v0j = CUV[P.cj0 - 1][1];
    } else {
        edge_i = corners_to_edge(P.ci0, P.ci1);
        edge_j = corners_to_edge(P.cj0, P.cj1);
    }
    const int nreg = P.kind == 1 ? 8 : 5;
    const int nq4 = nq * nq * nq * nq;
    const int total = nreg * nq4;
    double s = 0.0, c = 0.0;
    unsigned long long naabb = 0, ntri = 0;
    for (int idx = threadIdx.x; idx < total; idx += 64) {
        int region = idx / nq4 + 1;
        int rem = idx % nq4;
        int ia = rem / (nq * nq * nq), ib = (rem / (nq * nq)) % nq, ic = (rem / nq) % nq, id = rem % nq;
        double r = p01[ia], wr = w01[ia], e1 = p01[ib], we1 = w01[ib];
        double e2 = p01[ic], we2 = w01[ic], e3 = p01[id], we3 = w01[id];
        double u, v, ss, tt;
        if (P.kind == 1) vertex_map(region, r, e1, e2, e3, u0i, v0i, u0j, v0j, u, v, ss, tt);
        else {
            edge_map(region, r, e1, e2, e3, u, v, ss, tt);
            rotate_to_edge(u, v, edge_i);
            rotate_to_edge(ss, tt, edge_j);
        }
        if (u < 0 || u > 1 || v < 0 || v > 1 || ss < 0 || ss > 1 || tt < 0 || tt > 1) continue;
        double xi[3], ni[3], ai, xj[3], nj[3], aj;
        eval_point(0, a.coords, a.nodes8 + 8 * (int64_t)P.i, 2 * u - 1.0, 2 * v - 1.0, xi, ni, &ai);
        ai = ai * 4.0;
        eval_point(0, a.coords, a.nodes8 + 8 * (int64_t)P.j, 2 * ss - 1.0, 2 * tt - 1.0, xj, nj, &aj);
        aj = aj * 4.0;
        Ray64 R;
        double K = vf_exact(xi[0], xi[1], xi[2], ni[0], ni[1], ni[2], xj[0], xj[1], xj[2], nj[0], nj[1], nj[2], R);
        if (K == 0.0) continue;
        if (P.root >= 0 && blocked64(a.nodes64, a.tris64, P.root, R, naabb, ntri)) continue;
        double jac = (r * r) * r;
        neumaier(s, c, ((((((wr * we1) * we2) * we3) * jac) * K) * ai) * aj);
    }
    double v = wave_sum_d(s + c);
    if (threadIdx.x == 0) out[pidx] = v;
}

// Write Duffy raw values into F (before the fill converts raw to F) and add
// their row / group contributions in a fixed order per row.
__global__ __launch_bounds__(256) void k_duffy_apply(PairArgs a, const DuffyPair* __restrict__ pairs,
                                                     const double* __restrict__ val, const int32_t* __restrict__ rptr,
                                                     const int32_t* __restrict__ rlist, double* __restrict__ rowsum,
                                                     double* __restrict__ W) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= a.N) return;
    double Ar = a.area[r];
    double ts = 0.0, tc = 0.0;
    for (int k = rptr[r]; k < rptr[r + 1]; k++) {
        int pidx = rlist[k];
        const DuffyPair P = pairs[pidx];
        double raw = val[pidx];
        if (P.i == P.j) {  // self pair (self_vf): the diagonal, added by k_diag_apply
            a.diag[r] = raw;
            if (a.F) a.F[(int64_t)r + a.N * (int64_t)r] = raw;
            continue;
        }
        double f = raw / Ar;
        neumaier(ts, tc, f);
        if (P.i == r) {  // partner j > r: upper part, group of j
            W[(int64_t)a.elem_grp[P.j] * a.NS + r] += f;
            if (a.F) a.F[(int64_t)P.i + a.N * (int64_t)P.j] = raw;
        }
    }
    rowsum[r] += ts + tc;
}

// Self view factors: F_ii = raw_ii / A_i joins the row sum (the group sums
// take the diagonal on the host).
__global__ __launch_bounds__(256) void k_diag_apply(const double* __restrict__ diag, const double* __restrict__ area,
                                                    int64_t N, double* __restrict__ rowsum) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= N) return;
    rowsum[r] += diag[r] / area[r];
}

// ---------------------------------------------------------------------------
// Fill: F[i,j] = raw/A_i and F[j,i] = raw/A_j for i < j (Assembly.jl:211),
// F[i,i] = raw_ii/A_i with self view factors, else 0.  One 64x64 tile per
// workgroup, transposed half through LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fill(double* __restrict__ F, const double* __restrict__ area, int64_t N,
                                              const int2* __restrict__ tiles, int self_vf) {
    __shared__ double T[64][65];
    int2 tl = tiles[blockIdx.x];
    int64_t i0 = (int64_t)tl.x * 64, j0 = (int64_t)tl.y * 64;
    int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    int64_t i = i0 + tx;
    double Ai = i < N ? area[i] : 1.0;
    for (int r = ty; r < 64; r += 4) {
        int64_t j = j0 + r;
        double out = 0.0;
        if (i < N && j < N) {
            if (i < j) {
                double raw = F[i + N * j];
                F[i + N * j] = raw / Ai;
                out = raw / area[j];
            } else if (i == j) {
                F[i + N * j] = self_vf ? F[i + N * j] / Ai : 0.0;
            }
        }
        T[tx][r] = out;  // value for F[j, i]
    }
    __syncthreads();
    // write F[j, i] for i < j: column i = i0 + c, rows j0 + lane
    int64_t j = j0 + tx;
    for (int c = ty; c < 64; c += 4) {
        int64_t ii = i0 + c;
        if (ii < N && j < N && ii < j) F[j + N * ii] = T[c][tx];
    }
}

// ---------------------------------------------------------------------------
// Row reduction: rowsum[r] and W[h][r] = sum_{j in group h, j > r} F_rj.
//   upper part: part[c][r] over chunks c with chunk_last(c) > 64*strip(r)
//   lower part: partT[s][r] over strips s <= (chunk_end(chunk(r)) - 2) / 64
// Two stages: RVF_RSEG segments per row (chunk / strip index mod RVF_RSEG)
// summed with Neumaier into Tseg / Wseg, then a fixed-order final sum.
// ---------------------------------------------------------------------------
#define RVF_RSEG 16

__global__ __launch_bounds__(256) void k_reduce_seg(const double* __restrict__ part, const double* __restrict__ partT,
                                                    const Chunk* __restrict__ chunks, int nchunks,
                                                    const int32_t* __restrict__ chunk_of, int64_t N, int NS, int G,
                                                    double* __restrict__ Tseg, double* __restrict__ Wseg) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int seg = blockIdx.y;
    if (r >= N) return;
    int64_t sr = r / RVF_STRIP;
    int lo = 0, hi = nchunks;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (chunks[mid].end - 1 > sr * RVF_STRIP) hi = mid; else lo = mid + 1;
    }
    double ts = 0.0, tc = 0.0;
    int curg = -1;
    double gs = 0.0, gc = 0.0;
    double* Wr = Wseg + (int64_t)seg * G * NS + r;
    int c0 = lo + ((seg - lo % RVF_RSEG) + RVF_RSEG) % RVF_RSEG;  // first c >= lo with c % RSEG == seg
    for (int c = c0; c < nchunks; c += RVF_RSEG) {
        int g = chunks[c].grp;
        if (g != curg) {
            if (curg >= 0) Wr[(int64_t)curg * NS] += gs + gc;
            curg = g; gs = 0.0; gc = 0.0;
        }
        double v = part[((int64_t)chunks[c].first_tile + sr) * RVF_STRIP + (r - sr * RVF_STRIP)];
        neumaier(gs, gc, v);
        neumaier(ts, tc, v);
    }
    if (curg >= 0) Wr[(int64_t)curg * NS] += gs + gc;
    int cr = chunk_of[r];
    int64_t smax = chunks[cr].end >= 2 ? (int64_t)(chunks[cr].end - 2) / RVF_STRIP : -1;
    for (int64_t s2 = seg; s2 <= smax; s2 += RVF_RSEG) neumaier(ts, tc, partT[s2 * NS + r]);
    Tseg[(int64_t)seg * NS + r] = ts + tc;
}

__global__ __launch_bounds__(256) void k_reduce_fin(const double* __restrict__ Tseg, const double* __restrict__ Wseg,
                                                    int64_t N, int NS, int G, double* __restrict__ rowsum,
                                                    double* __restrict__ W) {
    int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= N) return;
    double ts = 0.0, tc = 0.0;
    for (int sg = 0; sg < RVF_RSEG; sg++) neumaier(ts, tc, Tseg[(int64_t)sg * NS + r]);
    rowsum[r] = ts + tc;
    for (int h = 0; h < G; h++) {
        double ws = 0.0, wc = 0.0;
        for (int sg = 0; sg < RVF_RSEG; sg++) neumaier(ws, wc, Wseg[((int64_t)sg * G + h) * NS + r]);
        W[(int64_t)h * NS + r] = ws + wc;
    }
}

// ---------------------------------------------------------------------------
// Kernel entry points.  Direct launches map one tile per workgroup; the
// persistent batched form drains a device-resident queue of (request, tile)
// work items from many meshes in one launch: each wavefront takes the next
// item with one atomic and exits when the queue is empty.
// ---------------------------------------------------------------------------
template <int KIND, bool WS = false>
__device__ __forceinline__ void run_tile(const PairArgs& a, int tile, double* lds) {
    if (KIND == RVF_K_FP64_PC4) tile_fp64<4>(a, tile);
    else if (KIND == RVF_K_FP64_PC9) tile_fp64<9>(a, tile);
    else if (KIND == RVF_K_FP64_PC16) tile_fp64<16>(a, tile);
    else if (KIND == RVF_K_FP64_OCC_LANE) tile_fp64_occ(a, tile, lds);
    else if (KIND == RVF_K_MIX_PC4) tile_mixed<4, true>(a, tile);
    else if (KIND == RVF_K_MIX_PC9) tile_mixed<9, true>(a, tile);
    else if (KIND == RVF_K_MIX_PC16) tile_mixed<16, true>(a, tile);
    else if (KIND == RVF_K_MIX_PC16C) tile_mixed<16, false>(a, tile);
    else if (KIND == RVF_K_MIX_OCC_LANE) tile_mixed_occ(a, tile);
    else if (KIND == RVF_K_FP64_QUEUE1) tile_occ_queue<false, 1, false, WS>(a, tile, lds);
    else if (KIND == RVF_K_FP64_QUEUE2) tile_occ_queue<false, 2, false, WS>(a, tile, lds);
    else if (KIND == RVF_K_FP64_QUEUE3) tile_occ_queue<false, 3, false, WS>(a, tile, lds);
    else if (KIND == RVF_K_MIX_QUEUE2) tile_occ_queue<true, 2, false, WS>(a, tile, lds);
    else if (KIND == RVF_K_MIX_QUEUE3) tile_occ_queue<true, 3, false, WS>(a, tile, lds);
    else if (KIND == RVF_K_MIX_QUEUE4) tile_occ_queue<true, 4, false, WS>(a, tile, lds);
    else if (KIND == RVF_K_MIX_QUEUE6) tile_occ_queue<true, 6, false, WS>(a, tile, lds);
    else if (KIND == RVF_K_FP64_SPLIT1) tile_occ_queue<false, 1, true, WS>(a, tile, lds);
    else if (KIND == RVF_K_FP64_SPLIT2) tile_occ_queue<false, 2, true, WS>(a, tile, lds);
    else if (KIND == RVF_K_FP64_SPLIT3) tile_occ_queue<false, 3, true, WS>(a, tile, lds);
    else if (KIND == RVF_K_MIX_SPLIT1) tile_occ_queue<true, 1, true, WS>(a, tile, lds);
    else if (KIND == RVF_K_MIX_SPLIT2) tile_occ_queue<true, 2, true, WS>(a, tile, lds);
    else if (KIND == RVF_K_MIX_SPLIT3) tile_occ_queue<true, 3, true, WS>(a, tile, lds);
    else if (KIND == RVF_K_FP64_2D) tile_fp64<8, true>(a, tile);
    else if (KIND == RVF_K_MIX_2D) tile_mixed<8, true, true>(a, tile);
    else if (KIND == RVF_K_FP64_OCC2D) tile_fp64_occ<true>(a, tile, lds);
    else if (KIND == RVF_K_MIX_OCC2D) tile_mixed_occ2d(a, tile);
    else tile_mc(a, tile);
}

// Zeroes the tile claim counter and the eight per-XCD counters that follow it
// (64 bytes apart, from c + 2).
__global__ __launch_bounds__(64) void k_zero_ctr(int* __restrict__ c) {
    if (threadIdx.x < 9) c[threadIdx.x == 0 ? 0 : 2 + (threadIdx.x - 1) * 16] = 0;
}

// kernel tags: k_pair<tag> names the variant in profiles
#define RVF_KIND_TAG(e, t) struct t { static constexpr int kind = RVF_K_##e; };
RVF_PAIR_KINDS(RVF_KIND_TAG)
#undef RVF_KIND_TAG

// Resident waves per SIMD the compiler must allow (1 = no constraint).  The
// mixed split traversal waits on node fetches more than on arithmetic; at
// 128 VGPRs four of its waves share a SIMD instead of three (its few spills
// are in the per-column setup, outside the traversal loop).  Others held
// just past a threshold, measured on the bench workloads: the FP64 queue
// kernel (135 VGPRs, three waves) at 128: 664.5 -> 657.0 ms on the occluded
// cube; Monte Carlo (232, two waves) at 168: 701 -> 653 ms on the cylinder;
// FP64 points-in-registers (214) at 168: 23.7 -> 23.5 ms on the unoccluded
// cube, Duffy pairs 119.0 -> 117.4 ms on the cylinder; curves (FP64 200 at
// 168, mixed 136 at 128): 2.42 -> 2.29 and 0.56 -> 0.52 ms on the circle.
// The mixed chunked-points kernel (196) at 168: 37.1 -> 30.3 ms on the cube
// at nquad 5.  Slower, so left alone: the mixed points-in-registers kernels
// at one more wave (9 points, 148 VGPRs at 128: 4.12 -> 4.25 ms; 16 points,
// 232 at 168: 13.6 -> 15.3 ms at nquad 4).
constexpr int k_pair_waves(int kind) {
    return (kind == RVF_K_MIX_SPLIT3 || kind == RVF_K_MIX_SPLIT2 || kind == RVF_K_FP64_QUEUE1 || kind == RVF_K_MIX_2D)
               ? 4
           : (kind == RVF_K_MC || kind == RVF_K_FP64_PC9 || kind == RVF_K_FP64_2D || kind == RVF_K_MIX_PC16C) ? 3
                                                                                                               : 1;
}

// Tile costs of the wave-queue kernels vary by orders of magnitude with where
// the pairs sit relative to the occluders.  Their launches are persistent: as
// many waves as fit the GPU claim tiles from a counter until none is left, so
// no XCD idles while another still holds a run of expensive tiles (the
// hardware hands a grid's workgroups to the XCDs in a fixed round robin).
// Measured on the occluded 13,824-element cube: 265 -> 250 ms at 16 columns
// per tile, 234 ms at 4 (FP64 730 -> 671 ms).
//
// A whole launch claims its tiles per XCD: the persistent grid's workgroup b
// runs on XCD b mod 8 (the dispatcher's round robin), and XCD x first drains
// the tiles of its own contiguous eighth of the strips (xorder[xseg[x] ..
// xseg[x+1]), from its own counter), then helps the other XCDs in turn.  An
// XCD's waves then share the point data of fewer strips and the tree nodes
// their rays pass in their own L2 (one per XCD).  Measured on the occluded
// cube: 207 -> 194 ms (mixed; the FP64 kernels keep one counter, with which
// they ran 666.1 ms against 668.5 in one run); the same partition with strips
// dealt round robin, 195 ms.  A tile's results do not depend on which wave
// runs it.
constexpr bool k_pair_dynamic(int kind) { return rvf_queue_kind(kind); }

template <class TAG>
__device__ __forceinline__ void pair_body(const PairArgs& a, double* lds) {
    if constexpr (!k_pair_dynamic(TAG::kind)) {
        run_tile<TAG::kind>(a, a.tile0 + (int)blockIdx.x * a.tile_step, lds);
    } else {
        // the wave's statistics: zeroed here, summed over all its tiles,
        // flushed once below
        constexpr int KD = TAG::kind;
        constexpr bool MIXQ = (KD >= RVF_K_MIX_QUEUE2 && KD <= RVF_K_MIX_QUEUE6) || rvf_mixed_split_kind(KD);
        uint64_t* wst = occ_wst<MIXQ, rvf_split_kind(KD)>(a, lds);
        if (threadIdx.x < RVF_QSTATS) wst[threadIdx.x] = 0;
        __syncthreads();
        // every lane takes part in the claim (lane 0 adds 1, the others 0) and
        // the index is read back with readfirstlane, so the loop and its exit
        // are wave-uniform; each claim advances the counter, so every wave
        // reaches w >= ntiles
        if (a.xorder) {
            // every counter only grows, so each XCD's range ends for every wave
            int x = (int)(blockIdx.x & 7u);
            for (int k = 0; k < 8; k++, x = (x + 1) & 7) {
                const int base = a.xseg[x], n = a.xseg[x + 1] - base;
                int* ctr = a.tile_ctr + 2 + x * 16;
                for (;;) {
                    const int w = __builtin_amdgcn_readfirstlane(atomicAdd(ctr, threadIdx.x == 0 ? 1 : 0));
                    if (w >= n) break;
                    run_tile<TAG::kind, true>(a, a.xorder[base + w], lds);
                }
            }
        } else {
            int w = __builtin_amdgcn_readfirstlane(atomicAdd(a.tile_ctr, threadIdx.x == 0 ? 1 : 0));
            while (w < a.ntiles) {
                run_tile<TAG::kind, true>(a, a.tile0 + w * a.tile_step, lds);
                w = __builtin_amdgcn_readfirstlane(atomicAdd(a.tile_ctr, threadIdx.x == 0 ? 1 : 0));
            }
        }
        __syncthreads();
        occ_flush<MIXQ>(a, wst);
    }
}

template <class TAG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(k_pair_waves(TAG::kind)))) void k_pair(PairArgs a) {
    extern __shared__ double lds[];
    pair_body<TAG>(a, lds);
}

// The mixed split kernels also built for three resident waves per SIMD (room
// for every value in registers: no spills) instead of their default four:
// an occupancy the autotuner can choose (rvf_plan_tune_waves).
template <class TAG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_pair_w3(PairArgs a) {
    extern __shared__ double lds[];
    pair_body<TAG>(a, lds);
}

template <class TAG>
__global__ __launch_bounds__(64) void k_pair_batched(const PairArgs* __restrict__ reqs, const int2* __restrict__ work,
                                                     const int* __restrict__ ctl) {
    // ctl[0]: queue head (zeroed before the launch), ctl[1]: number of items.
    // The claim is made by every lane without a lane-dependent branch (lane 0
    // adds 1, the others 0) and read back with readfirstlane, so the item index
    // and the loop exit are wave-uniform: the loop never runs with part of the
    // wave, which the tile code's cross-lane operations require.
    extern __shared__ double lds[];
    int* head = const_cast<int*>(ctl);
    const int nwork = ctl[1];
    for (;;) {
        const int w = __builtin_amdgcn_readfirstlane(atomicAdd(head, threadIdx.x == 0 ? 1 : 0));
        if (w >= nwork) break;
        const int2 it = work[w];
        const PairArgs& a = reqs[it.x];
        run_tile<TAG::kind>(a, it.y, lds);
    }
}

// ---------------------------------------------------------------------------
// Utility kernels: zero fill (graph-capturable replacement for memset) and a
// plain 16-byte-vector copy (reference point for HBM traffic counters).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_zero(uint4* __restrict__ p, int64_t n16, uint32_t* __restrict__ tail,
                                              int64_t ntail) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n16; k += stride) p[k] = make_uint4(0, 0, 0, 0);
    if (i < ntail) tail[i] = 0;
}

__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = i; k < n16; k += stride) dst[k] = src[k];
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side of this translation unit)
// ---------------------------------------------------------------------------
// Persistent launch: as many workgroups as fit the GPU (at least one), the
// claim counters zeroed by a kernel on the same stream (graph capture records
// no memset).
template <class KERN>
static void launch_persistent(KERN kern, const PairArgs* a, size_t lds, hipStream_t st) {
    int dev = 0, ncu = 0, nb = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, 64, lds);
    const int grid = std::max(1, std::min(a->ntiles, std::max(nb, 1) * std::max(ncu, 1)));
    hipLaunchKernelGGL(k_zero_ctr, dim3(1), dim3(64), 0, st, a->tile_ctr);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64), lds, st, *a);
}

template <class TAG>
static void launch_direct(const PairArgs* a, size_t lds, hipStream_t st) {
    if constexpr (k_pair_dynamic(TAG::kind)) {
        if constexpr (rvf_mixed_split_kind(TAG::kind)) {
            if (a->waves == 3) {
                launch_persistent(k_pair_w3<TAG>, a, lds, st);
                return;
            }
        }
        launch_persistent(k_pair<TAG>, a, lds, st);
    } else {
        hipLaunchKernelGGL(k_pair<TAG>, dim3((unsigned)a->ntiles), dim3(64), lds, st, *a);
    }
}
template <class TAG>
static void launch_batched(const PairArgs* reqs, const int2* work, const int* ctl, int grid, size_t lds,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_pair_batched<TAG>, dim3((unsigned)grid), dim3(64), lds, st, reqs, work, ctl);
}
template <class TAG>
static int batched_blocks(size_t lds) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_pair_batched<TAG>, 64, lds) != hipSuccess) return 0;
    return nb;
}

extern "C" {
int rvf_launch_precompute(const PrecompArgs* a, hipStream_t st) {
    int64_t nb = (a->N + 63) / 64;
    if (nb == 0) return 0;
    hipLaunchKernelGGL(k_precompute, dim3((unsigned)nb), dim3(64), 0, st, *a);
    return (int)hipGetLastError();
}

int rvf_pair_kind(int mixed, int occ, int pc, int mc) {
    if (mc) return RVF_K_MC;
    if (!mixed) return occ ? RVF_K_FP64_OCC_LANE : pc <= 4 ? RVF_K_FP64_PC4 : pc <= 9 ? RVF_K_FP64_PC9 : RVF_K_FP64_PC16;
    return occ ? RVF_K_MIX_OCC_LANE
               : pc <= 4 ? RVF_K_MIX_PC4 : pc <= 9 ? RVF_K_MIX_PC9 : pc <= 16 ? RVF_K_MIX_PC16 : RVF_K_MIX_PC16C;
}

#define RVF_KIND_CASE(e, t) case RVF_K_##e: RVF_KIND_CALL(t); break;

int rvf_launch_pair(const PairArgs* a, int kind, size_t lds, hipStream_t st) {
    if (a->ntiles == 0) return 0;
    if (kind < 0 || kind >= RVF_K_COUNT) return (int)hipErrorInvalidValue;
#define RVF_KIND_CALL(t) launch_direct<t>(a, lds, st)
    switch (kind) { RVF_PAIR_KINDS(RVF_KIND_CASE) }
#undef RVF_KIND_CALL
    return (int)hipGetLastError();
}

// Persistent batched pair launch: `grid` workgroups drain the (request, tile)
// items of `work`; ctl = {head, count} in device memory (head zeroed first).
int rvf_launch_pair_batched(int kind, const PairArgs* reqs, const int2* work, const int* ctl, int grid, size_t lds,
                            hipStream_t st) {
    if (grid <= 0) return 0;
    if (kind < 0 || kind >= RVF_K_COUNT) return (int)hipErrorInvalidValue;
#define RVF_KIND_CALL(t) launch_batched<t>(reqs, work, ctl, grid, lds, st)
    switch (kind) { RVF_PAIR_KINDS(RVF_KIND_CASE) }
#undef RVF_KIND_CALL
    return (int)hipGetLastError();
}

// Resident workgroups per CU of the batched kernel of `kind` with `lds` bytes.
int rvf_batched_blocks_per_cu(int kind, size_t lds) {
    if (kind < 0 || kind >= RVF_K_COUNT) return 0;
    int nb = 0;
#define RVF_KIND_CALL(t) nb = batched_blocks<t>(lds)
    switch (kind) { RVF_PAIR_KINDS(RVF_KIND_CASE) }
#undef RVF_KIND_CALL
    return nb;
}

int rvf_launch_mc_area(const PairArgs* a, double* area, double* area_orig, hipStream_t st) {
    if (a->N > 0)
        hipLaunchKernelGGL(k_mc_area, dim3((unsigned)((a->N + 63) / 64)), dim3(64), 0, st, *a, area, area_orig);
    return (int)hipGetLastError();
}

int rvf_launch_duffy(const PairArgs* a, const DuffyPair* pairs, int npairs, const double* gl01, int nq, double* val,
                     const int32_t* rptr, const int32_t* rlist, double* rowsum, double* W, int phase,
                     hipStream_t st) {
    if (npairs == 0) return 0;
    if (phase == 0)
        hipLaunchKernelGGL(k_duffy, dim3((unsigned)npairs), dim3(64), 0, st, *a, pairs, npairs, gl01, nq, val);
    else
        hipLaunchKernelGGL(k_duffy_apply, dim3((unsigned)((a->N + 255) / 256)), dim3(256), 0, st, *a, pairs, val,
                           rptr, rlist, rowsum, W);
    return (int)hipGetLastError();
}

int rvf_launch_zero(void* p, int64_t bytes, hipStream_t st) {
    if (bytes <= 0) return 0;
    int64_t n16 = bytes / 16, ntail = (bytes % 16) / 4;
    unsigned nb = (unsigned)std::min<int64_t>(std::max<int64_t>((n16 + 255) / 256, 1), 4096);
    hipLaunchKernelGGL(k_zero, dim3(nb), dim3(256), 0, st, (uint4*)p, n16, (uint32_t*)((char*)p + n16 * 16), ntail);
    return (int)hipGetLastError();
}

int rvf_launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st) {
    int64_t n16 = bytes / 16;
    if (n16 <= 0) return 0;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, st, (const uint4*)src, (uint4*)dst, n16);
    return (int)hipGetLastError();
}

int rvf_launch_fill(double* F, const double* area, int64_t N, const int2* tiles, int ntiles, int self_vf,
                    hipStream_t st) {
    if (ntiles == 0) return 0;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)ntiles), dim3(256), 0, st, F, area, N, tiles, self_vf);
    return (int)hipGetLastError();
}

int rvf_launch_diag(const double* diag, const double* area, int64_t N, double* rowsum, hipStream_t st) {
    if (N == 0) return 0;
    hipLaunchKernelGGL(k_diag_apply, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, diag, area, N, rowsum);
    return (int)hipGetLastError();
}

int rvf_launch_reduce(const double* part, const double* partT, const Chunk* chunks, int nchunks,
                      const int32_t* chunk_of, int64_t N, int NS, int G, double* rowsum, double* W,
                      double* Tseg, double* Wseg, hipStream_t st) {
    if (N == 0) return 0;
    unsigned nb = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_reduce_seg, dim3(nb, RVF_RSEG), dim3(256), 0, st, part, partT, chunks, nchunks, chunk_of, N,
                       NS, G, Tseg, Wseg);
    hipLaunchKernelGGL(k_reduce_fin, dim3(nb), dim3(256), 0, st, Tseg, Wseg, N, NS, G, rowsum, W);
    return (int)hipGetLastError();
}
}
