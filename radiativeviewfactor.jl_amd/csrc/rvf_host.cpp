// Host side of librvf_hip.so: C ABI, mesh planning, BVH construction and
// solve orchestration.  Compiled with -ffp-contract=off: the quadrature tables
// and the BVH boxes must be bit-identical to the reference's (Float64, no FMA).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rvf.h"
#include "rvf_common.h"

// device launch wrappers (rvf_kernels.hip)

extern "C" int rvf_launch_precompute(const PrecompArgs* a, hipStream_t st);
extern "C" int rvf_pair_kind(int mixed, int occ, int pc, int mc);
extern "C" int rvf_launch_pair(const PairArgs* a, int kind, size_t lds, hipStream_t st);

extern "C" int rvf_launch_mc_area(const PairArgs* a, double* area, double* area_orig, hipStream_t st);
extern "C" int rvf_launch_duffy(const PairArgs* a, const DuffyPair* pairs, int npairs, const double* gl01, int nq,
                                double* val, const int32_t* rptr, const int32_t* rlist, double* rowsum, double* W,
                                int phase, hipStream_t st);
extern "C" int rvf_launch_zero(void* p, int64_t bytes, hipStream_t st);
extern "C" int rvf_launch_copy(const void* src, void* dst, int64_t bytes, hipStream_t st);
extern "C" int rvf_launch_fill(double* F, const double* area, int64_t N, const int2* tiles, int ntiles,
                               int self_vf, hipStream_t st);
extern "C" int rvf_launch_diag(const double* diag, const double* area, int64_t N, double* rowsum, hipStream_t st);
extern "C" int rvf_launch_reduce(const double* part, const double* partT, const Chunk* chunks, int nchunks,
                                 const int32_t* chunk_of, int64_t N, int NS, int G, double* rowsum, double* W,
                                 double* Tseg, double* Wseg, hipStream_t st);
#define RVF_RSEG 16

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return fail(RVF_E_HIP, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

// ---------------------------------------------------------------------------
// device buffer
// ---------------------------------------------------------------------------
struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    bool own = true;  // false: a view of memory someone else owns (never freed here)
    ~DBuf() { release(); }
    void release() {
        if (p && own) (void)hipFree(p);
        p = nullptr; n = 0; own = true;
    }
    int alloc(size_t bytes) {
        release();
        if (bytes == 0) bytes = 16;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) { p = nullptr; return fail(RVF_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)); }
        n = bytes;
        return 0;
    }
    void view(void* q, size_t bytes) {
        release();
        p = q; n = q ? bytes : 0; own = false;
    }
    template <class T> T* as() const { return (T*)p; }
};

// Host <-> device copies ordered on one stream only.  A synchronous hipMemcpy
// also waits for work on the device's other streams (measured in the request
// server: up to 1.5 s behind another request's pair kernel), so every copy is
// an asynchronous one on the context stream followed by a wait on that stream.
static int copy_sync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t st) {
    if (!bytes) return 0;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, kind, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

static int upload(DBuf& b, const void* src, size_t bytes, hipStream_t st) {
    int rc = b.alloc(bytes);
    if (rc) return rc;
    return src ? copy_sync(b.p, src, bytes, hipMemcpyHostToDevice, st) : 0;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct rvf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[16] = {};
};

// ---------------------------------------------------------------------------
// Quadrature tables (src/Quadrature.jl:35-109; src/ViewFactorKernel.jl:42-71),
// evaluated from the same closed forms with correctly rounded sqrt/division.
// ---------------------------------------------------------------------------
static void legendre_nodes(int n, double* p, double* w) {
    // n > 5: the reference uses Golub-Welsch (LAPACK eigen); here Newton on
    // P_n in long double, rounded once.
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
        long double p0 = 1, p1 = x;
        for (int m = 2; m <= n; m++) {
            long double p2 = ((2 * m - 1) * x * p1 - (m - 1) * p0) / m;
            p0 = p1; p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1);
        p[n - 1 - k] = (double)x;
        w[n - 1 - k] = (double)(2.0L / ((1 - x * x) * dp * dp));
    }
}

static int gl1d(int n, double* p, double* w) {
    if (n == 1) { p[0] = 0.0; w[0] = 2.0; return 1; }
    if (n == 2) { p[0] = -1.0 / std::sqrt(3.0); p[1] = 1.0 / std::sqrt(3.0); w[0] = w[1] = 1.0; return 2; }
    if (n == 3) {
        double a = std::sqrt(3.0 / 5.0);
        p[0] = -a; p[1] = 0.0; p[2] = a; w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0;
        return 3;
    }
    if (n == 4) {
        double a = std::sqrt((3.0 + 2.0 * std::sqrt(6.0 / 5.0)) / 7.0);
        double b = std::sqrt((3.0 - 2.0 * std::sqrt(6.0 / 5.0)) / 7.0);
        double wa = (18.0 - std::sqrt(30.0)) / 36.0, wb = (18.0 + std::sqrt(30.0)) / 36.0;
        p[0] = -a; p[1] = -b; p[2] = b; p[3] = a; w[0] = wa; w[1] = wb; w[2] = wb; w[3] = wa;
        return 4;
    }
    if (n == 5) {
        double a = std::sqrt(5.0 + 2.0 * std::sqrt(10.0 / 7.0)) / 3.0;
        double b = std::sqrt(5.0 - 2.0 * std::sqrt(10.0 / 7.0)) / 3.0;
        double wa = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0, wb = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
        p[0] = -a; p[1] = -b; p[2] = 0.0; p[3] = b; p[4] = a;
        w[0] = wa; w[1] = wb; w[2] = 128.0 / 225.0; w[3] = wb; w[4] = wa;
        return 5;
    }
    if (n < 1 || n > 8) return -1;
    legendre_nodes(n, p, w);
    return n;
}

static int quad_rule(int nquad, double* pts, double* wts) {
    double p[8], w[8];
    int n = gl1d(nquad, p, w);
    if (n < 0) return -1;
    int k = 0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) { pts[2 * k] = p[i]; pts[2 * k + 1] = p[j]; wts[k] = w[i] * w[j]; k++; }
    return k;
}

static int tri_rule(int nquad, double* pts, double* wts) {
    if (nquad <= 1) { pts[0] = pts[1] = 1.0 / 3.0; wts[0] = 0.5; return 1; }
    if (nquad == 2) {
        const double x[3] = {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}, y[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
        for (int k = 0; k < 3; k++) { pts[2 * k] = x[k]; pts[2 * k + 1] = y[k]; wts[k] = 0.5 / 3.0; }
        return 3;
    }
    if (nquad == 3) {
        double a1 = 0.101286507323456, b1 = 1.0 - 2.0 * a1, a2 = 0.470142064105115, b2 = 1.0 - 2.0 * a2;
        const double x[7] = {a1, b1, a1, a2, b2, a2, 1.0 / 3.0}, y[7] = {a1, a1, b1, a2, a2, b2, 1.0 / 3.0};
        double w1 = 0.125939180544827 / 2.0, w2 = 0.132394440720100 / 2.0, w3 = 0.225 / 2.0;
        const double w[7] = {w1, w1, w1, w2, w2, w2, w3};
        for (int k = 0; k < 7; k++) { pts[2 * k] = x[k]; pts[2 * k + 1] = y[k]; wts[k] = w[k]; }
        return 7;
    }
    double cen = 1.0 / 3.0;
    double a1 = 0.260345966079038, b1 = 1.0 - 2.0 * a1, a2 = 0.065130102902216, b2 = 1.0 - 2.0 * a2;
    double a3 = 0.048690315425316, b3 = 0.312865496004875, c3 = (1.0 - a3) - b3;
    const double x[13] = {a1, b1, a1, a2, b2, a2, a3, b3, a3, c3, b3, c3, cen};
    const double y[13] = {a1, a1, b1, a2, a2, b2, b3, a3, c3, a3, c3, b3, cen};
    double wc = -0.149570044467670 / 2.0, wA = 0.175615257433204 / 2.0, wB = 0.053347235608839 / 2.0,
           wC = 0.077113760890257 / 2.0;
    const double w[13] = {wA, wA, wA, wB, wB, wB, wC, wC, wC, wC, wC, wC, wc};
    for (int k = 0; k < 13; k++) { pts[2 * k] = x[k]; pts[2 * k + 1] = y[k]; wts[k] = w[k]; }
    return 13;
}

// ---------------------------------------------------------------------------
// mesh
// ---------------------------------------------------------------------------
// Slot layout of a solve: slot k of the pair tiling holds original element
// slot_orig[k].  Identity for unoccluded solves; occluded quadrature solves
// order elements by (group, Morton code of the centroid) so that a strip's 64
// lanes cover a compact patch and their rays form coherent bundles.
struct Layout {
    bool built = false;
    std::vector<int32_t> slot_orig, orig_slot, slot_grp, chunk_of;
    std::vector<Chunk> chunks;
    std::vector<int2> tiles;
    std::vector<int32_t> xorder;  // tile indices grouped by XCD (contiguous eighths of the strips)
    int xseg[9] = {0};
    DBuf d_slot_orig, d_slot_grp, d_chunks, d_chunk_of, d_tiles, d_xorder;
};

struct rvf_mesh {
    rvf_ctx* ctx = nullptr;
    int64_t N = 0, Nn = 0;
    int NS = 0;
    std::vector<int8_t> fam;
    std::vector<int32_t> nodes8;   // N*8 1-based
    std::vector<int32_t> grp_rank; // per element
    std::vector<int32_t> tags;     // sorted group tags
    std::vector<double> hcoords;
    std::vector<int2> fill_tiles;
    double Sm = 1.0;               // max |coordinate|
    // layouts by key: bit 0 spatial order, bit 1 tiles for the pairs (i, i)
    // too, bits 2.. columns per tile
    mutable std::map<int, Layout> lay;
    mutable std::mutex lay_mu;     // plans on one mesh may be created from several threads
    bool dim2 = false;             // curve mesh (Line3 / Line2, mesh_dim 1)
    DBuf coords, d_fam, d_nodes8, d_grp, d_fill_tiles;
};

static inline uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Host part of a layout: slot order, group-pure column chunks of at most `cw`
// columns and the tile list (chunk-major: each chunk's strips consecutively).
// Fails, before the tile list is built, when it would hold 2^31 tiles or more
// (tile indices are 32-bit; about N^2 / (128 cw) tiles).
static int layout_host(const rvf_mesh* m, bool spatial, bool self, int cw, Layout& L) {
    int64_t N = m->N, NS = m->NS;
    std::vector<int32_t> order(N);
    for (int64_t e = 0; e < N; e++) order[e] = (int32_t)e;
    if (spatial) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int64_t k = 0; k < m->Nn; k++)
            for (int c = 0; c < 3; c++) { lo[c] = std::min(lo[c], m->hcoords[3 * k + c]); hi[c] = std::max(hi[c], m->hcoords[3 * k + c]); }
        std::vector<uint64_t> key(N);
        for (int64_t e = 0; e < N; e++) {
            int nc = m->fam[e] >= 4 ? 2 : (m->fam[e] == 0 || m->fam[e] == 2) ? 4 : 3;
            uint32_t q[3];
            for (int c = 0; c < 3; c++) {
                double x = 0;
                for (int a = 0; a < nc; a++) x += m->hcoords[3 * (int64_t)(m->nodes8[8 * e + a] - 1) + c];
                x /= nc;
                double u = hi[c] > lo[c] ? (x - lo[c]) / (hi[c] - lo[c]) : 0.0;
                q[c] = (uint32_t)std::min(1023.0, std::max(0.0, u * 1024.0));
            }
            uint64_t morton = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
            key[e] = ((uint64_t)m->grp_rank[e] << 32) | morton;
        }
        std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    }
    L.slot_orig.assign(NS, -1);
    L.orig_slot.assign(N, 0);
    L.slot_grp.assign(NS, 0);
    for (int64_t k = 0; k < N; k++) {
        L.slot_orig[k] = order[k];
        L.orig_slot[order[k]] = (int32_t)k;
        L.slot_grp[k] = m->grp_rank[order[k]];
    }
    // group-pure column chunks of <= cw slots
    L.chunk_of.assign(NS, 0);
    L.chunks.clear();
    for (int64_t e = 0; e < N;) {
        int64_t b = e;
        int g = L.slot_grp[e];
        while (e < N && e - b < cw && L.slot_grp[e] == g) e++;
        Chunk c; c.begin = (int32_t)b; c.end = (int32_t)e; c.grp = g; c.first_tile = 0;
        for (int64_t k = b; k < e; k++) L.chunk_of[k] = (int32_t)L.chunks.size();
        L.chunks.push_back(c);
    }
    // tiles (strip s, chunk c) exist iff 64 s < last(c); with self pairs
    // iff 64 s <= last(c)
    int64_t ntiles = 0;
    for (const Chunk& c : L.chunks) ntiles += (c.end - 1 + (self ? 1 : 0) + RVF_STRIP - 1) / RVF_STRIP;
    if (ntiles >= ((int64_t)1 << 31))
        return fail(RVF_E_ARG, "mesh too large for one plan: " + std::to_string(ntiles) +
                                   " pair tiles (at most 2^31 - 1; about N^2 / (128 x columns per tile))");
    L.tiles.clear();
    L.tiles.reserve((size_t)ntiles);
    for (size_t c = 0; c < L.chunks.size(); c++) {
        L.chunks[c].first_tile = (int32_t)L.tiles.size();
        int64_t last = L.chunks[c].end - 1 + (self ? 1 : 0);
        for (int64_t s = 0; s * RVF_STRIP < last; s++) L.tiles.push_back(make_int2((int)s, (int)c));
    }
    // the wave-queue kernels' claim order: XCD x takes the tiles of strips
    // [x nstrips / 8, (x + 1) nstrips / 8), in tile-list order
    const int64_t nst = std::max<int64_t>(1, NS / RVF_STRIP);
    std::vector<int64_t> per(9, 0);
    for (const int2& t : L.tiles) per[std::min<int64_t>(7, (int64_t)t.x * 8 / nst) + 1]++;
    for (int x = 0; x < 8; x++) per[x + 1] += per[x];
    for (int x = 0; x < 9; x++) L.xseg[x] = (int)per[x];
    L.xorder.assign(L.tiles.size(), 0);
    for (size_t t = 0; t < L.tiles.size(); t++)
        L.xorder[per[std::min<int64_t>(7, (int64_t)L.tiles[t].x * 8 / nst)]++] = (int32_t)t;
    return 0;
}

// Layouts are cached per mesh and (spatial order, self-pair tiles, columns).
static int get_layout(const rvf_mesh* m, bool spatial, bool self, int cw, const Layout** out) {
    std::lock_guard<std::mutex> lk(m->lay_mu);
    Layout& L = m->lay[(spatial ? 1 : 0) | (self ? 2 : 0) | (cw << 2)];
    if (!L.built) {
        const int64_t NS = m->NS;
        int rc = layout_host(m, spatial, self, cw, L);
        if (rc) return rc;
        const hipStream_t st = m->ctx->stream;
        if ((rc = upload(L.d_slot_orig, L.slot_orig.data(), sizeof(int32_t) * NS, st)) ||
            (rc = upload(L.d_slot_grp, L.slot_grp.data(), sizeof(int32_t) * NS, st)) ||
            (rc = upload(L.d_chunks, L.chunks.data(), sizeof(Chunk) * L.chunks.size(), st)) ||
            (rc = upload(L.d_chunk_of, L.chunk_of.data(), sizeof(int32_t) * NS, st)) ||
            (rc = upload(L.d_tiles, L.tiles.data(), sizeof(int2) * L.tiles.size(), st)) ||
            (rc = upload(L.d_xorder, L.xorder.data(), sizeof(int32_t) * L.xorder.size(), st)))
            return rc;
        L.built = true;
    }
    *out = &L;
    return 0;
}

// Families 0..3 (surface meshes) or 4..5 (curve meshes), never both: nodes[f]
// holds NNF[f] x n_per_family[f] node ids for family f.
static int mesh_upload(rvf_ctx* ctx, const double* coords, int64_t n_nodes, int64_t n_elems, const int8_t* family,
                       const int32_t* elem_node_idx, const int32_t* elem_group, const int32_t* const nodes[6],
                       const int64_t n_per_family[6], bool curve, rvf_mesh** out) {
    if (!ctx || !out || n_elems < 0 || n_nodes < 0) return fail(RVF_E_ARG, "rvf_mesh_upload: bad arguments");
    if (n_elems > ((int64_t)1 << 30)) return fail(RVF_E_ARG, "rvf_mesh_upload: too many elements");
    HIPCHK(hipSetDevice(ctx->device));
    const int NNF[6] = {8, 6, 4, 3, 3, 2};
    rvf_mesh* m = new rvf_mesh();
    m->ctx = ctx; m->N = n_elems; m->Nn = n_nodes; m->dim2 = curve;
    m->fam.assign(family, family + n_elems);
    m->nodes8.assign(8 * n_elems, 0);
    double Sm = 0.0;
    for (int64_t k = 0; k < 3 * n_nodes; k++) {
        if (!std::isfinite(coords[k])) { delete m; return fail(RVF_E_ARG, "rvf_mesh_upload: non-finite coordinate"); }
        Sm = std::max(Sm, std::fabs(coords[k]));
    }
    m->Sm = Sm > 0 ? Sm : 1.0;
    for (int64_t e = 0; e < n_elems; e++) {
        int f = family[e];
        if (curve ? (f < 4 || f > 5) : (f < 0 || f > 3)) {
            delete m;
            return fail(RVF_E_ARG, std::string(curve ? "curve mesh: element family must be Line3 (4) or Line2 (5), got "
                                                     : "GPU backend does not support element family code ") +
                                       std::to_string(f));
        }
        int64_t col = (int64_t)elem_node_idx[e] - 1;
        if (col < 0 || col >= n_per_family[f] || !nodes[f]) { delete m; return fail(RVF_E_ARG, "rvf_mesh_upload: elem_node_idx out of range"); }
        for (int a = 0; a < NNF[f]; a++) {
            int32_t nd = nodes[f][col * NNF[f] + a];
            if (nd < 1 || nd > n_nodes) { delete m; return fail(RVF_E_ARG, "rvf_mesh_upload: node index out of range"); }
            m->nodes8[8 * e + a] = nd;
        }
    }
    // groups: sorted distinct tags (Results.jl:59) -> ranks
    std::vector<int32_t> t(elem_group, elem_group + n_elems);
    std::sort(t.begin(), t.end());
    t.erase(std::unique(t.begin(), t.end()), t.end());
    m->tags = t;
    m->grp_rank.resize(n_elems);
    for (int64_t e = 0; e < n_elems; e++)
        m->grp_rank[e] = (int32_t)(std::lower_bound(t.begin(), t.end(), elem_group[e]) - t.begin());
    m->NS = (int)(((n_elems + RVF_STRIP - 1) / RVF_STRIP) * RVF_STRIP);
    m->hcoords.assign(coords, coords + 3 * n_nodes);
    int64_t nb = m->NS / RVF_STRIP;
    for (int64_t bj = 0; bj < nb; bj++)
        for (int64_t bi = 0; bi <= bj; bi++) m->fill_tiles.push_back(make_int2((int)bi, (int)bj));
    int rc;
    const hipStream_t st = ctx->stream;
    if ((rc = upload(m->coords, coords, sizeof(double) * 3 * n_nodes, st)) ||
        (rc = upload(m->d_fam, m->fam.data(), n_elems, st)) ||
        (rc = upload(m->d_nodes8, m->nodes8.data(), sizeof(int32_t) * 8 * n_elems, st)) ||
        (rc = upload(m->d_grp, m->grp_rank.data(), sizeof(int32_t) * n_elems, st)) ||
        (rc = upload(m->d_fill_tiles, m->fill_tiles.data(), sizeof(int2) * m->fill_tiles.size(), st))) {
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

int rvf_mesh_upload(rvf_ctx* ctx, const double* coords, int64_t n_nodes, int64_t n_elems, const int8_t* family,
                    const int32_t* elem_node_idx, const int32_t* elem_group, const int32_t* const nodes[4],
                    const int64_t n_per_family[4], rvf_mesh** out) {
    if (!nodes || !n_per_family) return fail(RVF_E_ARG, "rvf_mesh_upload: bad arguments");
    const int32_t* nd[6] = {nodes[0], nodes[1], nodes[2], nodes[3], nullptr, nullptr};
    const int64_t cnt[6] = {n_per_family[0], n_per_family[1], n_per_family[2], n_per_family[3], 0, 0};
    return mesh_upload(ctx, coords, n_nodes, n_elems, family, elem_node_idx, elem_group, nd, cnt, false, out);
}

int rvf_curve_mesh_upload(rvf_ctx* ctx, const double* coords, int64_t n_nodes, int64_t n_elems, const int8_t* family,
                          const int32_t* elem_node_idx, const int32_t* elem_group, const int32_t* line3_nodes,
                          int64_t n_line3, const int32_t* line2_nodes, int64_t n_line2, rvf_mesh** out) {
    const int32_t* nd[6] = {nullptr, nullptr, nullptr, nullptr, line3_nodes, line2_nodes};
    const int64_t cnt[6] = {0, 0, 0, 0, n_line3, n_line2};
    return mesh_upload(ctx, coords, n_nodes, n_elems, family, elem_node_idx, elem_group, nd, cnt, true, out);
}

void rvf_mesh_free(rvf_mesh* m) { delete m; }

int rvf_mesh_info(const rvf_mesh* m, int64_t* n, int32_t* G, int32_t* tags) {
    if (!m) return fail(RVF_E_ARG, "null mesh");
    if (n) *n = m->N;
    if (G) *G = (int32_t)m->tags.size();
    if (tags) for (size_t k = 0; k < m->tags.size(); k++) tags[k] = m->tags[k];
    return 0;
}

// ---------------------------------------------------------------------------
// BVH: build_bvh / _build_recursive! (src/BVH.jl:71-148), bit-identical.
// ---------------------------------------------------------------------------
static inline double jmin(double x, double y) {  // Julia Base.min for floats
    double d = x - y;
    double r = std::signbit(d) ? x : y;
    return (std::isnan(x) || std::isnan(y)) ? d : r;
}
static inline double jmax(double x, double y) {
    double d = x - y;
    double r = std::signbit(d) ? y : x;
    return (std::isnan(x) || std::isnan(y)) ? d : r;
}
static inline bool jisless(double a, double b) {
    if (std::isnan(a)) return false;
    if (std::isnan(b)) return true;
    if (a < b) return true;
    if (a > b) return false;
    return std::signbit(a) && !std::signbit(b);
}

struct HostNode { double lo[3], hi[3]; int left, right, start, count; };  // 1-based like the reference

struct Builder {
    const double* soup;   // merged (3,3,T), or a segment soup (3,2,T) with nv = 2
    int av, ac;           // AABB read strides (vertex, coord)
    int nv = 3;           // vertices per item
    std::vector<int> idx; // 1-based permutation
    std::vector<HostNode> nodes;
    std::vector<double> key;
    double rd(int64_t t, int v, int c) const { return soup[3 * nv * t + v * av + c * ac]; }
    void tri_box(int64_t t, double* lo, double* hi) const {
        for (int c = 0; c < 3; c++) lo[c] = hi[c] = rd(t, 0, c);
        for (int v = 1; v < nv; v++)
            for (int c = 0; c < 3; c++) { double x = rd(t, v, c); lo[c] = jmin(lo[c], x); hi[c] = jmax(hi[c], x); }
    }
    int build(int start, int stop) {  // 1-based inclusive
        double lo[3], hi[3];
        tri_box(idx[start - 1] - 1, lo, hi);
        for (int i = start + 1; i <= stop; i++) {
            double l2[3], h2[3];
            tri_box(idx[i - 1] - 1, l2, h2);
            for (int c = 0; c < 3; c++) { lo[c] = jmin(lo[c], l2[c]); hi[c] = jmax(hi[c], h2[c]); }
        }
        int count = stop - start + 1;
        HostNode nd;
        std::memcpy(nd.lo, lo, sizeof lo); std::memcpy(nd.hi, hi, sizeof hi);
        if (count <= 8) {
            nd.left = nd.right = 0; nd.start = start; nd.count = count;
            nodes.push_back(nd);
            return (int)nodes.size();
        }
        double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
        int axis = 0;
        for (int c = 1; c < 3; c++) if (jisless(ext[axis], ext[c])) axis = c;
        int mid = (start + stop) / 2;
        // stable sort by centroid: ((0 + x1) + x2) + x3, / nv (BVH.jl:100-109)
        std::vector<std::pair<double, int>> kv(count);
        for (int i = start; i <= stop; i++) {
            int64_t t = idx[i - 1] - 1;
            double s = 0.0;
            for (int v = 0; v < nv; v++) s += rd(t, v, axis);
            kv[i - start] = std::make_pair(s / (double)nv, idx[i - 1]);
        }
        std::stable_sort(kv.begin(), kv.end(),
                         [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                             return jisless(a.first, b.first);
                         });
        for (int i = start; i <= stop; i++) idx[i - 1] = kv[i - start].second;
        nd.left = nd.right = nd.start = nd.count = 0;
        nodes.push_back(nd);
        int me = (int)nodes.size();
        int l = build(start, mid);
        int r = build(mid + 1, stop);
        nodes[me - 1].left = l; nodes[me - 1].right = r;
        return me;
    }
};

struct Tree {
    std::vector<int> groups;          // active set (sorted)
    std::vector<double> soup;         // merged soup
    std::vector<int> idx;             // 1-based permutation
    std::vector<HostNode> nodes;
    std::vector<int> miss;            // 1-based, 0 = end
    int64_t node_base = 0, tri_base = 0;
};

static void assign_miss(Tree& t, int node, int next) {  // GPUBVH.jl:69-82
    t.miss[node - 1] = next;
    const HostNode& nd = t.nodes[node - 1];
    if (nd.left != 0) {
        int l = nd.left, r = nd.right;
        assign_miss(t, l, r);
        assign_miss(t, r, next);
    }
}

// Traversal tree of the mixed split kernels (no reference counterpart: the
// mixed mode's decisions come from exact FP64 triangle tests, so its FP32
// traversal may use any tree whose padded boxes bound the triangles).  Binned
// surface-area heuristic over the triangles as Moller-Trumbore reads them,
// leaves of at most 4: a triangle task costs several node steps, so small
// leaves pay.  Preorder layout with miss links, like the reference trees.
struct SahBuilder {
    static constexpr int kMaxLeaf = RVF_TLEAF_MAX, kBins = 16;
    static constexpr double kStep = 1.0, kTri = 3.0;  // relative costs
    std::vector<double> lo, hi, cen;  // [3*t]
    std::vector<int> idx;             // local triangle order
    std::vector<HostNode> nodes;      // 1-based links; leaves: start (1-based), count
    static double area(const double* l, const double* h) {
        const double x = h[0] - l[0], y = h[1] - l[1], z = h[2] - l[2];
        return (x < 0 || y < 0 || z < 0) ? 0.0 : 2.0 * (x * y + y * z + z * x);
    }
    int build(int s, int e, int depth) {  // idx[s, e)
        double blo[3], bhi[3], clo[3], chi[3];
        for (int c = 0; c < 3; c++) { blo[c] = clo[c] = INFINITY; bhi[c] = chi[c] = -INFINITY; }
        for (int k = s; k < e; k++) {
            const int t = idx[k];
            for (int c = 0; c < 3; c++) {
                blo[c] = std::min(blo[c], lo[3 * t + c]); bhi[c] = std::max(bhi[c], hi[3 * t + c]);
                clo[c] = std::min(clo[c], cen[3 * t + c]); chi[c] = std::max(chi[c], cen[3 * t + c]);
            }
        }
        HostNode nd;
        std::memcpy(nd.lo, blo, sizeof blo); std::memcpy(nd.hi, bhi, sizeof bhi);
        nd.left = nd.right = 0; nd.start = s + 1; nd.count = e - s;
        nodes.push_back(nd);
        const int me = (int)nodes.size();
        const int n = e - s;
        if (n <= 1) return me;
        int bax = -1, bbin = -1;
        double best = INFINITY;
        if (depth < 40) {  // deeper: object median, which bounds the depth
            const double An = std::max(area(blo, bhi), 1e-300);
            for (int ax = 0; ax < 3; ax++) {
                const double ext = chi[ax] - clo[ax];
                if (!(ext > 0)) continue;
                int cnt[kBins] = {0};
                double bl[kBins][3], bh[kBins][3];
                for (int b = 0; b < kBins; b++)
                    for (int c = 0; c < 3; c++) { bl[b][c] = INFINITY; bh[b][c] = -INFINITY; }
                for (int k = s; k < e; k++) {
                    const int t = idx[k];
                    const int b = std::min(kBins - 1, (int)((cen[3 * t + ax] - clo[ax]) / ext * kBins));
                    cnt[b]++;
                    for (int c = 0; c < 3; c++) {
                        bl[b][c] = std::min(bl[b][c], lo[3 * t + c]); bh[b][c] = std::max(bh[b][c], hi[3 * t + c]);
                    }
                }
                double rA[kBins];
                int rN[kBins];
                double al[3] = {INFINITY, INFINITY, INFINITY}, ah[3] = {-INFINITY, -INFINITY, -INFINITY};
                int acc = 0;
                for (int b = kBins - 1; b > 0; b--) {
                    acc += cnt[b];
                    for (int c = 0; c < 3; c++) { al[c] = std::min(al[c], bl[b][c]); ah[c] = std::max(ah[c], bh[b][c]); }
                    rA[b] = area(al, ah); rN[b] = acc;
                }
                for (int c = 0; c < 3; c++) { al[c] = INFINITY; ah[c] = -INFINITY; }
                acc = 0;
                for (int b = 1; b < kBins; b++) {  // split before bin b
                    acc += cnt[b - 1];
                    for (int c = 0; c < 3; c++) {
                        al[c] = std::min(al[c], bl[b - 1][c]); ah[c] = std::max(ah[c], bh[b - 1][c]);
                    }
                    if (acc == 0 || rN[b] == 0) continue;
                    const double cost = kStep + kTri * (area(al, ah) * acc + rA[b] * rN[b]) / An;
                    if (cost < best) { best = cost; bax = ax; bbin = b; }
                }
            }
            if (n <= kMaxLeaf && !(best < kTri * n)) return me;  // a leaf is cheaper
        }
        int mid;
        if (bax >= 0) {
            const double ext = chi[bax] - clo[bax];
            auto it = std::partition(idx.begin() + s, idx.begin() + e, [&](int t) {
                return std::min(kBins - 1, (int)((cen[3 * t + bax] - clo[bax]) / ext * kBins)) < bbin;
            });
            mid = (int)(it - idx.begin());
        } else {
            int ax = 0;
            for (int c = 1; c < 3; c++) if (chi[c] - clo[c] > chi[ax] - clo[ax]) ax = c;
            mid = s + n / 2;
            std::nth_element(idx.begin() + s, idx.begin() + mid, idx.begin() + e,
                             [&](int x, int y) { return cen[3 * x + ax] < cen[3 * y + ax]; });
        }
        if (mid <= s || mid >= e) mid = s + n / 2;
        nodes[me - 1].start = 0; nodes[me - 1].count = 0;
        const int l = build(s, mid, depth + 1);
        const int r = build(mid, e, depth + 1);
        nodes[me - 1].left = l; nodes[me - 1].right = r;
        return me;
    }
};

static void assign_miss_nodes(const std::vector<HostNode>& nodes, std::vector<int>& miss, int node, int next) {
    miss[node - 1] = next;
    const HostNode& nd = nodes[node - 1];
    if (nd.left != 0) {
        assign_miss_nodes(nodes, miss, nd.left, nd.right);
        assign_miss_nodes(nodes, miss, nd.right, next);
    }
}

struct rvf_obs {
    rvf_ctx* ctx = nullptr;
    int G = 0;
    int soup_read = 0;
    bool seg = false;            // segment soup of a curve mesh
    std::vector<Tree> trees;
    std::vector<int32_t> root;   // G*G absolute node index or -1
    std::vector<int32_t> troot;  // G*G traversal-tree root or -1
    int64_t n_nodes = 0, n_tris = 0, n_tnodes = 0;
    float Sm = 1.0f;
    DBuf d_root, d_n64, d_t64, d_n32, d_t32, d_parent, d_items32, d_items64;
    DBuf d_troot, d_tn32, d_tt32, d_tmap;
};

static float f32_down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}
static float f32_up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

int rvf_obs_build(rvf_ctx* ctx, const rvf_mesh* mesh, const double* soup, const int32_t* tri_group, int64_t T,
                  const int32_t* obs_groups, int32_t n_obs, int32_t soup_read, rvf_obs** out) {
    if (!ctx || !mesh || !out || T < 0 || n_obs < 0 || soup_read < 0 || soup_read > 3)
        return fail(RVF_E_ARG, "rvf_obs_build: bad arguments");
    // curve meshes are obstructed by segment soups (3,2,T) (MeshIO.jl:641-645), surfaces by triangles
    const bool seg = soup_read == RVF_SOUP_SEGMENTS;
    if (seg != mesh->dim2)
        return fail(RVF_E_ARG, seg ? "rvf_obs_build: segment soups obstruct curve meshes only"
                                   : "rvf_obs_build: a curve mesh needs a segment soup (RVF_SOUP_SEGMENTS)");
    const int nv = seg ? 2 : 3, SW = 3 * nv;  // vertices and doubles per item
    HIPCHK(hipSetDevice(ctx->device));
    for (int64_t k = 0; k < SW * T; k++)
        if (!std::isfinite(soup[k])) return fail(RVF_E_ARG, "rvf_obs_build: non-finite soup coordinate");
    rvf_obs* o = new rvf_obs();
    o->ctx = ctx; o->soup_read = soup_read; o->seg = seg;
    int G = (int)mesh->tags.size();
    o->G = G;
    o->root.assign((size_t)G * G, -1);
    // strides: coord c of vertex v at v*sv + c*sc
    int aav, aac, mtv, mtc;
    if (soup_read == RVF_SOUP_REFERENCE) { aac = 1; aav = 3; mtv = 1; mtc = 3; }
    else if (soup_read == RVF_SOUP_XYZ_VERTEX || seg) { aac = 1; aav = 3; mtc = 1; mtv = 3; }
    else { aav = 1; aac = 3; mtv = 1; mtc = 3; }
    double Sm = mesh->Sm;
    for (int64_t k = 0; k < SW * T; k++) Sm = std::max(Sm, std::fabs(soup[k]));
    o->Sm = (float)(Sm * (1 + 1e-6));
    std::map<std::vector<int>, int> seen;
    std::vector<int> act;
    for (int a = 0; a < G; a++)
        for (int b = 0; b < G; b++) {
            act.clear();
            for (int k = 0; k < n_obs; k++)
                if (obs_groups[k] != mesh->tags[a] && obs_groups[k] != mesh->tags[b]) act.push_back(obs_groups[k]);
            std::sort(act.begin(), act.end());
            if (act.empty()) continue;
            auto it = seen.find(act);
            if (it != seen.end()) { o->root[a * G + b] = it->second; continue; }
            Tree tr;
            tr.groups = act;
            for (int g : act)
                for (int64_t t = 0; t < T; t++)
                    if (tri_group[t] == g) tr.soup.insert(tr.soup.end(), soup + SW * t, soup + SW * t + SW);
            int64_t nt = (int64_t)tr.soup.size() / SW;
            if (nt == 0) { seen[act] = -1; continue; }
            if (nt >= ((int64_t)1 << 27)) { delete o; return fail(RVF_E_ARG, "rvf_obs_build: too many triangles"); }
            Builder bd;
            bd.soup = tr.soup.data(); bd.av = aav; bd.ac = aac; bd.nv = nv;
            bd.idx.resize(nt);
            for (int64_t t = 0; t < nt; t++) bd.idx[t] = (int)(t + 1);
            bd.build(1, (int)nt);
            tr.idx = bd.idx; tr.nodes = bd.nodes;
            tr.miss.assign(tr.nodes.size(), 0);
            assign_miss(tr, 1, 0);
            {   // the wave-cooperative traversal keeps its stack in one entry per lane
                std::vector<int> depth(tr.nodes.size(), 0);
                int dmax = 0;
                for (size_t k = 0; k < tr.nodes.size(); k++)
                    if (tr.nodes[k].left) {
                        depth[tr.nodes[k].left - 1] = depth[k] + 1;
                        depth[tr.nodes[k].right - 1] = depth[k] + 1;
                        dmax = std::max(dmax, depth[k] + 1);
                    }
                if (dmax + 2 > 64) { delete o; return fail(RVF_E_ARG, "rvf_obs_build: BVH deeper than 62 levels"); }
            }
            tr.node_base = o->n_nodes; tr.tri_base = o->n_tris;
            o->n_nodes += (int64_t)tr.nodes.size();
            o->n_tris += nt;
            int id = (int)o->trees.size();
            o->trees.push_back(std::move(tr));
            seen[act] = (int)o->trees[id].node_base;
            o->root[a * G + b] = (int)o->trees[id].node_base;
        }
    // flatten
    std::vector<BvhNode64> n64(std::max<int64_t>(o->n_nodes, 1));
    std::vector<BvhNode32> n32(std::max<int64_t>(o->n_nodes, 1));
    std::vector<Tri64> t64(std::max<int64_t>(o->n_tris, 1));
    std::vector<Tri32> t32(std::max<int64_t>(o->n_tris, 1));
    std::vector<int32_t> parent(std::max<int64_t>(o->n_nodes, 1), -1);
    if (o->n_nodes >= ((int64_t)1 << 27)) { delete o; return fail(RVF_E_ARG, "rvf_obs_build: too many nodes"); }
    const double u = 5.9604644775390625e-08;
    const float pad = (float)(128.0 * u * Sm);
    for (const Tree& tr : o->trees) {
        for (size_t k = 0; k < tr.nodes.size(); k++) {
            const HostNode& h = tr.nodes[k];
            BvhNode64& d = n64[tr.node_base + k];
            for (int c = 0; c < 3; c++) { d.lo[c] = h.lo[c]; d.hi[c] = h.hi[c]; }
            d.miss = tr.miss[k] ? (int32_t)(tr.node_base + tr.miss[k] - 1) : -1;
            d.start = h.count ? (int32_t)(tr.tri_base + h.start - 1) : 0;
            d.count = h.count;
            d.pad = h.left ? (int32_t)(tr.node_base + h.right - 1) : 0;  // right child
            BvhNode32& f = n32[tr.node_base + k];
            for (int c = 0; c < 3; c++) { f.lo[c] = f32_down(h.lo[c]) - pad; f.hi[c] = f32_up(h.hi[c]) + pad; }
            f.miss = d.miss;
            f.start_count = h.count ? ((d.start << 4) | d.count) : (d.pad << 4);
            if (h.left) {
                parent[tr.node_base + h.left - 1] = (int32_t)(tr.node_base + k);
                parent[tr.node_base + h.right - 1] = (int32_t)(tr.node_base + k);
            }
            if (h.left && h.left != (int)k + 2) { delete o; return fail(RVF_E_STATE, "bvh: non-preorder layout"); }
        }
        int64_t nt = (int64_t)tr.idx.size();
        for (int64_t k = 0; k < nt; k++) {
            int64_t t = tr.idx[k] - 1;
            const double* s = tr.soup.data() + SW * t;
            double v[3][3];
            for (int vv = 0; vv < 3; vv++)  // a segment (v0, v1) is stored as v0, e1 = v1 - v0, e2 = 0
                for (int c = 0; c < 3; c++) v[vv][c] = vv < nv ? s[vv * mtv + c * mtc] : s[c * mtc];
            Tri64& d = t64[tr.tri_base + k];
            for (int c = 0; c < 3; c++) {
                d.v0[c] = v[0][c]; d.e1[c] = v[1][c] - v[0][c]; d.e2[c] = v[2][c] - v[0][c];
            }
            d.pad = 0;
            Tri32& f = t32[tr.tri_base + k];
            double E1 = 0, E2 = 0;
            for (int c = 0; c < 3; c++) {
                f.v0[c] = (float)d.v0[c]; f.e1[c] = (float)d.e1[c]; f.e2[c] = (float)d.e2[c];
                E1 += std::fabs(d.e1[c]); E2 += std::fabs(d.e2[c]);
            }
            f.sa = (float)(E1 * E2 * (1 + 1e-6));
            f.E1 = f32_up(E1 * (1 + 1e-6));
            f.E2 = f32_up(E2 * (1 + 1e-6));
        }
    }
    // traversal trees of the mixed split kernels, one per reference tree
    std::vector<TNode32> tn32;
    std::vector<Tri32> tt32(std::max<int64_t>(o->n_tris, 1));
    std::vector<int32_t> tmap(std::max<int64_t>(o->n_tris, 1), 0);
    std::map<int64_t, int32_t> troot_of;  // reference root -> traversal root
    for (const Tree& tr : o->trees) {
        const int64_t nt = (int64_t)tr.idx.size();
        SahBuilder sb;
        sb.lo.resize(3 * nt); sb.hi.resize(3 * nt); sb.cen.resize(3 * nt);
        for (int64_t k = 0; k < nt; k++) {
            const Tri64& d = t64[tr.tri_base + k];
            for (int c = 0; c < 3; c++) {
                const double a = d.v0[c], b = d.v0[c] + d.e1[c], e = d.v0[c] + d.e2[c];
                sb.lo[3 * k + c] = std::min(a, std::min(b, e));
                sb.hi[3 * k + c] = std::max(a, std::max(b, e));
                sb.cen[3 * k + c] = 0.5 * (sb.lo[3 * k + c] + sb.hi[3 * k + c]);
            }
        }
        sb.idx.resize(nt);
        for (int64_t k = 0; k < nt; k++) sb.idx[k] = (int)k;
        sb.build(0, (int)nt, 0);
        std::vector<int> miss(sb.nodes.size(), 0);
        assign_miss_nodes(sb.nodes, miss, 1, 0);
        const int64_t base = (int64_t)tn32.size();
        if (base + (int64_t)sb.nodes.size() >= ((int64_t)1 << 27)) { delete o; return fail(RVF_E_ARG, "rvf_obs_build: too many nodes"); }
        for (size_t k = 0; k < sb.nodes.size(); k++) {
            const HostNode& h = sb.nodes[k];
            TNode32 f;
            for (int c = 0; c < 3; c++) { f.lo[c] = f32_down(h.lo[c]) - pad; f.hi[c] = f32_up(h.hi[c]) + pad; }
            f.exit = miss[k] ? (int32_t)(base + miss[k] - 1) : INT32_MAX;
            if (h.count > 1) { delete o; return fail(RVF_E_STATE, "bvh: traversal leaf with more than one triangle"); }
            f.link = h.count ? ~(int32_t)(tr.tri_base + h.start - 1) : (int32_t)(base + h.right - 1);
            if (h.left && h.left != (int)k + 2) { delete o; return fail(RVF_E_STATE, "bvh: non-preorder layout"); }
            tn32.push_back(f);
        }
        for (int64_t m = 0; m < nt; m++) {
            tt32[tr.tri_base + m] = t32[tr.tri_base + sb.idx[m]];
            tmap[tr.tri_base + m] = (int32_t)(tr.tri_base + sb.idx[m]);
        }
        troot_of[tr.node_base] = (int32_t)base;
    }
    o->n_tnodes = (int64_t)tn32.size();
    if (tn32.empty()) tn32.resize(1);
    o->troot.assign((size_t)G * G, -1);
    for (size_t k = 0; k < o->root.size(); k++)
        if (o->root[k] >= 0) o->troot[k] = troot_of.at(o->root[k]);
    // wave-queue items: nodes [0, M), then triangles [M, M+T)
    const int64_t M = o->n_nodes, NI = std::max<int64_t>(o->n_nodes + o->n_tris, 1);
    if (NI >= ((int64_t)1 << 30)) { delete o; return fail(RVF_E_ARG, "rvf_obs_build: too many BVH items"); }
    std::vector<Item32> i32((size_t)NI);
    std::vector<Item64> i64((size_t)NI);
    std::memset(i32.data(), 0, sizeof(Item32) * i32.size());
    std::memset(i64.data(), 0, sizeof(Item64) * i64.size());
    for (const Tree& tr : o->trees) {
        for (size_t k = 0; k < tr.nodes.size(); k++) {
            const int64_t n = tr.node_base + (int64_t)k;
            const BvhNode64& d = n64[n];
            const BvhNode32& f = n32[n];
            const int32_t pass = d.count ? (int32_t)(M + d.start) : (int32_t)(n + 1);
            for (int c = 0; c < 3; c++) {
                i64[n].v[c] = d.lo[c]; i64[n].v[3 + c] = d.hi[c];
                i32[n].v[c] = f.lo[c]; i32[n].v[3 + c] = f.hi[c];
            }
            i64[n].link_x = i32[n].link_x = pass;
            i64[n].link_y = i32[n].link_y = d.miss;
            for (int c = 0; c < d.count; c++) {
                const int64_t t = d.start + c, it = M + t;
                const bool last = c == d.count - 1;
                for (int e = 0; e < 3; e++) {
                    i64[it].v[e] = t64[t].v0[e]; i64[it].v[3 + e] = t64[t].e1[e]; i64[it].v[6 + e] = t64[t].e2[e];
                    i32[it].v[e] = t32[t].v0[e]; i32[it].v[3 + e] = t32[t].e1[e]; i32[it].v[6 + e] = t32[t].e2[e];
                }
                i64[it].link_x = i32[it].link_x = last ? d.miss : (int32_t)(it + 1);
                i64[it].link_y = i32[it].link_y = (int32_t)n;
            }
        }
    }
    int rc;
    const hipStream_t st = ctx->stream;
    if ((rc = upload(o->d_items32, i32.data(), sizeof(Item32) * i32.size(), st)) ||
        (rc = upload(o->d_items64, i64.data(), sizeof(Item64) * i64.size(), st)) ||
        (rc = upload(o->d_root, o->root.data(), sizeof(int32_t) * std::max(1, G * G), st)) ||
        (rc = upload(o->d_n64, n64.data(), sizeof(BvhNode64) * n64.size(), st)) ||
        (rc = upload(o->d_n32, n32.data(), sizeof(BvhNode32) * n32.size(), st)) ||
        (rc = upload(o->d_t64, t64.data(), sizeof(Tri64) * t64.size(), st)) ||
        (rc = upload(o->d_t32, t32.data(), sizeof(Tri32) * t32.size(), st)) ||
        (rc = upload(o->d_parent, parent.data(), sizeof(int32_t) * parent.size(), st)) ||
        (rc = upload(o->d_troot, o->troot.data(), sizeof(int32_t) * std::max(1, G * G), st)) ||
        (rc = upload(o->d_tn32, tn32.data(), sizeof(TNode32) * tn32.size(), st)) ||
        (rc = upload(o->d_tt32, tt32.data(), sizeof(Tri32) * tt32.size(), st)) ||
        (rc = upload(o->d_tmap, tmap.data(), sizeof(int32_t) * tmap.size(), st))) {
        delete o;
        return rc;
    }
    *out = o;
    return 0;
}

void rvf_obs_free(rvf_obs* o) { delete o; }

int rvf_obs_info(const rvf_obs* o, int32_t* n_trees, int64_t* n_nodes, int64_t* n_tris) {
    if (!o) return fail(RVF_E_ARG, "null obs");
    if (n_trees) *n_trees = (int32_t)o->trees.size();
    if (n_nodes) *n_nodes = o->n_nodes;
    if (n_tris) *n_tris = o->n_tris;
    return 0;
}

int rvf_obs_export(const rvf_obs* o, int32_t tree, int64_t cap, double* box, int32_t* meta, int32_t* tri_idx,
                   int64_t* n_nodes, int64_t* n_tris) {
    if (!o || tree < 0 || tree >= (int)o->trees.size()) return fail(RVF_E_ARG, "rvf_obs_export: bad tree");
    const Tree& t = o->trees[tree];
    if (n_nodes) *n_nodes = (int64_t)t.nodes.size();
    if (n_tris) *n_tris = (int64_t)t.idx.size();
    if ((int64_t)t.nodes.size() > cap || (int64_t)t.idx.size() > cap) return fail(RVF_E_ARG, "rvf_obs_export: cap too small");
    for (size_t k = 0; k < t.nodes.size(); k++) {
        for (int c = 0; c < 3; c++) { box[6 * k + c] = t.nodes[k].lo[c]; box[6 * k + 3 + c] = t.nodes[k].hi[c]; }
        meta[5 * k] = t.nodes[k].left; meta[5 * k + 1] = t.nodes[k].right;
        meta[5 * k + 2] = t.nodes[k].start; meta[5 * k + 3] = t.nodes[k].count; meta[5 * k + 4] = t.miss[k];
    }
    for (size_t k = 0; k < t.idx.size(); k++) tri_idx[k] = t.idx[k];
    return 0;
}

int rvf_bvh_build_host(const double* soup, int64_t T, int32_t soup_read, int64_t cap, double* box, int32_t* meta,
                       int32_t* tri_idx, int64_t* n_nodes) {
    if (!soup || T <= 0 || soup_read < 0 || soup_read > 3) return fail(RVF_E_ARG, "rvf_bvh_build_host: bad arguments");
    Builder bd;
    bd.soup = soup;
    if (soup_read == RVF_SOUP_VERTEX_XYZ) { bd.av = 1; bd.ac = 3; } else { bd.ac = 1; bd.av = 3; }
    if (soup_read == RVF_SOUP_SEGMENTS) bd.nv = 2;
    bd.idx.resize(T);
    for (int64_t t = 0; t < T; t++) bd.idx[t] = (int)(t + 1);
    bd.build(1, (int)T);
    Tree tr;
    tr.nodes = bd.nodes; tr.idx = bd.idx;
    tr.miss.assign(tr.nodes.size(), 0);
    assign_miss(tr, 1, 0);
    if (n_nodes) *n_nodes = (int64_t)tr.nodes.size();
    if ((int64_t)tr.nodes.size() > cap) return fail(RVF_E_ARG, "rvf_bvh_build_host: cap too small");
    for (size_t k = 0; k < tr.nodes.size(); k++) {
        for (int c = 0; c < 3; c++) { box[6 * k + c] = tr.nodes[k].lo[c]; box[6 * k + 3 + c] = tr.nodes[k].hi[c]; }
        meta[5 * k] = tr.nodes[k].left; meta[5 * k + 1] = tr.nodes[k].right;
        meta[5 * k + 2] = tr.nodes[k].start; meta[5 * k + 3] = tr.nodes[k].count; meta[5 * k + 4] = tr.miss[k];
    }
    for (int64_t t = 0; t < T; t++) tri_idx[t] = tr.idx[t];
    return 0;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct rvf_plan {
    rvf_ctx* ctx = nullptr;
    const rvf_mesh* mesh = nullptr;
    const rvf_obs* obs = nullptr;
    rvf_opts opts{};
    int nq_quad = 0, nq_tri = 0, nq_line = 0, maxpts = 0;
    DBuf rule, P64, Q64, P32, Q32, anchor, area, elem_np, bad, F, part, partT, rowsum, W, audit_c, audit_h, counters;
    DBuf Tseg, Wseg, area_orig, diag;
    hipGraphExec_t gexec = nullptr;
    const Layout* L = nullptr;
    // Duffy singular pairs
    int npairs = 0;
    DBuf d_pairs, d_gl01, d_dval, d_rptr, d_rlist;
    hipEvent_t ev[8] = {};
    rvf_stats st{};
    bool ran = false;
    int refill_min = 0;  // 0: the kernel's default (default_refill)
    int cw = 0;      // columns per tile of the current layout
    int cw_req = 0;  // columns asked for by rvf_plan_tune_tiles (0: the kernel's default)
    int waves = 0;   // rvf_plan_tune_waves: 3 = the mixed split kernels' three-wave build, 0 = default
    bool scratch = false;  // buffers are pieces of the static scratch array
};

static int occ_kind(int variant, bool mixed, int64_t n_tris, int64_t pts);
static const char* occ_kind_error(int variant, int64_t n_tris, bool mixed, int64_t pts);
static int64_t plan_pts(const rvf_plan* p);
static int pair_kind(const rvf_plan* p);

// Columns per tile.  The wave-queue kernels claim tiles dynamically and run
// best on narrow ones (the tail of a launch is one tile's time; measured on
// the occluded 13,824-element cube, mixed 250 / 240 / 234 ms at 16 / 8 / 4
// columns, later 160.4 / 158.0 / 158.2 ms at 4 / 2 / 1 and FP64 656 / 651 /
// 649 ms; on the occluded plates mixed 205.3 / 204.3 / 205.9 and FP64 1111 /
// 1103 / 1100 ms): two for the mixed split kernels, one for the FP64 queue
// kernels, four for the other wave-queue kernels; the other kernels amortise
// their per-tile set-up over RVF_CHUNK columns.
static int default_columns(int kind) {
    if (!rvf_queue_kind(kind)) return RVF_CHUNK;
    if (rvf_mixed_split_kind(kind)) return 2;
    if (kind >= RVF_K_FP64_QUEUE1 && kind <= RVF_K_FP64_QUEUE3) return 1;
    return RVF_CHUNK_QUEUE;
}
static int plan_columns(const rvf_plan* p) {
    return p->cw_req ? p->cw_req : default_columns(pair_kind(p));
}
// Lanes with an empty slot before a wave-queue refill.  Measured on the full
// occluded cube, mixed split 3: 12, 16, 20, 24, 28, 32, 40, 48 -> 182.3,
// 180.2, 179.1, 178.4, 178.4, 180.9, 183.7, 188.1 ms.
static int default_refill(int kind) { return rvf_mixed_split_kind(kind) ? 24 : 16; }
// spatially compact strips for occluded quadrature solves (coherent ray bundles)
static bool plan_spatial(const rvf_plan* p) {
    return p->obs && !p->obs->trees.empty() && p->opts.method == RVF_METHOD_QUAD;
}
static size_t part_bytes(const Layout* L) { return sizeof(double) * std::max<size_t>(L->tiles.size(), 1) * RVF_STRIP; }

// Device buffers of a plan come either from hipMalloc (rvf_plan_create) or
// from consecutive 256-byte-aligned pieces of the library's static scratch
// array (rvf_compute_view_factors_noalloc).  In the second form `used` counts
// the bytes the plan needs even past `cap`, so a refusal can say how many.
struct BufSource {
    char* base = nullptr;  // null: hipMalloc
    size_t cap = 0, used = 0;
    bool over = false;
    int get(DBuf& b, size_t bytes) {
        if (!base) return b.alloc(bytes);
        if (bytes == 0) bytes = 16;
        const size_t off = (used + 255) & ~(size_t)255;
        used = off + bytes;
        if (used > cap) { over = true; b.view(nullptr, 0); return 0; }
        b.view(base + off, bytes);
        return 0;
    }
};

// Sets up a plan: quadrature tables, slot layout, Duffy pair list and every
// device buffer of the solve.  With src.base set (the no-allocation form) the
// layout goes into *own_layout and the plan's device buffers, layout and
// uploads included, into the scratch pieces; F is then the caller's F_dev.
static int plan_init(rvf_ctx* ctx, const rvf_mesh* mesh, const rvf_obs* obs, const rvf_opts* opts, rvf_plan* p,
                     BufSource& src, Layout* own_layout, double* F_dev) {
    if (opts->nquad < 1 || opts->nquad > 8) return fail(RVF_E_ARG, "nquad must be in 1..8");
    if (opts->method < 0 || opts->method > 2) return fail(RVF_E_ARG, "unknown method");
    if (opts->precision < 0 || opts->precision > 1) return fail(RVF_E_ARG, "unknown precision");
    if (opts->method == RVF_METHOD_MC && opts->n_samples < 1) return fail(RVF_E_ARG, "n_samples must be >= 1");
    if (obs && (obs->G != (int)mesh->tags.size() || obs->seg != mesh->dim2))
        return fail(RVF_E_ARG, "obstruction set built for another mesh");
    if (obs && !obs->trees.empty() && opts->method != RVF_METHOD_MC &&
        occ_kind(opts->variant, opts->precision == RVF_MIXED, obs->n_tris, 0) < 0)
        return fail(RVF_E_ARG, std::string("rvf_plan_create: ") +
                                   occ_kind_error(opts->variant, obs->n_tris, opts->precision == RVF_MIXED, 0));
    const bool scratch = src.base != nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    p->ctx = ctx; p->mesh = mesh; p->obs = obs; p->opts = *opts;
    std::vector<double> rule(RVF_RULE_SIZE, 0.0);
    p->nq_quad = quad_rule(opts->nquad, rule.data(), rule.data() + 2 * 64);
    p->nq_tri = tri_rule(opts->nquad, rule.data() + 3 * 64, rule.data() + 3 * 64 + 26);
    {   // curve elements: 1-D Gauss-Legendre points stored as (xi, 0) (ViewFactorKernel.jl:284-301)
        double lp[8], lw[8];
        p->nq_line = gl1d(opts->nquad, lp, lw);
        for (int k = 0; k < p->nq_line; k++) {
            rule[RVF_RULE_LINE + 2 * k] = lp[k];
            rule[RVF_RULE_LINE + 16 + k] = lw[k];
        }
    }
    bool anyq = false, anyt = false;
    for (int8_t f : mesh->fam) { if (f == 0 || f == 2) anyq = true; else anyt = true; }
    p->maxpts = mesh->dim2 ? p->nq_line : std::max(anyq ? p->nq_quad : 0, anyt ? p->nq_tri : 0);
    if (p->maxpts == 0) p->maxpts = 1;
    if (obs && !obs->trees.empty() && opts->method != RVF_METHOD_MC && !mesh->dim2 &&
        occ_kind(opts->variant, opts->precision == RVF_MIXED, obs->n_tris, plan_pts(p)) < 0)
        return fail(RVF_E_ARG, std::string("rvf_plan_create: ") +
                                   occ_kind_error(opts->variant, obs->n_tris, opts->precision == RVF_MIXED, plan_pts(p)));
    int64_t N = mesh->N, NS = mesh->NS;
    int G = (int)mesh->tags.size();
    const bool spatial = plan_spatial(p);
    const bool self = (opts->flags & RVF_SELF_VF) != 0;
    p->scratch = scratch;
    p->cw = plan_columns(p);
    int rc;
    if (scratch) {
        if ((rc = layout_host(mesh, spatial, self, p->cw, *own_layout))) return rc;
        p->L = own_layout;
    } else if ((rc = get_layout(mesh, spatial, self, p->cw, &p->L))) {
        return rc;
    }
    size_t nstrip = NS / RVF_STRIP;
    // Duffy: singular Quad8 pairs (singularity_type, DuffyKernel.jl:95-115)
    std::vector<DuffyPair> pairs;
    std::vector<int32_t> rptr, rlist;
    double g01[16] = {0};
    if (opts->method == RVF_METHOD_DUFFY) {
        std::map<int32_t, std::vector<int32_t>> by_corner;
        for (int64_t e = 0; e < N; e++)
            if (mesh->fam[e] == 0)
                for (int q = 0; q < 4; q++) by_corner[mesh->nodes8[8 * e + q]].push_back((int32_t)e);
        std::vector<int32_t> cand;
        for (int64_t i = 0; i < N; i++) {
            if (mesh->fam[i] != 0) continue;
            cand.clear();
            for (int q = 0; q < 4; q++)
                for (int32_t j : by_corner[mesh->nodes8[8 * i + q]])
                    if (j > i || (self && j == i)) cand.push_back(j);
            std::sort(cand.begin(), cand.end());
            cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
            for (int32_t j : cand) {
                DuffyPair d{};
                d.i = (int32_t)i; d.j = j;
                int n = 0, ci[2] = {0, 0}, cj[2] = {0, 0};
                for (int a = 0; a < 4; a++)
                    for (int b = 0; b < 4; b++)
                        if (mesh->nodes8[8 * i + a] == mesh->nodes8[8 * (int64_t)j + b]) {
                            if (n < 2) { ci[n] = a + 1; cj[n] = b + 1; }
                            n++;
                        }
                d.kind = n == 1 ? 1 : 2;
                d.ci0 = ci[0]; d.ci1 = ci[1]; d.cj0 = cj[0]; d.cj1 = cj[1];
                d.root = obs && !obs->trees.empty() ? obs->root[(size_t)mesh->grp_rank[i] * G + mesh->grp_rank[j]] : -1;
                pairs.push_back(d);
            }
        }
        p->npairs = (int)pairs.size();
        rptr.assign(N + 1, 0);
        for (const DuffyPair& d : pairs) { rptr[d.i + 1]++; if (d.j != d.i) rptr[d.j + 1]++; }
        for (int64_t r = 0; r < N; r++) rptr[r + 1] += rptr[r];
        rlist.assign(rptr[N], 0);
        std::vector<int32_t> fillp(rptr.begin(), rptr.end() - 1);
        for (int k = 0; k < (int)pairs.size(); k++) {
            rlist[fillp[pairs[k].i]++] = k;
            if (pairs[k].j != pairs[k].i) rlist[fillp[pairs[k].j]++] = k;
        }
        double p1[8], w1[8];
        int nq = gl1d(opts->nquad, p1, w1);
        for (int k = 0; k < nq; k++) { g01[k] = (p1[k] + 1.0) / 2.0; g01[8 + k] = w1[k] / 2.0; }
        p->st.singular_pairs = p->npairs;
    }
    // device buffers
    if ((rc = src.get(p->rule, sizeof(double) * rule.size())) ||
        (rc = src.get(p->P64, sizeof(double) * 8 * p->maxpts * NS)) ||
        (rc = src.get(p->Q64, sizeof(double) * 8 * p->maxpts * NS)) ||
        (rc = src.get(p->P32, sizeof(float) * 8 * p->maxpts * NS)) ||
        (rc = src.get(p->Q32, sizeof(float) * 8 * p->maxpts * NS)) ||
        (rc = src.get(p->anchor, sizeof(double) * 4 * NS)) || (rc = src.get(p->area, sizeof(double) * NS)) ||
        (rc = src.get(p->area_orig, sizeof(double) * std::max<int64_t>(N, 1))) ||
        (rc = src.get(p->elem_np, sizeof(int32_t) * NS)) || (rc = src.get(p->bad, sizeof(int))) ||
        (rc = src.get(p->part, part_bytes(p->L))) ||
        (rc = src.get(p->partT, sizeof(double) * std::max<size_t>(nstrip, 1) * NS)) ||
        (rc = src.get(p->rowsum, sizeof(double) * NS)) || (rc = src.get(p->W, sizeof(double) * std::max(G, 1) * NS)) ||
        (self && (rc = src.get(p->diag, sizeof(double) * NS))) ||
        (rc = src.get(p->counters, sizeof(unsigned long long) * RVF_NCOUNTERS)) ||
        (rc = src.get(p->Tseg, sizeof(double) * RVF_RSEG * NS)) ||
        (rc = src.get(p->Wseg, sizeof(double) * RVF_RSEG * std::max(G, 1) * NS)))
        return rc;
    if (opts->flags & RVF_WANT_F) {
        if (scratch) p->F.view(F_dev, sizeof(double) * N * N);
        else if ((rc = p->F.alloc(sizeof(double) * N * N))) return rc;
    }
    if (opts->method == RVF_METHOD_DUFFY &&
        ((rc = src.get(p->d_pairs, sizeof(DuffyPair) * pairs.size())) || (rc = src.get(p->d_gl01, sizeof g01)) ||
         (rc = src.get(p->d_dval, sizeof(double) * pairs.size())) ||
         (rc = src.get(p->d_rptr, sizeof(int32_t) * rptr.size())) ||
         (rc = src.get(p->d_rlist, sizeof(int32_t) * rlist.size()))))
        return rc;
    Layout* own = scratch ? own_layout : nullptr;
    if (own &&
        ((rc = src.get(own->d_slot_orig, sizeof(int32_t) * NS)) || (rc = src.get(own->d_slot_grp, sizeof(int32_t) * NS)) ||
         (rc = src.get(own->d_chunks, sizeof(Chunk) * own->chunks.size())) ||
         (rc = src.get(own->d_chunk_of, sizeof(int32_t) * NS)) ||
         (rc = src.get(own->d_tiles, sizeof(int2) * own->tiles.size())) ||
         (rc = src.get(own->d_xorder, sizeof(int32_t) * own->xorder.size()))))
        return rc;
    if ((opts->flags & RVF_AUDIT) && obs) {
        if (scratch) return fail(RVF_E_ARG, "RVF_AUDIT needs device allocations");
        if ((rc = p->audit_c.alloc(sizeof(uint32_t) * N * N)) || (rc = p->audit_h.alloc(sizeof(uint32_t) * N * N)))
            return rc;
        HIPCHK(hipMemsetAsync(p->audit_c.p, 0, p->audit_c.n, ctx->stream));
        HIPCHK(hipMemsetAsync(p->audit_h.p, 0, p->audit_h.n, ctx->stream));
    }
    if (src.over)
        return fail(RVF_E_NOMEM, "this solve needs " + std::to_string(src.used) + " bytes of scratch; the static scratch array holds " +
                                     std::to_string(src.cap) + " bytes (use rvf_compute_view_factors)");
    // uploads and zero fills, ordered on the context stream
    auto h2d = [&](DBuf& b, const void* h, size_t bytes) -> int {
        if (!bytes) return 0;
        if (scratch) HIPCHK(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, ctx->stream));
        else return copy_sync(b.p, h, bytes, hipMemcpyHostToDevice, ctx->stream);
        return 0;
    };
    if (scratch) {
        // the array keeps the previous call's contents: clear this solve's part
        if ((rc = rvf_launch_zero(src.base, (int64_t)src.used, ctx->stream)))
            return fail(RVF_E_HIP, std::string("zero launch: ") + hipGetErrorString((hipError_t)rc));
    }
    if ((rc = h2d(p->rule, rule.data(), sizeof(double) * rule.size()))) return rc;
    if (opts->method == RVF_METHOD_DUFFY &&
        ((rc = h2d(p->d_pairs, pairs.data(), sizeof(DuffyPair) * pairs.size())) || (rc = h2d(p->d_gl01, g01, sizeof g01)) ||
         (rc = h2d(p->d_rptr, rptr.data(), sizeof(int32_t) * rptr.size())) ||
         (rc = h2d(p->d_rlist, rlist.data(), sizeof(int32_t) * rlist.size()))))
        return rc;
    if (own &&
        ((rc = h2d(own->d_slot_orig, own->slot_orig.data(), sizeof(int32_t) * NS)) ||
         (rc = h2d(own->d_slot_grp, own->slot_grp.data(), sizeof(int32_t) * NS)) ||
         (rc = h2d(own->d_chunks, own->chunks.data(), sizeof(Chunk) * own->chunks.size())) ||
         (rc = h2d(own->d_chunk_of, own->chunk_of.data(), sizeof(int32_t) * NS)) ||
         (rc = h2d(own->d_tiles, own->tiles.data(), sizeof(int2) * own->tiles.size())) ||
         (rc = h2d(own->d_xorder, own->xorder.data(), sizeof(int32_t) * own->xorder.size()))))
        return rc;
    if (scratch) {
        // the host arrays above are released on return: finish the copies first
        HIPCHK(hipStreamSynchronize(ctx->stream));
    } else {
        // zero point arrays once (padding lanes / points read but never accumulated)
        HIPCHK(hipMemsetAsync(p->P64.p, 0, p->P64.n, ctx->stream));
        HIPCHK(hipMemsetAsync(p->P32.p, 0, p->P32.n, ctx->stream));
        HIPCHK(hipMemsetAsync(p->anchor.p, 0, p->anchor.n, ctx->stream));
        HIPCHK(hipMemsetAsync(p->area.p, 0, p->area.n, ctx->stream));
        HIPCHK(hipMemsetAsync(p->elem_np.p, 0, p->elem_np.n, ctx->stream));
        for (int k = 0; k < 8; k++) HIPCHK(hipEventCreate(&p->ev[k]));
    }
    p->st.n_elems = N;
    p->st.unique_pairs = self ? N * (N + 1) / 2 : N * (N - 1) / 2;
    p->st.tiles = (int64_t)p->L->tiles.size();
    return 0;
}

int rvf_plan_create(rvf_ctx* ctx, const rvf_mesh* mesh, const rvf_obs* obs, const rvf_opts* opts, rvf_plan** out) {
    if (!ctx || !mesh || !opts || !out) return fail(RVF_E_ARG, "rvf_plan_create: bad arguments");
    rvf_plan* p = new rvf_plan();
    BufSource src;
    const int rc = plan_init(ctx, mesh, obs, opts, p, src, nullptr, nullptr);
    if (rc) { rvf_plan_free(p); return rc; }
    *out = p;
    return 0;
}

static PairArgs pair_args(const rvf_plan* p) {
    const rvf_mesh* m = p->mesh;
    PairArgs a{};
    const Layout* L = p->L;
    a.tiles = L->d_tiles.as<int2>(); a.ntiles = (int)L->tiles.size(); a.tile_step = 1;
    a.slot_orig = L->d_slot_orig.as<int32_t>();
    a.NS = m->NS; a.maxpts = p->maxpts; a.N = m->N;
    a.chunks = L->d_chunks.as<Chunk>();
    // per-XCD claims for the mixed wave-queue kernels (the FP64 ones measured
    // slower with them: 666.1 -> 668.5 ms on the occluded cube, one run)
    a.xorder = p->opts.precision == RVF_MIXED ? L->d_xorder.as<int32_t>() : nullptr;
    for (int x = 0; x < 9; x++) a.xseg[x] = L->xseg[x];
    a.elem_np = p->elem_np.as<int32_t>(); a.elem_grp = L->d_slot_grp.as<int32_t>(); a.area = p->area.as<double>();
    a.P64 = p->P64.as<double>(); a.Q64 = p->Q64.as<double>(); a.P32 = p->P32.as<float>(); a.Q32 = p->Q32.as<float>();
    a.anchor = p->anchor.as<double>();
    a.F = p->F.as<double>(); a.part = p->part.as<double>(); a.partT = p->partT.as<double>();
    a.self_vf = p->diag.p ? 1 : 0; a.diag = p->diag.as<double>();
    a.G = (int)m->tags.size();
    a.dim2 = m->dim2 ? 1 : 0;
    if (p->obs && !p->obs->trees.empty()) {
        a.tree_root = p->obs->d_root.as<int32_t>();
        a.nodes64 = p->obs->d_n64.as<BvhNode64>(); a.tris64 = p->obs->d_t64.as<Tri64>();
        a.nodes32 = p->obs->d_n32.as<BvhNode32>(); a.tris32 = p->obs->d_t32.as<Tri32>();
        a.parent = p->obs->d_parent.as<int32_t>();
        a.items32 = p->obs->d_items32.as<Item32>(); a.items64 = p->obs->d_items64.as<Item64>();
        a.ttree_root = p->obs->d_troot.as<int32_t>(); a.tnodes32 = p->obs->d_tn32.as<TNode32>();
        a.ttris32 = p->obs->d_tt32.as<Tri32>(); a.tmap = p->obs->d_tmap.as<int32_t>();
        a.n_nodes = (int)p->obs->n_nodes;
        a.refill_min = p->refill_min ? p->refill_min : default_refill(pair_kind(p));
        a.verify = p->obs->soup_read == RVF_SOUP_REFERENCE ? 1 : 0;
        a.cull_limit = 256;  // box tests per pair; measured: 64 leaves 2 % of the node steps, more gains nothing
        a.hint = 1;  // (a second hint per pair saved 8 % of the node steps but cost more than it saved)
        a.waves = rvf_mixed_split_kind(pair_kind(p)) ? p->waves : 0;
        a.Sm = p->obs->Sm;
    } else {
        a.Sm = 1.0f;
    }
    a.audit_count = p->audit_c.as<uint32_t>(); a.audit_hash = p->audit_h.as<uint32_t>();
    a.counters = p->counters.as<unsigned long long>();
    a.tile_ctr = (int*)(a.counters + 15);  // slots 0..10 hold the statistics; per-XCD counters follow
    a.coords = m->coords.as<double>(); a.fam = m->d_fam.as<int8_t>(); a.nodes8 = m->d_nodes8.as<int32_t>();
    a.duffy = p->opts.method == RVF_METHOD_DUFFY ? 1 : 0;
    a.mc_nsamp = p->opts.n_samples;
    a.mc_s = (int)std::floor(std::sqrt((double)p->opts.n_samples));
    a.mc_seed = p->opts.seed;
    return a;
}

static bool plan_occluded_quad(const rvf_plan* p) {
    return p->obs && !p->obs->trees.empty() && p->opts.method != RVF_METHOD_MC;
}

// Kernel kind for an occluded quadrature plan's variant, -1 if invalid.  The
// split variants address triangles with RVF_TQ_TRI_BITS bits of a task word:
// with 2^19 or more triangles over all trees they are invalid.  The mixed
// split variants also index the row points (pts = 8 maxpts NS FP32 values,
// 0 while unknown) with 32-bit offsets.  Past either bound the mixed default
// falls back to the queue form.
static constexpr int64_t RVF_SPLIT_PTS_MAX = ((int64_t)1 << 32) - 128;
static int64_t plan_pts(const rvf_plan* p) { return 8 * (int64_t)p->maxpts * p->mesh->NS; }
static bool split_ok(bool mixed, int64_t n_tris, int64_t pts) {
    return n_tris < ((int64_t)1 << RVF_TQ_TRI_BITS) && (!mixed || pts < RVF_SPLIT_PTS_MAX);
}
static int default_variant(bool mixed, int64_t n_tris, int64_t pts) {
    if (!mixed) return RVF_VARIANT_QUEUE(1);
    return split_ok(mixed, n_tris, pts) ? RVF_VARIANT_SPLIT(3) : RVF_VARIANT_QUEUE(3);
}
static int occ_kind(int variant, bool mixed, int64_t n_tris, int64_t pts) {
    if (variant == RVF_VARIANT_DEFAULT) variant = default_variant(mixed, n_tris, pts);
    if (variant == RVF_VARIANT_LANE) return mixed ? RVF_K_MIX_OCC_LANE : RVF_K_FP64_OCC_LANE;
    if (variant > 20) {  // split node / triangle-task form
        if (!split_ok(mixed, n_tris, pts)) return -1;
        const int S = variant - 20;
        if (!mixed) return S == 1 ? RVF_K_FP64_SPLIT1 : S == 2 ? RVF_K_FP64_SPLIT2 : S == 3 ? RVF_K_FP64_SPLIT3 : -1;
        return S == 1 ? RVF_K_MIX_SPLIT1 : S == 2 ? RVF_K_MIX_SPLIT2 : S == 3 ? RVF_K_MIX_SPLIT3 : -1;
    }
    const int S = variant - 10;
    if (!mixed) return S == 1 ? RVF_K_FP64_QUEUE1 : S == 2 ? RVF_K_FP64_QUEUE2 : S == 3 ? RVF_K_FP64_QUEUE3 : -1;
    return S == 2 ? RVF_K_MIX_QUEUE2 : S == 3 ? RVF_K_MIX_QUEUE3 : S == 4 ? RVF_K_MIX_QUEUE4
         : S == 6 ? RVF_K_MIX_QUEUE6 : -1;
}
static const char* occ_kind_error(int variant, int64_t n_tris, bool mixed, int64_t pts) {
    if (variant > 20 && variant <= 23 && n_tris >= ((int64_t)1 << RVF_TQ_TRI_BITS))
        return "split traversal variants need fewer than 2^19 obstruction triangles over all trees";
    if (variant > 20 && variant <= 23 && mixed && pts >= RVF_SPLIT_PTS_MAX)
        return "mixed split traversal variants need 8 x (points per element) x (elements) below 2^32";
    return "unknown traversal variant for this precision";
}

// Which pair-kernel variant a plan runs, and the LDS it needs.
static int pair_kind(const rvf_plan* p) {
    const bool mixed = p->opts.precision == RVF_MIXED;
    if (p->mesh->dim2 && p->opts.method != RVF_METHOD_MC) {  // curve meshes: one traversal form
        if (plan_occluded_quad(p)) return mixed ? RVF_K_MIX_OCC2D : RVF_K_FP64_OCC2D;
        return mixed ? RVF_K_MIX_2D : RVF_K_FP64_2D;
    }
    if (plan_occluded_quad(p)) return occ_kind(p->opts.variant, mixed, p->obs->n_tris, plan_pts(p));
    int pc = p->maxpts;
    if (!mixed && pc > 9) pc = 9;  // register chunks of 9 points
    return rvf_pair_kind(mixed, 0, pc, p->opts.method == RVF_METHOD_MC);
}
static size_t pair_lds(const rvf_plan* p) {
    if (!plan_occluded_quad(p)) return 0;
    const int kind = pair_kind(p);
    const bool mixed = p->opts.precision == RVF_MIXED;
    const size_t accq = mixed ? 0 : sizeof(double) * 64 * (size_t)p->maxpts;
    if (kind == RVF_K_FP64_OCC_LANE || kind == RVF_K_FP64_OCC2D) return accq;
    if (kind == RVF_K_MIX_OCC2D) return 0;
    if (kind == RVF_K_MIX_OCC_LANE) return 0;
    const size_t ring =
        rvf_split_kind(kind)
            ? (mixed ? (RVF_TQCAP_MIX + RVF_CQCAP) * sizeof(uint32_t) + 6 * 64 * sizeof(float) +
                           RVF_FRONTIER * 64 * sizeof(int32_t)
                     : RVF_TQCAP * sizeof(uint32_t))
            : 0;
    return accq + RVF_QCOMBO * (sizeof(uint64_t) + 2 * sizeof(uint32_t)) + 2 * 64 * sizeof(int32_t) + ring +
           RVF_QSTATS * sizeof(uint64_t);
}

// A solve is three stages on one stream: the prologue (zero fills, point
// precompute, Monte Carlo areas), the pair tiles (one launch here; the server
// batches the tiles of many plans instead), and the epilogue (Duffy pairs,
// row/group reductions, the F fill).  Zero fills use a kernel, never
// hipMemsetAsync, so the sequence can be captured and replayed as a graph.
// With tim, HIP events bracket each stage (not used while capturing).
static int enqueue_prologue(rvf_plan* p, hipStream_t st, bool tim) {
    const rvf_mesh* m = p->mesh;
    if (tim) HIPCHK(hipEventRecord(p->ev[0], st));
    int rc;
    if ((rc = rvf_launch_zero(p->bad.p, sizeof(int), st)) || (rc = rvf_launch_zero(p->counters.p, p->counters.n, st)) ||
        (rc = rvf_launch_zero(p->diag.p, p->diag.n, st)))
        return fail(RVF_E_HIP, std::string("zero launch: ") + hipGetErrorString((hipError_t)rc));
    PrecompArgs pa{};
    pa.coords = m->coords.as<double>(); pa.fam = m->d_fam.as<int8_t>(); pa.nodes8 = m->d_nodes8.as<int32_t>();
    pa.N = m->N; pa.NS = m->NS; pa.maxpts = p->maxpts; pa.nq_quad = p->nq_quad; pa.nq_tri = p->nq_tri;
    pa.nq_line = p->nq_line;
    pa.rule = p->rule.as<double>();
    pa.P64 = p->P64.as<double>(); pa.Q64 = p->Q64.as<double>(); pa.P32 = p->P32.as<float>(); pa.Q32 = p->Q32.as<float>();
    pa.anchor = p->anchor.as<double>(); pa.area = p->area.as<double>(); pa.elem_np = p->elem_np.as<int32_t>();
    pa.bad = p->bad.as<int>();
    pa.slot_orig = p->L->d_slot_orig.as<int32_t>();
    pa.area_orig = p->area_orig.as<double>();
    rc = rvf_launch_precompute(&pa, st);
    if (rc) return fail(RVF_E_HIP, std::string("precompute launch: ") + hipGetErrorString((hipError_t)rc));
    if (p->opts.method == RVF_METHOD_MC) {
        PairArgs a = pair_args(p);
        rc = rvf_launch_mc_area(&a, p->area.as<double>(), p->area_orig.as<double>(), st);
        if (rc) return fail(RVF_E_HIP, std::string("mc area launch: ") + hipGetErrorString((hipError_t)rc));
    }
    if (tim) HIPCHK(hipEventRecord(p->ev[1], st));
    return 0;
}

static int enqueue_pairs(rvf_plan* p, hipStream_t st) {
    PairArgs a = pair_args(p);
    const int rc = rvf_launch_pair(&a, pair_kind(p), pair_lds(p), st);
    if (rc) return fail(RVF_E_HIP, std::string("pair launch: ") + hipGetErrorString((hipError_t)rc));
    return 0;
}

static int enqueue_epilogue(rvf_plan* p, hipStream_t st, bool tim) {
    const rvf_mesh* m = p->mesh;
    int rc;
    if (tim) HIPCHK(hipEventRecord(p->ev[2], st));
    PairArgs a = pair_args(p);
    if (p->npairs) {
        rc = rvf_launch_duffy(&a, p->d_pairs.as<DuffyPair>(), p->npairs, p->d_gl01.as<double>(), p->opts.nquad,
                              p->d_dval.as<double>(), nullptr, nullptr, nullptr, nullptr, 0, st);
        if (rc) return fail(RVF_E_HIP, std::string("duffy launch: ") + hipGetErrorString((hipError_t)rc));
    }
    if (tim) HIPCHK(hipEventRecord(p->ev[3], st));
    if ((rc = rvf_launch_zero(p->Wseg.p, p->Wseg.n, st)))
        return fail(RVF_E_HIP, std::string("zero launch: ") + hipGetErrorString((hipError_t)rc));
    rc = rvf_launch_reduce(p->part.as<double>(), p->partT.as<double>(), p->L->d_chunks.as<Chunk>(),
                           (int)p->L->chunks.size(), p->L->d_chunk_of.as<int32_t>(), m->N, m->NS, (int)m->tags.size(),
                           p->rowsum.as<double>(), p->W.as<double>(), p->Tseg.as<double>(), p->Wseg.as<double>(), st);
    if (rc) return fail(RVF_E_HIP, std::string("reduce launch: ") + hipGetErrorString((hipError_t)rc));
    if (p->npairs) {
        rc = rvf_launch_duffy(&a, p->d_pairs.as<DuffyPair>(), p->npairs, nullptr, 0, p->d_dval.as<double>(),
                              p->d_rptr.as<int32_t>(), p->d_rlist.as<int32_t>(), p->rowsum.as<double>(),
                              p->W.as<double>(), 1, st);
        if (rc) return fail(RVF_E_HIP, std::string("duffy apply launch: ") + hipGetErrorString((hipError_t)rc));
    }
    if (p->diag.p) {
        rc = rvf_launch_diag(p->diag.as<double>(), p->area.as<double>(), m->N, p->rowsum.as<double>(), st);
        if (rc) return fail(RVF_E_HIP, std::string("diagonal launch: ") + hipGetErrorString((hipError_t)rc));
    }
    if (tim) HIPCHK(hipEventRecord(p->ev[4], st));
    if (p->F.p) {
        rc = rvf_launch_fill(p->F.as<double>(), p->area_orig.as<double>(), m->N, m->d_fill_tiles.as<int2>(),
                             (int)m->fill_tiles.size(), p->diag.p ? 1 : 0, st);
        if (rc) return fail(RVF_E_HIP, std::string("fill launch: ") + hipGetErrorString((hipError_t)rc));
    }
    if (tim) HIPCHK(hipEventRecord(p->ev[5], st));
    return 0;
}


static int enqueue_solve(rvf_plan* p, hipStream_t st, bool tim) {
    int rc;
    if ((rc = enqueue_prologue(p, st, tim)) || (rc = enqueue_pairs(p, st)) || (rc = enqueue_epilogue(p, st, tim)))
        return rc;
    return 0;
}

// --- internal interface for the request server (rvf_server.cpp) ---
int rvfi_plan_prologue(rvf_plan* p, hipStream_t st) { return enqueue_prologue(p, st, false); }
int rvfi_plan_epilogue(rvf_plan* p, hipStream_t st) {
    const int rc = enqueue_epilogue(p, st, false);
    if (!rc) p->ran = true;  // results can be fetched once the stream has drained
    return rc;
}
static int plan_collect(rvf_plan* p);
// the server calls this once an event after the request's epilogue has
// completed (the context stream may still be busy with other requests)
int rvfi_plan_collect(rvf_plan* p) { return plan_collect(p); }
int rvfi_plan_batch(const rvf_plan* p, PairArgs* a, int* kind, size_t* lds) {
    *a = pair_args(p);
    *kind = pair_kind(p);
    *lds = pair_lds(p);
    return 0;
}
int rvfi_plan_columns(const rvf_plan* p) { return p->cw; }
size_t rvfi_plan_bytes(const rvf_plan* p) {
    const DBuf* bufs[] = {&p->rule, &p->P64, &p->Q64, &p->P32, &p->Q32, &p->anchor, &p->area, &p->elem_np, &p->bad,
                          &p->F, &p->part, &p->partT, &p->rowsum, &p->W, &p->audit_c, &p->audit_h, &p->counters,
                          &p->Tseg, &p->Wseg, &p->area_orig, &p->d_pairs, &p->d_gl01, &p->d_dval, &p->d_rptr,
                          &p->d_rlist};
    size_t s = 0;
    for (const DBuf* b : bufs) s += b->n;
    return s;
}
hipStream_t rvfi_ctx_stream(const rvf_ctx* c) { return c->stream; }
int rvfi_ctx_device(const rvf_ctx* c) { return c->device; }
int rvfi_fail(int code, const std::string& msg) { return fail(code, msg); }

int rvf_plan_kernel(const rvf_plan* p, char* name, int32_t cap) {
    if (!p || !name || cap < 1) return fail(RVF_E_ARG, "rvf_plan_kernel: bad arguments");
    static const char* const names[] = {
#define RVF_KIND_NAME(e, t) "k_pair<" #t ">",
        RVF_PAIR_KINDS(RVF_KIND_NAME)
#undef RVF_KIND_NAME
    };
    const int k = pair_kind(p);
    if (k < 0 || k >= RVF_K_COUNT) return fail(RVF_E_STATE, "rvf_plan_kernel: no pair kernel");
    // the three-wave build of the mixed split kernels: k_pair_w3<tag>
    if (rvf_mixed_split_kind(k) && p->waves == 3) std::snprintf(name, (size_t)cap, "k_pair_w3%s", names[k] + 6);
    else std::snprintf(name, (size_t)cap, "%s", names[k]);
    return 0;
}

// Moves a plan to the layout of plan_columns(p) (cached on the mesh) and
// grows its row-partial buffer when the new layout has more tiles.  The
// caller has drained the context stream.
static int plan_relayout(rvf_plan* p) {
    const int cw = plan_columns(p);
    if (cw == p->cw) return 0;
    if (p->scratch) return fail(RVF_E_STATE, "a plan in the static scratch array keeps its layout");
    const Layout* L = nullptr;
    int rc = get_layout(p->mesh, plan_spatial(p), (p->opts.flags & RVF_SELF_VF) != 0, cw, &L);
    if (rc) return rc;
    if (p->part.n < part_bytes(L) && (rc = p->part.alloc(part_bytes(L)))) return rc;
    p->L = L;
    p->cw = cw;
    p->st.tiles = (int64_t)L->tiles.size();
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }  // re-captured on the next run
    return 0;
}

int rvf_plan_tune(rvf_plan* p, int32_t variant, int32_t refill_min) {
    if (!p) return fail(RVF_E_ARG, "rvf_plan_tune: null plan");
    if (refill_min < 0 || refill_min > 64) return fail(RVF_E_ARG, "rvf_plan_tune: refill_min must be in 0..64");
    if (p->obs && !p->obs->trees.empty() && p->opts.method != RVF_METHOD_MC &&
        occ_kind(variant, p->opts.precision == RVF_MIXED, p->obs->n_tris, plan_pts(p)) < 0)
        return fail(RVF_E_ARG, std::string("rvf_plan_tune: ") +
                                   occ_kind_error(variant, p->obs->n_tris, p->opts.precision == RVF_MIXED, plan_pts(p)));
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    const int v0 = p->opts.variant;
    p->opts.variant = variant;
    const int rc = plan_relayout(p);
    if (rc) { p->opts.variant = v0; return rc; }
    p->refill_min = refill_min;
    if (!rvf_mixed_split_kind(pair_kind(p))) p->waves = 0;  // the three-wave build exists for these alone
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }  // re-captured on the next run
    return 0;
}

int rvf_plan_tune_waves(rvf_plan* p, int32_t waves) {
    if (!p) return fail(RVF_E_ARG, "rvf_plan_tune_waves: null plan");
    const int k = pair_kind(p);
    const bool split = rvf_mixed_split_kind(k);
    if (waves == 4 && split) waves = 0;  // their default build
    if (waves != 0 && !(waves == 3 && split))
        return fail(RVF_E_ARG, "rvf_plan_tune_waves: waves must be 0 (the kernel's build), or 3 or 4 for the mixed "
                               "split traversal kernels");
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    p->waves = waves;
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }  // re-captured on the next run
    return 0;
}

int rvf_plan_tune_tiles(rvf_plan* p, int32_t columns) {
    if (!p) return fail(RVF_E_ARG, "rvf_plan_tune_tiles: null plan");
    if (columns < 0 || columns > RVF_CHUNK)
        return fail(RVF_E_ARG, "rvf_plan_tune_tiles: columns must be in 0.." + std::to_string(RVF_CHUNK));
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    const int c0 = p->cw_req;
    p->cw_req = columns;
    const int rc = plan_relayout(p);
    if (rc) p->cw_req = c0;
    return rc;
}

int rvf_plan_autotune(rvf_plan* p, int64_t sample_tiles, rvf_tune_entry* ent, int32_t cap, int32_t* n,
                      int32_t* variant, int32_t* refill_min) {
    if (!p || sample_tiles < 0 || cap < 0 || (cap > 0 && !ent) || !n)
        return fail(RVF_E_ARG, "rvf_plan_autotune: bad arguments");
    *n = 0;
    if (variant) *variant = p->opts.variant;
    if (refill_min) *refill_min = p->refill_min ? p->refill_min : default_refill(pair_kind(p));
    if (!plan_occluded_quad(p) || p->mesh->dim2) return 0;  // curve meshes have one traversal form
    if (p->scratch) return fail(RVF_E_STATE, "rvf_plan_autotune: a plan in the static scratch array is not tuned");
    rvf_ctx* c = p->ctx;
    hipStream_t st = c->stream;
    HIPCHK(hipSetDevice(c->device));
    int rc = enqueue_prologue(p, st, false);  // the point tables the pair kernel reads
    if (rc) return rc;
    const int64_t nt0 = (int64_t)p->L->tiles.size();
    if (nt0 == 0) return 0;
    // Samples are fractions of the problem: sample_tiles (0 = 8192) tiles of
    // the plan's layout as it stands, so that configurations with other tile
    // widths time the same share of the pairs.
    const double frac = std::min(1.0, (double)(sample_tiles ? sample_tiles : 8192) / (double)nt0);
    const double big = std::min(1.0, 4.0 * frac);
    const bool mixed = p->opts.precision == RVF_MIXED;
    const int v0 = p->opts.variant, r0 = p->refill_min, c0 = p->cw_req, w0 = p->waves;
    // Times the pair kernel over a share f of the tiles taken at a fixed
    // stride through the tile list (all tiles when f = 1).  Tile costs vary
    // with where the pair sits relative to the occluders, so a strided sample
    // follows the whole problem's mix where contiguous ranges do not.
    int64_t timed = 0;
    auto measure = [&](int v, int refill, int cols, double f, float* ms, int waves = 0) -> int {
        p->opts.variant = v;
        p->refill_min = refill;
        p->cw_req = cols;
        p->waves = waves;
        int e = plan_relayout(p);
        if (e) return e;
        const int64_t nt = (int64_t)p->L->tiles.size();
        const int64_t want = f >= 1.0 ? nt : std::max<int64_t>(1, (int64_t)(f * (double)nt));
        PairArgs a = pair_args(p);
        const int kind = pair_kind(p);
        const size_t lds = pair_lds(p);
        const int64_t step = want >= nt ? 1 : nt / want;
        const int64_t cnt = want >= nt ? nt : want;  // (cnt - 1) * step < nt
        a.tile0 = 0;
        a.tile_step = (int)step;
        a.xorder = nullptr;  // a strided sample claims from one counter
        a.ntiles = (int)std::min<int64_t>(cnt, 256);  // first launch of a kernel loads its code
        e = rvf_launch_pair(&a, kind, lds, st);
        HIPCHK(hipEventRecord(p->ev[6], st));
        a.ntiles = (int)cnt;
        if (!e) e = rvf_launch_pair(&a, kind, lds, st);
        if (e) return fail(RVF_E_HIP, std::string("autotune launch: ") + hipGetErrorString((hipError_t)e));
        HIPCHK(hipEventRecord(p->ev[7], st));
        HIPCHK(hipEventSynchronize(p->ev[7]));
        HIPCHK(hipEventElapsedTime(ms, p->ev[6], p->ev[7]));
        timed = cnt;
        return 0;
    };
    const int64_t ntri = p->obs->n_tris;
    std::vector<int> cands = {RVF_VARIANT_LANE};
    for (int S = 1; S <= 6; S++)
        if (occ_kind(RVF_VARIANT_QUEUE(S), mixed, ntri, plan_pts(p)) >= 0) cands.push_back(RVF_VARIANT_QUEUE(S));
    for (int S = 1; S <= 3; S++)
        if (occ_kind(RVF_VARIANT_SPLIT(S), mixed, ntri, plan_pts(p)) >= 0) cands.push_back(RVF_VARIANT_SPLIT(S));
    // a measured configuration as it ran (refill 0 stands for the kernel's default)
    auto add = [&](int v, int refill, float ms, int waves = 0) {
        const int r = v == RVF_VARIANT_LANE ? 0 : refill ? refill : default_refill(pair_kind(p));
        if (*n < cap) ent[*n] = rvf_tune_entry{v, r, timed, (double)ms, p->cw, waves};
        ++*n;
    };
    // 1. rank every variant (at its default tile width) on the sample
    std::vector<std::pair<float, int>> rank;
    for (int v : cands) {
        float ms = 0;
        if ((rc = measure(v, 0, 0, frac, &ms))) break;
        add(v, 0, ms);
        rank.push_back({ms, v});
    }
    // 2. re-time the three fastest on the whole problem (or a four times
    //    larger sample), then the refill thresholds of the winner, then its
    //    tile widths.  The library default (variant, refill and width at
    //    their defaults) is timed twice there: another configuration replaces the
    //    current choice only if faster by more than twice their spread (and
    //    at least 1 %), so run-to-run noise does not swap the kernel that runs.
    const int vdef = default_variant(mixed, ntri, plan_pts(p));
    int best_v = vdef, best_r = 0, best_c = 0;
    float best = 3.4e38f, margin = 0.01f;
    if (!rc) {
        float d1 = 0, d2 = 0;
        if (!(rc = measure(vdef, 0, 0, big, &d1))) {
            add(vdef, 0, d1);
            if (!(rc = measure(vdef, 0, 0, big, &d2))) {
                add(vdef, 0, d2);
                best = std::min(d1, d2);
                margin = std::max(0.01f, 2.0f * std::fabs(d1 - d2) / best);
            }
        }
    }
    auto better = [&](float ms) { return ms < best * (1.0f - margin); };
    if (!rc) {
        std::sort(rank.begin(), rank.end());
        for (size_t k = 0, done = 0; k < rank.size() && done < 3; k++) {
            if (rank[k].second == vdef) continue;
            float ms = 0;
            if ((rc = measure(rank[k].second, 0, 0, big, &ms))) break;
            add(rank[k].second, 0, ms);
            done++;
            if (better(ms)) { best = ms; best_v = rank[k].second; best_r = 0; }
        }
    }
    if (!rc && best_v != RVF_VARIANT_LANE) {
        const int bv = best_v;
        p->opts.variant = bv;
        const int rdef = default_refill(pair_kind(p));
        for (int refill : {8, 16, 24, 32}) {
            if (refill == rdef) continue;
            float ms = 0;
            if ((rc = measure(bv, refill, 0, big, &ms))) break;
            add(bv, refill, ms);
            if (better(ms)) { best = ms; best_r = refill; }
        }
    }
    if (!rc) {
        p->opts.variant = best_v;
        const int cdef = plan_columns(p);  // the choice's default width (cw_req is 0 here)
        for (int cols : {1, 2, RVF_CHUNK_QUEUE, 2 * RVF_CHUNK_QUEUE, RVF_CHUNK}) {
            if (cols == cdef) continue;
            float ms = 0;
            if ((rc = measure(best_v, best_r, cols, big, &ms))) break;
            add(best_v, best_r, ms);
            if (better(ms)) { best = ms; best_c = cols; }
        }
    }
    // 3. occupancy: the mixed split kernels' three-wave build (no spills)
    int best_w = 0;
    if (!rc) {
        p->opts.variant = best_v;
        if (rvf_mixed_split_kind(pair_kind(p))) {
            float ms = 0;
            if (!(rc = measure(best_v, best_r, best_c, big, &ms, 3))) {
                add(best_v, best_r, ms, 3);
                if (better(ms)) { best = ms; best_w = 3; }
            }
        }
    }
    if (*n > cap) *n = cap;
    if (rc) {
        p->opts.variant = v0; p->refill_min = r0; p->cw_req = c0; p->waves = w0;
        (void)plan_relayout(p);
        return rc;
    }
    p->opts.variant = best_v;
    p->refill_min = best_r;
    p->cw_req = best_c;
    p->waves = best_w;
    if ((rc = plan_relayout(p))) return rc;
    if (variant) *variant = best_v;
    if (refill_min) *refill_min = best_v == RVF_VARIANT_LANE ? 0 : best_r ? best_r : default_refill(pair_kind(p));
    if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }  // re-captured on the next run
    return 0;
}

int rvf_plan_run(rvf_plan* p) {
    if (!p) return fail(RVF_E_ARG, "null plan");
    rvf_ctx* c = p->ctx;
    hipStream_t st = c->stream;
    HIPCHK(hipSetDevice(c->device));
    bool tim = (p->opts.flags & RVF_TIMING) != 0;
    if (p->opts.flags & RVF_GRAPH) {
        // steady-state solves: capture the launch sequence once, replay it
        if (!p->gexec) {
            hipGraph_t g;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int rc = enqueue_solve(p, st, false);
            hipError_t e = hipStreamEndCapture(st, &g);
            if (rc) return rc;
            if (e != hipSuccess) return fail(RVF_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { p->gexec = nullptr; return fail(RVF_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e)); }
        }
        if (tim) HIPCHK(hipEventRecord(p->ev[0], st));
        HIPCHK(hipGraphLaunch(p->gexec, st));
        if (tim) {
            // per-stage events are not recorded inside the graph: the whole solve is timed
            for (int k = 1; k <= 5; k++) HIPCHK(hipEventRecord(p->ev[k], st));
        }
    } else {
        int rc = enqueue_solve(p, st, tim);
        if (rc) return rc;
    }
    p->ran = true;
    return 0;
}

// Reads the degenerate-element flag and the traversal counters of a solve
// whose work has completed.
static int plan_collect(rvf_plan* p) {
    int bad = 0;
    const hipStream_t st = p->ctx->stream;
    int rc = copy_sync(&bad, p->bad.p, sizeof(int), hipMemcpyDeviceToHost, st);
    if (rc) return rc;
    if (bad) return fail(RVF_E_ARG, "degenerate element: zero or non-finite area element at a quadrature point");
    unsigned long long cnt[16];
    if ((rc = copy_sync(cnt, p->counters.p, sizeof cnt, hipMemcpyDeviceToHost, st))) return rc;
    p->st.retests = (int64_t)cnt[0];
    p->st.rays = (int64_t)cnt[1];
    p->st.aabb_tests = (int64_t)cnt[2];
    p->st.tri_tests = (int64_t)cnt[3];
    p->st.trav_iters = (int64_t)cnt[4];
    p->st.slot_steps = (int64_t)cnt[5];
    p->st.task_batches = (int64_t)cnt[6];
    p->st.exact_batches = (int64_t)cnt[7];
    p->st.sure_hits = (int64_t)cnt[8];
    p->st.culled_rays = (int64_t)cnt[9];
    p->st.culled_waves = (int64_t)cnt[10];
    return 0;
}

int rvf_plan_sync(rvf_plan* p) {
    if (!p) return fail(RVF_E_ARG, "null plan");
    HIPCHK(hipSetDevice(p->ctx->device));
    HIPCHK(hipStreamSynchronize(p->ctx->stream));
    int rc = plan_collect(p);
    if (rc) return rc;
    if (p->opts.flags & RVF_TIMING) {
        float ms;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[0], p->ev[1])); p->st.precompute_ms = ms;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[1], p->ev[2])); p->st.pair_ms = ms;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[2], p->ev[3])); p->st.duffy_ms = ms;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[3], p->ev[4])); p->st.reduce_ms = ms;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[4], p->ev[5])); p->st.fill_ms = ms;
        HIPCHK(hipEventElapsedTime(&ms, p->ev[0], p->ev[5])); p->st.total_ms = ms;
    }
    return 0;
}

int rvf_host_alloc(size_t bytes, void** out) {
    if (!out || bytes == 0) return fail(RVF_E_ARG, "rvf_host_alloc: bad arguments");
    *out = nullptr;
    const hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *out = nullptr;
        return fail(RVF_E_NOMEM, std::string("rvf_host_alloc: ") + hipGetErrorString(e));
    }
    return 0;
}

int rvf_host_free(void* ptr) {
    if (ptr) HIPCHK(hipHostFree(ptr));
    return 0;
}

int rvf_plan_fetch(rvf_plan* p, double* F, double* A, double* rowsum, double* F_group, double* A_group,
                   int32_t* group_tags, uint32_t* audit_count, uint32_t* audit_hash) {
    if (!p || !p->ran) return fail(RVF_E_STATE, "rvf_plan_fetch: plan has not run");
    const rvf_mesh* m = p->mesh;
    int64_t N = m->N;
    int G = (int)m->tags.size();
    HIPCHK(hipSetDevice(p->ctx->device));
    const hipStream_t st = p->ctx->stream;
    auto d2h = [&](void* dst, const void* src, size_t bytes) { return copy_sync(dst, src, bytes, hipMemcpyDeviceToHost, st); };
    int rc;
    if (F) {
        if (!p->F.p) return fail(RVF_E_STATE, "F was not requested (RVF_WANT_F)");
        if ((rc = d2h(F, p->F.p, sizeof(double) * N * N))) return rc;
    }
    const Layout* L = p->L;
    std::vector<double> hA(N), hAo(N), hW((size_t)G * m->NS);
    if ((rc = d2h(hA.data(), p->area.p, sizeof(double) * N))) return rc;      // slot order
    if ((rc = d2h(hAo.data(), p->area_orig.p, sizeof(double) * N))) return rc; // element order
    if (A) std::memcpy(A, hAo.data(), sizeof(double) * N);
    if (rowsum) {
        std::vector<double> rs(N);
        if ((rc = d2h(rs.data(), p->rowsum.p, sizeof(double) * N))) return rc;
        for (int64_t k = 0; k < N; k++) rowsum[L->slot_orig[k]] = rs[k];
    }
    if (F_group || A_group) {
        if ((rc = d2h(hW.data(), p->W.p, sizeof(double) * G * m->NS))) return rc;
        // A_g F_gh = sum_{i in g} A_i sum_{j in h} F_ij (Results.jl:56-82); with W[h][k] =
        // sum over slots l > k in group h of F, the pairs below the slot diagonal
        // come from the symmetric raw integral: sum_{l in h} A_l W[g][l]
        std::vector<double> Ag(G, 0.0), R((size_t)G * G, 0.0);
        for (int64_t i = 0; i < N; i++) Ag[m->grp_rank[i]] += hAo[i];
        for (int64_t k = 0; k < N; k++) {
            int g = L->slot_grp[k];
            for (int h = 0; h < G; h++) {
                double v = hA[k] * hW[(size_t)h * m->NS + k];
                R[(size_t)g * G + h] += v;
                R[(size_t)h * G + g] += v;
            }
        }
        if (p->diag.p) {  // self pairs: raw_ii = A_i F_ii, once
            std::vector<double> hd(N);
            if ((rc = d2h(hd.data(), p->diag.p, sizeof(double) * N))) return rc;
            for (int64_t k = 0; k < N; k++) R[(size_t)L->slot_grp[k] * (G + 1)] += hd[k];
        }
        if (A_group) for (int g = 0; g < G; g++) A_group[g] = Ag[g];
        if (F_group)
            for (int g = 0; g < G; g++)
                for (int h = 0; h < G; h++) F_group[g + (size_t)G * h] = R[(size_t)g * G + h] / Ag[g];
    }
    if (group_tags) for (int g = 0; g < G; g++) group_tags[g] = m->tags[g];
    if (audit_count || audit_hash) {
        if (!p->audit_c.p) return fail(RVF_E_STATE, "audit was not requested (RVF_AUDIT with an obstruction set)");
        if (audit_count) if ((rc = d2h(audit_count, p->audit_c.p, sizeof(uint32_t) * N * N))) return rc;
        if (audit_hash) if ((rc = d2h(audit_hash, p->audit_h.p, sizeof(uint32_t) * N * N))) return rc;
    }
    return 0;
}

int rvf_plan_stats(const rvf_plan* p, rvf_stats* st) {
    if (!p || !st) return fail(RVF_E_ARG, "null plan");
    *st = p->st;
    st->tile_columns = p->cw;
    st->waves = p->waves;
    return 0;
}

int rvf_plan_device_ptr(const rvf_plan* p, int which, void** ptr) {
    if (!p || !ptr) return fail(RVF_E_ARG, "null plan");
    *ptr = which == 0 ? p->F.p : which == 1 ? p->area_orig.p : which == 2 ? p->rowsum.p : nullptr;
    return 0;
}

void rvf_plan_free(rvf_plan* p) {
    if (!p) return;
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    for (int k = 0; k < 8; k++) if (p->ev[k]) (void)hipEventDestroy(p->ev[k]);
    delete p;
}

int rvf_compute_view_factors(rvf_ctx* ctx, const rvf_mesh* mesh, const rvf_obs* obs, const rvf_opts* opts,
                             double* F, double* A, double* F_group, double* A_group, int32_t* group_tags,
                             rvf_stats* st) {
    rvf_opts o = *opts;
    if (F) o.flags |= RVF_WANT_F;
    rvf_plan* p = nullptr;
    int rc = rvf_plan_create(ctx, mesh, obs, &o, &p);
    if (rc) return rc;
    if ((rc = rvf_plan_run(p)) || (rc = rvf_plan_sync(p)) ||
        (rc = rvf_plan_fetch(p, F, A, nullptr, F_group, A_group, group_tags, nullptr, nullptr))) {
        rvf_plan_free(p);
        return rc;
    }
    if (st) *st = p->st;
    rvf_plan_free(p);
    return 0;
}

// ---------------------------------------------------------------------------
// Solve without device allocation (no reference counterpart): for callers that
// cannot call hipMalloc.  Every intermediate of the solve lives in `scratch`,
// one array per GPU for the life of the process; its device address is looked
// up once per device when a context is created (scratch_ptr), and host code
// uses only that address.  Calls on one GPU share the array: calls that
// overlap on one GPU race, so callers serialise them.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[RVF_NOALLOC_SCRATCH_BYTES];

#define RVF_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[RVF_MAX_DEVICES] = {};
static hipError_t g_scratch_err[RVF_MAX_DEVICES] = {};

// With `device` current: record the device address of `scratch` on it.
static void scratch_lookup(int device) {
    if (device < 0 || device >= RVF_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        g_scratch_err[device] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        (void)hipGetLastError();
        return;
    }
    g_scratch_ptr[device] = ptr;
    g_scratch_err[device] = hipSuccess;
}

static int scratch_ptr(int device, char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= RVF_MAX_DEVICES) return fail(RVF_E_ARG, "device index beyond the scratch table");
    if (!g_scratch_ptr[device])
        return fail(RVF_E_HIP, std::string("static scratch array unavailable on this device: ") +
                                   hipGetErrorString(g_scratch_err[device]));
    *out = (char*)g_scratch_ptr[device];
    return 0;
}

int rvf_compute_view_factors_noalloc(rvf_ctx* ctx, const rvf_mesh* mesh, const rvf_obs* obs, const rvf_opts* opts,
                                     double* F_dev, double* A, double* F_group, double* A_group, int32_t* group_tags,
                                     rvf_stats* st) {
    if (!ctx || !mesh || !opts) return fail(RVF_E_ARG, "rvf_compute_view_factors_noalloc: bad arguments");
    if (opts->flags & (RVF_AUDIT | RVF_GRAPH | RVF_TIMING))
        return fail(RVF_E_ARG, "rvf_compute_view_factors_noalloc: RVF_AUDIT, RVF_GRAPH and RVF_TIMING need device "
                               "allocations (use rvf_plan_create)");
    char* base = nullptr;
    int rc = scratch_ptr(ctx->device, &base);
    if (rc) return rc;
    rvf_opts o = *opts;
    o.flags &= ~RVF_WANT_F;
    if (F_dev) o.flags |= RVF_WANT_F;
    rvf_plan p;    // buffers are views into the scratch array and F_dev
    Layout lay;    // the slot layout, built for this call
    BufSource src;
    src.base = base;
    src.cap = (size_t)RVF_NOALLOC_SCRATCH_BYTES;
    rc = plan_init(ctx, mesh, obs, &o, &p, src, &lay, F_dev);
    if (!rc) rc = enqueue_solve(&p, ctx->stream, false);
    if (!rc) {
        p.ran = true;
        rc = rvf_plan_sync(&p);
    } else {
        // leave nothing of this call running on the shared array
        (void)hipStreamSynchronize(ctx->stream);
    }
    if (!rc) rc = rvf_plan_fetch(&p, nullptr, A, nullptr, F_group, A_group, group_tags, nullptr, nullptr);
    if (!rc && st) *st = p.st;
    return rc;
}

// ---------------------------------------------------------------------------
// context + misc
// ---------------------------------------------------------------------------
int rvf_ctx_create(int device, rvf_ctx** out) {
    if (!out) return fail(RVF_E_ARG, "null out");
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(RVF_E_ARG, "no such device");
    HIPCHK(hipSetDevice(device));
    scratch_lookup(device);  // per device, with the device current
    rvf_ctx* c = new rvf_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return fail(RVF_E_HIP, "hipStreamCreate failed"); }
    for (int k = 0; k < 16; k++) (void)hipEventCreate(&c->ev[k]);
    *out = c;
    return 0;
}

void rvf_ctx_destroy(rvf_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < 16; k++) if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* rvf_last_error(void) { return g_err.c_str(); }
int rvf_version(void) { return 3; }  // 3: rvf_plan_tune_waves, rvf_stats.waves, rvf_tune_entry.waves

int rvf_quad_rule(int32_t nquad, int32_t tri, double* pts, double* wts) {
    if (!pts || !wts) return fail(RVF_E_ARG, "null output");
    int n = tri ? tri_rule(nquad, pts, wts) : quad_rule(nquad, pts, wts);
    if (n < 0) return fail(RVF_E_ARG, "nquad out of range");
    return n;
}

int rvf_copy_probe(rvf_ctx* c, int64_t bytes, int32_t reps, float* ms) {
    if (!c || bytes < 16 || reps < 1) return fail(RVF_E_ARG, "rvf_copy_probe: bad arguments");
    HIPCHK(hipSetDevice(c->device));
    DBuf a, b;
    int rc;
    if ((rc = a.alloc(bytes)) || (rc = b.alloc(bytes))) return rc;
    if ((rc = rvf_launch_zero(a.p, bytes, c->stream))) return fail(RVF_E_HIP, "zero launch failed");
    HIPCHK(hipEventRecord(c->ev[14], c->stream));
    for (int k = 0; k < reps; k++)
        if ((rc = rvf_launch_copy(a.p, b.p, bytes, c->stream))) return fail(RVF_E_HIP, "copy launch failed");
    HIPCHK(hipEventRecord(c->ev[15], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[15]));
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, c->ev[14], c->ev[15]));
    if (ms) *ms = t / reps;
    return 0;
}

int rvf_ctx_event_record(rvf_ctx* c, int32_t slot) {
    if (!c || slot < 0 || slot >= 16) return fail(RVF_E_ARG, "bad event slot");
    HIPCHK(hipEventRecord(c->ev[slot], c->stream));
    return 0;
}

int rvf_ctx_event_elapsed(rvf_ctx* c, int32_t a, int32_t b, float* ms) {
    if (!c || a < 0 || a >= 16 || b < 0 || b >= 16 || !ms) return fail(RVF_E_ARG, "bad event slot");
    HIPCHK(hipEventSynchronize(c->ev[b]));
    HIPCHK(hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return 0;
}

int rvf_ctx_sync(rvf_ctx* c) {
    if (!c) return fail(RVF_E_ARG, "null ctx");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}
