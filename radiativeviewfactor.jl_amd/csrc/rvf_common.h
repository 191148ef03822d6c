// Shared device/host definitions for the MI355X view-factor library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define RVF_STRIP 64          // i-elements per pair tile (one per lane of a wavefront)
#define RVF_CHUNK 16          // max j-elements per pair tile (group-pure column chunk)
#define RVF_CHUNK_QUEUE 4     // default columns per tile of the wave-queue kernels (dynamic tile claims)
#define RVF_MAXPTS 64         // quadrature points per element (nquad <= 8 quads)
#define RVF_RULE_LINE (3 * 64 + 39)  // offset of the 1-D rule in the precompute rule table
#define RVF_RULE_SIZE (RVF_RULE_LINE + 24)

// Flattened BVH node, FP64 (64 B).  Preorder layout: an interior node's left
// child is the next node; `miss` is the DFS successor (-1 = end of tree).
struct __align__(16) BvhNode64 {
    double lo[3];
    double hi[3];
    int32_t miss;
    int32_t start;   // first triangle (absolute index into the triangle arrays)
    int32_t count;   // 0 for interior nodes
    int32_t pad;     // interior nodes: right child (absolute)
};

// FP32 node (32 B) with conservatively padded bounds.
struct __align__(16) BvhNode32 {
    float lo[3];
    int32_t miss;
    float hi[3];
    int32_t start_count;  // leaves: (start << 4) | count (count <= 15); interior: right child << 4
};

// Node of the mixed traversal trees (32 B), padded FP32 bounds.  Leaves hold
// one triangle: link = ~triangle (negative); interior: link = right child
// (the left child is the next node).  exit = the preorder index after the
// node's subtree, INT32_MAX on the tree's right spine, so a walk that
// leaves a subtree always lands at or past that subtree's own exit.
struct __align__(16) TNode32 {
    float lo[3];
    int32_t exit;
    float hi[3];
    int32_t link;
};

// Triangle for Moller-Trumbore, FP64: v0, e1 = v1 - v0, e2 = v2 - v0 as the
// reference's reading of the soup yields them (72 B, padded to 80).
struct __align__(16) Tri64 {
    double v0[3];
    double e1[3];
    double e2[3];
    double pad;
};

// FP32 triangle (48 B): v0, e1, e2 and edge magnitudes for error bounds.
struct __align__(16) Tri32 {
    float v0[3];
    float sa;        // |e1|_1 * |e2|_1, scales the |a| < 1e-10 band
    float e1[3];
    float E1;        // |e1|_1 of the FP64 edge, rounded up
    float e2[3];
    float E2;        // |e2|_1 of the FP64 edge, rounded up
};

// Traversal items for the wave-queue kernels: the nodes of every tree
// (indices [0, M)) followed by the triangles ([M, M+T)) in one array, so one
// traversal step is one record fetch whatever its kind.
//   node:     box lo[3], hi[3]; link.x = item after a passed slab test (left
//             child, or the leaf's first triangle); link.y = miss link.
//   triangle: v0[3], e1[3], e2[3]; link.x = next item (next triangle of the
//             leaf, or the leaf's miss link); link.y = the leaf node.
// Links are -1 at the end of a tree.  FP32 boxes are conservatively padded.
struct __align__(16) Item32 {
    float v[9];
    int32_t link_x, link_y;
    int32_t pad;
};
struct __align__(16) Item64 {
    double v[9];
    int32_t link_x, link_y;
};

// Pair-kernel kinds: (enumerator, kernel tag).  Each kind is a separate
// kernel k_pair<tag>, so profiles name the variant that ran.
#define RVF_PAIR_KINDS(X)                                                                                    \
    X(FP64_PC4, fp64_pc4) X(FP64_PC9, fp64_pc9) X(FP64_PC16, fp64_pc16) X(FP64_OCC_LANE, fp64_occ_lane)     \
    X(MIX_PC4, mixed_pc4) X(MIX_PC9, mixed_pc9) X(MIX_PC16, mixed_pc16) X(MIX_PC16C, mixed_pc16_chunked)    \
    X(MIX_OCC_LANE, mixed_occ_lane) X(MC, monte_carlo)                                                      \
    X(FP64_QUEUE1, fp64_occ_queue1) X(FP64_QUEUE2, fp64_occ_queue2) X(FP64_QUEUE3, fp64_occ_queue3)        \
    X(MIX_QUEUE2, mixed_occ_queue2) X(MIX_QUEUE3, mixed_occ_queue3) X(MIX_QUEUE4, mixed_occ_queue4)        \
    X(MIX_QUEUE6, mixed_occ_queue6) X(FP64_SPLIT1, fp64_occ_split1) X(FP64_SPLIT2, fp64_occ_split2)         \
    X(FP64_SPLIT3, fp64_occ_split3) X(MIX_SPLIT1, mixed_occ_split1) X(MIX_SPLIT2, mixed_occ_split2)         \
    X(MIX_SPLIT3, mixed_occ_split3) X(FP64_2D, fp64_curve) X(MIX_2D, mixed_curve)                           \
    X(FP64_OCC2D, fp64_curve_occ) X(MIX_OCC2D, mixed_curve_occ)
enum RvfPairKind {
#define RVF_KIND_ENUM(e, t) RVF_K_##e,
    RVF_PAIR_KINDS(RVF_KIND_ENUM)
#undef RVF_KIND_ENUM
    RVF_K_COUNT
};
// The wave-queue traversal kinds (queue and split forms, both precisions) are
// consecutive in the list: their launches are persistent, claiming tiles
// dynamically, and their default tiles are RVF_CHUNK_QUEUE columns wide.
constexpr bool rvf_queue_kind(int kind) { return kind >= RVF_K_FP64_QUEUE1 && kind <= RVF_K_MIX_SPLIT3; }
static_assert(RVF_K_MIX_SPLIT3 - RVF_K_FP64_QUEUE1 == 12, "wave-queue kinds must stay consecutive");
// the split kinds (an LDS triangle-task ring), consecutive too
constexpr bool rvf_split_kind(int kind) { return kind >= RVF_K_FP64_SPLIT1 && kind <= RVF_K_MIX_SPLIT3; }
static_assert(RVF_K_MIX_SPLIT3 - RVF_K_FP64_SPLIT1 == 5, "split kinds must stay consecutive");
// the mixed split kinds, consecutive too (their refill default differs)
constexpr bool rvf_mixed_split_kind(int kind) { return kind >= RVF_K_MIX_SPLIT1 && kind <= RVF_K_MIX_SPLIT3; }
static_assert(RVF_K_MIX_SPLIT3 - RVF_K_MIX_SPLIT1 == 2, "mixed split kinds must stay consecutive");
#define RVF_QCOMBO 128   // (p, q) combinations per wave-queue batch
#define RVF_TQCAP 2048   // triangle tasks per wavefront ring (FP64 split variants)
#define RVF_TQCAP_MIX 256   // the same for the mixed split variants (single-triangle leaves)
#define RVF_TLEAF_MAX 1  // triangles per leaf of the mixed traversal trees
#define RVF_FRONTIER 16  // mixed split variants: subtrees a pair's rays start from (per lane, LDS)
#define RVF_CQCAP 128    // mixed split variants: candidates for the exact FP64 test
#define RVF_QSTATS 10    // per-wave statistics of the wave-queue kernels (LDS, u64)
#define RVF_NCOUNTERS 80 // a plan's u64 counter slots: statistics, the tile claim counter (slot 15), 8 per-XCD claim counters
// A split-variant triangle task packs (tag << RVF_TQ_TRI_BITS) | triangle into
// 32 bits (tag = combination << 6 | lane < 2^13), so the split variants serve
// obstruction sets of fewer than 2^19 triangles over all their trees.
#define RVF_TQ_TRI_BITS 19

// Column chunk: a run of <= RVF_CHUNK consecutive elements of one group.  The
// tiles of a chunk are consecutive in the tile list (strips 0, 1, ...), from
// first_tile on, so a row's partial of chunk c sits at tile first_tile + strip.
struct Chunk {
    int32_t begin;
    int32_t end;     // exclusive
    int32_t grp;     // group rank
    int32_t first_tile;
};

struct PairArgs {
    const int2* tiles;        // (strip, chunk)
    int ntiles;               // tiles of this launch (grid of a direct launch)
    int tile0;                // first tile of a direct launch (tile = tile0 + block * tile_step)
    int tile_step;            // tile stride of a direct launch (1 but for sampled timing)
    int pad_ts;
    int* tile_ctr;            // wave-queue kernels: tile claim counter (zeroed before each launch)
    int NS;                   // slots = N rounded up to RVF_STRIP
    int maxpts;               // per-element point stride in Q arrays
    int64_t N;                // elements
    const Chunk* chunks;
    const int32_t* xorder;    // wave-queue kernels, whole launches: tile indices grouped by XCD, or null
    int xseg[9];              // XCD x's tiles are xorder[xseg[x] .. xseg[x+1])
    const int32_t* slot_orig; // slot -> original element index (-1 = padding)
    const int32_t* elem_np;   // points per slot
    const int32_t* elem_grp;  // group rank per slot
    const double* area;       // per slot
    // FP64 point data
    const double* P64;        // point-major [(p*8+c)*NS + e]: x y z nx ny nz w dA
    const double* Q64;        // element-major [(e*maxpts + p)*8 + c]
    // FP32 point data (mixed)
    const float* P32;         // point-major [(p*8+c)*NS + e]: dx dy dz nx ny nz wdA 0
    const float* Q32;         // element-major [(e*maxpts + p)*8 + c]
    const double* anchor;     // [e*4 + c]
    // outputs
    double* F;                // N*N col-major: raw_ij at [i + N*j], i < j (nullable)
    double* part;             // [ntiles * 64]: per tile and lane, sum_{j in chunk, j > i} F_ij
    double* partT;            // [nstrips * NS]: sum_{i in strip, i < j} F_ji
    // obstruction
    int G;                    // radiating groups
    const int32_t* tree_root; // [G*G] root node or -1
    const BvhNode64* nodes64;
    const Tri64* tris64;
    const BvhNode32* nodes32;
    const Tri32* tris32;
    const int32_t* parent;    // per node, absolute; -1 at a root
    const Item32* items32;    // wave-queue traversal records (nodes then triangles)
    const Item64* items64;
    // mixed split traversal: a tree of its own over the same triangles (binned
    // SAH, small leaves), FP32 padded boxes; triangles in its leaf order and
    // their indices in the reference trees' order (exact test, leaf check)
    const int32_t* ttree_root; // [G*G] root node or -1
    const TNode32* tnodes32;
    const Tri32* ttris32;
    const int32_t* tmap;
    int n_nodes;              // M: items below M are nodes
    int refill_min;           // wave-queue kernels: refill a slot once this many lanes have it empty
    int verify;               // mixed: confirm FP64 hits by the leaf's exact slab chain
    int cull_limit;           // mixed occluded: box tests a pair's capsule walk may spend (1: root box only)
    int hint;                 // mixed split traversal: try the pair's last certain blocker before walking
    int waves;                // mixed split kinds, direct launches: 3 = the build for three waves per SIMD, else 0
    float Sm;                 // mixed: scene coordinate magnitude (FP32 error bounds)
    // audit (nullable): per pair blocked-ray count / order-free hash at [i + N*j]
    uint32_t* audit_count;
    uint32_t* audit_hash;
    // element topology (Duffy singular-pair skip, Monte Carlo sampling)
    const double* coords;     // 3 x Nn
    const int8_t* fam;        // N
    const int32_t* nodes8;    // 8 x N, 1-based
    int duffy;                // skip Quad8 pairs sharing a corner (done by k_duffy)
    int dim2;                 // curve mesh (mesh_dim 1): 2-D kernel, segment occluders in nodes64 / tris64
    int self_vf;              // 1: the pairs (i, i) too, raw value to diag (not to part / partT)
    double* diag;             // [NS] raw_ii per slot (self_vf)
    int mc_nsamp, mc_s;       // Monte Carlo samples per pair, strata per side
    uint64_t mc_seed;
    // [0] FP64 re-tests (mixed), [1] rays cast, [2] AABB tests, [3] triangle
    // tests (FP64: canonical stackless order, GPUBVH.jl:202-285)
    unsigned long long* counters;
};

// Inputs/outputs of the per-element precompute kernel.
struct PrecompArgs {
    const double* coords;
    const int8_t* fam;
    const int32_t* nodes8;     // N*8, 1-based
    int64_t N;
    int NS, maxpts;
    int nq_quad, nq_tri, nq_line;
    const double* rule;        // [quad pts 2*64][quad w 64][tri pts 2*13][tri w 13][line pts (xi,0) 2*8][line w 8]
    double* P64; double* Q64; float* P32; float* Q32; double* anchor; double* area;
    int32_t* elem_np;
    int* bad;                  // set to 1 if any dA is 0 or not finite
    const int32_t* slot_orig;  // slot -> original element
    double* area_orig;         // areas in original element order
};

// One singular Quad8 pair for the Duffy / Sauter-Schwab kernel.
struct DuffyPair {
    int32_t i, j;        // i < j
    int32_t kind;        // 1 = common vertex, 2 = common edge
    int32_t ci0, ci1, cj0, cj1;  // shared local corners (1-based), first two in scan order
    int32_t root;        // obstruction tree or -1
};
