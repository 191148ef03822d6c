/*
 * rvf.h — C ABI of the MI355X element-pair view-factor library
 * (librvf_hip.so).  Plain pointers and sizes only; no C++ exceptions cross
 * it.  Every entry point returns 0 on success and a negative RVF_E* code on
 * failure; rvf_last_error() then holds a message (thread-local).  The
 * library never calls abort() or exit().
 *
 * Each entry point names the reference interface it replaces
 * (rot4te/RadiativeViewFactor.jl v0.5.0, file:line).
 */
#ifndef RVF_H
#define RVF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RVF_OK 0
#define RVF_E_ARG (-1)       /* invalid argument / mesh */
#define RVF_E_HIP (-2)       /* HIP runtime error (incl. device faults) */
#define RVF_E_NOMEM (-3)     /* device allocation failed / admission refused */
#define RVF_E_STATE (-4)     /* object used in a wrong state */
#define RVF_E_CANCELLED (-5) /* serving: request cancelled */
#define RVF_E_TIMEOUT (-6)   /* serving: wait timed out */

/* family codes: src/GPUKernels.jl:344-375 */
#define RVF_QUAD8 0
#define RVF_TRI6 1
#define RVF_QUAD4 2
#define RVF_TRI3 3

/* methods: compute_view_factors keywords monte_carlo / use_duffy
 * (src/Assembly.jl:64-73) */
#define RVF_METHOD_QUAD 0
#define RVF_METHOD_MC 1
#define RVF_METHOD_DUFFY 2

#define RVF_FP64 0   /* reference arithmetic, bit-faithful */
#define RVF_MIXED 1  /* FP32 bulk, FP64 accumulation and refinement */

/* soup reading (src/BVH.jl:81-109 vs 232-240, quirk Q1) */
#define RVF_SOUP_REFERENCE 0    /* AABB reads [xyz,vertex], Moller-Trumbore reads [vertex,xyz] */
#define RVF_SOUP_XYZ_VERTEX 1   /* geometric reading of a [xyz,vertex,tri] soup (load_mesh) */
#define RVF_SOUP_VERTEX_XYZ 2   /* geometric reading of a [vertex,xyz,tri] soup */
#define RVF_SOUP_SEGMENTS 3     /* curve meshes: segment soup (3,2,T) [xyz,endpoint,seg] (src/MeshIO.jl:641-645) */

#define RVF_WANT_F 1     /* materialise the N x N element matrix F on the device */
#define RVF_AUDIT 2      /* per-pair blocked-ray count + hash (occluded solves) */
#define RVF_TIMING 4     /* record per-kernel HIP event timings */
#define RVF_GRAPH 8      /* capture the solve's launch sequence once and replay it as a hipGraph */
#define RVF_SELF_VF 16   /* self view factors: integrate the pairs (i, i) too (src/Assembly.jl:173,193) */

typedef struct rvf_ctx rvf_ctx;
typedef struct rvf_mesh rvf_mesh;
typedef struct rvf_obs rvf_obs;
typedef struct rvf_plan rvf_plan;

typedef struct rvf_opts {
    int32_t nquad;      /* Gauss points per direction (1..8) */
    int32_t method;     /* RVF_METHOD_* */
    int32_t precision;  /* RVF_FP64 / RVF_MIXED */
    int32_t n_samples;  /* Monte Carlo samples per pair */
    uint64_t seed;      /* Monte Carlo stream seed */
    int32_t flags;      /* RVF_WANT_F | RVF_AUDIT | RVF_TIMING | RVF_GRAPH | RVF_SELF_VF */
    int32_t variant;    /* occluded quadrature traversal: RVF_VARIANT_* (0 = library default) */
} rvf_opts;

/* Occluded-quadrature traversal variants (rvf_opts.variant).  Results do not
 * depend on the variant: FP64 visibility decisions and sums are identical. */
#define RVF_VARIANT_DEFAULT 0
#define RVF_VARIANT_LANE 1        /* each lane traces its own (p, q) rays, one at a time */
#define RVF_VARIANT_QUEUE(S) (10 + (S))  /* wave-shared ray queue, S rays in flight per lane:
                                           FP64 S = 1..3, mixed S = 2, 3, 4, 6 */
#define RVF_VARIANT_SPLIT(S) (20 + (S))  /* as QUEUE, slots walk nodes only; triangle tests run
                                           wave-wide from an LDS task ring: S = 1..3 */

typedef struct rvf_stats {
    double precompute_ms, pair_ms, duffy_ms, fill_ms, reduce_ms, total_ms;
    int64_t n_elems, unique_pairs, tiles;
    int64_t rays;       /* rays cast (occluded solves) */
    int64_t retests;    /* FP64 ray-triangle re-tests (mixed mode) */
    int64_t singular_pairs;
    int64_t aabb_tests; /* FP64 mode: box tests in the stackless order */
    int64_t tri_tests;  /* FP64 mode: triangle tests in the stackless order */
    /* wave-queue traversal (occluded quadrature): loop iterations per wavefront
     * summed, slot steps that advanced a ray, triangle-task and exact-test
     * batches (split variants) */
    int64_t trav_iters, slot_steps, task_batches, exact_batches;
    /* mixed mode, split traversal: triangles the FP32 classifier proved hit
     * (no FP64 re-test needed) */
    int64_t sure_hits;
    /* occluded quadrature, wave-queue kernels: rays of pairs culled before any
     * traversal (their rays cannot reach an occluder); counted in `rays` */
    int64_t culled_rays;
    /* occluded quadrature, wave-queue kernels: (wavefront, column element)
     * passes that skipped the trace because every pair of the wave was culled */
    int64_t culled_waves;
    /* columns (column elements) per pair tile of the plan's layout */
    int64_t tile_columns;
    /* pair-kernel build the plan launches (rvf_plan_tune_waves): 3 = the mixed
     * split kernels' three-waves-per-SIMD build, 0 = the kernel's default */
    int64_t waves;
} rvf_stats;

/* --- lifecycle (no reference counterpart: the reference is one Julia process) */
int rvf_ctx_create(int device, rvf_ctx** out);
void rvf_ctx_destroy(rvf_ctx* ctx);
const char* rvf_last_error(void);
int rvf_version(void);

/* --- mesh upload: replaces build_gpu_arrays (src/GPUKernels.jl:336-412).
 * coords: 3 x n_nodes column-major; family/elem_node_idx/elem_group: per
 * element; nodes[f]: (8|6|4|3) x n_per_family[f] column-major 1-based node ids,
 * elem_node_idx is 1-based into its family's matrix (as the reference builds
 * them). */
int rvf_mesh_upload(rvf_ctx* ctx, const double* coords, int64_t n_nodes, int64_t n_elems,
                    const int8_t* family, const int32_t* elem_node_idx, const int32_t* elem_group,
                    const int32_t* const nodes[4], const int64_t n_per_family[4], rvf_mesh** out);
/* curve meshes (mesh_dim 1): family 4 = Line3 [left, right, midpoint], 5 =
 * Line2; line3_nodes 3 x n_line3, line2_nodes 2 x n_line2 column-major.  The
 * reference computes these on the CPU only (src/Assembly.jl:83-86; kernel
 * src/ViewFactorKernel.jl:169-180, segment occlusion src/BVH.jl:254-322);
 * obstruction soups of a curve mesh are segment soups (RVF_SOUP_SEGMENTS). */
int rvf_curve_mesh_upload(rvf_ctx* ctx, const double* coords, int64_t n_nodes, int64_t n_elems,
                          const int8_t* family, const int32_t* elem_node_idx, const int32_t* elem_group,
                          const int32_t* line3_nodes, int64_t n_line3, const int32_t* line2_nodes, int64_t n_line2,
                          rvf_mesh** out);
void rvf_mesh_free(rvf_mesh* mesh);
int rvf_mesh_info(const rvf_mesh* mesh, int64_t* n_elems, int32_t* n_groups, int32_t* group_tags);

/* --- obstruction: replaces build_flat_bvh_from_mesh (src/GPUBVH.jl:146-174)
 * with the CPU path's semantics (src/Assembly.jl:119-143): one BVH per
 * distinct active set obstruction_groups \ {g_i, g_j}, built bit-identically
 * to build_bvh (src/BVH.jl:71-148).  soup: (3,3,T) column-major, triangles of
 * each group contiguous in ascending group order, tri_group: tag per triangle.
 * Curve meshes: soup_read RVF_SOUP_SEGMENTS and a (3,2,T) segment soup. */
int rvf_obs_build(rvf_ctx* ctx, const rvf_mesh* mesh, const double* soup, const int32_t* tri_group, int64_t T,
                  const int32_t* obstruction_groups, int32_t n_obs, int32_t soup_read, rvf_obs** out);
void rvf_obs_free(rvf_obs* obs);
/* inspection: number of distinct trees / nodes / triangles */
int rvf_obs_info(const rvf_obs* obs, int32_t* n_trees, int64_t* n_nodes, int64_t* n_tris);
/* export tree k in the reference's layout: box (6 x nodes: lo, hi), meta
 * (5 x nodes: left, right, tri_start, tri_count, miss; 1-based, 0 = none),
 * tri_idx (1-based permutation into the active set's merged soup). */
int rvf_obs_export(const rvf_obs* obs, int32_t tree, int64_t cap, double* box, int32_t* meta, int32_t* tri_idx,
                   int64_t* n_nodes, int64_t* n_tris);

/* host-only build of one tree from a soup (no device needed), same layout as
 * rvf_obs_export; for checking the builder against build_bvh (src/BVH.jl:71-148). */
int rvf_bvh_build_host(const double* soup, int64_t T, int32_t soup_read, int64_t cap, double* box, int32_t* meta,
                       int32_t* tri_idx, int64_t* n_nodes);

/* --- solve: replaces the GPU hook target compute_view_factors_gpu
 * (src/GPUAssembly.jl:32-105) through launch_vf_kernel! / launch_mc_kernel!
 * (src/GPUKernels.jl:454-500, src/GPUMCKernels.jl:295-328).  A plan owns the
 * device buffers; rvf_plan_run enqueues all kernels on the context stream. */
int rvf_plan_create(rvf_ctx* ctx, const rvf_mesh* mesh, const rvf_obs* obs /*nullable*/, const rvf_opts* opts,
                    rvf_plan** out);
int rvf_plan_run(rvf_plan* plan);
int rvf_plan_sync(rvf_plan* plan);
/* Results, host copies (each pointer nullable):
 *   F N x N column-major (requires RVF_WANT_F), A N, rowsum N,
 *   F_group G x G column-major over sorted tags, A_group G, group_tags G,
 *   audit_count / audit_hash N x N (upper triangle i<j at i + N*j). */
int rvf_plan_fetch(rvf_plan* plan, double* F, double* A, double* rowsum, double* F_group, double* A_group,
                   int32_t* group_tags, uint32_t* audit_count, uint32_t* audit_hash);
int rvf_plan_stats(const rvf_plan* plan, rvf_stats* st);
/* Page-locked host memory (no counterpart in the reference: CUDA.jl copies F
 * into pageable memory, src/GPUAssembly.jl:83-84).  rvf_plan_fetch into such
 * a buffer is one DMA transfer at PCIe speed instead of a staged copy; reuse
 * it across solves.  rvf_host_free releases it (null is ignored). */
int rvf_host_alloc(size_t bytes, void** out);
int rvf_host_free(void* ptr);
/* device pointer of an output (0 = F, 1 = A, 2 = rowsum) for device-side use */
int rvf_plan_device_ptr(const rvf_plan* plan, int which, void** ptr);
/* Tuning knobs of an existing plan (no counterpart in the reference; used by
 * the autotuner): traversal variant (RVF_VARIANT_*, 0 = default) and the
 * wave-queue refill threshold (lanes with an empty slot before a refill,
 * 1..64; 0 = the variant's default: 24 for the mixed split variants, 16
 * otherwise).  F, A and every visibility decision do not depend on either. */
int rvf_plan_tune(rvf_plan* plan, int32_t variant, int32_t refill_min);
/* Columns (column elements) per pair tile, 1..16; 0 = the pair kernel's
 * default (4 for the wave-queue traversal variants, whose launches claim
 * tiles dynamically; 16 otherwise).  No reference counterpart.  F, A and
 * every visibility decision do not depend on it; row sums and group sums are
 * added in another order (differences of a few ulps). */
int rvf_plan_tune_tiles(rvf_plan* plan, int32_t columns);
/* Occupancy of the pair kernel (no reference counterpart): resident waves per
 * SIMD its build is compiled for.  0 = the kernel's default; the mixed split
 * traversal kernels (default 4 waves at 128 VGPRs) also have a build for 3
 * (every value in registers, no spills), chosen with 3 (4 = their default).
 * Other values, or 3 / 4 with another pair kernel, return RVF_E_ARG.  The
 * setting goes back to 0 when rvf_plan_tune moves to another kernel.  F, A and
 * every visibility decision do not depend on it. */
int rvf_plan_tune_waves(rvf_plan* plan, int32_t waves);
/* Name of the pair kernel the plan launches (as profilers show it), e.g.
 * "k_pair<mixed_occ_split3>" (no reference counterpart; for measurement). */
int rvf_plan_kernel(const rvf_plan* plan, char* name, int32_t cap);
/* One measured configuration of rvf_plan_autotune. */
typedef struct {
    int32_t variant;     /* RVF_VARIANT_* */
    int32_t refill_min;  /* wave-queue refill threshold (0 where it does not apply) */
    int64_t tiles;       /* tiles timed */
    double ms;           /* pair-kernel time over those tiles */
    int32_t columns;     /* columns per tile of the layout timed */
    int32_t waves;       /* rvf_plan_tune_waves setting timed (0 = the kernel's default) */
} rvf_tune_entry;
/* Autotuner (no reference counterpart): runs the plan's prologue, then times
 * its pair kernel under every traversal variant valid for its precision (each
 * at its default tile width) over a strided share of the tiles (sample_tiles
 * of the plan's current layout, 0 = 8192; the same share of the pairs for
 * every width), times the library default twice and the three fastest others
 * once over the whole problem (or four times the share when the problem is
 * larger), tries the refill thresholds of 8, 16, 24 and 32 other than the
 * choice's default for it, then its tile widths of 4, 8 and 16 columns, then
 * (mixed split kernels) their three-wave build, and keeps the default unless
 * another configuration is faster by more than twice the default's
 * run-to-run spread (at least 1 %), applied as rvf_plan_tune,
 * rvf_plan_tune_tiles and rvf_plan_tune_waves would (*variant, *refill_min;
 * the width and build in rvf_stats.tile_columns and .waves).  Writes up to
 * cap measurements to entries and their count to *n.  Plans without
 * obstruction trees have a single pair kernel: *n = 0 and nothing changes.
 * F, A and every visibility decision of later runs are the same whichever
 * configuration is kept. */
int rvf_plan_autotune(rvf_plan* plan, int64_t sample_tiles, rvf_tune_entry* entries, int32_t cap, int32_t* n,
                      int32_t* variant, int32_t* refill_min);
void rvf_plan_free(rvf_plan* plan);

/* One-shot: the body of the Julia hook (compute_view_factors_gpu,
 * src/GPUAssembly.jl:32-105) — upload, plan, run, fetch. */
int rvf_compute_view_factors(rvf_ctx* ctx, const rvf_mesh* mesh, const rvf_obs* obs, const rvf_opts* opts,
                             double* F, double* A, double* F_group, double* A_group, int32_t* group_tags,
                             rvf_stats* st);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  rvf_compute_view_factors without any device allocation:
 * the solve's intermediates (point tables, slot layout, tiles, partial and
 * row sums, counters, Duffy pair lists) live in one static device array of
 * RVF_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; F, when
 * F_dev is non-null, is written to that caller-owned device buffer (N x N
 * doubles, column-major), the other results to host memory as
 * rvf_compute_view_factors writes them.  A solve whose intermediates exceed
 * the array (at nquad 3 about 5,000 elements with occluders in mixed precision,
 * 3,700 in FP64, 10,700 without occluders)
 * returns RVF_E_NOMEM with the size it needs; RVF_AUDIT, RVF_GRAPH and
 * RVF_TIMING are refused (RVF_E_ARG).
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * race and must be serialised by the caller.  Returns when the solve is done. */
#define RVF_NOALLOC_SCRATCH_BYTES (64ll << 20)
int rvf_compute_view_factors_noalloc(rvf_ctx* ctx, const rvf_mesh* mesh, const rvf_obs* obs, const rvf_opts* opts,
                                     double* F_dev, double* A, double* F_group, double* A_group,
                                     int32_t* group_tags, rvf_stats* st);

/* --- quadrature tables as the device sees them (gauss_legendre_2d,
 * src/Quadrature.jl:96-109; tri_quad_rule, src/ViewFactorKernel.jl:42-71) */
int rvf_quad_rule(int32_t nquad, int32_t tri, double* pts, double* wts);

/* --- timing on the context stream */
int rvf_ctx_event_record(rvf_ctx* ctx, int32_t slot);
int rvf_ctx_event_elapsed(rvf_ctx* ctx, int32_t slot_a, int32_t slot_b, float* ms);
int rvf_ctx_sync(rvf_ctx* ctx);

/* --- request server: continuous batching of solves on one GPU (no reference
 * counterpart: the reference solves one mesh per call, src/Assembly.jl:64-96).
 * Clients on any thread submit solves; a worker thread runs rounds in which one
 * persistent batched kernel per kernel kind drains (request, tile) items from
 * every running request, and admits pending requests in arrival order while
 * their estimated device footprint fits hbm_budget.  A request's result is
 * identical to its standalone rvf_compute_view_factors result. */
/* --- mesh files (host only; no GPU) -------------------------------------- */
/* Reads a mesh file into host arrays, as load_mesh (src/MeshIO.jl:94-145) and
 * load_vtu (ext/RadiativeViewFactorReadVTKExt.jl:71-157) do: Gmsh .msh in
 * format 2.2 or 4.1 (ASCII or binary), or an XML VTK unstructured grid (.vtu,
 * or .vtk in XML form; ascii, base64 or appended data, raw or zlib).
 * Elements of the supported families in every physical group of dimension
 * surface_dim: 2 = surfaces (Quad9 read as Quad8), 1 = curves (Line2, Line3;
 * a .msh file's line elements are turned to face the surface next to their
 * curve, as _orient_line3_normals!, src/MeshIO.jl:469-599); without physical
 * groups a single "default" group over every entity of that dimension; VTK
 * groups from a per-cell integer array (group_field, or NULL to try
 * CellEntityIds, gmsh:physical, RegionId, MaterialIds, region, group).
 * reverse_normals swaps nodes as _reverse_all_normals! (src/MeshIO.jl:613-632). */
typedef struct rvf_meshfile rvf_meshfile;
int rvf_meshfile_read(const char* path, int32_t surface_dim, int32_t reverse_normals, const char* group_field,
                      rvf_meshfile** out);
/* flags: 1 = default group synthesised, 2 = Quad9 centre nodes dropped,
 * 4 = the requested VTK group array was absent (single default group),
 * 8 = some line elements had no curve entity or adjacent surface to orient by,
 * 16 = some line elements were reoriented */
int rvf_meshfile_info(const rvf_meshfile* mf, int64_t* n_nodes, int64_t* n_elems, int32_t* n_groups,
                      int32_t* flags);
/* coords 3 x Nn column-major; family N (0 Quad8, 1 Tri6, 2 Quad4, 3 Tri3, 4 Line3, 5 Line2);
 * nodes 8 x N column-major, 1-based, unused slots 0; group N (tags) */
int rvf_meshfile_arrays(const rvf_meshfile* mf, double* coords, int8_t* family, int32_t* nodes, int32_t* group);
/* group k (ascending tag order): tag and name; returns the name's length */
int rvf_meshfile_group(const rvf_meshfile* mf, int32_t k, int32_t* tag, char* name, int32_t cap);
void rvf_meshfile_free(rvf_meshfile* mf);

typedef struct rvf_server rvf_server;
typedef struct rvf_server_opts {
    int64_t hbm_budget;   /* device bytes for admitted requests (0 = 64 GiB) */
    int32_t max_active;   /* requests admitted at once (0 = 64) */
    int32_t round_tiles;  /* pair tiles per request per round of the batched kernel (0 = 4096) */
    int32_t direct_tiles; /* requests with at least this many tiles get one launch of their whole
                             tile range on a stream of their own, smaller ones rounds of the
                             batched kernel (0 = 512, -1 = never) */
    int32_t reserved;
} rvf_server_opts;
typedef struct rvf_request_stats {
    double t_submit_ms, t_admit_ms, t_start_ms, t_done_ms;  /* since the server started */
    int64_t footprint_bytes, tiles;
} rvf_request_stats;
typedef struct rvf_server_counters {
    int64_t submitted, completed, failed, cancelled, rounds, tiles;
    int64_t peak_bytes, used_bytes, pending, active;
    /* worker time: launching (direct ranges and batched rounds), admitting
     * (upload, BVH, plan), waiting for the device, retiring (result copies) */
    double launch_ms, admit_ms, wait_ms, retire_ms;
    int64_t batched_tiles;  /* of `tiles`, those run by the persistent batched kernel */
} rvf_server_counters;
/* request states returned by rvf_server_poll / rvf_server_wait */
#define RVF_REQ_PENDING 0
#define RVF_REQ_RUNNING 1
#define RVF_REQ_DONE 2
#define RVF_REQ_FAILED 3
#define RVF_REQ_CANCELLED 4
int rvf_server_create(rvf_ctx* ctx, const rvf_server_opts* opts /*nullable*/, rvf_server** out);
void rvf_server_destroy(rvf_server* server);  /* cancels pending requests, finishes running ones */
/* mesh arguments as rvf_mesh_upload, obstruction arguments as rvf_obs_build
 * (soup nullable: no obstruction); copied, so the caller may free them at once.
 * opts.flags RVF_WANT_F keeps F for rvf_server_result. */
int rvf_server_submit(rvf_server* server, const double* coords, int64_t n_nodes, int64_t n_elems,
                      const int8_t* family, const int32_t* elem_node_idx, const int32_t* elem_group,
                      const int32_t* const nodes[4], const int64_t n_per_family[4], const double* soup,
                      const int32_t* tri_group, int64_t T, const int32_t* obs_groups, int32_t n_obs,
                      int32_t soup_read, const rvf_opts* opts, int64_t* ticket);
int rvf_server_poll(rvf_server* server, int64_t ticket);                     /* RVF_REQ_* */
int rvf_server_wait(rvf_server* server, int64_t ticket, int32_t timeout_ms); /* -1 = no limit */
/* results as rvf_plan_fetch (each pointer nullable) plus the request's timeline */
int rvf_server_result(rvf_server* server, int64_t ticket, double* F, double* A, double* rowsum, double* F_group,
                      double* A_group, int32_t* group_tags, rvf_request_stats* stats);
int rvf_server_cancel(rvf_server* server, int64_t ticket);
int rvf_server_release(rvf_server* server, int64_t ticket);  /* forget a finished request */
int rvf_server_stats(rvf_server* server, rvf_server_counters* stats);

/* measurement aid (no reference counterpart): copy `bytes` device-to-device
 * `reps` times with the library's 16-byte-vector copy kernel; average ms.
 * Used to calibrate HBM traffic counters against a kernel of known size. */
int rvf_copy_probe(rvf_ctx* ctx, int64_t bytes, int32_t reps, float* ms);

#ifdef __cplusplus
}
#endif
#endif
