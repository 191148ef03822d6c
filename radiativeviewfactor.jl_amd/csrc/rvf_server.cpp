// Request server: continuous batching of view-factor solves on one GPU.
//
// Clients submit solves (mesh, optional obstruction soup, options) from any
// thread; one worker thread owns the device streams.  The worker loops:
//   1. admits pending requests in arrival order while their estimated HBM
//      footprint fits the budget and a slot is free: uploads the mesh, builds
//      the BVHs, creates the plan and enqueues its prologue (point precompute)
//      on the context stream;
//   2. a large request (at least direct_tiles tiles) gets its whole tile range
//      as a direct launch, then its epilogue, on the slot stream with the
//      fewest tiles in flight, ordered after its prologue by an event; the
//      slot streams run concurrently, so
//      requests that arrive while others run start at once and share the
//      device at workgroup granularity;
//   3. small requests share rounds: per pair-kernel kind, one persistent
//      batched kernel drains a device list of up to round_tiles (request,
//      tile) items per request in one launch on the context stream, followed
//      by the epilogues of the requests it finished;
//   4. retires requests whose epilogue event has completed: copies their
//      results to host memory and frees their device memory, which lets
//      further requests in.
// A request's tiles are the ones its standalone plan launches, computed by
// the same device code, and its reductions run per request, so a result is
// bit-identical whether it was solved alone or in a mix.
//
// No counterpart in the reference, which solves one mesh per call
// (src/Assembly.jl:64-96); the entry points are declared in include/rvf.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rvf.h"
#include "rvf_common.h"

// internal interface of rvf_host.cpp
int rvfi_plan_prologue(rvf_plan* p, hipStream_t st);
int rvfi_plan_epilogue(rvf_plan* p, hipStream_t st);
int rvfi_plan_collect(rvf_plan* p);
int rvfi_plan_batch(const rvf_plan* p, PairArgs* a, int* kind, size_t* lds);
int rvfi_plan_columns(const rvf_plan* p);
size_t rvfi_plan_bytes(const rvf_plan* p);
hipStream_t rvfi_ctx_stream(const rvf_ctx* c);
int rvfi_ctx_device(const rvf_ctx* c);
int rvfi_fail(int code, const std::string& msg);
extern "C" int rvf_launch_pair(const PairArgs* a, int kind, size_t lds, hipStream_t st);
extern "C" int rvf_launch_pair_batched(int kind, const PairArgs* reqs, const int2* work, const int* ctl, int grid,
                                       size_t lds, hipStream_t st);
extern "C" int rvf_batched_blocks_per_cu(int kind, size_t lds);
extern "C" int rvf_launch_zero(void* p, int64_t bytes, hipStream_t st);

namespace {

enum { REQ_PENDING = 0, REQ_RUNNING = 1, REQ_DONE = 2, REQ_FAILED = 3, REQ_CANCELLED = 4 };

struct Req {
    int64_t id = 0;
    int state = REQ_PENDING;
    bool cancel = false;
    // inputs (host copies, freed once the device objects exist)
    std::vector<double> coords;
    int64_t Nn = 0, N = 0;
    std::vector<int8_t> fam;
    std::vector<int32_t> eni, grp;
    std::vector<int32_t> nodes[4];
    int64_t nfam[4] = {0, 0, 0, 0};
    std::vector<double> soup;
    std::vector<int32_t> tri_group, obs_groups;
    int64_t T = 0;
    int32_t soup_read = 0;
    bool has_obs = false;
    rvf_opts opts{};
    // device side
    rvf_mesh* mesh = nullptr;
    rvf_obs* obs = nullptr;
    rvf_plan* plan = nullptr;
    PairArgs args{};
    int kind = 0, slot = -1, ntiles = 0, next_tile = 0;
    int cols = RVF_CHUNK;              // columns per tile of the request's layout
    int stream = -1;                   // slot stream of a direct request
    bool epilogue = false;             // epilogue enqueued: retire once done_ev completes
    hipEvent_t ready_ev = nullptr;     // prologue done (context stream)
    hipEvent_t done_ev = nullptr;      // epilogue done
    size_t lds = 0, est_bytes = 0;
    // results
    std::vector<double> F, A, rowsum, Fg, Ag;
    std::vector<int32_t> tags;
    std::string err;
    rvf_request_stats st{};
};

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

struct rvf_server {
    rvf_ctx* ctx = nullptr;
    rvf_server_opts o{};
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    int64_t next_id = 1;
    std::deque<Req*> pending;
    std::vector<Req*> active;          // admitted, owned by the worker
    std::map<int64_t, Req*> all;
    size_t used = 0;                   // estimated bytes of admitted requests
    std::vector<int> free_slots;
    // Device objects of finished requests wait here for the reaper thread.
    // Freeing device memory waits for all work queued on the device and holds
    // up the worker's launches meanwhile, so the reaper frees only when the
    // memory is wanted: an admission waits on the budget, the retained bytes
    // pass a quarter of the budget, or the server stops.
    struct Dead { rvf_plan* plan; rvf_obs* obs; rvf_mesh* mesh; size_t bytes; };
    std::deque<Dead> graveyard;
    size_t dead_bytes = 0;
    bool mem_wanted = false;           // an admission waits for memory
    std::condition_variable cv_reap;
    bool reap_stop = false;
    std::thread reaper;
    // device batch buffers
    PairArgs* d_args = nullptr;        // [max_active]
    int2* d_work = nullptr;            // [work_cap]
    int* d_ctl = nullptr;              // [2 * RVF_K_COUNT]: head, count per kind
    int2* h_work = nullptr;            // pinned staging
    int* h_ctl = nullptr;
    PairArgs* h_args = nullptr;        // pinned staging of d_args (a slot's entry is rewritten only after
                                       // its previous request retired)
    static constexpr int kStreams = 4;
    hipStream_t streams[kStreams] = {};
    int64_t stream_tiles[kStreams] = {};  // tiles launched on each stream and not yet retired (worker only)
    int64_t work_cap = 0;
    int ncu = 0;
    std::chrono::steady_clock::time_point t0;
    rvf_server_counters st{};

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

// Obstruction triangles the library will hold: one tree (in both traversal
// forms) per distinct set of obstruction groups that a pair of element groups
// sees, namely the obstruction groups other than the pair's own (the same
// rule as rvf_obs_build), so a triangle is held once per set it belongs to.
int64_t held_triangles(const Req& r) {
    if (!r.has_obs || r.T == 0 || r.obs_groups.empty()) return 0;
    std::vector<int32_t> tags(r.grp.begin(), r.grp.end());
    std::sort(tags.begin(), tags.end());
    tags.erase(std::unique(tags.begin(), tags.end()), tags.end());
    std::map<int32_t, int64_t> per_group;
    for (int32_t g : r.tri_group) per_group[g]++;
    std::set<std::vector<int32_t>> seen;
    std::vector<int32_t> act;
    int64_t held = 0;
    for (int32_t a : tags)
        for (int32_t b : tags) {
            act.clear();
            for (int32_t g : r.obs_groups)
                if (g != a && g != b) act.push_back(g);
            std::sort(act.begin(), act.end());
            if (act.empty() || !seen.insert(act).second) continue;
            for (int32_t g : act) {
                auto it = per_group.find(g);
                if (it != per_group.end()) held += it->second;
            }
        }
    return held;
}

// Device bytes a request will hold: mesh, obstruction trees, point tables,
// row/column partials, F and audit arrays (upper estimate).
size_t estimate_bytes(const Req& r) {
    const int64_t nq = r.opts.nquad;
    const int64_t tri_pts = nq <= 1 ? 1 : nq == 2 ? 3 : nq == 3 ? 7 : 13;
    const int64_t maxpts = std::max<int64_t>(nq * nq, tri_pts);
    const int64_t NS = cdiv(std::max<int64_t>(r.N, 1), RVF_STRIP) * RVF_STRIP;
    // row partials: one 64-lane record per tile; tiles of chunks of the
    // request's default width (w columns: occluded solves 2 mixed, 1 FP64;
    // 4 bounds every other kind) number about N^2 / (128 w) plus one per
    // chunk, and each of up to 64 extra chunks at group boundaries adds at
    // most one tile per strip
    const int64_t w = r.T > 0 ? (r.opts.precision == RVF_MIXED ? 2 : 1) : RVF_CHUNK_QUEUE;
    const int64_t nchunks = cdiv(r.N, w) + 64;
    const int64_t tiles = cdiv(r.N * r.N, 128 * w) + nchunks + 64 * (NS / RVF_STRIP);
    size_t b = 0;
    b += (size_t)r.Nn * 24 + (size_t)r.N * 64;                          // mesh
    b += (size_t)held_triangles(r) * 512;                                // trees, triangles, items
    b += (size_t)NS * maxpts * 8 * (8 + 8 + 4 + 4);                      // P64 Q64 P32 Q32
    b += (size_t)NS * 72;                                                // anchors, areas, sums, diagonal
    b += (size_t)tiles * (RVF_STRIP * 8 + 12) + (size_t)(NS / RVF_STRIP) * NS * 8;  // part, tile lists, partT
    if (r.opts.flags & RVF_WANT_F) b += (size_t)r.N * r.N * 8;
    if (r.opts.flags & RVF_AUDIT) b += (size_t)r.N * r.N * 8;
    return b + b / 8 + (1u << 20);
}

// Ends a request: its device work must have completed.  Hands its device
// objects to the reaper (which returns the memory to the budget) and frees
// its slot.
void finish(rvf_server* s, Req* r, int state) {
    if (r->stream >= 0) { s->stream_tiles[r->stream] -= r->ntiles; r->stream = -1; }
    if (r->ready_ev) { (void)hipEventDestroy(r->ready_ev); r->ready_ev = nullptr; }
    if (r->done_ev) { (void)hipEventDestroy(r->done_ev); r->done_ev = nullptr; }
    std::lock_guard<std::mutex> g(s->mu);
    if (r->plan || r->obs || r->mesh) {
        s->graveyard.push_back({r->plan, r->obs, r->mesh, r->est_bytes});
        s->dead_bytes += r->est_bytes;
        s->cv_reap.notify_all();
    } else {
        s->used -= std::min(s->used, r->est_bytes);
    }
    r->plan = nullptr; r->obs = nullptr; r->mesh = nullptr;
    if (r->slot >= 0) s->free_slots.push_back(r->slot);
    r->slot = -1;
    r->state = state;
    r->st.t_done_ms = s->now_ms();
    if (state == REQ_DONE) s->st.completed++;
    else if (state == REQ_FAILED) s->st.failed++;
    else s->st.cancelled++;
    s->active.erase(std::remove(s->active.begin(), s->active.end(), r), s->active.end());
    s->cv_done.notify_all();
}

// A direct request goes to the slot stream with the fewest tiles in flight,
// so a short request does not queue behind a long one.
int pick_stream(const rvf_server* s) {
    int best = 0;
    for (int k = 1; k < rvf_server::kStreams; k++)
        if (s->stream_tiles[k] < s->stream_tiles[best]) best = k;
    return best;
}

// Large requests get direct launches of their tile ranges; smaller ones are
// drained by the batched kernel.
// direct_tiles counts tiles of RVF_CHUNK columns: narrower tiles count pro rata
bool is_direct(const rvf_server* s, const Req* r) {
    return s->o.direct_tiles > 0 && (int64_t)r->ntiles * r->cols >= (int64_t)s->o.direct_tiles * RVF_CHUNK;
}

// Upload, build and enqueue the prologue of an admitted request.  The
// prologue goes on the context stream, where plan creation queued its zero
// fills of the same buffers; ready_ev, recorded after it, orders it before
// the request's tiles on a slot stream.
int admit(rvf_server* s, Req* r) {
    hipStream_t st = rvfi_ctx_stream(s->ctx);
    const int32_t* nodes[4] = {r->nodes[0].data(), r->nodes[1].data(), r->nodes[2].data(), r->nodes[3].data()};
    int rc = rvf_mesh_upload(s->ctx, r->coords.data(), r->Nn, r->N, r->fam.data(), r->eni.data(), r->grp.data(),
                             nodes, r->nfam, &r->mesh);
    if (!rc && r->has_obs)
        rc = rvf_obs_build(s->ctx, r->mesh, r->soup.data(), r->tri_group.data(), r->T, r->obs_groups.data(),
                           (int32_t)r->obs_groups.size(), r->soup_read, &r->obs);
    if (!rc) rc = rvf_plan_create(s->ctx, r->mesh, r->obs, &r->opts, &r->plan);
    if (!rc) rc = rvfi_plan_batch(r->plan, &r->args, &r->kind, &r->lds);
    if (!rc) rc = rvfi_plan_prologue(r->plan, st);
    if (!rc) {
        hipError_t e = hipEventCreateWithFlags(&r->ready_ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(r->ready_ev, st);
        if (e != hipSuccess) rc = rvfi_fail(RVF_E_HIP, std::string("server: events: ") + hipGetErrorString(e));
    }
    if (!rc) {
        r->ntiles = r->args.ntiles;
        r->cols = rvfi_plan_columns(r->plan);
        s->h_args[r->slot] = r->args;  // pinned: the copy is asynchronous
        hipError_t e = hipMemcpyAsync(s->d_args + r->slot, s->h_args + r->slot, sizeof(PairArgs),
                                      hipMemcpyHostToDevice, st);
        if (e != hipSuccess) rc = rvfi_fail(RVF_E_HIP, std::string("server: args upload: ") + hipGetErrorString(e));
    }
    if (rc) {
        r->err = rvf_last_error();
        return rc;
    }
    // host copies of the inputs are no longer needed
    std::vector<double>().swap(r->coords);
    std::vector<double>().swap(r->soup);
    for (auto& v : r->nodes) std::vector<int32_t>().swap(v);
    return 0;
}

void reaper_loop(rvf_server* s) {
    (void)hipSetDevice(rvfi_ctx_device(s->ctx));
    for (;;) {
        rvf_server::Dead d;
        {
            std::unique_lock<std::mutex> g(s->mu);
            s->cv_reap.wait(g, [&] {
                return s->reap_stop || (!s->graveyard.empty() &&
                                        (s->mem_wanted || s->dead_bytes > (size_t)s->o.hbm_budget / 4));
            });
            if (s->graveyard.empty()) break;
            d = s->graveyard.front();
            s->graveyard.pop_front();
            s->dead_bytes -= std::min(s->dead_bytes, d.bytes);
        }
        rvf_plan_free(d.plan);
        rvf_obs_free(d.obs);
        rvf_mesh_free(d.mesh);
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->used -= std::min(s->used, d.bytes);
        }
        s->cv_work.notify_all();
    }
}

void worker_loop(rvf_server* s) {
    (void)hipSetDevice(rvfi_ctx_device(s->ctx));
    hipStream_t st = rvfi_ctx_stream(s->ctx);
    std::vector<Req*> running;  // admitted, not yet retired
    hipEvent_t round_ev = nullptr;
    (void)hipEventCreateWithFlags(&round_ev, hipEventDisableTiming);
    bool round_open = false;    // a batched round is in flight
    for (;;) {
        const double t0 = s->now_ms();
        int rc = 0;
        // ---- 1. admissions ----
        std::vector<Req*> admitted;
        {
            std::unique_lock<std::mutex> g(s->mu);
            if (s->stop && running.empty() && s->pending.empty()) break;
            while (!s->pending.empty() && !s->free_slots.empty()) {
                Req* r = s->pending.front();
                if (r->cancel) {
                    s->pending.pop_front();
                    r->state = REQ_CANCELLED;
                    r->st.t_done_ms = s->now_ms();
                    s->st.cancelled++;
                    s->cv_done.notify_all();
                    continue;
                }
                if (r->est_bytes > (size_t)s->o.hbm_budget) {
                    s->pending.pop_front();
                    r->state = REQ_FAILED;
                    r->err = "request needs more device memory than the server's budget";
                    r->st.t_done_ms = s->now_ms();
                    s->st.failed++;
                    s->cv_done.notify_all();
                    continue;
                }
                if (s->used + r->est_bytes > (size_t)s->o.hbm_budget) {  // wait for memory
                    s->mem_wanted = true;
                    s->cv_reap.notify_all();
                    break;
                }
                s->mem_wanted = false;
                s->pending.pop_front();
                r->slot = s->free_slots.back();
                s->free_slots.pop_back();
                s->used += r->est_bytes;
                s->st.peak_bytes = std::max<int64_t>(s->st.peak_bytes, (int64_t)s->used);
                r->state = REQ_RUNNING;
                r->st.t_admit_ms = s->now_ms();
                s->active.push_back(r);
                admitted.push_back(r);
            }
            if (admitted.empty() && running.empty()) {
                if (s->stop && s->pending.empty()) break;
                s->cv_work.wait(g);
                continue;
            }
        }
        for (Req* r : admitted) {
            if (admit(s, r)) {
                (void)hipStreamSynchronize(st);  // a partial prologue may still use its buffers
                finish(s, r, REQ_FAILED);
                continue;
            }
            r->st.tiles = r->ntiles;
            running.push_back(r);
        }
        const double t1 = s->now_ms();
        // ---- 2. large requests: whole tile range + epilogue on the slot stream ----
        for (Req* r : running) {
            if (rc || r->cancel || r->epilogue || !is_direct(s, r)) continue;
            r->stream = pick_stream(s);
            s->stream_tiles[r->stream] += r->ntiles;
            hipStream_t rs = s->streams[r->stream];
            hipError_t e = hipStreamWaitEvent(rs, r->ready_ev, 0);
            if (e != hipSuccess) { rc = rvfi_fail(RVF_E_HIP, std::string("server: wait event: ") + hipGetErrorString(e)); break; }
            PairArgs a = r->args;
            a.tile0 = 0;
            a.ntiles = r->ntiles;
            rc = rvf_launch_pair(&a, r->kind, r->lds, rs);
            if (rc) { rc = rvfi_fail(RVF_E_HIP, std::string("server: pair launch: ") + hipGetErrorString((hipError_t)rc)); break; }
            r->next_tile = r->ntiles;
            r->st.t_start_ms = s->now_ms();
            if ((rc = rvfi_plan_epilogue(r->plan, rs))) break;
            e = hipEventRecord(r->done_ev, rs);
            if (e != hipSuccess) { rc = rvfi_fail(RVF_E_HIP, std::string("server: record: ") + hipGetErrorString(e)); break; }
            r->epilogue = true;
            std::lock_guard<std::mutex> g(s->mu);
            s->st.tiles += r->ntiles;
        }
        // ---- 3. small requests: one batched round at a time ----
        if (!rc && !round_open) {
            struct Chunk { Req* r; int t0, t1; };
            std::vector<Chunk> chunks[RVF_K_COUNT];
            std::vector<Req*> completing;
            int64_t nitems = 0;
            for (Req* r : running) {
                if (r->cancel || r->epilogue || is_direct(s, r)) continue;
                if (r->next_tile < r->ntiles) {
                    const int n = std::min(r->ntiles - r->next_tile, s->o.round_tiles);
                    for (int t = r->next_tile; t < r->next_tile + n; t += 32)
                        chunks[r->kind].push_back({r, t, std::min(t + 32, r->next_tile + n)});
                    r->next_tile += n;
                    nitems += n;
                    if (r->st.t_start_ms == 0) r->st.t_start_ms = s->now_ms();
                }
                if (r->next_tile >= r->ntiles) completing.push_back(r);
            }
            if (nitems > 0 || !completing.empty()) {
                // this round's requests were all admitted before: their prologues
                // precede it on the context stream
                int64_t off = 0;
                for (int k = 0; k < RVF_K_COUNT && !rc; k++) {
                    if (chunks[k].empty()) continue;
                    std::map<Req*, std::vector<Chunk>> byreq;
                    for (const Chunk& c : chunks[k]) byreq[c.r].push_back(c);
                    const int64_t start = off;
                    size_t lds = 0;
                    for (bool more = true; more;) {  // interleave the requests' chunks
                        more = false;
                        for (auto& kv : byreq) {
                            if (kv.second.empty()) continue;
                            const Chunk c = kv.second.front();
                            kv.second.erase(kv.second.begin());
                            for (int t = c.t0; t < c.t1; t++) s->h_work[off++] = make_int2(c.r->slot, t);
                            lds = std::max(lds, c.r->lds);
                            more = true;
                        }
                    }
                    const int cnt = (int)(off - start);
                    s->h_ctl[2 * k] = 0;
                    s->h_ctl[2 * k + 1] = cnt;
                    hipError_t e = hipMemcpyAsync(s->d_work + start, s->h_work + start, sizeof(int2) * cnt,
                                                  hipMemcpyHostToDevice, st);
                    if (e == hipSuccess)
                        e = hipMemcpyAsync(s->d_ctl + 2 * k, s->h_ctl + 2 * k, 2 * sizeof(int), hipMemcpyHostToDevice, st);
                    if (e != hipSuccess) { rc = rvfi_fail(RVF_E_HIP, std::string("server: work upload: ") + hipGetErrorString(e)); break; }
                    const int per_cu = std::max(1, rvf_batched_blocks_per_cu(k, lds));
                    const int grid = (int)std::min<int64_t>(cnt, (int64_t)s->ncu * per_cu);
                    rc = rvf_launch_pair_batched(k, s->d_args, s->d_work + start, s->d_ctl + 2 * k, grid, lds, st);
                    if (rc) rc = rvfi_fail(RVF_E_HIP, std::string("server: batched launch: ") + hipGetErrorString((hipError_t)rc));
                }
                for (Req* r : completing) {
                    if (rc) break;
                    if ((rc = rvfi_plan_epilogue(r->plan, st))) break;
                    hipError_t e = hipEventRecord(r->done_ev, st);
                    if (e != hipSuccess) { rc = rvfi_fail(RVF_E_HIP, std::string("server: record: ") + hipGetErrorString(e)); break; }
                    r->epilogue = true;
                }
                if (!rc) {
                    hipError_t e = hipEventRecord(round_ev, st);
                    if (e != hipSuccess) rc = rvfi_fail(RVF_E_HIP, std::string("server: record: ") + hipGetErrorString(e));
                    round_open = true;
                }
                std::lock_guard<std::mutex> g(s->mu);
                s->st.rounds++;
                s->st.tiles += nitems;
                s->st.batched_tiles += nitems;
            }
        }
        const double t2 = s->now_ms();
        if (rc) {  // a launch failed: fail every running request
            const std::string msg = rvf_last_error();
            (void)hipDeviceSynchronize();
            for (Req* r : running) { r->err = msg; finish(s, r, REQ_FAILED); }
            running.clear();
            round_open = false;
            continue;
        }
        // ---- 4. wait for progress: the open round, or any request's epilogue ----
        bool progressed = false;
        while (!progressed) {
            if (round_open && hipEventQuery(round_ev) == hipSuccess) { round_open = false; progressed = true; }
            for (Req* r : running)
                if (r->epilogue && hipEventQuery(r->done_ev) == hipSuccess) progressed = true;
            if (progressed) break;
            {
                std::lock_guard<std::mutex> g(s->mu);
                if (!s->pending.empty() && !s->free_slots.empty() &&
                    s->used + s->pending.front()->est_bytes <= (size_t)s->o.hbm_budget)
                    break;  // a new request can be admitted now
                if (s->stop && running.empty()) break;
            }
            bool any_open = round_open;
            for (Req* r : running) any_open |= r->epilogue;
            if (!any_open) break;  // nothing in flight: go launch
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
        const double t3 = s->now_ms();
        // ---- retire finished requests ----
        std::vector<Req*> keep;
        for (Req* r : running) {
            if (r->cancel && !r->epilogue && !round_open) { finish(s, r, REQ_CANCELLED); continue; }
            if (!r->epilogue || hipEventQuery(r->done_ev) != hipSuccess) {
                keep.push_back(r);
                continue;
            }
            int frc = rvfi_plan_collect(r->plan);
            if (!frc) {
                const int64_t N = r->N;
                int32_t G = 0;
                rvf_mesh_info(r->mesh, nullptr, &G, nullptr);
                if (r->opts.flags & RVF_WANT_F) r->F.resize((size_t)N * N);
                r->A.resize(N); r->rowsum.resize(N);
                r->Fg.resize((size_t)G * G); r->Ag.resize(G); r->tags.resize(G);
                frc = rvf_plan_fetch(r->plan, r->F.empty() ? nullptr : r->F.data(), r->A.data(), r->rowsum.data(),
                                     r->Fg.data(), r->Ag.data(), r->tags.data(), nullptr, nullptr);
            }
            if (frc) r->err = rvf_last_error();
            finish(s, r, frc ? REQ_FAILED : REQ_DONE);
        }
        running.swap(keep);
        std::lock_guard<std::mutex> g(s->mu);
        s->st.admit_ms += t1 - t0;
        s->st.launch_ms += t2 - t1;
        s->st.wait_ms += t3 - t2;
        s->st.retire_ms += s->now_ms() - t3;
    }
    if (round_ev) (void)hipEventDestroy(round_ev);
}

}  // namespace

extern "C" {

int rvf_server_create(rvf_ctx* ctx, const rvf_server_opts* opts, rvf_server** out) {
    if (!ctx || !out) return rvfi_fail(RVF_E_ARG, "rvf_server_create: bad arguments");
    rvf_server* s = new rvf_server();
    s->ctx = ctx;
    if (opts) s->o = *opts;
    if (s->o.hbm_budget <= 0) s->o.hbm_budget = (int64_t)64 << 30;
    if (s->o.max_active <= 0) s->o.max_active = 64;
    if (s->o.round_tiles <= 0) s->o.round_tiles = 4096;
    if (s->o.direct_tiles == 0) s->o.direct_tiles = 512;
    s->o.max_active = std::min(s->o.max_active, 1024);
    hipError_t e = hipSetDevice(rvfi_ctx_device(ctx));
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, rvfi_ctx_device(ctx));
    s->ncu = e == hipSuccess ? prop.multiProcessorCount : 0;
    s->work_cap = (int64_t)s->o.max_active * s->o.round_tiles;
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_args, sizeof(PairArgs) * s->o.max_active);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_work, sizeof(int2) * s->work_cap);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_ctl, sizeof(int) * 2 * RVF_K_COUNT);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_work, sizeof(int2) * s->work_cap, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_ctl, sizeof(int) * 2 * RVF_K_COUNT, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&s->h_args, sizeof(PairArgs) * s->o.max_active, hipHostMallocDefault);
    for (int k = 0; k < rvf_server::kStreams && e == hipSuccess; k++)
        e = hipStreamCreateWithFlags(&s->streams[k], hipStreamNonBlocking);
    if (e != hipSuccess) {
        rvf_server_destroy(s);
        return rvfi_fail(RVF_E_HIP, std::string("rvf_server_create: ") + hipGetErrorString(e));
    }
    for (int k = s->o.max_active - 1; k >= 0; k--) s->free_slots.push_back(k);
    s->t0 = std::chrono::steady_clock::now();
    s->reaper = std::thread(reaper_loop, s);
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return 0;
}

void rvf_server_destroy(rvf_server* s) {
    if (!s) return;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
            for (Req* r : s->pending) r->cancel = true;
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    if (s->reaper.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->reap_stop = true;
        }
        s->cv_reap.notify_all();
        s->reaper.join();
    }
    for (auto& kv : s->all) delete kv.second;
    if (s->d_args) (void)hipFree(s->d_args);
    if (s->d_work) (void)hipFree(s->d_work);
    if (s->d_ctl) (void)hipFree(s->d_ctl);
    if (s->h_work) (void)hipHostFree(s->h_work);
    if (s->h_args) (void)hipHostFree(s->h_args);
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
    for (int k = 0; k < rvf_server::kStreams; k++)
        if (s->streams[k]) (void)hipStreamDestroy(s->streams[k]);
    delete s;
}

int rvf_server_submit(rvf_server* s, const double* coords, int64_t n_nodes, int64_t n_elems, const int8_t* family,
                      const int32_t* elem_node_idx, const int32_t* elem_group, const int32_t* const nodes[4],
                      const int64_t n_per_family[4], const double* soup, const int32_t* tri_group, int64_t T,
                      const int32_t* obs_groups, int32_t n_obs, int32_t soup_read, const rvf_opts* opts,
                      int64_t* ticket) {
    if (!s || !coords || n_nodes < 1 || n_elems < 1 || !family || !elem_node_idx || !elem_group || !nodes ||
        !n_per_family || !opts || !ticket || T < 0 || n_obs < 0)
        return rvfi_fail(RVF_E_ARG, "rvf_server_submit: bad arguments");
    if (opts->method == RVF_METHOD_MC) return rvfi_fail(RVF_E_ARG, "rvf_server_submit: Monte Carlo is served by rvf_compute_view_factors");
    Req* r = new Req();
    r->Nn = n_nodes; r->N = n_elems;
    r->coords.assign(coords, coords + 3 * n_nodes);
    r->fam.assign(family, family + n_elems);
    r->eni.assign(elem_node_idx, elem_node_idx + n_elems);
    r->grp.assign(elem_group, elem_group + n_elems);
    static const int nn[4] = {8, 6, 4, 3};
    for (int f = 0; f < 4; f++) {
        r->nfam[f] = n_per_family[f];
        if (n_per_family[f] < 0 || (n_per_family[f] > 0 && !nodes[f])) { delete r; return rvfi_fail(RVF_E_ARG, "rvf_server_submit: bad node arrays"); }
        if (n_per_family[f]) r->nodes[f].assign(nodes[f], nodes[f] + nn[f] * n_per_family[f]);
    }
    if (soup && T > 0 && n_obs > 0) {
        if (!tri_group || !obs_groups) { delete r; return rvfi_fail(RVF_E_ARG, "rvf_server_submit: bad obstruction arrays"); }
        r->has_obs = true;
        r->soup.assign(soup, soup + 9 * T);
        r->tri_group.assign(tri_group, tri_group + T);
        r->obs_groups.assign(obs_groups, obs_groups + n_obs);
        r->T = T;
        r->soup_read = soup_read;
    }
    r->opts = *opts;
    r->opts.flags &= ~(RVF_GRAPH | RVF_TIMING);
    r->est_bytes = estimate_bytes(*r);
    r->st.footprint_bytes = (int64_t)r->est_bytes;
    std::lock_guard<std::mutex> g(s->mu);
    r->id = s->next_id++;
    r->st.t_submit_ms = s->now_ms();
    s->all[r->id] = r;
    s->pending.push_back(r);
    s->st.submitted++;
    *ticket = r->id;
    s->cv_work.notify_all();
    return 0;
}

int rvf_server_poll(rvf_server* s, int64_t ticket) {
    if (!s) return rvfi_fail(RVF_E_ARG, "rvf_server_poll: null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return rvfi_fail(RVF_E_ARG, "rvf_server_poll: unknown ticket");
    return it->second->state;
}

int rvf_server_wait(rvf_server* s, int64_t ticket, int32_t timeout_ms) {
    if (!s) return rvfi_fail(RVF_E_ARG, "rvf_server_wait: null server");
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return rvfi_fail(RVF_E_ARG, "rvf_server_wait: unknown ticket");
    Req* r = it->second;
    auto done = [&] { return r->state >= REQ_DONE; };
    if (timeout_ms < 0) s->cv_done.wait(g, done);
    else if (!s->cv_done.wait_for(g, std::chrono::milliseconds(timeout_ms), done))
        return rvfi_fail(RVF_E_TIMEOUT, "rvf_server_wait: timed out");
    return r->state;
}

int rvf_server_result(rvf_server* s, int64_t ticket, double* F, double* A, double* rowsum, double* F_group,
                      double* A_group, int32_t* group_tags, rvf_request_stats* stats) {
    if (!s) return rvfi_fail(RVF_E_ARG, "rvf_server_result: null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return rvfi_fail(RVF_E_ARG, "rvf_server_result: unknown ticket");
    Req* r = it->second;
    if (stats) *stats = r->st;
    if (r->state == REQ_FAILED) return rvfi_fail(RVF_E_STATE, "request failed: " + r->err);
    if (r->state == REQ_CANCELLED) return rvfi_fail(RVF_E_CANCELLED, "request was cancelled");
    if (r->state != REQ_DONE) return rvfi_fail(RVF_E_STATE, "request not finished");
    if (F) {
        if (r->F.empty()) return rvfi_fail(RVF_E_STATE, "F was not requested (RVF_WANT_F)");
        std::memcpy(F, r->F.data(), sizeof(double) * r->F.size());
    }
    if (A) std::memcpy(A, r->A.data(), sizeof(double) * r->A.size());
    if (rowsum) std::memcpy(rowsum, r->rowsum.data(), sizeof(double) * r->rowsum.size());
    if (F_group) std::memcpy(F_group, r->Fg.data(), sizeof(double) * r->Fg.size());
    if (A_group) std::memcpy(A_group, r->Ag.data(), sizeof(double) * r->Ag.size());
    if (group_tags) std::memcpy(group_tags, r->tags.data(), sizeof(int32_t) * r->tags.size());
    return 0;
}

int rvf_server_cancel(rvf_server* s, int64_t ticket) {
    if (!s) return rvfi_fail(RVF_E_ARG, "rvf_server_cancel: null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return rvfi_fail(RVF_E_ARG, "rvf_server_cancel: unknown ticket");
    it->second->cancel = true;
    s->cv_work.notify_all();
    return 0;
}

int rvf_server_release(rvf_server* s, int64_t ticket) {
    if (!s) return rvfi_fail(RVF_E_ARG, "rvf_server_release: null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return rvfi_fail(RVF_E_ARG, "rvf_server_release: unknown ticket");
    if (it->second->state < REQ_DONE) return rvfi_fail(RVF_E_STATE, "rvf_server_release: request still in flight");
    delete it->second;
    s->all.erase(it);
    return 0;
}

int rvf_server_stats(rvf_server* s, rvf_server_counters* out) {
    if (!s || !out) return rvfi_fail(RVF_E_ARG, "rvf_server_stats: bad arguments");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    out->used_bytes = (int64_t)s->used;
    out->pending = (int64_t)s->pending.size();
    out->active = (int64_t)s->active.size();
    return 0;
}

}  // extern "C"
