// Serving loop: several trained SOMs resident in HBM, BMU and activation-map requests of uneven
// sizes batched continuously into shared launches.
//
//   submit()  -> pending queue (FIFO)
//   admission -> a request is admitted when its working memory fits in the free part of the budget
//                (one device slab, first-fit)
//   upload    -> rows move to the device on the copy stream in pieces of whole tiles (at most 16 MB),
//                at most 3 pieces in flight, taken round-robin from the requests being uploaded: a
//                large request does not hold the small ones behind its whole upload, and its first
//                tiles are computed while its later pieces are still on the way
//   batching  -> every iteration, the uploaded tiles (512 rows) of all requests are shared out
//                round-robin into one launch per dimension class; the launch is a hipGraph captured
//                per (dimension, tile-count bucket) with a fixed tile-table buffer, so a replay is a
//                table upload plus one graph launch
//   finishing -> a request whose last tile has been launched gets its activation reduction and the
//                copy of its result to the host enqueued; its completion event frees its memory
//
// Requests never share accumulators: BMUs are per row, activation maps are exact fixed-point sums
// scaled from the request's own size, so a result is bit-identical alone and in any mix.
#include "common.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "launch.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int MAX_MODELS = 64;
constexpr int BUCKETS[] = {16, 64, 256, 1024, 4096};
constexpr int NBUCKETS = sizeof(BUCKETS) / sizeof(BUCKETS[0]);
constexpr int MAX_TILES = 4096;
constexpr int STAGING_SLOTS = 8;
constexpr size_t ALIGN = 256;
constexpr size_t ACT_TABLE_BUDGET = (size_t)32 << 20;
// Uploads move in pieces of at most UPLOAD_PIECE bytes (whole tiles), at most MAX_INFLIGHT of them on
// the copy stream at a time, taken round-robin from the requests being uploaded.
constexpr size_t UPLOAD_PIECE = (size_t)16 << 20;
constexpr int MAX_INFLIGHT = 3;
// Tile launches of one dimension class queued on the compute stream at a time. While one runs, newly
// uploaded tiles gather into the next launch instead of queueing as many small launches behind it (a
// 16 MB piece of D = 32 rows is only 256 tiles, which leaves half of the workgroup slots empty).
constexpr int MAX_LAUNCHES_INFLIGHT = 1;

inline size_t align_up(size_t v) { return (v + ALIGN - 1) / ALIGN * ALIGN; }

// First-fit allocator over one device slab (offsets only).
struct Slab {
    char *base = nullptr;
    size_t capacity = 0, used = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    void init(char *b, size_t cap)
    {
        base = b;
        capacity = cap;
        free_.clear();
        free_[0] = cap;
        used = 0;
    }
    bool alloc(size_t bytes, size_t *off)
    {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                size_t rest = it->second - bytes;
                size_t o = it->first;
                free_.erase(it);
                if (rest) free_[o + bytes] = rest;
                used += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = align_up(bytes);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

enum State { QUEUED = 0, UPLOADING = 1, READY = 2, FINISHING = 3, DONE = 4, FAILED = 5 };

struct Request {
    int64_t ticket = 0;
    int model = 0, kind = 0;
    const float *x_host = nullptr;
    const double *rho_host = nullptr;
    int64_t n = 0;
    void *out_host = nullptr;
    State state = QUEUED;
    size_t off = 0, bytes = 0;
    float *x = nullptr;
    double *rho = nullptr;
    int32_t *c1 = nullptr;
    double *act = nullptr;
    unsigned long long *table = nullptr;
    unsigned long long *bounds = nullptr;
    int *shifts = nullptr;
    int nwg = 1;
    int64_t next_row = 0;     // first row not yet handed to a launch
    int64_t up_rows = 0;      // rows whose copy is queued on the copy stream
    int64_t ready_rows = 0;   // rows on the device (their pieces seen complete)
    std::deque<std::pair<int64_t, hipEvent_t>> pieces;  // queued copies in order: (end row, event)
    hipEvent_t ev_done = nullptr;
    Clock::time_point t_submit, t_done;
};

struct GraphSlot {
    bool built = false;
    hipGraphExec_t exec = nullptr;
    BmuTile *tiles = nullptr;       // device table baked into the graph
    int64_t *fb_list = nullptr;     // fallback entries
    unsigned int *fb_count = nullptr;
};

}  // namespace

struct somhip_server {
    somhip_ctx *ctx = nullptr;
    int arithmetic = SOMHIP_RERANK_REF32;
    hipStream_t compute = nullptr, copy = nullptr;
    char *slab_mem = nullptr;
    Slab slab;
    std::vector<somhip_model *> models;
    std::vector<int> model_dpad;
    BmuModelDev *models_dev = nullptr;
    GraphSlot graphs[2][NBUCKETS];
    BmuTile *staging = nullptr;  // pinned host, STAGING_SLOTS * MAX_TILES
    hipEvent_t staging_ev[STAGING_SLOTS] = {};
    int staging_next = 0;
    std::vector<hipEvent_t> ev_pool;  // completed piece and launch events for reuse
    std::deque<hipEvent_t> launch_evs[2];  // per dimension class: events after the queued tile launches
    int inflight = 0;                 // pieces queued and not yet seen complete
    size_t up_cursor = 0;             // round-robin position among the requests being uploaded

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<int64_t, std::shared_ptr<Request>> all;
    std::vector<std::shared_ptr<Request>> active;  // scheduler-thread owned
    int64_t next_ticket = 1;
    bool stop = false;
    std::string error;
    std::thread worker;

    // statistics
    std::atomic<int64_t> completed{0}, launches{0}, tiles_launched{0};

    int build_graph(int dclass, int b);
    int queue_piece(Request &R);
    bool poll_pieces(Request &R, hipError_t *err);
    int launch_tiles(int dclass, const std::vector<BmuTile> &tl);
    int finish(Request &R);
    void loop();
    void fail(const std::string &msg)
    {
        std::lock_guard<std::mutex> g(mu);
        if (error.empty()) error = msg;
        for (auto &kv : all)
            if (kv.second->state != DONE) kv.second->state = FAILED;
        cv_done.notify_all();
    }
};

namespace {
int dclass_of(int dpad) { return dpad == 16 ? 0 : 1; }
}  // namespace

int somhip_server::build_graph(int dc, int b)
{
    GraphSlot &G = graphs[dc][b];
    const int grid = BUCKETS[b];
    SOM_HIP_TRY(hipMalloc((void **)&G.tiles, sizeof(BmuTile) * grid));
    SOM_HIP_TRY(hipMalloc((void **)&G.fb_list, sizeof(int64_t) * (size_t)grid * SOM_TILE_ROWS));
    SOM_HIP_TRY(hipMalloc((void **)&G.fb_count, sizeof(unsigned int)));
    // empty tiles are skipped by the kernels
    std::vector<BmuTile> empty(grid);
    for (auto &t : empty) t = BmuTile{};
    SOM_HIP_TRY(hipMemcpy(G.tiles, empty.data(), sizeof(BmuTile) * grid, hipMemcpyHostToDevice));
    hipGraph_t graph = nullptr;
    SOM_HIP_TRY(hipStreamBeginCapture(compute, hipStreamCaptureModeThreadLocal));
    // the fallback counter is cleared by a kernel node (no memset node in the graph)
    int rc = somhip_launch_zero_u32(compute, G.fb_count);
    if (rc == SOMHIP_OK)
        rc = somhip_launch_bmu_tiles(compute, dc == 0 ? 16 : 32, arithmetic, 0, models_dev, G.tiles, grid, G.fb_list,
                                     G.fb_count);
    hipError_t e = hipStreamEndCapture(compute, &graph);
    if (rc != SOMHIP_OK) return rc;
    SOM_HIP_TRY(e);
    SOM_HIP_TRY(hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0));
    SOM_HIP_TRY(hipGraphDestroy(graph));
    G.built = true;
    return SOMHIP_OK;
}

int somhip_server::launch_tiles(int dc, const std::vector<BmuTile> &tl)
{
    const int nt = (int)tl.size();
    int b = 0;
    while (b < NBUCKETS - 1 && BUCKETS[b] < nt) b++;
    GraphSlot &G = graphs[dc][b];
    if (!G.built) {
        int rc = build_graph(dc, b);
        if (rc) return rc;
    }
    // stage the table in pinned memory; pad with empty tiles up to the bucket
    const int slot = staging_next;
    staging_next = (staging_next + 1) % STAGING_SLOTS;
    SOM_HIP_TRY(hipEventSynchronize(staging_ev[slot]));
    BmuTile *st = staging + (size_t)slot * MAX_TILES;
    std::copy(tl.begin(), tl.end(), st);
    for (int i = nt; i < BUCKETS[b]; i++) st[i] = BmuTile{};
    SOM_HIP_TRY(hipMemcpyAsync(G.tiles, st, sizeof(BmuTile) * BUCKETS[b], hipMemcpyHostToDevice, compute));
    SOM_HIP_TRY(hipEventRecord(staging_ev[slot], compute));
    SOM_HIP_TRY(hipGraphLaunch(G.exec, compute));
    hipEvent_t e = nullptr;
    if (!ev_pool.empty()) {
        e = ev_pool.back();
        ev_pool.pop_back();
    } else {
        SOM_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    launch_evs[dc].push_back(e);  // owned by the queue from here, even if the record fails
    SOM_HIP_TRY(hipEventRecord(e, compute));
    launches++;
    tiles_launched += nt;
    return SOMHIP_OK;
}

int somhip_server::finish(Request &R)
{
    somhip_model *M = models[R.model];
    if (R.kind == SOMHIP_SERVE_ACTIVATION) {
        int rc = somhip_launch_absmax(compute, NO_ROWS, 0, 16, R.rho, 0, R.rho ? R.n : 0, R.bounds, ctx->num_cus);
        if (!rc) rc = somhip_launch_shifts(compute, R.bounds, 0, 0, R.rho != nullptr, R.n, R.shifts);
        if (!rc) SOM_HIP_TRY(hipMemsetAsync(R.act, 0, sizeof(double) * M->mn, compute));
        if (!rc)
            rc = somhip_launch_scatter(compute, R.c1, NO_ROWS, 0, R.rho, 0, R.n, R.shifts, M->mn, 1, R.nwg, R.table,
                                       R.act, ctx->lds_per_block);
        if (rc) return rc;
        SOM_HIP_TRY(hipMemcpyAsync(R.out_host, R.act, sizeof(double) * M->mn, hipMemcpyDeviceToHost, compute));
    } else if (R.n > 0) {
        SOM_HIP_TRY(hipMemcpyAsync(R.out_host, R.c1, sizeof(int32_t) * R.n, hipMemcpyDeviceToHost, compute));
    }
    SOM_HIP_TRY(hipEventRecord(R.ev_done, compute));
    R.state = FINISHING;
    return SOMHIP_OK;
}

// Queues the next piece of R's rows (and weights) on the copy stream.
int somhip_server::queue_piece(Request &R)
{
    somhip_model *M = models[R.model];
    const size_t row_bytes = sizeof(float) * M->dim + (R.rho ? sizeof(double) : 0);
    const int64_t tiles = std::max<int64_t>(1, (int64_t)(UPLOAD_PIECE / row_bytes) / SOM_TILE_ROWS);
    const int64_t r0 = R.up_rows, r1 = std::min(R.n, r0 + tiles * SOM_TILE_ROWS);
    hipEvent_t e = nullptr;
    if (!ev_pool.empty()) {
        e = ev_pool.back();
        ev_pool.pop_back();
    } else {
        SOM_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    R.pieces.emplace_back(r1, e);  // owned by the request from here, even if a copy below fails
    SOM_HIP_TRY(hipMemcpyAsync(R.x + r0 * M->dim, R.x_host + r0 * M->dim, sizeof(float) * (r1 - r0) * M->dim,
                               hipMemcpyHostToDevice, copy));
    if (R.rho)
        SOM_HIP_TRY(hipMemcpyAsync(R.rho + r0, R.rho_host + r0, sizeof(double) * (r1 - r0), hipMemcpyHostToDevice,
                                   copy));
    SOM_HIP_TRY(hipEventRecord(e, copy));
    R.up_rows = r1;
    inflight++;
    return SOMHIP_OK;
}

// Advances R.ready_rows over its completed pieces; returns whether any completed.
bool somhip_server::poll_pieces(Request &R, hipError_t *err)
{
    bool any = false;
    while (!R.pieces.empty()) {
        const hipError_t q = hipEventQuery(R.pieces.front().second);
        if (q == hipErrorNotReady) break;
        if (q != hipSuccess) {
            *err = q;
            return any;
        }
        R.ready_rows = R.pieces.front().first;
        ev_pool.push_back(R.pieces.front().second);
        R.pieces.pop_front();
        inflight--;
        any = true;
    }
    return any;
}

void somhip_server::loop()
{
    if (hipSetDevice(ctx->device) != hipSuccess) {
        fail("hipSetDevice failed in the serving thread");
        return;
    }
    for (;;) {
        bool progress = false;
        {
            std::unique_lock<std::mutex> g(mu);
            if (stop && pending.empty() && active.empty()) break;
            if (!error.empty()) break;
            // admission: FIFO while the head fits in the free budget
            while (!pending.empty()) {
                auto R = pending.front();
                size_t off;
                if (!slab.alloc(R->bytes, &off)) break;
                pending.pop_front();
                R->off = off;
                char *p = slab.base + off;
                somhip_model *M = models[R->model];
                auto take = [&](size_t sz) { char *q = p; p += align_up(sz); return q; };
                R->x = (float *)take(sizeof(float) * R->n * M->dim);
                R->c1 = (int32_t *)take(sizeof(int32_t) * std::max<int64_t>(R->n, 1));
                if (R->rho_host) R->rho = (double *)take(sizeof(double) * R->n);
                if (R->kind == SOMHIP_SERVE_ACTIVATION) {
                    R->act = (double *)take(sizeof(double) * M->mn);
                    R->bounds = (unsigned long long *)take(512);
                    R->shifts = (int *)take(512);
                    R->table = (unsigned long long *)take(sizeof(unsigned long long) * (size_t)R->nwg * M->mn);
                }
                R->state = R->n > 0 ? UPLOADING : READY;
                active.push_back(R);
                progress = true;
            }
        }
        // pieces that completed
        for (auto &R : active) {
            hipError_t q = hipSuccess;
            if (poll_pieces(*R, &q)) progress = true;
            if (q != hipSuccess) {
                fail(std::string("upload event: ") + hipGetErrorString(q));
                return;
            }
            if (R->state == UPLOADING && R->ready_rows == R->n) R->state = READY;
        }
        // more pieces, round-robin over the requests with rows left to upload
        while (inflight < MAX_INFLIGHT && !active.empty()) {
            Request *pick = nullptr;
            for (size_t k = 0; k < active.size(); k++) {
                Request *R = active[(up_cursor + k) % active.size()].get();
                if (R->state == UPLOADING && R->up_rows < R->n) {
                    pick = R;
                    up_cursor = (up_cursor + k + 1) % active.size();
                    break;
                }
            }
            if (!pick) break;
            if (queue_piece(*pick)) {
                fail(std::string("upload: ") + somhip_last_error());
                return;
            }
            progress = true;
        }
        // batching: share tiles round-robin among the ready requests of each dimension class
        for (int dc = 0; dc < 2; dc++) {
            // requests with uploaded rows not yet launched (pieces are whole tiles, so a tile below
            // ready_rows is on the device), and finished requests (every row launched)
            std::vector<Request *> ready, done_rows;
            for (auto &R : active) {
                if (dclass_of(model_dpad[R->model]) != dc) continue;
                if ((R->state == UPLOADING || R->state == READY) && R->next_row < R->ready_rows) ready.push_back(R.get());
                else if (R->state == READY && R->next_row >= R->n) done_rows.push_back(R.get());
            }
            if (ready.empty() && done_rows.empty()) continue;
            // launches of this class that have completed free their slots
            while (!launch_evs[dc].empty()) {
                const hipError_t q = hipEventQuery(launch_evs[dc].front());
                if (q == hipErrorNotReady) break;
                if (q != hipSuccess) {
                    fail(std::string("launch event: ") + hipGetErrorString(q));
                    return;
                }
                ev_pool.push_back(launch_evs[dc].front());
                launch_evs[dc].pop_front();
            }
            if ((int)launch_evs[dc].size() >= MAX_LAUNCHES_INFLIGHT) ready.clear();  // they wait for a slot
            std::vector<BmuTile> tl;
            tl.reserve(MAX_TILES);
            std::vector<Request *> finished;
            const int share = ready.empty() ? 0 : std::max(16, MAX_TILES / (int)ready.size());
            for (int pass = 0; pass < 2 && (int)tl.size() < MAX_TILES; pass++) {
                for (Request *R : ready) {
                    int quota = pass == 0 ? share : MAX_TILES;
                    while (R->next_row < R->ready_rows && quota > 0 && (int)tl.size() < MAX_TILES) {
                        BmuTile t;
                        t.x = R->x;
                        t.n = R->n;
                        t.row0 = R->next_row;
                        t.c1 = R->c1;
                        t.c2 = nullptr;
                        t.d1 = nullptr;
                        t.model = R->model;
                        t.default_index = 0;
                        tl.push_back(t);
                        R->next_row += SOM_TILE_ROWS;
                        quota--;
                    }
                }
            }
            for (Request *R : ready)
                if (R->state == READY && R->next_row >= R->n) finished.push_back(R);
            for (Request *R : done_rows) finished.push_back(R);
            if (!tl.empty()) {
                int rc = launch_tiles(dc, tl);
                if (rc) {
                    fail(std::string("launch: ") + somhip_last_error());
                    return;
                }
            }
            for (Request *R : finished) {
                int rc = finish(*R);
                if (rc) {
                    fail(std::string("finish: ") + somhip_last_error());
                    return;
                }
            }
            if (!tl.empty() || !finished.empty()) progress = true;
        }
        // completions
        for (size_t i = 0; i < active.size();) {
            auto &R = active[i];
            if (R->state == FINISHING) {
                hipError_t q = hipEventQuery(R->ev_done);
                if (q == hipSuccess) {
                    std::lock_guard<std::mutex> g(mu);
                    R->t_done = Clock::now();
                    R->state = DONE;
                    slab.release(R->off, R->bytes);
                    completed++;
                    active.erase(active.begin() + i);
                    cv_done.notify_all();
                    progress = true;
                    continue;
                } else if (q != hipErrorNotReady) {
                    fail(std::string("completion event: ") + hipGetErrorString(q));
                    return;
                }
            }
            i++;
        }
        if (!progress) {
            if (!active.empty()) {
                // work in flight: poll again at once (a timed wait sleeps tens of microseconds past its
                // timeout, which small requests pay several times over)
                std::this_thread::yield();
                continue;
            }
            std::unique_lock<std::mutex> g(mu);
            if (pending.empty()) cv_work.wait_for(g, std::chrono::microseconds(2000));
        }
    }
}

extern "C" {

int somhip_host_alloc(size_t bytes, void **ptr)
{
    if (!ptr) return SOMHIP_ERR_ARG;
    SOM_HIP_TRY(hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return SOMHIP_OK;
}

int somhip_host_free(void *ptr)
{
    if (!ptr) return SOMHIP_OK;
    SOM_HIP_TRY(hipHostFree(ptr));
    return SOMHIP_OK;
}

int somhip_server_create(somhip_ctx *C, size_t memory_budget, int arithmetic, somhip_server **out)
{
    if (!C || !out) {
        somhip_set_error("null argument");
        return SOMHIP_ERR_ARG;
    }
    if (arithmetic != SOMHIP_RERANK_REF32 && arithmetic != SOMHIP_RERANK_FP64) {
        somhip_set_error("invalid arithmetic");
        return SOMHIP_ERR_ARG;
    }
    if (memory_budget < ((size_t)1 << 20)) {
        somhip_set_error("memory budget below 1 MiB");
        return SOMHIP_ERR_ARG;
    }
    *out = nullptr;
    SOM_HIP_TRY(hipSetDevice(C->device));
    auto S = std::make_unique<somhip_server>();
    S->ctx = C;
    S->arithmetic = arithmetic;
    SOM_HIP_TRY(hipStreamCreateWithFlags(&S->compute, hipStreamNonBlocking));
    SOM_HIP_TRY(hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking));
    SOM_HIP_TRY(hipMalloc((void **)&S->slab_mem, memory_budget));
    S->slab.init(S->slab_mem, memory_budget);
    SOM_HIP_TRY(hipMalloc((void **)&S->models_dev, sizeof(BmuModelDev) * MAX_MODELS));
    SOM_HIP_TRY(hipHostMalloc((void **)&S->staging, sizeof(BmuTile) * MAX_TILES * STAGING_SLOTS, hipHostMallocDefault));
    for (int i = 0; i < STAGING_SLOTS; i++) {
        SOM_HIP_TRY(hipEventCreateWithFlags(&S->staging_ev[i], hipEventDisableTiming));
        SOM_HIP_TRY(hipEventRecord(S->staging_ev[i], S->compute));
    }
    somhip_server *raw = S.release();
    raw->worker = std::thread([raw] { raw->loop(); });
    *out = raw;
    return SOMHIP_OK;
}

int somhip_server_destroy(somhip_server *S)
{
    if (!S) return SOMHIP_OK;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->ctx->device);
    (void)hipStreamSynchronize(S->compute);
    (void)hipStreamSynchronize(S->copy);
    for (auto &kv : S->all) {
        for (auto &pc : kv.second->pieces) (void)hipEventDestroy(pc.second);
        if (kv.second->ev_done) (void)hipEventDestroy(kv.second->ev_done);
    }
    for (auto e : S->ev_pool) (void)hipEventDestroy(e);
    for (auto &q : S->launch_evs)
        for (auto e : q) (void)hipEventDestroy(e);
    for (int dc = 0; dc < 2; dc++)
        for (int b = 0; b < NBUCKETS; b++) {
            GraphSlot &G = S->graphs[dc][b];
            if (G.exec) (void)hipGraphExecDestroy(G.exec);
            if (G.tiles) (void)hipFree(G.tiles);
            if (G.fb_list) (void)hipFree(G.fb_list);
            if (G.fb_count) (void)hipFree(G.fb_count);
        }
    for (int i = 0; i < STAGING_SLOTS; i++)
        if (S->staging_ev[i]) (void)hipEventDestroy(S->staging_ev[i]);
    if (S->staging) (void)hipHostFree(S->staging);
    if (S->models_dev) (void)hipFree(S->models_dev);
    if (S->slab_mem) (void)hipFree(S->slab_mem);
    (void)hipStreamDestroy(S->compute);
    (void)hipStreamDestroy(S->copy);
    delete S;
    return SOMHIP_OK;
}

int somhip_server_add_model(somhip_server *S, somhip_model *M, int *model_id)
{
    if (!S || !M || !model_id) {
        somhip_set_error("null argument");
        return SOMHIP_ERR_ARG;
    }
    if (M->ctx != S->ctx) {
        somhip_set_error("model belongs to another context");
        return SOMHIP_ERR_ARG;
    }
    if (M->f64) {
        somhip_set_error("serving takes float32 models");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    if (M->dpad > 32) {
        somhip_set_error("serving supports dim <= 32");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> g(S->mu);
    if ((int)S->models.size() >= MAX_MODELS) {
        somhip_set_error("too many models");
        return SOMHIP_ERR_ARG;
    }
    if (!S->all.empty()) {
        somhip_set_error("add models before submitting requests");
        return SOMHIP_ERR_STATE;
    }
    if (M->pblk * 32 >= 65536) {
        somhip_set_error("serving supports at most 65504 neuron slots (about 65,000 neurons)");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    int rc = somhip_prepare(M);
    if (rc) return rc;
    SOM_HIP_TRY(hipStreamSynchronize(M->ctx->stream));
    BmuModelDev d = somhip_model_dev(M);
    const int id = (int)S->models.size();
    SOM_HIP_TRY(hipMemcpy(S->models_dev + id, &d, sizeof(d), hipMemcpyHostToDevice));
    S->models.push_back(M);
    S->model_dpad.push_back(M->dpad);
    *model_id = id;
    return SOMHIP_OK;
}

int somhip_server_submit(somhip_server *S, int model_id, int kind, const float *x_host, const double *rho_host,
                         int64_t n, void *out_host, int64_t *ticket)
{
    if (!S || !ticket || !out_host || n < 0 || (n > 0 && !x_host)) {
        somhip_set_error("invalid submit arguments");
        return SOMHIP_ERR_ARG;
    }
    if (kind != SOMHIP_SERVE_WINNERS && kind != SOMHIP_SERVE_ACTIVATION) {
        somhip_set_error("invalid request kind");
        return SOMHIP_ERR_ARG;
    }
    auto R = std::make_shared<Request>();
    R->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (!S->error.empty()) {
            somhip_set_error("server failed: " + S->error);
            return SOMHIP_ERR_HIP;
        }
        if (model_id < 0 || model_id >= (int)S->models.size()) {
            somhip_set_error("unknown model id");
            return SOMHIP_ERR_ARG;
        }
    }
    somhip_model *M = S->models[model_id];
    R->model = model_id;
    R->kind = kind;
    R->x_host = x_host;
    R->rho_host = rho_host;
    R->n = n;
    R->out_host = out_host;
    if (kind == SOMHIP_SERVE_ACTIVATION)
        R->nwg = somhip_scatter_nwg(std::max<int64_t>(n, 1), M->mn, 1, S->ctx->num_cus, ACT_TABLE_BUDGET);
    size_t b = align_up(sizeof(float) * n * M->dim) + align_up(sizeof(int32_t) * std::max<int64_t>(n, 1));
    if (rho_host) b += align_up(sizeof(double) * n);
    if (kind == SOMHIP_SERVE_ACTIVATION)
        b += align_up(sizeof(double) * M->mn) + 2 * align_up(512) +
             align_up(sizeof(unsigned long long) * (size_t)R->nwg * M->mn);
    R->bytes = b;
    if (b > S->slab.capacity) {
        somhip_set_error("request exceeds the serving memory budget");
        return SOMHIP_ERR_ARG;
    }
    SOM_HIP_TRY(hipEventCreateWithFlags(&R->ev_done, hipEventDisableTiming));
    {
        std::lock_guard<std::mutex> g(S->mu);
        R->ticket = S->next_ticket++;
        S->pending.push_back(R);
        S->all[R->ticket] = R;
        *ticket = R->ticket;
    }
    S->cv_work.notify_one();
    return SOMHIP_OK;
}

int somhip_server_wait(somhip_server *S, int64_t ticket, double timeout_s, int *done, double *latency_ms)
{
    if (!S || !done) return SOMHIP_ERR_ARG;
    std::unique_lock<std::mutex> g(S->mu);
    auto it = S->all.find(ticket);
    if (it == S->all.end()) {
        somhip_set_error("unknown ticket");
        return SOMHIP_ERR_ARG;
    }
    auto R = it->second;
    auto pred = [&] { return R->state == DONE || R->state == FAILED; };
    if (timeout_s > 0) S->cv_done.wait_for(g, std::chrono::duration<double>(timeout_s), pred);
    if (R->state == FAILED) {
        somhip_set_error("server failed: " + S->error);
        return SOMHIP_ERR_HIP;
    }
    *done = R->state == DONE ? 1 : 0;
    if (*done) {
        if (latency_ms)
            *latency_ms = std::chrono::duration<double, std::milli>(R->t_done - R->t_submit).count();
        if (R->ev_done) (void)hipEventDestroy(R->ev_done);
        R->ev_done = nullptr;
        S->all.erase(it);
    }
    return SOMHIP_OK;
}

int somhip_server_stats(somhip_server *S, int64_t *completed, int64_t *launches, int64_t *tiles, int64_t *in_flight,
                        size_t *memory_in_use)
{
    if (!S) return SOMHIP_ERR_ARG;
    std::lock_guard<std::mutex> g(S->mu);
    if (completed) *completed = S->completed.load();
    if (launches) *launches = S->launches.load();
    if (tiles) *tiles = S->tiles_launched.load();
    if (in_flight) *in_flight = (int64_t)S->pending.size() + (int64_t)S->active.size();
    if (memory_in_use) *memory_in_use = S->slab.used;
    return SOMHIP_OK;
}

}  // extern "C"
