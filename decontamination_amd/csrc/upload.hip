// Chunk uploader: overlaps the host-to-device copy of chunk k+1 with the kernels of chunk k.
//
// The reference copies every chunk to the device inside each kernel launch, synchronously
// (decontamination/jit/__init__.py:372-373), once per chunk of the generator protocol
// (decontamination/algo/__init__.py:21-27, som_batch.py:295-342). Here chunks go through a ring of
// device slots filled on a copy stream of the uploader's own: the copy into slot s waits (on the
// device) for the kernels that read slot s the previous time round, and the context's stream waits
// for the copy, so the chunk loop on the host only ever blocks inside the copy itself.
//
// Pageable host memory goes through a ring of pinned staging pieces: host threads copy piece i+1
// while the DMA engine drains piece i, so a chunk moves at the pinned DMA rate (about 58 GB/s on the
// MI355X hosts, against about 29 GB/s for the runtime's own pageable path). Pinned memory (from
// somhip_host_alloc) is copied directly.
#include "common.hpp"

#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>

namespace {

constexpr int STAGE_PIECES = 4;                  // pinned staging pieces in the ring
constexpr size_t STAGE_PIECE = (size_t)32 << 20; // bytes per piece
constexpr int COPY_THREADS = 8;                  // host threads per piece copy (including the caller)
constexpr size_t STAGE_MIN = (size_t)4 << 20;    // smaller pageable copies go straight to the runtime

// Process-wide staging: the pinned pieces and the helper threads that fill them (a memcpy of one
// piece is split in COPY_THREADS parts; the caller takes part 0). Created on first use and kept until
// the process ends, so any number of uploaders shares 128 MB of pinned memory; they take turns.
struct CopyPool {
    std::vector<std::thread> th;
    std::mutex use;  // held by one uploader's staged copy at a time
    std::mutex m;
    char *stage[STAGE_PIECES] = {};
    somhip_uploader *last[STAGE_PIECES] = {};  // uploader whose DMA read the piece last (null: none pending)
    int next_piece = 0;
    std::condition_variable go, fin;
    char *dst = nullptr;
    const char *src = nullptr;
    size_t n = 0;
    long gen = 0;
    int pending = 0;

    void part(int k, char *d, const char *s, size_t bytes)
    {
        const size_t per = (bytes / COPY_THREADS + 63) & ~(size_t)63;
        const size_t a = std::min(bytes, per * k), b = k == COPY_THREADS - 1 ? bytes : std::min(bytes, per * (k + 1));
        if (b > a) memcpy(d + a, s + a, b - a);
    }
    void start()
    {
        for (int k = 1; k < COPY_THREADS; k++)
            th.emplace_back([this, k] {
                long seen = 0;
                for (;;) {
                    char *d;
                    const char *s;
                    size_t bytes;
                    {
                        std::unique_lock<std::mutex> l(m);
                        go.wait(l, [&] { return gen != seen; });
                        seen = gen;
                        d = dst; s = src; bytes = n;
                    }
                    part(k, d, s, bytes);
                    std::lock_guard<std::mutex> l(m);
                    if (--pending == 0) fin.notify_one();
                }
            });
    }
    // caller holds `use`
    void copy(char *d, const char *s, size_t bytes)
    {
        {
            std::lock_guard<std::mutex> l(m);
            dst = d; src = s; n = bytes;
            pending = COPY_THREADS - 1;
            gen++;
        }
        go.notify_all();
        part(0, d, s, bytes);
        std::unique_lock<std::mutex> l(m);
        fin.wait(l, [&] { return pending == 0; });
    }
};

CopyPool &copy_pool()
{
    static CopyPool *pool = [] {
        CopyPool *p = new CopyPool();  // never destroyed: its threads wait on a condition at exit
        p->start();
        return p;
    }();
    return *pool;
}

}  // namespace

struct somhip_uploader {
    somhip_ctx *ctx = nullptr;
    hipEvent_t drained[STAGE_PIECES] = {};  // per staging piece: recorded on `copy` after this uploader's DMA
    bool staged = false;                    // has used the staging pieces
    hipStream_t copy = nullptr;
    int slots = 0;
    int cur = -1;                   // slot handed out by the last put (-1: none yet)
    std::vector<DevBuf> x, rho;     // per slot
    std::vector<hipEvent_t> done;   // per slot: recorded on the compute stream after its last reader
    std::vector<hipEvent_t> ready;  // per slot: recorded on the copy stream after its copy
    std::vector<char> done_valid;
};

namespace {

void uploader_release(somhip_uploader *U)
{
    if (U->copy) (void)hipStreamSynchronize(U->copy);
    if (U->ctx) (void)hipStreamSynchronize(U->ctx->stream);
    if (U->staged) {
        CopyPool &P = copy_pool();
        std::lock_guard<std::mutex> turn(P.use);
        for (int p = 0; p < STAGE_PIECES; p++)
            if (P.last[p] == U) P.last[p] = nullptr;  // its DMAs have finished with the stream
    }
    for (auto e : U->drained) if (e) (void)hipEventDestroy(e);
    for (auto &b : U->x) b.release();
    for (auto &b : U->rho) b.release();
    for (auto e : U->done) if (e) (void)hipEventDestroy(e);
    for (auto e : U->ready) if (e) (void)hipEventDestroy(e);
    if (U->copy) (void)hipStreamDestroy(U->copy);
}

bool is_pinned(const void *p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // pageable memory is an expected miss, not an error to keep
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

// Queues the copy of `bytes` from host `src` to device `dst` on the copy stream. Pageable sources of
// STAGE_MIN bytes or more go through the staging ring; returns once the host bytes have all been read
// (the runtime's own pageable copy also returns only then).
int stage_copy(somhip_uploader *U, char *dst, const char *src, size_t bytes)
{
    if (!bytes) return SOMHIP_OK;
    if (bytes < STAGE_MIN || is_pinned(src)) {
        SOM_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, U->copy));
        return SOMHIP_OK;
    }
    CopyPool &P = copy_pool();
    std::lock_guard<std::mutex> turn(P.use);
    for (int p = 0; p < STAGE_PIECES; p++)
        if (!P.stage[p]) SOM_HIP_TRY(hipHostMalloc((void **)&P.stage[p], STAGE_PIECE, hipHostMallocPortable));
    for (size_t off = 0; off < bytes; off += STAGE_PIECE) {
        const size_t len = std::min(STAGE_PIECE, bytes - off);
        const int p = P.next_piece;
        P.next_piece = (p + 1) % STAGE_PIECES;
        if (P.last[p]) SOM_HIP_TRY(hipEventSynchronize(P.last[p]->drained[p]));  // its last DMA is done
        P.last[p] = nullptr;
        P.copy(P.stage[p], src + off, len);
        SOM_HIP_TRY(hipMemcpyAsync(dst + off, P.stage[p], len, hipMemcpyHostToDevice, U->copy));
        SOM_HIP_TRY(hipEventRecord(U->drained[p], U->copy));
        P.last[p] = U;
        U->staged = true;
    }
    return SOMHIP_OK;
}

}  // namespace

extern "C" {

int somhip_uploader_create(somhip_ctx *C, int slots, somhip_uploader **out)
{
    if (!C || !out || slots < 2 || slots > 8) {
        somhip_set_error("somhip_uploader_create: need a context and 2 <= slots <= 8");
        return SOMHIP_ERR_ARG;
    }
    *out = nullptr;
    SOM_HIP_TRY(hipSetDevice(C->device));
    somhip_uploader *U = new somhip_uploader();
    U->ctx = C;
    U->slots = slots;
    U->x.resize(slots);
    U->rho.resize(slots);
    U->done.assign(slots, nullptr);
    U->ready.assign(slots, nullptr);
    U->done_valid.assign(slots, 0);
    hipError_t e = hipStreamCreateWithFlags(&U->copy, hipStreamNonBlocking);
    for (int s = 0; e == hipSuccess && s < slots; s++) {
        e = hipEventCreateWithFlags(&U->done[s], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&U->ready[s], hipEventDisableTiming);
    }
    for (int p = 0; e == hipSuccess && p < STAGE_PIECES; p++)
        e = hipEventCreateWithFlags(&U->drained[p], hipEventDisableTiming);
    if (e != hipSuccess) {
        somhip_set_error(std::string("somhip_uploader_create: ") + hipGetErrorString(e));
        uploader_release(U);
        delete U;
        return SOMHIP_ERR_HIP;
    }
    *out = U;
    return SOMHIP_OK;
}

int somhip_uploader_destroy(somhip_uploader *U)
{
    if (!U) return SOMHIP_OK;
    (void)hipSetDevice(U->ctx->device);
    uploader_release(U);
    delete U;
    return SOMHIP_OK;
}

int somhip_uploader_put(somhip_uploader *U, const void *x_host, size_t x_bytes, const void *rho_host,
                        size_t rho_bytes, void **x_dev, void **rho_dev)
{
    if (!U || !x_dev || (x_bytes && !x_host) || (rho_bytes && !rho_host) || (rho_bytes && !rho_dev)) {
        somhip_set_error("somhip_uploader_put: invalid arguments");
        return SOMHIP_ERR_ARG;
    }
    somhip_ctx *C = U->ctx;
    SOM_HIP_TRY(hipSetDevice(C->device));
    // everything queued on the context's stream so far is the last reader of the previous slot
    if (U->cur >= 0) {
        SOM_HIP_TRY(hipEventRecord(U->done[U->cur], C->stream));
        U->done_valid[U->cur] = 1;
    }
    const int s = (U->cur + 1) % U->slots;
    if (U->done_valid[s]) {
        // growing a slot frees it: its readers must have finished first
        if (x_bytes > U->x[s].bytes || rho_bytes > U->rho[s].bytes) SOM_HIP_TRY(hipEventSynchronize(U->done[s]));
        SOM_HIP_TRY(hipStreamWaitEvent(U->copy, U->done[s], 0));
    }
    int rc;
    if ((rc = U->x[s].ensure(x_bytes ? x_bytes : 1)) || (rc = U->rho[s].ensure(rho_bytes ? rho_bytes : 1))) return rc;
    if ((rc = stage_copy(U, (char *)U->x[s].p, (const char *)x_host, x_bytes)) ||
        (rc = stage_copy(U, (char *)U->rho[s].p, (const char *)rho_host, rho_bytes)))
        return rc;
    SOM_HIP_TRY(hipEventRecord(U->ready[s], U->copy));
    SOM_HIP_TRY(hipStreamWaitEvent(C->stream, U->ready[s], 0));
    U->cur = s;
    *x_dev = U->x[s].p;
    if (rho_dev) *rho_dev = rho_bytes ? U->rho[s].p : nullptr;
    return SOMHIP_OK;
}

}  // extern "C"
