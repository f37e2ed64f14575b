// Chunk uploader: overlaps the host-to-device copy of chunk k+1 with the kernels of chunk k.
//
// The reference copies every chunk to the device inside each kernel launch, synchronously
// (decontamination/jit/__init__.py:372-373), once per chunk of the generator protocol
// (decontamination/algo/__init__.py:21-27, som_batch.py:295-342). Here chunks go through a ring of
// device slots filled on a copy stream of the uploader's own: the copy into slot s waits (on the
// device) for the kernels that read slot s the previous time round, and the context's stream waits
// for the copy, so the chunk loop on the host only ever blocks inside the copy itself.
#include "common.hpp"

struct somhip_uploader {
    somhip_ctx *ctx = nullptr;
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
    for (auto &b : U->x) b.release();
    for (auto &b : U->rho) b.release();
    for (auto e : U->done) if (e) (void)hipEventDestroy(e);
    for (auto e : U->ready) if (e) (void)hipEventDestroy(e);
    if (U->copy) (void)hipStreamDestroy(U->copy);
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
    if (x_bytes) SOM_HIP_TRY(hipMemcpyAsync(U->x[s].p, x_host, x_bytes, hipMemcpyHostToDevice, U->copy));
    if (rho_bytes) SOM_HIP_TRY(hipMemcpyAsync(U->rho[s].p, rho_host, rho_bytes, hipMemcpyHostToDevice, U->copy));
    SOM_HIP_TRY(hipEventRecord(U->ready[s], U->copy));
    SOM_HIP_TRY(hipStreamWaitEvent(C->stream, U->ready[s], 0));
    U->cur = s;
    *x_dev = U->x[s].p;
    if (rho_dev) *rho_dev = rho_bytes ? U->rho[s].p : nullptr;
    return SOMHIP_OK;
}

}  // extern "C"
