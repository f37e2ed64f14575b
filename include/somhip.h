/*
 * libsomhip: MI355X-native (gfx950) batch self-organising-map hot path.
 *
 * C ABI, plain pointers and sizes only. Every device pointer argument refers to memory on the
 * context's device (from somhip_malloc or any other HIP allocation on that device); every host
 * pointer argument is borrowed for the duration of the call. All functions return SOMHIP_OK (0) or
 * a negative error code; somhip_last_error() gives the message for the calling thread. No
 * exception crosses the ABI.
 *
 * Threading: one context owns one HIP stream and is not re-entrant. Work is enqueued on that stream;
 * functions that return data to the host synchronise the stream first.
 *
 * Reference interface replaced (paths relative to the reference repository,
 * lpsc-euclid/decontamination):
 *   - the launch protocol Kernel.__getitem__((enable_gpu, tpb, sizes))(*args)
 *     (decontamination/jit/__init__.py:334-412) and DeviceArray
 *     (decontamination/jit/__init__.py:96-296) -> contexts, models and device buffers below;
 *   - each Numba-CUDA kernel of the SOM path -> one entry point below, cited at its declaration.
 */
#ifndef SOMHIP_H
#define SOMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define SOMHIP_OK 0
#define SOMHIP_ERR_ARG (-1)         /* invalid argument (maps to ValueError) */
#define SOMHIP_ERR_HIP (-2)         /* HIP runtime error (maps to RuntimeError) */
#define SOMHIP_ERR_UNSUPPORTED (-3) /* shape/option outside what the kernels support */
#define SOMHIP_ERR_NODEVICE (-4)    /* no usable gfx950 device */
#define SOMHIP_ERR_STATE (-5)       /* call out of order (e.g. update before any accumulate) */

/* Topologies (decontamination/algo/som_abstract.py:81-111; None means hexagonal, :54). */
#define SOMHIP_TOPOLOGY_SQUARE 0
#define SOMHIP_TOPOLOGY_HEXAGONAL 1

/* BMU search precision. */
#define SOMHIP_PREC_EXACT 0       /* direct-difference scan on the vector ALUs, rerank arithmetic below */
#define SOMHIP_PREC_FP32 1        /* plain FP32 MFMA distance GEMM |w|^2 - 2 x.w, argmin on its values */
#define SOMHIP_PREC_BF16_RERANK 2 /* 3-term split-BF16 MFMA GEMM, provable candidate window, exact rerank;
                                    maps beyond 65,504 neuron slots run PREC_EXACT (same winners) */

/* Arithmetic of the exact distance (PREC_EXACT and the rerank of PREC_BF16_RERANK). */
#define SOMHIP_RERANK_REF32 0 /* reference: FP32 difference and square, FP64 sum (algo/__init__.py:89-98) */
#define SOMHIP_RERANK_FP64 1  /* FP64 difference, square and sum */

typedef struct somhip_ctx somhip_ctx;
typedef struct somhip_model somhip_model;
typedef struct somhip_accum somhip_accum;

/* Message of the last failing call on this thread ("" if none). */
const char *somhip_last_error(void);

/* Library version string. */
const char *somhip_version(void);

/* Number of visible HIP devices (0 when none). Replaces GPU_OPTIMIZATION_AVAILABLE
 * (decontamination/jit/__init__.py:37-43). */
int somhip_device_count(int *count);

/* Context: device + stream. Replaces the implicit numba default device/stream. */
int somhip_ctx_create(int device, somhip_ctx **out);
int somhip_ctx_destroy(somhip_ctx *ctx);
int somhip_ctx_synchronize(somhip_ctx *ctx);
/* Use an external hipStream_t (NULL restores the context's own stream). */
int somhip_ctx_set_stream(somhip_ctx *ctx, void *hip_stream);
int somhip_ctx_get_stream(somhip_ctx *ctx, void **hip_stream);

/* Per-kernel timing with HIP events on the context's stream (off by default). Enabling resets the
 * totals; get_timing synchronises and returns the summed duration and launch count of one kernel. */
#define SOMHIP_TIMER_BMU_EXACT 0
#define SOMHIP_TIMER_BMU_FP32 1
#define SOMHIP_TIMER_BMU_BF16 2
#define SOMHIP_TIMER_BMU_FALLBACK 3
#define SOMHIP_TIMER_PREP 4
#define SOMHIP_TIMER_ERRORS 5
#define SOMHIP_TIMER_ABSMAX 6
#define SOMHIP_TIMER_SCATTER 7
#define SOMHIP_TIMER_UPDATE 8
#define SOMHIP_TIMER_ITER 9
int somhip_ctx_set_timing(somhip_ctx *ctx, int enable);
int somhip_ctx_get_timing(somhip_ctx *ctx, int kernel, double *total_ms, int64_t *launches);

/* Device buffers: DeviceArray / device_array_{empty,zeros} / cu.to_device / copy_to_host
 * (decontamination/jit/__init__.py:96-296). Copies are stream-ordered; d2h synchronises. */
int somhip_malloc(somhip_ctx *ctx, size_t bytes, void **dptr);
int somhip_free(somhip_ctx *ctx, void *dptr);
int somhip_memcpy_h2d(somhip_ctx *ctx, void *dst, const void *src, size_t bytes);
int somhip_memcpy_d2h(somhip_ctx *ctx, void *dst, const void *src, size_t bytes);
int somhip_memcpy_d2d(somhip_ctx *ctx, void *dst, const void *src, size_t bytes);
int somhip_memset(somhip_ctx *ctx, void *dptr, int value, size_t bytes);

/* Chunk uploader for the generator protocol (decontamination/algo/__init__.py:21-27; chunk loops
 * som_batch.py:295-342, som_abstract.py:562-593, 632-664), replacing the synchronous per-launch
 * copies of decontamination/jit/__init__.py:372-373. A ring of `slots` (2..8) device slots is filled
 * on the uploader's own copy stream: put() copies one chunk (x and optional rho) into the next
 * slot, after the kernels that read that slot last time round have finished, and makes the
 * context's stream wait for the copy. The copy of chunk k+1 thus overlaps the kernels of chunk k.
 * The returned device pointers stay valid until `slots` further puts. Pageable chunks of 4 MB or
 * more are staged through pinned pieces by host threads, overlapped with their DMA; the 128 MB of
 * pieces are allocated once per process and shared by all uploaders, which take turns. Pageable host
 * memory may be reused once put() returns; memory from somhip_host_alloc is copied directly and
 * must stay untouched until the context's stream has passed the kernels that read the chunk. */
typedef struct somhip_uploader somhip_uploader;
int somhip_uploader_create(somhip_ctx *ctx, int slots, somhip_uploader **out);
int somhip_uploader_destroy(somhip_uploader *up);
int somhip_uploader_put(somhip_uploader *up, const void *x_host, size_t x_bytes, const void *rho_host,
                        size_t rho_bytes, void **x_dev, void **rho_dev);

/* Model: weights (mn, dim) FP32 and topography (mn, 2) FP32 resident on the device.
 * Replaces SOM_Abstract._weights/_topography being copied to the device on every launch
 * (decontamination/jit/__init__.py:372-373). topography_host must be the reference's FP32
 * topography (som_abstract.py:103-111); dim <= 64 (padded to 16, 32 or 64). */
int somhip_model_create(somhip_ctx *ctx, int m, int n, int dim, int topology, const float *topography_host,
                        somhip_model **out);
int somhip_model_destroy(somhip_model *model);
/* Copy weights in (from host if on_device == 0, else from a device pointer). */
int somhip_model_set_weights(somhip_model *model, const float *w, int on_device);
/* Copy weights out to a host buffer (mn*dim floats; synchronises). */
int somhip_model_get_weights(somhip_model *model, float *w_host);
/* Device pointer of the (mn, dim) weights. */
int somhip_model_weights_ptr(somhip_model *model, float **dptr);

/* Winners: out[i] = first index of the minimum distance of x_i, 0 when no distance is finite.
 * Replaces _find_bmus_kernel + _find_bmu_xpu (decontamination/algo/som_abstract.py:736-783).
 * x: device (n, dim) FP32; out: device int32 (n). */
int somhip_find_bmus(somhip_model *model, const float *x, int64_t n, int precision, int arithmetic, int32_t *out);

/* Winners, as somhip_find_bmus, without allocating: no hipMalloc/hipFree or other allocation inside the
 * call (kernel timers are suspended for its duration). No counterpart in the reference: it is for callers
 * that cannot allocate, such as code under HIP graph capture and serving loops. The BF16 path's fallback
 * list lives in a fixed device array of 256 + 8 * SOMHIP_NOALLOC_MAX_ROWS bytes (16 MB) that every device
 * holds for the life of the process, cleared by a kernel at the start of each call; n above
 * SOMHIP_NOALLOC_MAX_ROWS is rejected with SOMHIP_ERR_ARG (split the input). All calls on one device share
 * that array: calls that overlap on one device (several contexts or streams at once) race. When the weights
 * changed since the last search, the call first rebuilds the model's derived data with the prep kernels,
 * in the model's own buffers. */
#define SOMHIP_NOALLOC_MAX_ROWS (1 << 21)
int somhip_find_bmus_noalloc(somhip_model *model, const float *x, int64_t n, int precision, int arithmetic,
                             int32_t *out);

/* Activation map: result[bmu(x_i)] += rho_i for rho_i > 0 (rho == NULL means ones), accumulated
 * into the caller's device FP64 array (mn). Per call the sum is exact (fixed-point integers), so
 * the result does not depend on launch shape or batching.
 * Replaces _count_bmus_kernel (decontamination/algo/som_abstract.py:699-732). */
int somhip_count_bmus(somhip_model *model, const float *x, const double *rho, int64_t n, int precision,
                      int arithmetic, double *result);

/* Errors: errors[0] += sum rho*sqrt(d1), errors[1] += sum rho*[|t1-t2|^2 > r^2], errors[2] += sum rho
 * over samples with rho > 0 and two distinct valid BMUs (r^2 = 2 square, 1 hexagonal), accumulated
 * into the caller's device FP64 array (3); deterministic fixed-order sums.
 * Replaces _compute_errors_kernel + _compute_errors_xpu (decontamination/algo/som_abstract.py:787-878). */
int somhip_compute_errors(somhip_model *model, const float *x, const double *rho, int64_t n, int precision,
                          int arithmetic, double *errors);

/* Galaxy number density from the footprint winners: the two activation maps of
 * Decontamination_SOM.compute_gndm (catalog: rho = coverage-corrected density, footprint: ones;
 * rows with rho > 0 count, the _count_bmus_kernel rule) cast to FP32, gnd = catalog / footprint in
 * FP32 where footprint > 0 and NaN elsewhere, and gndm[i] = gnd[winners[i]] (NaN for an index outside
 * [0, mn)). Replaces the two get_activation_map passes and _compute_gndm
 * (decontamination/decontamination/decontamination_som.py:600-667): the winners of the footprint
 * are the BMUs both maps would search again. Device pointers: winners int32 (n) and gndm FP32 (n),
 * both 16-byte aligned; rho FP64 (n); catalog_map, footprint_map, gnd FP32 (mn). */
int somhip_gndm(somhip_ctx *ctx, const int32_t *winners, const double *rho, int64_t n, int mn, float *catalog_map,
                float *footprint_map, float *gnd, float *gndm);

/* Weighted covariance sums of one chunk for Covariance.compute (decontamination/algo/covariance.py,
 * _update_welford_cov_sums and _finalize_welford_cov), the pass that initialises SOM_PCA. A row
 * counts when its weight is finite and > 0 and all its values are finite (rho == NULL means ones).
 * x: device (n, dim), FP32 or (x_is_f64) FP64; 1 <= dim <= 64. out_host (FP64, 1 + dim + dim*dim):
 * total weight W, mean (dim), co-moment M2 = sum w (x - mean)(x - mean)^T (dim x dim). Two passes over
 * the chunk with fixed-order reductions (deterministic); chunks merge with Chan's update, and the
 * covariance is M2 / (W - ddof). Synchronises. */
int somhip_covariance_chunk(somhip_ctx *ctx, const void *x, int x_is_f64, const double *rho, int64_t n, int dim,
                            double *out_host);

/* Training accumulator: per-neuron sums S_c = sum rho_i x_i and c_c = sum rho_i over the samples whose
 * BMU is c (epoch mode), or the neighbourhood-weighted numerator/denominator (vector-count mode).
 * Replaces the numerator/denominator DeviceArrays of SOM_Batch.train
 * (decontamination/algo/som_batch.py:289-291,394-396). */
int somhip_accum_create(somhip_model *model, somhip_accum **out);
int somhip_accum_destroy(somhip_accum *acc);
int somhip_accum_reset(somhip_accum *acc);

/* Epoch-mode accumulation of one chunk: BMUs of x under the current weights, then exact fixed-point
 * per-BMU sums of rho*x and rho (samples with rho <= 0 or no valid BMU are skipped, as in
 * som_batch.py:512-514,538-540). If errors != NULL, the errors of the current weights over this
 * chunk are accumulated into it exactly as somhip_compute_errors would (fused, no second pass).
 * Replaces SOM_Batch._train_step1_epoch_kernel + _train_step2_xpu
 * (decontamination/algo/som_batch.py:94-150,509-566) in the factorised form num = H.S, den = H.c. */
int somhip_batch_epoch_accumulate(somhip_accum *acc, const float *x, const double *rho, int64_t n, int precision,
                                  int arithmetic, double *errors);

/* Epoch-mode update: sigma = sigma0 / (1 + 2*cur_epoch/n_epochs) clamped >= 1e-12, then
 * num = H.S, den = H.c with H_jc = exp(-|t_j - t_c|^2 / (2 sigma^2)) in FP64, and
 * w_j = num_j / den_j if den_j > 1e-12 else 0 written to the model (and to history_out, a device
 * (mn, dim) FP32 buffer, if not NULL). Replaces the per-thread neighbourhood scatter
 * (som_batch.py:548-566) and the host division (som_batch.py:354-359). */
int somhip_batch_epoch_update(somhip_accum *acc, int64_t cur_epoch, int64_t n_epochs, float sigma0,
                              float *history_out);

/* Vector-count-mode accumulation of one chunk (sigma_i = sigma0/(1 + 2*(cur_vector+i)/n_vectors)
 * per sample, so H is not factorisable): num_j += rho_i h_ji x_i, den_j += rho_i h_ji in FP64,
 * deterministic. Replaces SOM_Batch._train_step1_iter_kernel (som_batch.py:154-210). */
int somhip_batch_iter_accumulate(somhip_accum *acc, const float *x, const double *rho, int64_t n,
                                 int64_t cur_vector, int64_t n_vectors, float sigma0, int precision,
                                 int arithmetic);
/* Vector-count-mode division (som_batch.py:481-486). */
int somhip_batch_iter_update(somhip_accum *acc, float *history_out);

/* ------------------------------------------------------------------------------------------------
 * float64 SOMs: the reference's dtype=np.float64 (som_abstract.py:46-68, som_batch.py:289-291,
 * 311-315), where weights, topography, data (`vectors.astype(dtype)`), sigma and the sums are FP64.
 * A float64 model holds FP64 weights and topography; its entry points take FP64 device rows and
 * search with the exact FP64 scan (the reference's square_distance_xpu on float64 operands:
 * FP64 difference, square and sequential sum), so they have no precision/arithmetic arguments.
 * The per-BMU sums are the same exact fixed-point sums of the FP64 products rho*x; the density is not
 * rounded. The update computes the neighbourhood from FP64 topography differences and writes FP64
 * weights. Each replaces the float32 entry point of the same name (same outputs and errors); the
 * float32 entry points reject float64 models and the reverse, and serving takes float32 models only.
 * ------------------------------------------------------------------------------------------------ */
int somhip_model_create_f64(somhip_ctx *ctx, int m, int n, int dim, int topology, const double *topography_host,
                            somhip_model **out);
int somhip_model_set_weights_f64(somhip_model *model, const double *w, int on_device);
int somhip_model_get_weights_f64(somhip_model *model, double *w_host);
int somhip_find_bmus_f64(somhip_model *model, const double *x, int64_t n, int32_t *out);
int somhip_count_bmus_f64(somhip_model *model, const double *x, const double *rho, int64_t n, double *result);
int somhip_compute_errors_f64(somhip_model *model, const double *x, const double *rho, int64_t n, double *errors);
int somhip_batch_epoch_accumulate_f64(somhip_accum *acc, const double *x, const double *rho, int64_t n,
                                      double *errors);
int somhip_batch_epoch_update_f64(somhip_accum *acc, int64_t cur_epoch, int64_t n_epochs, double sigma0,
                                  double *history_out);
int somhip_batch_iter_accumulate_f64(somhip_accum *acc, const double *x, const double *rho, int64_t n,
                                     int64_t cur_vector, int64_t n_vectors, double sigma0);
int somhip_batch_iter_update_f64(somhip_accum *acc, double *history_out);

/* ------------------------------------------------------------------------------------------------
 * Serving. No counterpart in the reference (it has no serving loop); built on the same kernels as
 * somhip_find_bmus / somhip_count_bmus with PREC_BF16_RERANK.
 *
 * A server keeps models resident, admits requests while their working memory fits in
 * `memory_budget` bytes of device memory (one slab allocated at creation), uploads their rows on a
 * copy stream, and batches the 512-row tiles of every uploaded request into shared launches
 * (hipGraphs captured per dimension class and tile-count bucket). Results are bit-identical to the
 * single-request entry points, alone or in any mix. A scheduler thread owns the server's streams.
 * Host buffers passed to submit must stay valid until the request completes; pinned buffers from
 * somhip_host_alloc avoid staging copies.
 * ------------------------------------------------------------------------------------------------ */
#define SOMHIP_SERVE_WINNERS 0    /* out_host: int32[n] BMU indices (default 0) */
#define SOMHIP_SERVE_ACTIVATION 1 /* out_host: double[mn] activation map (rho == NULL means ones) */

typedef struct somhip_server somhip_server;

int somhip_host_alloc(size_t bytes, void **ptr);
int somhip_host_free(void *ptr);
int somhip_server_create(somhip_ctx *ctx, size_t memory_budget, int arithmetic, somhip_server **out);
/* Waits for every submitted request, then frees the server. */
int somhip_server_destroy(somhip_server *server);
/* Register a model (its weights are prepared once; models must be added before the first submit). */
int somhip_server_add_model(somhip_server *server, somhip_model *model, int *model_id);
int somhip_server_submit(somhip_server *server, int model_id, int kind, const float *x_host, const double *rho_host,
                         int64_t n, void *out_host, int64_t *ticket);
/* Waits up to timeout_s (0 = poll) for a ticket; on completion sets *done = 1 and the latency from
 * submit to the result being in out_host, and forgets the ticket. */
int somhip_server_wait(somhip_server *server, int64_t ticket, double timeout_s, int *done, double *latency_ms);
int somhip_server_stats(somhip_server *server, int64_t *completed, int64_t *launches, int64_t *tiles,
                        int64_t *in_flight, size_t *memory_in_use);

#ifdef __cplusplus
}
#endif

#endif /* SOMHIP_H */
