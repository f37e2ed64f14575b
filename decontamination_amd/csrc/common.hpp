// Shared device helpers and internal host structures for libsomhip (gfx950 / CDNA4 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include <string>
#include <vector>

#include "../../include/somhip.h"

// ------------------------------------------------------------------------------------------------
// Vector types for the MFMA builtins.
// ------------------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

#define SOM_FLOAT_EPS 1.0e-12

// 32x32 MFMA accumulator layout (v_mfma_f32_32x32x{2f32,16_bf16}): lane l, register r holds
// D[row][col] with col = l % 32 and row = 8*(r/4) + 4*(l/32) + (r%4).
__device__ __forceinline__ int acc_row(int r, int lane) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// Round-to-nearest-even float -> bf16 bits (NaN kept quiet).
__host__ __device__ __forceinline__ unsigned short f32_to_bf16_bits(float f)
{
    unsigned int u = __builtin_bit_cast(unsigned int, f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__host__ __device__ __forceinline__ float bf16_bits_to_f32(unsigned short b)
{
    return __builtin_bit_cast(float, (unsigned int)b << 16);
}

// The reference's square_distance_xpu (decontamination/algo/__init__.py:89-98): FP32 difference,
// FP32 square (numba types float32 ** int as float32), FP64 sequential accumulation.
// The explicit _rn intrinsics keep the compiler from contracting or reassociating.
__device__ __forceinline__ double sqd_ref32(const float *a, const float *b, int dim)
{
    double r = 0.0;
    for (int k = 0; k < dim; k++) {
        float d = __fsub_rn(a[k], b[k]);
        float s = __fmul_rn(d, d);
        r = __dadd_rn(r, (double)s);
    }
    return r;
}

// FP64 direct difference (the FP64 oracle's arithmetic).
__device__ __forceinline__ double sqd_fp64(const float *a, const float *b, int dim)
{
    double r = 0.0;
    for (int k = 0; k < dim; k++) {
        double d = __dsub_rn((double)a[k], (double)b[k]);
        r = __dadd_rn(r, __dmul_rn(d, d));
    }
    return r;
}

// The same direct difference on FP64 operands (float64 SOMs: data, weights and topography in FP64).
__device__ __forceinline__ double sqd_f64(const double *a, const double *b, int dim)
{
    double r = 0.0;
    for (int k = 0; k < dim; k++) {
        double d = __dsub_rn(a[k], b[k]);
        r = __dadd_rn(r, __dmul_rn(d, d));
    }
    return r;
}

// square_distance_xpu of two grid positions in the topography's own type
__device__ __forceinline__ double sqd_topo(const float *a, const float *b) { return sqd_ref32(a, b, 2); }
__device__ __forceinline__ double sqd_topo(const double *a, const double *b) { return sqd_f64(a, b, 2); }

template <int MODE>
__device__ __forceinline__ double sqd_mode(const float *a, const float *b, int dim)
{
    if constexpr (MODE == SOMHIP_RERANK_FP64) return sqd_fp64(a, b, dim);
    else return sqd_ref32(a, b, dim);
}

// Reference top-2 update rule (decontamination/algo/som_abstract.py:832-853) on (distance, index)
// candidates visited in any order: the result equals the in-order scan when ties are broken by the
// lower index. Distances that are NaN or >= 1e99 never enter (the reference starts from 1e99).
struct Top2 {
    double d1, d2;
    int i1, i2;
    __device__ __forceinline__ void init() { d1 = d2 = 1.0e99; i1 = i2 = -1; }
    __device__ __forceinline__ static bool less(double da, int ia, double db, int ib)
    {
        return da < db || (da == db && (unsigned)ia < (unsigned)ib);
    }
    __device__ __forceinline__ void push(double d, int idx)
    {
        if (!(d < 1.0e99)) return;  // NaN or too large
        if (idx == i1 || idx == i2) return;
        if (less(d, idx, d1, i1)) {
            d2 = d1; i2 = i1; d1 = d; i1 = idx;
        } else if (less(d, idx, d2, i2)) {
            d2 = d; i2 = idx;
        }
    }
};

// ------------------------------------------------------------------------------------------------
// Descriptors shared by the BMU kernels and the serving loop.
// ------------------------------------------------------------------------------------------------
// Everything the BF16 kernel needs to know about one prepared model.
struct BmuModelDev {
    const bf16x8 *frag;     // bf16 A fragments (permuted slots)
    const f32x4 *norms4;    // |w'|^2 per permuted slot
    const int *perm;        // slot -> neuron
    const float *w;         // FP32 weights (mn, dim)
    const float *centre;    // weight centroid (dpad)
    const float *wmax;      // max_j |w'_jk| (dpad)
    const float *scal;      // [0] max_j |w'_j|^2
    const float *wslot;     // FP32 weights (as given) per permuted slot, (slots, dpad), zero padded
    int nblk;               // 32-slot blocks of the permuted order
    int nblk_a;             // leading blocks that hold the coarse sub-lattice (== nblk: none)
    int mn, dim;
    int ncol;               // map columns (neuron index = row * ncol + column)
};

// One unit of serving work: rows [row0, row0 + SOM_TILE_ROWS) of one request.
struct BmuTile {
    const float *x;     // request samples (n, dim)
    int64_t n;          // request rows (0 = empty tile)
    int64_t row0;       // first row of this tile
    int32_t *c1;        // request outputs
    int32_t *c2;
    double *d1;
    int model;          // index into the model table
    int default_index;
};

constexpr int SOM_TILE_ROWS = 512;     // rows per BF16 workgroup (8 waves x 2 groups x 32)
constexpr int FB_TILE_SHIFT = 40;      // serving fallback entry = (tile << 40) | row

// ------------------------------------------------------------------------------------------------
// Host-side structures.
// ------------------------------------------------------------------------------------------------
void somhip_set_error(const std::string &msg);

#define SOM_HIP_TRY(expr)                                                                       \
    do {                                                                                        \
        hipError_t _e = (expr);                                                                 \
        if (_e != hipSuccess) {                                                                 \
            somhip_set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                \
            return SOMHIP_ERR_HIP;                                                              \
        }                                                                                       \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need)
    {
        if (need <= bytes) return SOMHIP_OK;
        if (p) SOM_HIP_TRY(hipFree(p));
        p = nullptr;
        bytes = 0;
        SOM_HIP_TRY(hipMalloc(&p, need));
        bytes = need;
        return SOMHIP_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// Per-call device workspace (BMU outputs, reduction partials, fixed-point scales and tables).
struct Workspace {
    DevBuf c1, c2, d1, part, bounds, shifts, table, xmax, maps, bucket;
    // Training: c1b holds the BMUs of the previous epoch_accumulate call, whose samples were c1_x / c1_n;
    // the BF16 search groups each tile's rows by them (a stale buffer only changes the grouping's speed).
    DevBuf c1b;
    const void *c1_x = nullptr;
    int64_t c1_n = -1;
    long long c1_uploads = -1;  // the model's upload count when c1 was written
    void release()
    {
        c1.release(); c2.release(); d1.release(); part.release(); bounds.release(); shifts.release(); table.release();
        xmax.release(); maps.release(); bucket.release(); c1b.release();
        c1_x = nullptr;
        c1_n = -1;
    }
};

// Optional per-kernel timing with HIP events recorded on the context's stream.
struct KernelTimer {
    bool enabled = false;
    struct Pending {
        int id;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double total_ms[16] = {0};
    long long launches[16] = {0};
};

enum TimerId {
    TK_BMU_EXACT = 0, TK_BMU_FP32, TK_BMU_BF16, TK_BMU_FALLBACK, TK_PREP, TK_ERRORS, TK_ABSMAX, TK_SCATTER,
    TK_UPDATE, TK_ITER, TK_COUNT
};

struct somhip_ctx {
    int device = 0;
    KernelTimer timer;
    hipStream_t stream = nullptr;      // stream in use (own or external)
    hipStream_t own_stream = nullptr;  // stream created by the context
    // growable device scratch for the BMU fallback list
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    int num_cus = 256;
    size_t lds_per_block = 65536;  // device limit of LDS bytes per workgroup
    Workspace ws;  // inference entry points
    void *noalloc_scratch = nullptr;  // this device's copy of the allocation-free entry point's scratch array
};

// Derived per-weights data prepared by the prep kernels before any BMU search.
struct somhip_model {
    somhip_ctx *ctx = nullptr;
    int m = 0, n = 0, mn = 0, dim = 0, dpad = 0, topology = SOMHIP_TOPOLOGY_HEXAGONAL;
    bool separable = false; // topography is the closed-form square / hexagonal grid (separable neighbourhood)
    int nblk = 0;          // number of 32-neuron blocks in natural order (mn rounded up to 32)
    int pblk = 0;          // number of 32-slot blocks in the permuted (BF16) order
    int pblk_a = 0;        // leading permuted blocks holding the coarse sub-lattice (== pblk: none)
    float *w = nullptr;    // (mn, dim) FP32 weights, row-major
    float *topo = nullptr; // (mn, 2) FP32 topography
    // float64 models (somhip_model_create_f64) hold their weights and topography here instead and have
    // none of the FP32 buffers or prepared data
    bool f64 = false;
    double *w64 = nullptr;
    double *topo64 = nullptr;
    long long uploads = 0; // weight uploads by the caller (set_weights); an epoch's own update does not count
    // prepared data (valid when prepared == true)
    bool prepared = false;
    float *centre = nullptr;        // (dpad) FP32 centroid of the weights
    float *wmax = nullptr;          // (dpad) max_j |w'_jk|
    float *scal = nullptr;          // [0] = max_j |w'_j|^2 (as float)
    float *norms_perm = nullptr;    // (pblk*32) |w'|^2 in permuted slot order (padding slots = +inf)
    float *norms_nat = nullptr;     // (nblk*32) |w'|^2 in natural order (padding slots = +inf)
    unsigned short *frag_bf = nullptr;  // bf16 A fragments [blk][kc][hi,lo][64 lanes][8]
    float *frag_f32 = nullptr;      // FP32 A fragments [blk][dpad/8][64 lanes][4]
    float *wslot = nullptr;         // FP32 weights (as given) per permuted slot [slot][dpad], zero padded
    int *perm = nullptr;            // (pblk*32) slot -> neuron (or -1 for padding)
};

// Timing helpers: bracket a launch with timer_begin / timer_end (no-ops unless enabled).
struct TimerScope {
    somhip_ctx *C;
    int id;
    hipEvent_t a = nullptr, b = nullptr;
    TimerScope(somhip_ctx *c, int i);
    ~TimerScope();
};
int somhip_timer_collect(somhip_ctx *C);

// Internal launchers (return SOMHIP_OK or an error code).
int somhip_prepare(somhip_model *mdl);
// The BMU precision a launch with `precision` runs on this model (BF16_RERANK maps beyond 65,504
// neuron slots run the exact kernel, which returns the same winners).
int somhip_bmu_precision(const somhip_model *mdl, int precision);
// xmax_part (nullable): per-wave rows of dpad floats with max |x_k| over finite values, for the
// fixed-point scales of training; somhip_bmu_xmax_rows gives the number of rows written.
// fb_list / fb_count (nullable, both or neither): the BF16 path's fallback workspace (n int64 entries and
// one counter); when absent the context's grow-only workspace is used (and allocated if too small).
// prev_c1 (nullable, must not alias c1): earlier BMUs of the same rows; the BF16 search groups each
// tile's rows by them, which changes only its speed, never its results.
int somhip_launch_bmu(somhip_model *mdl, const float *x, int64_t n, int precision, int arithmetic, int need_c2,
                      int default_index, int32_t *c1, int32_t *c2, double *d1, float *xmax_part,
                      int64_t *fb_list = nullptr, unsigned int *fb_count = nullptr,
                      const int32_t *prev_c1 = nullptr);
int64_t somhip_bmu_xmax_rows(int precision, int dpad, int64_t n);
// Exact FP64 search of a float64 model (data and weights in FP64); same output contract.
int somhip_launch_bmu_f64(somhip_model *mdl, const double *x, int64_t n, int need_c2, int default_index, int32_t *c1,
                          int32_t *c2, double *d1);
BmuModelDev somhip_model_dev(const somhip_model *M);
// Serving form of the BF16 search: `grid` tiles from `tiles` (n == 0 tiles are skipped), then the exact
// fallback over the flagged (tile, row) entries with a fixed grid; fb_count must be zero on entry.
int somhip_launch_bmu_tiles(hipStream_t st, int dpad, int arithmetic, int need_c2, const BmuModelDev *models,
                            const BmuTile *tiles, int grid, int64_t *fb_list, unsigned int *fb_count);
int somhip_launch_zero_u32(hipStream_t st, unsigned int *p);
