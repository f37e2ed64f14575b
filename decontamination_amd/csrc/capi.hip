// extern "C" entry points of libsomhip (declared in include/somhip.h).
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "launch.hpp"

namespace {

thread_local std::string g_last_error;

// f64: the entry point is one of the float64 (_f64) forms
inline int check_model(const somhip_model *M, bool f64 = false)
{
    if (!M || !M->ctx) {
        somhip_set_error("null model");
        return SOMHIP_ERR_ARG;
    }
    if (M->f64 != f64) {
        somhip_set_error(M->f64 ? "float64 model: use the _f64 entry points"
                                : "float32 model: the _f64 entry points take float64 models");
        return SOMHIP_ERR_ARG;
    }
    return SOMHIP_OK;
}

inline int check_prec(int precision, int arithmetic)
{
    if (precision < SOMHIP_PREC_EXACT || precision > SOMHIP_PREC_BF16_RERANK) {
        somhip_set_error("invalid precision " + std::to_string(precision));
        return SOMHIP_ERR_ARG;
    }
    if (arithmetic != SOMHIP_RERANK_REF32 && arithmetic != SOMHIP_RERANK_FP64) {
        somhip_set_error("invalid arithmetic " + std::to_string(arithmetic));
        return SOMHIP_ERR_ARG;
    }
    return SOMHIP_OK;
}

// Slot order of the BF16 path.
//
// Base order: slot p -> neuron (p*S) mod mn, S coprime to mn and chosen so that no multiple k*S
// (k < KSPREAD) lands on a grid neighbour offset (+-1, +-n, +-n+-1): a neuron and its grid
// neighbours then sit in different 32-slot blocks, which keeps the second of the per-block minima
// close to the true second-best distance.
//
// Coarse section: on maps of at least 1024 neurons the sub-lattice {(a, b) : a % 3 == 0, b % 3 == 0}
// (neuron a*n + b) goes first, in base order, padded to a whole number of 8-block chunks; the other
// neurons follow in base order. The BMU kernel scans the coarse section first to get a threshold
// before it scans the rest. Padding slots map to -1.
struct PermLayout {
    std::vector<int> perm;
    int pblk = 0, pblk_a = 0;
};

int spread_stride(int m, int n, int kspread)
{
    const int mn = m * n;
    if (mn <= 64) return 1;
    auto bad = [&](int s) {
        if (std::gcd(s, mn) != 1) return true;
        for (int k = 1; k < kspread; k++) {
            long long o = ((long long)k * s) % mn;
            long long d = std::min(o, (long long)mn - o);
            if (d <= 1 || std::llabs(d - n) <= 1) return true;
        }
        return false;
    };
    const int start = std::max(2, (int)(0.6180339887 * mn));
    for (int s = start; s < mn; s++)
        if (!bad(s)) return s;
    for (int s = start - 1; s > 1; s--)
        if (!bad(s)) return s;
    return 0;
}

PermLayout make_perm(int m, int n)
{
    const int mn = m * n;
    int S = spread_stride(m, n, 40);
    if (!S) S = spread_stride(m, n, 32);
    if (!S) S = 1;
    std::vector<int> base(mn);
    for (int p = 0; p < mn; p++) base[p] = (int)(((long long)p * S) % mn);
    PermLayout L;
    constexpr int COARSE_STEP = 3, CHUNK_BLOCKS = 8;
    if (mn >= 1024) {
        std::vector<int> coarse, rest;
        for (int j : base) ((j / n) % COARSE_STEP == 0 && (j % n) % COARSE_STEP == 0 ? coarse : rest).push_back(j);
        const int na = ((int)coarse.size() + 32 * CHUNK_BLOCKS - 1) / (32 * CHUNK_BLOCKS) * CHUNK_BLOCKS;
        const int nr = ((int)rest.size() + 31) / 32;
        L.pblk_a = na;
        L.pblk = na + nr;
        L.perm.assign((size_t)L.pblk * 32, -1);
        std::copy(coarse.begin(), coarse.end(), L.perm.begin());
        std::copy(rest.begin(), rest.end(), L.perm.begin() + (size_t)na * 32);
    } else {
        L.pblk = L.pblk_a = (mn + 31) / 32;
        L.perm.assign((size_t)L.pblk * 32, -1);
        std::copy(base.begin(), base.end(), L.perm.begin());
    }
    return L;
}

}  // namespace

void somhip_set_error(const std::string &msg) { g_last_error = msg; }

static hipEvent_t timer_event(somhip_ctx *C)
{
    if (!C->timer.pool.empty()) {
        hipEvent_t e = C->timer.pool.back();
        C->timer.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimerScope::TimerScope(somhip_ctx *c, int i) : C(c), id(i)
{
    if (!C || !C->timer.enabled) return;
    a = timer_event(C);
    b = timer_event(C);
    if (a) (void)hipEventRecord(a, C->stream);
}

TimerScope::~TimerScope()
{
    if (!C || !C->timer.enabled || !a || !b) return;
    (void)hipEventRecord(b, C->stream);
    C->timer.pending.push_back({id, a, b});
}

int somhip_timer_collect(somhip_ctx *C)
{
    for (auto &p : C->timer.pending) {
        SOM_HIP_TRY(hipEventSynchronize(p.b));
        float ms = 0.f;
        SOM_HIP_TRY(hipEventElapsedTime(&ms, p.a, p.b));
        C->timer.total_ms[p.id] += ms;
        C->timer.launches[p.id] += 1;
        C->timer.pool.push_back(p.a);
        C->timer.pool.push_back(p.b);
    }
    C->timer.pending.clear();
    return SOMHIP_OK;
}

// Scratch of somhip_find_bmus_noalloc: a fallback counter (first 256 bytes) and a list of up to
// SOMHIP_NOALLOC_MAX_ROWS int64 entries. Every device holds its own copy of this array for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress returns on that device
// (kept in the context), never through the array's name.
static __device__ __align__(256) unsigned char scratch[256 + sizeof(int64_t) * (size_t)SOMHIP_NOALLOC_MAX_ROWS];

// Suspends the context's kernel timers for the duration of a call (event creation and the pending list
// would allocate).
struct TimingOff {
    somhip_ctx *C;
    bool was;
    explicit TimingOff(somhip_ctx *c) : C(c), was(c->timer.enabled) { C->timer.enabled = false; }
    ~TimingOff() { C->timer.enabled = was; }
};

// sigma = sigma0 * asymptotic_decay(e, E) >= FLOAT_EPS (som_batch.py:107-110, algo/__init__.py:64-67)
static double epoch_sigma(double sigma0, int64_t cur_epoch, int64_t n_epochs)
{
    const double sigma = sigma0 * (1.0 / (1.0 + 2.0 * (double)cur_epoch / (double)n_epochs));
    return sigma < SOM_FLOAT_EPS ? SOM_FLOAT_EPS : sigma;
}

struct somhip_accum {
    somhip_model *model = nullptr;
    int mode = 0;  // 0 empty, 1 epoch sums, 2 vector-count numerator/denominator
    double *acc = nullptr;  // (mn, dim+1) FP64
    Workspace ws;
    DevBuf upd;
};

extern "C" {

const char *somhip_last_error(void) { return g_last_error.c_str(); }

const char *somhip_version(void) { return "somhip 0.1.0 (gfx950)"; }

int somhip_device_count(int *count)
{
    if (!count) return SOMHIP_ERR_ARG;
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    *count = (e == hipSuccess) ? c : 0;
    return SOMHIP_OK;
}

int somhip_ctx_create(int device, somhip_ctx **out)
{
    if (!out) {
        somhip_set_error("null output pointer");
        return SOMHIP_ERR_ARG;
    }
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        somhip_set_error("no HIP device visible");
        return SOMHIP_ERR_NODEVICE;
    }
    if (device < 0 || device >= count) {
        somhip_set_error("device index out of range");
        return SOMHIP_ERR_ARG;
    }
    SOM_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    SOM_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0) {
        somhip_set_error(std::string("libsomhip is built for gfx950; device is ") + prop.gcnArchName);
        return SOMHIP_ERR_NODEVICE;
    }
    somhip_ctx *C = new somhip_ctx();
    C->device = device;
    C->num_cus = prop.multiProcessorCount;
    C->lds_per_block = prop.sharedMemPerBlock;
    hipError_t e = hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete C;
        somhip_set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        return SOMHIP_ERR_HIP;
    }
    C->own_stream = C->stream;
    // this device's copy of the allocation-free entry point's scratch array (the device is current)
    e = hipGetSymbolAddress(&C->noalloc_scratch, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipStreamDestroy(C->own_stream);
        delete C;
        somhip_set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return SOMHIP_ERR_HIP;
    }
    *out = C;
    return SOMHIP_OK;
}

int somhip_ctx_destroy(somhip_ctx *C)
{
    if (!C) return SOMHIP_OK;
    (void)hipSetDevice(C->device);
    (void)hipStreamSynchronize(C->stream);
    C->ws.release();
    for (auto &p : C->timer.pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
    for (auto e : C->timer.pool) (void)hipEventDestroy(e);
    if (C->scratch) (void)hipFree(C->scratch);
    if (C->own_stream) (void)hipStreamDestroy(C->own_stream);
    delete C;
    return SOMHIP_OK;
}

int somhip_ctx_synchronize(somhip_ctx *C)
{
    if (!C) return SOMHIP_ERR_ARG;
    SOM_HIP_TRY(hipStreamSynchronize(C->stream));
    return SOMHIP_OK;
}

int somhip_ctx_set_stream(somhip_ctx *C, void *s)
{
    if (!C) return SOMHIP_ERR_ARG;
    SOM_HIP_TRY(hipStreamSynchronize(C->stream));
    C->stream = s ? (hipStream_t)s : C->own_stream;
    return SOMHIP_OK;
}

int somhip_ctx_get_stream(somhip_ctx *C, void **s)
{
    if (!C || !s) return SOMHIP_ERR_ARG;
    *s = (void *)C->stream;
    return SOMHIP_OK;
}

int somhip_ctx_set_timing(somhip_ctx *C, int enable)
{
    if (!C) return SOMHIP_ERR_ARG;
    int rc = somhip_timer_collect(C);
    if (rc) return rc;
    C->timer.enabled = enable != 0;
    for (int i = 0; i < 16; i++) {
        C->timer.total_ms[i] = 0.0;
        C->timer.launches[i] = 0;
    }
    return SOMHIP_OK;
}

int somhip_ctx_get_timing(somhip_ctx *C, int kernel, double *total_ms, int64_t *launches)
{
    if (!C || kernel < 0 || kernel >= TK_COUNT || !total_ms || !launches) return SOMHIP_ERR_ARG;
    int rc = somhip_timer_collect(C);
    if (rc) return rc;
    *total_ms = C->timer.total_ms[kernel];
    *launches = C->timer.launches[kernel];
    return SOMHIP_OK;
}

int somhip_malloc(somhip_ctx *C, size_t bytes, void **dptr)
{
    if (!C || !dptr) return SOMHIP_ERR_ARG;
    SOM_HIP_TRY(hipSetDevice(C->device));
    SOM_HIP_TRY(hipMalloc(dptr, bytes ? bytes : 1));
    return SOMHIP_OK;
}

int somhip_free(somhip_ctx *C, void *dptr)
{
    if (!C) return SOMHIP_ERR_ARG;
    if (!dptr) return SOMHIP_OK;
    SOM_HIP_TRY(hipStreamSynchronize(C->stream));
    SOM_HIP_TRY(hipFree(dptr));
    return SOMHIP_OK;
}

int somhip_memcpy_h2d(somhip_ctx *C, void *dst, const void *src, size_t bytes)
{
    if (!C || (bytes && (!dst || !src))) return SOMHIP_ERR_ARG;
    if (!bytes) return SOMHIP_OK;
    SOM_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, C->stream));
    SOM_HIP_TRY(hipStreamSynchronize(C->stream));
    return SOMHIP_OK;
}

int somhip_memcpy_d2h(somhip_ctx *C, void *dst, const void *src, size_t bytes)
{
    if (!C || (bytes && (!dst || !src))) return SOMHIP_ERR_ARG;
    if (!bytes) return SOMHIP_OK;
    SOM_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, C->stream));
    SOM_HIP_TRY(hipStreamSynchronize(C->stream));
    return SOMHIP_OK;
}

int somhip_memcpy_d2d(somhip_ctx *C, void *dst, const void *src, size_t bytes)
{
    if (!C || (bytes && (!dst || !src))) return SOMHIP_ERR_ARG;
    if (!bytes) return SOMHIP_OK;
    SOM_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, C->stream));
    return SOMHIP_OK;
}

int somhip_memset(somhip_ctx *C, void *dptr, int value, size_t bytes)
{
    if (!C || (bytes && !dptr)) return SOMHIP_ERR_ARG;
    if (!bytes) return SOMHIP_OK;
    SOM_HIP_TRY(hipMemsetAsync(dptr, value, bytes, C->stream));
    return SOMHIP_OK;
}

}  // extern "C"

namespace {

int check_model_shape(somhip_ctx *C, int m, int n, int dim, int topology, const void *topo_host, somhip_model **out)
{
    if (!C || !out || !topo_host) {
        somhip_set_error("null argument");
        return SOMHIP_ERR_ARG;
    }
    *out = nullptr;
    if (m <= 0 || n <= 0 || dim <= 0) {
        somhip_set_error("m, n and dim must be positive");
        return SOMHIP_ERR_ARG;
    }
    if (dim > 64) {
        somhip_set_error("dim > 64 is not supported");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    if ((long long)m * n > (1 << 24)) {
        somhip_set_error("too many neurons");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    if (topology != SOMHIP_TOPOLOGY_SQUARE && topology != SOMHIP_TOPOLOGY_HEXAGONAL) {
        somhip_set_error("invalid topology");
        return SOMHIP_ERR_ARG;
    }
    return SOMHIP_OK;
}

// The reference's grid (som_abstract.py:81-111): neuron i*n + j at (i, j), or (i + 0.5 (j % 2),
// j sqrt(3)/2) rounded to the topography's type for hexagonal; only then does the neighbourhood
// factor into rows and columns.
template <class T>
bool closed_form_grid(const T *topo, int m, int n, int topology)
{
    if (m > 2048 || n > 2048) return false;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            const T ex = topology == SOMHIP_TOPOLOGY_SQUARE ? (T)i : (T)i + (T)0.5 * (T)(j % 2);
            const T ey = topology == SOMHIP_TOPOLOGY_SQUARE ? (T)j : (T)(j * 0.8660254037844386);
            const T *t = topo + 2 * ((size_t)i * n + j);
            if (!(t[0] == ex && t[1] == ey)) return false;
        }
    return true;
}

// The model's weights and topography in T: float for float32 models, double for float64 models.
template <class T> T *model_w(somhip_model *M);
template <> float *model_w<float>(somhip_model *M) { return M->w; }
template <> double *model_w<double>(somhip_model *M) { return M->w64; }
template <class T> const T *model_topo(const somhip_model *M);
template <> const float *model_topo<float>(const somhip_model *M) { return M->topo; }
template <> const double *model_topo<double>(const somhip_model *M) { return M->topo64; }

// The BMU search: FP32 rows with the requested precision and arithmetic, FP64 rows (float64 models)
// with the exact FP64 scan.
int search(somhip_model *M, const float *x, int64_t n, int precision, int arithmetic, int need_c2, int default_index,
           int32_t *c1, int32_t *c2, double *d1)
{
    return somhip_launch_bmu(M, x, n, precision, arithmetic, need_c2, default_index, c1, c2, d1, nullptr);
}
int search(somhip_model *M, const double *x, int64_t n, int, int, int need_c2, int default_index, int32_t *c1,
           int32_t *c2, double *d1)
{
    return somhip_launch_bmu_f64(M, x, n, need_c2, default_index, c1, c2, d1);
}

// Bodies shared by the float32 and float64 entry points (model and precision already checked).
template <class T>
int count_bmus_t(somhip_model *M, const T *x, const double *rho, int64_t n, int precision, int arithmetic,
                 double *result)
{
    if (n < 0 || (n > 0 && (!x || !result))) {
        somhip_set_error("invalid x/result/n");
        return SOMHIP_ERR_ARG;
    }
    if (n == 0) return SOMHIP_OK;
    somhip_ctx *C = M->ctx;
    Workspace &W = C->ws;
    const int nwg = somhip_scatter_nwg(n, M->mn, 1, C->num_cus, SCATTER_TABLE_BUDGET);
    int rc;
    if ((rc = W.c1.ensure(sizeof(int32_t) * n)) || (rc = W.bounds.ensure(256)) || (rc = W.shifts.ensure(256)) ||
        (rc = W.table.ensure((size_t)nwg * M->mn * 8)))
        return rc;
    int32_t *c1 = (int32_t *)W.c1.p;
    if ((rc = search(M, x, n, precision, arithmetic, 0, 0, c1, nullptr, nullptr))) return rc;
    {
        TimerScope ts(C, TK_ABSMAX);
        if ((rc = somhip_launch_absmax(C->stream, NO_ROWS, 0, 16, rho, 0, rho ? n : 0, (unsigned long long *)W.bounds.p,
                                       C->num_cus)))
            return rc;
        if ((rc = somhip_launch_shifts(C->stream, (const unsigned long long *)W.bounds.p, 0, 0, rho != nullptr, n,
                                       (int *)W.shifts.p)))
            return rc;
    }
    TimerScope ts(C, TK_SCATTER);
    return somhip_launch_scatter(C->stream, c1, NO_ROWS, 0, rho, 0, n, (const int *)W.shifts.p, M->mn, 1, nwg,
                                 (unsigned long long *)W.table.p, result, C->lds_per_block);
}

template <class T>
int compute_errors_t(somhip_model *M, const T *x, const double *rho, int64_t n, int precision, int arithmetic,
                     double *errors)
{
    if (n < 0 || (n > 0 && (!x || !errors))) {
        somhip_set_error("invalid x/errors/n");
        return SOMHIP_ERR_ARG;
    }
    if (n == 0) return SOMHIP_OK;
    somhip_ctx *C = M->ctx;
    Workspace &W = C->ws;
    const int nparts = somhip_errors_nparts(n);
    int rc;
    if ((rc = W.c1.ensure(sizeof(int32_t) * n)) || (rc = W.c2.ensure(sizeof(int32_t) * n)) ||
        (rc = W.d1.ensure(sizeof(double) * n)) || (rc = W.part.ensure(sizeof(double) * 3 * nparts)))
        return rc;
    int32_t *c1 = (int32_t *)W.c1.p, *c2 = (int32_t *)W.c2.p;
    double *d1 = (double *)W.d1.p;
    if ((rc = search(M, x, n, precision, arithmetic, 1, -1, c1, c2, d1))) return rc;
    const double penalty = M->topology == SOMHIP_TOPOLOGY_SQUARE ? 2.0 : 1.0;
    TimerScope ts(C, TK_ERRORS);
    return somhip_launch_errors(C->stream, c1, c2, d1, rho, n, model_topo<T>(M), penalty, (double *)W.part.p, nparts,
                                errors);
}

template <class T>
int epoch_update_t(somhip_accum *A, int64_t cur_epoch, int64_t n_epochs, double sigma0, T *history_out)
{
    somhip_model *M = A->model;
    if (A->mode != 1) {
        somhip_set_error("epoch update without epoch accumulation");
        return SOMHIP_ERR_STATE;
    }
    if (n_epochs <= 0 || cur_epoch < 0) {
        somhip_set_error("invalid epoch counters");
        return SOMHIP_ERR_ARG;
    }
    const double sigma = epoch_sigma(sigma0, cur_epoch, n_epochs);
    M->prepared = false;  // the derived data of the FP32 search follow the new weights
    int rc;
    TimerScope ts(M->ctx, TK_UPDATE);
    if (M->separable && somhip_update_separable_fits(M->m, M->n, M->dim)) {
        if ((rc = A->upd.ensure(sizeof(double) * 2 * (size_t)M->mn * (M->dim + 1)))) return rc;
        return somhip_launch_update_separable(M->ctx->stream, A->acc, model_topo<T>(M), M->m, M->n,
                                              M->topology == SOMHIP_TOPOLOGY_HEXAGONAL ? 1 : 0, M->dim, sigma,
                                              (double *)A->upd.p, model_w<T>(M), history_out);
    }
    const int nsplit = somhip_update_nsplit(M->mn);
    if ((rc = A->upd.ensure(sizeof(double) * (size_t)nsplit * M->mn * (M->dim + 1)))) return rc;
    return somhip_launch_update(M->ctx->stream, A->acc, model_topo<T>(M), M->mn, M->dim, sigma, (double *)A->upd.p,
                                nsplit, model_w<T>(M), history_out);
}

template <class T>
int iter_accumulate_t(somhip_accum *A, const T *x, const double *rho, int64_t n, int64_t cur_vector,
                      int64_t n_vectors, T sigma0, int precision, int arithmetic)
{
    somhip_model *M = A->model;
    if (A->mode == 1) {
        somhip_set_error("accumulator holds epoch sums; reset it first");
        return SOMHIP_ERR_STATE;
    }
    if (n < 0 || (n > 0 && !x) || n_vectors <= 0 || cur_vector < 0) {
        somhip_set_error("invalid x/n/vector counters");
        return SOMHIP_ERR_ARG;
    }
    A->mode = 2;
    if (n == 0) return SOMHIP_OK;
    somhip_ctx *C = M->ctx;
    Workspace &W = A->ws;
    const int nsplit = somhip_iter_nsplit(n, M->mn);
    int rc;
    if ((rc = W.c1.ensure(sizeof(int32_t) * n)) ||
        (rc = A->upd.ensure(sizeof(double) * (size_t)nsplit * M->mn * (M->dim + 1))))
        return rc;
    int32_t *c1 = (int32_t *)W.c1.p;
    W.c1_x = nullptr;  // c1 no longer holds the epoch path's BMUs
    if ((rc = search(M, x, n, precision, arithmetic, 0, -1, c1, nullptr, nullptr))) return rc;
    TimerScope ts(C, TK_ITER);
    return somhip_launch_iter(C->stream, c1, x, M->dim, M->dpad, rho, n, cur_vector, n_vectors, sigma0,
                              model_topo<T>(M), M->mn, (double *)A->upd.p, nsplit, A->acc);
}

template <class T>
int iter_update_t(somhip_accum *A, T *history_out)
{
    if (A->mode != 2) {
        somhip_set_error("vector-count update without vector-count accumulation");
        return SOMHIP_ERR_STATE;
    }
    somhip_model *M = A->model;
    M->prepared = false;
    return somhip_launch_divide(M->ctx->stream, A->acc, M->mn, M->dim, model_w<T>(M), history_out);
}

}  // namespace

extern "C" {

int somhip_model_create(somhip_ctx *C, int m, int n, int dim, int topology, const float *topo_host, somhip_model **out)
{
    int rc0 = check_model_shape(C, m, n, dim, topology, topo_host, out);
    if (rc0) return rc0;
    SOM_HIP_TRY(hipSetDevice(C->device));
    somhip_model *M = new somhip_model();
    M->ctx = C;
    M->m = m; M->n = n; M->mn = m * n; M->dim = dim;
    // padded dimension of the kernels' layouts: 16, 32, or 64 (the MFMA paths take K slices of 16)
    M->dpad = dim <= 16 ? 16 : dim <= 32 ? 32 : 64;
    M->topology = topology;
    M->nblk = (M->mn + 31) / 32;
    M->separable = closed_form_grid(topo_host, m, n, topology);
    PermLayout layout = make_perm(m, n);
    M->pblk = layout.pblk;
    M->pblk_a = layout.pblk_a;
    const size_t slots = (size_t)M->nblk * 32;
    const size_t pslots = (size_t)M->pblk * 32;
    auto fail = [&](hipError_t e, const char *what) {
        somhip_set_error(std::string(what) + ": " + hipGetErrorString(e));
        somhip_model_destroy(M);
        return SOMHIP_ERR_HIP;
    };
    hipError_t e;
#define SOM_ALLOC(ptr, bytes)                                  \
    if ((e = hipMalloc((void **)&(ptr), (bytes))) != hipSuccess) \
        return fail(e, "hipMalloc");
    SOM_ALLOC(M->w, sizeof(float) * M->mn * dim);
    SOM_ALLOC(M->topo, sizeof(float) * M->mn * 2);
    SOM_ALLOC(M->centre, sizeof(float) * M->dpad);
    SOM_ALLOC(M->wmax, sizeof(float) * M->dpad);
    SOM_ALLOC(M->scal, sizeof(float) * 4);
    SOM_ALLOC(M->norms_perm, sizeof(float) * pslots);
    SOM_ALLOC(M->norms_nat, sizeof(float) * slots);
    SOM_ALLOC(M->frag_bf, sizeof(unsigned short) * pslots * M->dpad * 2);
    SOM_ALLOC(M->frag_f32, sizeof(float) * slots * M->dpad);
    SOM_ALLOC(M->wslot, sizeof(float) * pslots * M->dpad);
    SOM_ALLOC(M->perm, sizeof(int) * pslots);
#undef SOM_ALLOC
    if ((e = hipMemcpyAsync(M->perm, layout.perm.data(), sizeof(int) * pslots, hipMemcpyHostToDevice, C->stream)) != hipSuccess)
        return fail(e, "hipMemcpy perm");
    if ((e = hipMemcpyAsync(M->topo, topo_host, sizeof(float) * M->mn * 2, hipMemcpyHostToDevice, C->stream)) != hipSuccess)
        return fail(e, "hipMemcpy topography");
    if ((e = hipMemsetAsync(M->w, 0, sizeof(float) * M->mn * dim, C->stream)) != hipSuccess) return fail(e, "hipMemset");
    if ((e = hipStreamSynchronize(C->stream)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    M->prepared = false;
    *out = M;
    return SOMHIP_OK;
}

int somhip_model_destroy(somhip_model *M)
{
    if (!M) return SOMHIP_OK;
    if (M->ctx) (void)hipStreamSynchronize(M->ctx->stream);
    void *ptrs[] = {M->w,         M->topo,    M->centre,   M->wmax,  M->scal, M->norms_perm,
                    M->norms_nat, M->frag_bf, M->frag_f32, M->wslot, M->perm, M->w64, M->topo64};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete M;
    return SOMHIP_OK;
}

int somhip_model_set_weights(somhip_model *M, const float *w, int on_device)
{
    int rc = check_model(M);
    if (rc) return rc;
    if (!w) {
        somhip_set_error("null weights");
        return SOMHIP_ERR_ARG;
    }
    SOM_HIP_TRY(hipMemcpyAsync(M->w, w, sizeof(float) * M->mn * M->dim,
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, M->ctx->stream));
    if (!on_device) SOM_HIP_TRY(hipStreamSynchronize(M->ctx->stream));
    M->prepared = false;
    M->uploads++;  // earlier BMUs kept by accumulators no longer describe this map's schedule
    return SOMHIP_OK;
}

int somhip_model_get_weights(somhip_model *M, float *w_host)
{
    int rc = check_model(M);
    if (rc) return rc;
    if (!w_host) return SOMHIP_ERR_ARG;
    SOM_HIP_TRY(hipMemcpyAsync(w_host, M->w, sizeof(float) * M->mn * M->dim, hipMemcpyDeviceToHost, M->ctx->stream));
    SOM_HIP_TRY(hipStreamSynchronize(M->ctx->stream));
    return SOMHIP_OK;
}

int somhip_model_weights_ptr(somhip_model *M, float **dptr)
{
    int rc = check_model(M);
    if (rc) return rc;
    if (!dptr) return SOMHIP_ERR_ARG;
    *dptr = M->w;
    return SOMHIP_OK;
}

int somhip_find_bmus(somhip_model *M, const float *x, int64_t n, int precision, int arithmetic, int32_t *out)
{
    int rc = check_model(M);
    if (rc) return rc;
    if ((rc = check_prec(precision, arithmetic))) return rc;
    if (n < 0 || (n > 0 && (!x || !out))) {
        somhip_set_error("invalid x/out/n");
        return SOMHIP_ERR_ARG;
    }
    return somhip_launch_bmu(M, x, n, precision, arithmetic, 0, 0, out, nullptr, nullptr, nullptr);
}

int somhip_find_bmus_noalloc(somhip_model *M, const float *x, int64_t n, int precision, int arithmetic, int32_t *out)
{
    int rc = check_model(M);
    if (rc) return rc;
    if ((rc = check_prec(precision, arithmetic))) return rc;
    if (n < 0 || (n > 0 && (!x || !out))) {
        somhip_set_error("invalid x/out/n");
        return SOMHIP_ERR_ARG;
    }
    if (n > SOMHIP_NOALLOC_MAX_ROWS) {
        somhip_set_error("somhip_find_bmus_noalloc takes at most " + std::to_string(SOMHIP_NOALLOC_MAX_ROWS) +
                         " rows per call (got " + std::to_string(n) + "); split the input");
        return SOMHIP_ERR_ARG;
    }
    somhip_ctx *C = M->ctx;
    if (!C->noalloc_scratch) {
        somhip_set_error("context has no scratch address");
        return SOMHIP_ERR_STATE;
    }
    TimingOff quiet(C);
    unsigned int *fb_count = (unsigned int *)C->noalloc_scratch;
    int64_t *fb_list = (int64_t *)((unsigned char *)C->noalloc_scratch + 256);
    return somhip_launch_bmu(M, x, n, precision, arithmetic, 0, 0, out, nullptr, nullptr, nullptr, fb_list, fb_count);
}

int somhip_count_bmus(somhip_model *M, const float *x, const double *rho, int64_t n, int precision, int arithmetic,
                      double *result)
{
    int rc = check_model(M);
    if (rc || (rc = check_prec(precision, arithmetic))) return rc;
    return count_bmus_t(M, x, rho, n, precision, arithmetic, result);
}

int somhip_compute_errors(somhip_model *M, const float *x, const double *rho, int64_t n, int precision,
                          int arithmetic, double *errors)
{
    int rc = check_model(M);
    if (rc || (rc = check_prec(precision, arithmetic))) return rc;
    return compute_errors_t(M, x, rho, n, precision, arithmetic, errors);
}

int somhip_accum_create(somhip_model *M, somhip_accum **out)
{
    int rc = check_model(M, M && M->f64);
    if (rc) return rc;
    if (!out) return SOMHIP_ERR_ARG;
    somhip_accum *A = new somhip_accum();
    A->model = M;
    hipError_t e = hipMalloc((void **)&A->acc, sizeof(double) * M->mn * (M->dim + 1));
    if (e != hipSuccess) {
        delete A;
        somhip_set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return SOMHIP_ERR_HIP;
    }
    *out = A;
    return somhip_accum_reset(A);
}

int somhip_accum_destroy(somhip_accum *A)
{
    if (!A) return SOMHIP_OK;
    (void)hipStreamSynchronize(A->model->ctx->stream);
    if (A->acc) (void)hipFree(A->acc);
    A->ws.release();
    A->upd.release();
    delete A;
    return SOMHIP_OK;
}

int somhip_accum_reset(somhip_accum *A)
{
    if (!A) return SOMHIP_ERR_ARG;
    somhip_model *M = A->model;
    SOM_HIP_TRY(hipMemsetAsync(A->acc, 0, sizeof(double) * M->mn * (M->dim + 1), M->ctx->stream));
    A->mode = 0;
    return SOMHIP_OK;
}

int somhip_batch_epoch_accumulate(somhip_accum *A, const float *x, const double *rho, int64_t n, int precision,
                                  int arithmetic, double *errors)
{
    if (!A) return SOMHIP_ERR_ARG;
    somhip_model *M = A->model;
    int rc = check_model(M);
    if (rc || (rc = check_prec(precision, arithmetic))) return rc;
    if (A->mode == 2) {
        somhip_set_error("accumulator holds vector-count sums; reset it first");
        return SOMHIP_ERR_STATE;
    }
    if (n < 0 || (n > 0 && !x)) {
        somhip_set_error("invalid x/n");
        return SOMHIP_ERR_ARG;
    }
    A->mode = 1;
    if (n == 0) return SOMHIP_OK;
    somhip_ctx *C = M->ctx;
    const int nv = M->dim + 1;
    const int nwg = somhip_scatter_nwg(n, M->mn, nv, C->num_cus, SCATTER_TABLE_BUDGET);
    const int nparts = somhip_errors_nparts(n);
    Workspace &W = A->ws;
    precision = somhip_bmu_precision(M, precision);  // the per-wave |x| rows depend on the kernel that runs
    const int64_t xrows = somhip_bmu_xmax_rows(precision, M->dpad, n);
    // the previous call's BMUs of the same samples (the previous epoch) guide the BF16 search's grouping
    // (not after an explicit weight upload: a new schedule starts without them, as a new train() does)
    const bool have_prev = precision == SOMHIP_PREC_BF16_RERANK && W.c1_x == (const void *)x && W.c1_n == n &&
                           W.c1_uploads == M->uploads;
    if (have_prev) std::swap(W.c1, W.c1b);
    if ((rc = W.c1.ensure(sizeof(int32_t) * n)) || (rc = W.c2.ensure(sizeof(int32_t) * n)) ||
        (rc = W.d1.ensure(sizeof(double) * n)) || (rc = W.part.ensure(sizeof(double) * 3 * nparts)) ||
        (rc = W.bounds.ensure(512)) || (rc = W.shifts.ensure(512)) ||
        (rc = W.table.ensure((size_t)nwg * M->mn * nv * 8)) ||
        (rc = W.xmax.ensure(sizeof(float) * (size_t)xrows * M->dpad)) ||
        (rc = W.bucket.ensure(sizeof(unsigned int) * (size_t)n)))
        return rc;
    int32_t *c1 = (int32_t *)W.c1.p, *c2 = (int32_t *)W.c2.p;
    double *d1 = (double *)W.d1.p;
    const bool need_err = errors != nullptr;
    // the BMU search also writes per-wave max |x_k| rows, from which the fixed-point scales follow
    W.c1_x = nullptr;  // c1 is being rewritten
    if ((rc = somhip_launch_bmu(M, x, n, precision, arithmetic, need_err ? 1 : 0, -1, c1, need_err ? c2 : nullptr,
                                need_err ? d1 : nullptr, (float *)W.xmax.p, nullptr, nullptr,
                                have_prev ? (const int32_t *)W.c1b.p : nullptr)))
        return rc;
    W.c1_x = x;
    W.c1_n = n;
    W.c1_uploads = M->uploads;
    if (need_err) {
        const double penalty = M->topology == SOMHIP_TOPOLOGY_SQUARE ? 2.0 : 1.0;
        TimerScope ts(C, TK_ERRORS);
        if ((rc = somhip_launch_errors(C->stream, c1, c2, d1, rho, n, M->topo, penalty, (double *)W.part.p, nparts,
                                       errors)))
            return rc;
    }
    {
        TimerScope ts(C, TK_ABSMAX);
        if (rho &&
            (rc = somhip_launch_absmax(C->stream, NO_ROWS, 0, 16, rho, 1, n, (unsigned long long *)W.bounds.p, C->num_cus)))
            return rc;
        if ((rc = somhip_launch_shifts_parts(C->stream, (const float *)W.xmax.p, xrows, M->dpad, M->dim,
                                             rho ? (const unsigned long long *)W.bounds.p : nullptr, n, (int *)W.shifts.p)))
            return rc;
    }
    TimerScope ts(C, TK_SCATTER);
    return somhip_launch_scatter(C->stream, c1, x, M->dim, rho, 1, n, (const int *)W.shifts.p, M->mn, nv, nwg,
                                 (unsigned long long *)W.table.p, A->acc, C->lds_per_block, (unsigned int *)W.bucket.p);
}

int somhip_batch_epoch_update(somhip_accum *A, int64_t cur_epoch, int64_t n_epochs, float sigma0, float *history_out)
{
    if (!A) return SOMHIP_ERR_ARG;
    int rc = check_model(A->model);
    if (rc) return rc;
    return epoch_update_t(A, cur_epoch, n_epochs, (double)sigma0, history_out);
}

int somhip_batch_iter_accumulate(somhip_accum *A, const float *x, const double *rho, int64_t n, int64_t cur_vector,
                                 int64_t n_vectors, float sigma0, int precision, int arithmetic)
{
    if (!A) return SOMHIP_ERR_ARG;
    int rc = check_model(A->model);
    if (rc || (rc = check_prec(precision, arithmetic))) return rc;
    return iter_accumulate_t(A, x, rho, n, cur_vector, n_vectors, sigma0, precision, arithmetic);
}

int somhip_batch_iter_update(somhip_accum *A, float *history_out)
{
    if (!A) return SOMHIP_ERR_ARG;
    int rc = check_model(A->model);
    if (rc) return rc;
    return iter_update_t(A, history_out);
}

// ------------------------------------------------------------------------------------------------
// float64 SOMs (the reference's dtype=np.float64): data, weights and topography in FP64; the BMU
// search is the exact FP64 scan, the sums are the same fixed-point scatter, the update writes FP64.
// ------------------------------------------------------------------------------------------------
int somhip_model_create_f64(somhip_ctx *C, int m, int n, int dim, int topology, const double *topo_host,
                            somhip_model **out)
{
    int rc = check_model_shape(C, m, n, dim, topology, topo_host, out);
    if (rc) return rc;
    SOM_HIP_TRY(hipSetDevice(C->device));
    somhip_model *M = new somhip_model();
    M->ctx = C;
    M->m = m; M->n = n; M->mn = m * n; M->dim = dim;
    M->dpad = (dim + 15) / 16 * 16;
    M->topology = topology;
    M->f64 = true;
    M->separable = closed_form_grid(topo_host, m, n, topology);
    hipError_t e;
    if ((e = hipMalloc((void **)&M->w64, sizeof(double) * M->mn * dim)) != hipSuccess ||
        (e = hipMalloc((void **)&M->topo64, sizeof(double) * M->mn * 2)) != hipSuccess ||
        (e = hipMemcpyAsync(M->topo64, topo_host, sizeof(double) * M->mn * 2, hipMemcpyHostToDevice, C->stream)) !=
            hipSuccess ||
        (e = hipMemsetAsync(M->w64, 0, sizeof(double) * M->mn * dim, C->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(C->stream)) != hipSuccess) {
        somhip_set_error(std::string("float64 model: ") + hipGetErrorString(e));
        somhip_model_destroy(M);
        return SOMHIP_ERR_HIP;
    }
    *out = M;
    return SOMHIP_OK;
}

int somhip_model_set_weights_f64(somhip_model *M, const double *w, int on_device)
{
    int rc = check_model(M, true);
    if (rc) return rc;
    if (!w) {
        somhip_set_error("null weights");
        return SOMHIP_ERR_ARG;
    }
    SOM_HIP_TRY(hipMemcpyAsync(M->w64, w, sizeof(double) * M->mn * M->dim,
                               on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, M->ctx->stream));
    if (!on_device) SOM_HIP_TRY(hipStreamSynchronize(M->ctx->stream));
    M->uploads++;
    return SOMHIP_OK;
}

int somhip_model_get_weights_f64(somhip_model *M, double *w_host)
{
    int rc = check_model(M, true);
    if (rc) return rc;
    if (!w_host) return SOMHIP_ERR_ARG;
    SOM_HIP_TRY(hipMemcpyAsync(w_host, M->w64, sizeof(double) * M->mn * M->dim, hipMemcpyDeviceToHost,
                               M->ctx->stream));
    SOM_HIP_TRY(hipStreamSynchronize(M->ctx->stream));
    return SOMHIP_OK;
}

int somhip_find_bmus_f64(somhip_model *M, const double *x, int64_t n, int32_t *out)
{
    int rc = check_model(M, true);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!x || !out))) {
        somhip_set_error("invalid x/out/n");
        return SOMHIP_ERR_ARG;
    }
    return somhip_launch_bmu_f64(M, x, n, 0, 0, out, nullptr, nullptr);
}

int somhip_count_bmus_f64(somhip_model *M, const double *x, const double *rho, int64_t n, double *result)
{
    int rc = check_model(M, true);
    if (rc) return rc;
    return count_bmus_t(M, x, rho, n, SOMHIP_PREC_EXACT, SOMHIP_RERANK_FP64, result);
}

int somhip_compute_errors_f64(somhip_model *M, const double *x, const double *rho, int64_t n, double *errors)
{
    int rc = check_model(M, true);
    if (rc) return rc;
    return compute_errors_t(M, x, rho, n, SOMHIP_PREC_EXACT, SOMHIP_RERANK_FP64, errors);
}

int somhip_batch_epoch_accumulate_f64(somhip_accum *A, const double *x, const double *rho, int64_t n, double *errors)
{
    if (!A) return SOMHIP_ERR_ARG;
    somhip_model *M = A->model;
    int rc = check_model(M, true);
    if (rc) return rc;
    if (A->mode == 2) {
        somhip_set_error("accumulator holds vector-count sums; reset it first");
        return SOMHIP_ERR_STATE;
    }
    if (n < 0 || (n > 0 && !x)) {
        somhip_set_error("invalid x/n");
        return SOMHIP_ERR_ARG;
    }
    A->mode = 1;
    if (n == 0) return SOMHIP_OK;
    somhip_ctx *C = M->ctx;
    const int nv = M->dim + 1;
    const int nwg = somhip_scatter_nwg(n, M->mn, nv, C->num_cus, SCATTER_TABLE_BUDGET);
    const int nparts = somhip_errors_nparts(n);
    Workspace &W = A->ws;
    if ((rc = W.c1.ensure(sizeof(int32_t) * n)) || (rc = W.c2.ensure(sizeof(int32_t) * n)) ||
        (rc = W.d1.ensure(sizeof(double) * n)) || (rc = W.part.ensure(sizeof(double) * 3 * nparts)) ||
        (rc = W.bounds.ensure(sizeof(unsigned long long) * (M->dim + 1))) ||
        (rc = W.shifts.ensure(sizeof(int) * (M->dim + 1))) ||
        (rc = W.table.ensure((size_t)nwg * M->mn * nv * 8)) || (rc = W.bucket.ensure(sizeof(unsigned int) * (size_t)n)))
        return rc;
    int32_t *c1 = (int32_t *)W.c1.p, *c2 = (int32_t *)W.c2.p;
    double *d1 = (double *)W.d1.p;
    const bool need_err = errors != nullptr;
    if ((rc = somhip_launch_bmu_f64(M, x, n, need_err ? 1 : 0, -1, c1, need_err ? c2 : nullptr,
                                    need_err ? d1 : nullptr)))
        return rc;
    if (need_err) {
        const double penalty = M->topology == SOMHIP_TOPOLOGY_SQUARE ? 2.0 : 1.0;
        TimerScope ts(C, TK_ERRORS);
        if ((rc = somhip_launch_errors(C->stream, c1, c2, d1, rho, n, (const double *)M->topo64, penalty,
                                       (double *)W.part.p, nparts, errors)))
            return rc;
    }
    {
        // the density stays FP64 (the reference casts it to dtype); scales from max |x_k| and max rho
        TimerScope ts(C, TK_ABSMAX);
        if ((rc = somhip_launch_absmax(C->stream, x, M->dim, M->dpad, rho, 0, n, (unsigned long long *)W.bounds.p,
                                       C->num_cus)) ||
            (rc = somhip_launch_shifts(C->stream, (const unsigned long long *)W.bounds.p, M->dim, 1, rho != nullptr, n,
                                       (int *)W.shifts.p)))
            return rc;
    }
    TimerScope ts(C, TK_SCATTER);
    return somhip_launch_scatter(C->stream, c1, x, M->dim, rho, 0, n, (const int *)W.shifts.p, M->mn, nv, nwg,
                                 (unsigned long long *)W.table.p, A->acc, C->lds_per_block, (unsigned int *)W.bucket.p);
}

int somhip_batch_epoch_update_f64(somhip_accum *A, int64_t cur_epoch, int64_t n_epochs, double sigma0,
                                  double *history_out)
{
    if (!A) return SOMHIP_ERR_ARG;
    int rc = check_model(A->model, true);
    if (rc) return rc;
    return epoch_update_t(A, cur_epoch, n_epochs, sigma0, history_out);
}

int somhip_batch_iter_accumulate_f64(somhip_accum *A, const double *x, const double *rho, int64_t n,
                                     int64_t cur_vector, int64_t n_vectors, double sigma0)
{
    if (!A) return SOMHIP_ERR_ARG;
    int rc = check_model(A->model, true);
    if (rc) return rc;
    return iter_accumulate_t(A, x, rho, n, cur_vector, n_vectors, sigma0, SOMHIP_PREC_EXACT, SOMHIP_RERANK_FP64);
}

int somhip_batch_iter_update_f64(somhip_accum *A, double *history_out)
{
    if (!A) return SOMHIP_ERR_ARG;
    int rc = check_model(A->model, true);
    if (rc) return rc;
    return iter_update_t(A, history_out);
}

}  // extern "C"
