/*
 * CPU oracle for the batch self-organising-map hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This file restates, in plain C, the reference's CPU kernels so the
 * HIP product can be checked against it. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load the library built from it; the product never calls it.
 *
 * Two arithmetic modes are provided for every function:
 *   ORC_REF32 - the reference's own arithmetic: FP32 weights/data, FP32 difference and FP32 square,
 *               FP64 accumulation of the squares (decontamination/algo/__init__.py:89-98; numba types
 *               float32 ** int as float32), FP32 accumulators for the training numerator/denominator
 *               (decontamination/algo/som_batch.py:289-291,564-566).
 *   ORC_FP64  - the same algorithm with every quantity promoted to double (the FP64 oracle).
 *
 * Reference citations (file:line under the reference repository):
 *   find_bmu            decontamination/algo/som_abstract.py:764-783  (_find_bmu_xpu)
 *   find_bmus           decontamination/algo/som_abstract.py:736-760  (_find_bmus_kernel)
 *   count_bmus          decontamination/algo/som_abstract.py:699-732  (_count_bmus_kernel)
 *   compute_errors      decontamination/algo/som_abstract.py:787-878  (_compute_errors_kernel/_xpu)
 *   train epoch step    decontamination/algo/som_batch.py:94-150,509-566
 *   train iter step     decontamination/algo/som_batch.py:154-210,509-566
 *   sigma decay         decontamination/algo/__init__.py:64-74, clamp som_batch.py:107-110,133-136
 *   welford update      decontamination/algo/covariance.py:30-84 (_update_welford_cov_sums)
 *
 * Pinned by: the reference's golden vectors (test/test_som_batch.py:30-89,
 * test/test_som_abstract.py:41-64), brute force on tiny inputs and closed forms (tests/test_oracle.py).
 *
 * The *_omp variants restate the reference's parallel CPU kernel (nb.prange over samples with
 * LLVM atomicrmw fadd on the FP32 accumulators, decontamination/jit/atomic.py:119); they are the
 * timed CPU baseline. Their summation order depends on thread timing, like the reference's.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#define ORC_REF32 0
#define ORC_FP64 1
#define ORC_FLOAT_EPS 1.0e-12

/* square_distance_xpu: result = 0.0 (float64); result += (a - b) ** 2 with float32 a, b. */
static inline double sqd_ref32(const float *a, const float *b, int dim)
{
    double result = 0.0;
    for (int k = 0; k < dim; k++) {
        float d = a[k] - b[k];
        float s = d * d;
        result += (double)s;
    }
    return result;
}

static inline double sqd_fp64(const float *a, const float *b, int dim)
{
    double result = 0.0;
    for (int k = 0; k < dim; k++) {
        double d = (double)a[k] - (double)b[k];
        result += d * d;
    }
    return result;
}

static inline double sqd(int mode, const float *a, const float *b, int dim)
{
    return mode == ORC_FP64 ? sqd_fp64(a, b, dim) : sqd_ref32(a, b, dim);
}

/* _find_bmu_xpu: strict '>' keeps the first index of the minimum; NaN distances are skipped;
 * distances >= 1e99 (i.e. +inf) never win; min_index starts at default_index. */
static int find_bmu(int mode, const float *w, const float *x, int mn, int dim, int default_index)
{
    double min_distance = 1.0e99;
    int min_index = default_index;
    for (int j = 0; j < mn; j++) {
        double d = sqd(mode, w + (size_t)j * dim, x, dim);
        if (d != d) continue;
        if (min_distance > d) {
            min_distance = d;
            min_index = j;
        }
    }
    return min_index;
}

/* _find_bmus_kernel: result[i] = bmu(x_i) with default index 0. */
void orc_find_bmus(int mode, const float *w, int mn, int dim, const float *x, int64_t n, int32_t *out)
{
    for (int64_t i = 0; i < n; i++) out[i] = find_bmu(mode, w, x + (size_t)i * dim, mn, dim, 0);
}

/* _count_bmus_kernel: result[bmu] += rho for rho > 0 (FP64 accumulator). */
void orc_count_bmus(int mode, const float *w, int mn, int dim, const float *x, const double *rho, int64_t n,
                    double *result)
{
    for (int64_t i = 0; i < n; i++) {
        double r = rho ? rho[i] : 1.0;
        if (r > 0.0) result[find_bmu(mode, w, x + (size_t)i * dim, mn, dim, 0)] += r;
    }
}

/* _compute_errors_xpu: top-2 BMUs with the reference's update rule, then
 * errors[0] += rho*sqrt(d1); errors[1] += rho if |t1-t2|^2 > penalty; errors[2] += rho. */
static void compute_errors_one(int mode, double *errors, const float *w, const float *topo, const float *x,
                               double density, double penalty, int mn, int dim, int32_t *c1_out, int32_t *c2_out)
{
    if (c1_out) { *c1_out = -1; *c2_out = -1; }
    if (!(density > 0.0)) return;
    double md2 = 1.0e99, md1 = 1.0e99;
    int mi2 = -1, mi1 = -1;
    for (int j = 0; j < mn; j++) {
        double d = sqd(mode, w + (size_t)j * dim, x, dim);
        if (d != d) continue;
        if (md1 > d) {
            if (j != mi1) { md2 = md1; mi2 = mi1; }
            md1 = d;
            mi1 = j;
        } else if (j != mi1 && md2 > d) {
            md2 = d;
            mi2 = j;
        }
    }
    if (mi1 < 0 || mi2 < 0) return;
    if (mi2 == mi1) return;
    if (c1_out) { *c1_out = mi1; *c2_out = mi2; }
    errors[0] += density * sqrt(md1);
    if (sqd(mode, topo + 2 * (size_t)mi1, topo + 2 * (size_t)mi2, 2) > penalty) errors[1] += density;
    errors[2] += density;
}

void orc_compute_errors(int mode, const float *w, const float *topo, int mn, int dim, const float *x,
                        const double *rho, int64_t n, double penalty, double *errors)
{
    for (int64_t i = 0; i < n; i++)
        compute_errors_one(mode, errors, w, topo, x + (size_t)i * dim, rho ? rho[i] : 1.0, penalty, mn, dim, 0, 0);
}

/* Per-sample top-2 (c1, c2) as the errors kernel sees them (-1 when the sample is skipped). */
void orc_top2(int mode, const float *w, const float *topo, int mn, int dim, const float *x, const double *rho,
              int64_t n, int32_t *c1, int32_t *c2)
{
    double scratch[3] = {0, 0, 0};
    for (int64_t i = 0; i < n; i++)
        compute_errors_one(mode, scratch, w, topo, x + (size_t)i * dim, rho ? rho[i] : 1.0, 2.0, mn, dim, c1 + i,
                           c2 + i);
}

/* asymptotic_decay_cpu and the FLOAT_EPS clamp. sigma0 is the float32-rounded constructor value. */
static inline double sigma_at(float sigma0, int64_t e, int64_t etot)
{
    double s = (double)sigma0 * (1.0 / (1.0 + 2.0 * (double)e / (double)etot));
    return s < ORC_FLOAT_EPS ? ORC_FLOAT_EPS : s;
}

/* _train_step2_xpu for one sample. REF32 adds into float accumulators in argument order (one of the
 * orders the reference's atomics allow); FP64 adds into double accumulators. */
static void train_step2(int mode, void *num, void *den, const float *w, const float *topo, const float *x,
                        float density32, double density64, double sigma, int mn, int dim)
{
    double density = mode == ORC_FP64 ? density64 : (double)density32;
    if (!(density > 0.0)) return;
    int bmu = find_bmu(mode, w, x, mn, dim, -1);
    if (bmu < 0) return;
    const float *tb = topo + 2 * (size_t)bmu;
    double inv_2sigma2 = 1.0 / (2.0 * sigma * sigma);
    for (int j = 0; j < mn; j++) {
        double h = density * exp(-sqd(mode, topo + 2 * (size_t)j, tb, 2) * inv_2sigma2);
        if (mode == ORC_FP64) {
            double *nj = (double *)num + (size_t)j * dim;
            for (int k = 0; k < dim; k++) nj[k] += h * (double)x[k];
            ((double *)den)[j] += h;
        } else {
            float *nj = (float *)num + (size_t)j * dim;
            for (int k = 0; k < dim; k++) nj[k] += (float)(h * (double)x[k]);
            ((float *)den)[j] += (float)(h * 1.0);
        }
    }
}

/* _train_step1_epoch_kernel (CPU branch): one sigma for the whole chunk. num/den are float (REF32)
 * or double (FP64) arrays of shape (mn, dim) and (mn,). rho32 / rho64 may be NULL (ones). */
void orc_train_epoch_accumulate(int mode, void *num, void *den, const float *w, const float *topo, const float *x,
                                const double *rho, int64_t n, int64_t cur_epoch, int64_t n_epochs, float sigma0,
                                int mn, int dim)
{
    double sigma = sigma_at(sigma0, cur_epoch, n_epochs);
    for (int64_t i = 0; i < n; i++) {
        double r = rho ? rho[i] : 1.0;
        train_step2(mode, num, den, w, topo, x + (size_t)i * dim, (float)r, r, sigma, mn, dim);
    }
}

/* _train_step1_iter_kernel (CPU branch): sigma depends on the global vector index cur_vector + i. */
void orc_train_iter_accumulate(int mode, void *num, void *den, const float *w, const float *topo, const float *x,
                               const double *rho, int64_t n, int64_t cur_vector, int64_t n_vectors, float sigma0,
                               int mn, int dim)
{
    for (int64_t i = 0; i < n; i++) {
        double r = rho ? rho[i] : 1.0;
        double sigma = sigma_at(sigma0, cur_vector + i, n_vectors);
        train_step2(mode, num, den, w, topo, x + (size_t)i * dim, (float)r, r, sigma, mn, dim);
    }
}

/* Host division of som_batch.py:354-359: w = num/den where den > FLOAT_EPS else 0 (FP32 arrays for
 * REF32, double for FP64; the result is written as float). */
void orc_divide(int mode, const void *num, const void *den, int mn, int dim, float *w_out)
{
    for (int j = 0; j < mn; j++) {
        for (int k = 0; k < dim; k++) {
            if (mode == ORC_FP64) {
                double d = ((const double *)den)[j];
                w_out[(size_t)j * dim + k] = d > ORC_FLOAT_EPS ? (float)(((const double *)num)[(size_t)j * dim + k] / d) : 0.0f;
            } else {
                float d = ((const float *)den)[j];
                w_out[(size_t)j * dim + k] = d > ORC_FLOAT_EPS ? ((const float *)num)[(size_t)j * dim + k] / d : 0.0f;
            }
        }
    }
}

/* Parallel restatement of the reference's CPU epoch kernel, used only as the timed CPU baseline:
 * prange over samples, FP32 atomic adds into the shared accumulators (atomicrmw fadd, monotonic). */
void orc_train_epoch_accumulate_omp(float *num, float *den, const float *w, const float *topo, const float *x,
                                    const double *rho, int64_t n, int64_t cur_epoch, int64_t n_epochs,
                                    float sigma0, int mn, int dim)
{
    double sigma = sigma_at(sigma0, cur_epoch, n_epochs);
    double inv_2sigma2 = 1.0 / (2.0 * sigma * sigma);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        float density = rho ? (float)rho[i] : 1.0f;
        if (!(density > 0.0f)) continue;
        const float *xi = x + (size_t)i * dim;
        int bmu = find_bmu(ORC_REF32, w, xi, mn, dim, -1);
        if (bmu < 0) continue;
        const float *tb = topo + 2 * (size_t)bmu;
        for (int j = 0; j < mn; j++) {
            double h = (double)density * exp(-sqd_ref32(topo + 2 * (size_t)j, tb, 2) * inv_2sigma2);
            float *nj = num + (size_t)j * dim;
            for (int k = 0; k < dim; k++) {
                float v = (float)(h * (double)xi[k]);
#pragma omp atomic
                nj[k] += v;
            }
            float v = (float)h;
#pragma omp atomic
            den[j] += v;
        }
    }
}

/* Parallel restatement of _compute_errors_kernel (CPU branch) for the CPU baseline. */
void orc_compute_errors_omp(const float *w, const float *topo, int mn, int dim, const float *x, const double *rho,
                            int64_t n, double penalty, double *errors)
{
    double e0 = 0, e1 = 0, e2 = 0;
#pragma omp parallel for schedule(static) reduction(+ : e0, e1, e2)
    for (int64_t i = 0; i < n; i++) {
        double loc[3] = {0, 0, 0};
        compute_errors_one(ORC_REF32, loc, w, topo, x + (size_t)i * dim, rho ? rho[i] : 1.0, penalty, mn, dim, 0, 0);
        e0 += loc[0];
        e1 += loc[1];
        e2 += loc[2];
    }
    errors[0] += e0;
    errors[1] += e1;
    errors[2] += e2;
}

/* Per-sample top-2 of the _compute_errors_xpu rule (som_abstract.py:832-853) with the distance of the
 * first, parallel over samples (each sample is independent, so the result equals the sequential
 * loop's). c1 = -1 (d1 = NaN) when no distance is finite, c2 = -1 when there is no second. Used by the
 * tests and the accuracy legs of bench.py to check full-size outputs. */
void orc_top2_omp(int mode, const float *w, int mn, int dim, const float *x, int64_t n, int32_t *c1, int32_t *c2,
                  double *d1)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        const float *xi = x + (size_t)i * dim;
        double md2 = 1.0e99, md1 = 1.0e99;
        int mi2 = -1, mi1 = -1;
        for (int j = 0; j < mn; j++) {
            double d = sqd(mode, w + (size_t)j * dim, xi, dim);
            if (d != d) continue;
            if (md1 > d) {
                if (j != mi1) { md2 = md1; mi2 = mi1; }
                md1 = d;
                mi1 = j;
            } else if (j != mi1 && md2 > d) {
                md2 = d;
                mi2 = j;
            }
        }
        c1[i] = mi1;
        c2[i] = mi2;
        d1[i] = mi1 >= 0 ? md1 : NAN;
    }
}

/* Parallel restatement of _count_bmus_kernel (som_abstract.py:699-732, nb.prange over samples with FP64
 * atomic adds) for the CPU baseline: per-thread FP64 maps summed at the end. */
void orc_count_bmus_omp(int mode, const float *w, int mn, int dim, const float *x, const double *rho, int64_t n,
                        double *result, double *thread_maps, int max_threads)
{
#pragma omp parallel num_threads(max_threads)
    {
        double *loc = thread_maps + (size_t)omp_get_thread_num() * mn;
        for (int j = 0; j < mn; j++) loc[j] = 0.0;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; i++) {
            double r = rho ? rho[i] : 1.0;
            if (r > 0.0) loc[find_bmu(mode, w, x + (size_t)i * dim, mn, dim, 0)] += r;
        }
    }
    for (int t = 0; t < max_threads; t++)
        for (int j = 0; j < mn; j++) result[j] += thread_maps[(size_t)t * mn + j];
}

/* Covariance.compute's sample loop (decontamination/algo/covariance.py:30-84, _update_welford_cov_sums):
 * sequential weighted Welford update in FP64 over one chunk, skipping rows whose weight is not
 * finite or <= 0 and rows with a non-finite value. x is FP64 (the reference casts each chunk with
 * astype(np.float64)); w may be NULL (ones). Updates *sum_w, mean (dim) and the upper triangle of
 * m2 (dim x dim, row-major); delta is scratch (dim). */
void orc_welford_update(double *sum_w, double *mean, double *m2, double *delta, const double *x, const double *w,
                        int64_t n, int dim)
{
    double sw = *sum_w;
    for (int64_t i = 0; i < n; i++) {
        const double wi = w ? w[i] : 1.0;
        if (!isfinite(wi) || wi <= 0.0) continue;
        const double *xi = x + i * dim;
        int ok = 1;
        for (int d = 0; d < dim; d++) ok &= isfinite(xi[d]) ? 1 : 0;
        if (!ok) continue;
        const double sw_new = sw + wi;
        for (int d = 0; d < dim; d++) delta[d] = xi[d] - mean[d];
        const double a = wi / sw_new;
        for (int d = 0; d < dim; d++) mean[d] += a * delta[d];
        for (int j = 0; j < dim; j++) {
            const double dj = delta[j];
            for (int k = j; k < dim; k++) m2[j * dim + k] += wi * dj * (xi[k] - mean[k]);
        }
        sw = sw_new;
    }
    *sum_w = sw;
}
