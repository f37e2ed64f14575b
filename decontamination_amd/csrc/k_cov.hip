// Weighted covariance of a chunk of samples (Covariance.compute, decontamination/algo/covariance.py:
// _update_welford_cov_sums and _finalize_welford_cov), the pass that initialises a SOM_PCA.
//
// The reference runs Welford's update sample by sample on one CPU thread. Here one chunk takes two
// HBM-bound passes with fixed-order reductions (deterministic):
//   pass 1: W = sum w, mean = sum w x / W over the valid rows;
//   pass 2: the corrected two-pass sums on d = x - mean: S = sum w d d^T (upper triangle, 4x4
//           register blocks over a 64-row LDS tile) and c = sum w d, then M2 = S - c c^T / W and
//           mean += c / W.
// A row is valid when its weight is finite and > 0 and all its values are finite (the reference's
// rule). Chunks are merged on the host with Chan's update.
#include "launch.hpp"

#include <algorithm>

namespace {

constexpr int COV_NT = 256;
constexpr int COV_TR = 64;  // rows per LDS tile in pass 2

template <typename T>
__device__ __forceinline__ bool finite_val(T v)
{
    return __builtin_isfinite(v);
}

__device__ __forceinline__ double row_weight(const double *rho, int64_t i)
{
    const double w = rho ? rho[i] : 1.0;
    return (__builtin_isfinite(w) && w > 0.0) ? w : 0.0;
}

// pass 1: per-workgroup [sum w, sum w x_k (k < dim)]
template <typename T, int DP>
__global__ __launch_bounds__(COV_NT) void cov_sums_kernel(const T *x, const double *rho, int64_t n, int dim,
                                                          int64_t rows_per_wg, double *part)
{
    __shared__ double red[COV_NT / 64][DP + 1];
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_wg;
    const int64_t r1 = min(n, r0 + rows_per_wg);
    double acc[DP + 1];
#pragma unroll
    for (int k = 0; k <= DP; k++) acc[k] = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += COV_NT) {
        double w = row_weight(rho, i);
        T v[DP];
        bool ok = true;
#pragma unroll
        for (int k = 0; k < DP; k++) {
            v[k] = k < dim ? x[i * dim + k] : (T)0;
            ok = ok && finite_val(v[k]);
        }
        if (!ok) w = 0.0;
        if (w > 0.0) {
            acc[0] += w;
#pragma unroll
            for (int k = 0; k < DP; k++) acc[k + 1] += w * (double)v[k];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= DP; k++) {
        double s = acc[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) red[wave][k] = s;
    }
    __syncthreads();
    // [0] = sum w, [1 + j] = sum w x_j
    for (int k = threadIdx.x; k <= dim; k += COV_NT) {
        double s = 0.0;
        for (int wv = 0; wv < COV_NT / 64; wv++) s += red[wv][k];
        part[(int64_t)blockIdx.x * (dim + 1) + k] = s;
    }
}

// one workgroup: stats = [W, mean (dim)] from the pass-1 partials, in partial order
__global__ void cov_mean_kernel(const double *part, int nparts, int dim, double *stats)
{
    for (int k = threadIdx.x; k <= dim; k += blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < nparts; p++) s += part[(int64_t)p * (dim + 1) + k];
        stats[k] = s;
    }
    __syncthreads();
    const double W = stats[0];
    __syncthreads();
    for (int k = threadIdx.x; k < dim; k += blockDim.x) stats[1 + k] = W > 0.0 ? stats[1 + k] / W : 0.0;
}

template <int DP>
struct CovBlocks {
    static constexpr int NB = DP / 4;
    static constexpr int B = NB * (NB + 1) / 2;         // upper-triangle 4x4 blocks
    static constexpr int P = B >= COV_NT ? 1 : COV_NT / B;  // row phases
    static constexpr int OUT = B * 16 + DP;             // partial entries per workgroup
};

// pass 2: per-workgroup [S blocks (B x 16), c (DP)]
template <typename T, int DP>
__global__ __launch_bounds__(COV_NT) void cov_m2_kernel(const T *x, const double *rho, int64_t n, int dim,
                                                        int64_t rows_per_wg, const double *stats, double *part)
{
    using CB = CovBlocks<DP>;
    static_assert(CB::B <= COV_NT, "one 4x4 block per thread at most");
    __shared__ double d[COV_TR][DP];
    __shared__ double wd[COV_TR][DP];
    __shared__ double wt[COV_TR];
    __shared__ int bad[COV_TR];
    __shared__ double red[COV_NT][17];
    const int t = threadIdx.x;
    const int b = t % CB::B, p = t / CB::B;
    const bool active = p < CB::P;
    // block b -> (bj, bk), bj <= bk, row-major over the upper triangle
    int bj = 0, bk = 0;
    {
        int rem = b;
        while (rem >= CB::NB - bj) { rem -= CB::NB - bj; bj++; }
        bk = bj + rem;
    }
    double acc[16];
#pragma unroll
    for (int e = 0; e < 16; e++) acc[e] = 0.0;
    double cacc = 0.0;
    const double *mean = stats + 1;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_wg;
    const int64_t r1 = min(n, r0 + rows_per_wg);
    for (int64_t base = r0; base < r1; base += COV_TR) {
        if (t < COV_TR) {
            const int64_t i = base + t;
            const double w = i < r1 ? row_weight(rho, i) : 0.0;
            wt[t] = w;
            bad[t] = w > 0.0 ? 0 : 1;
        }
        __syncthreads();
        for (int e = t; e < COV_TR * DP; e += COV_NT) {
            const int r = e / DP, k = e % DP;
            const int64_t i = base + r;
            double v = 0.0;
            if (k < dim && i < r1) {
                const T xv = x[i * dim + k];
                if (!finite_val(xv)) bad[r] = 1;
                v = (double)xv - mean[k];
            }
            d[r][k] = v;
        }
        __syncthreads();
        for (int e = t; e < COV_TR * DP; e += COV_NT) {
            const int r = e / DP, k = e % DP;
            if (bad[r]) {
                d[r][k] = 0.0;
                wd[r][k] = 0.0;
            } else {
                wd[r][k] = wt[r] * d[r][k];
            }
        }
        __syncthreads();
        if (active) {
            for (int r = p; r < COV_TR; r += CB::P) {
                double a[4], c[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    a[q] = wd[r][4 * bj + q];
                    c[q] = d[r][4 * bk + q];
                }
#pragma unroll
                for (int q = 0; q < 4; q++)
#pragma unroll
                    for (int s = 0; s < 4; s++) acc[4 * q + s] += a[q] * c[s];
            }
        }
        if (t < DP)
            for (int r = 0; r < COV_TR; r++) cacc += wd[r][t];
        __syncthreads();
    }
    // phases of one block summed in phase order
#pragma unroll
    for (int e = 0; e < 16; e++) red[t][e] = acc[e];
    __syncthreads();
    double *out = part + (int64_t)blockIdx.x * CB::OUT;
    if (active && p == 0) {
        for (int e = 0; e < 16; e++) {
            double s = 0.0;
            for (int q = 0; q < CB::P; q++) s += red[q * CB::B + b][e];
            out[b * 16 + e] = s;
        }
    }
    if (t < DP) out[CB::B * 16 + t] = cacc;
}

// one workgroup: stats = [W, corrected mean (dim), M2 (dim x dim)] from the pass-2 partials
template <int DP>
__global__ void cov_final_kernel(const double *part, int nparts, int dim, double *stats, double *tmp)
{
    using CB = CovBlocks<DP>;
    for (int e = threadIdx.x; e < CB::OUT; e += blockDim.x) {
        double s = 0.0;
        for (int q = 0; q < nparts; q++) s += part[(int64_t)q * CB::OUT + e];
        tmp[e] = s;
    }
    __syncthreads();
    const double W = stats[0];
    const double *c = tmp + CB::B * 16;
    double *m2 = stats + 1 + dim;
    for (int e = threadIdx.x; e < dim * dim; e += blockDim.x) {
        int j = e / dim, k = e % dim;
        if (j > k) { const int s = j; j = k; k = s; }
        const int bj = j / 4, bk = k / 4;
        // index of block (bj, bk) in the row-major upper triangle
        const int b = bj * CB::NB - bj * (bj - 1) / 2 + (bk - bj);
        const double s = tmp[b * 16 + (j % 4) * 4 + (k % 4)];
        m2[e] = W > 0.0 ? s - c[j] * c[k] / W : 0.0;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < dim; k += blockDim.x)
        if (W > 0.0) stats[1 + k] += c[k] / W;
}

template <typename T, int DP>
int cov_launch(somhip_ctx *C, const T *x, const double *rho, int64_t n, int dim, double *out_host)
{
    using CB = CovBlocks<DP>;
    hipStream_t st = C->stream;
    const int64_t groups = std::max<int64_t>(1, std::min<int64_t>((int64_t)C->num_cus * 2, (n + 1023) / 1024));
    const int64_t rows_per_wg = std::max<int64_t>(1, (n + groups - 1) / groups);
    const int G = (int)((n + rows_per_wg - 1) / rows_per_wg);
    const size_t stats_n = 1 + (size_t)dim + (size_t)dim * dim;
    Workspace &Wk = C->ws;
    int rc;
    if ((rc = Wk.part.ensure(sizeof(double) * (size_t)std::max(G, 1) * std::max(CB::OUT, DP + 1))) ||
        (rc = Wk.maps.ensure(sizeof(double) * (stats_n + CB::OUT))))
        return rc;
    double *part = (double *)Wk.part.p;
    double *stats = (double *)Wk.maps.p;
    double *tmp = stats + stats_n;
    if (n > 0) {
        hipLaunchKernelGGL((cov_sums_kernel<T, DP>), dim3(G), dim3(COV_NT), 0, st, x, rho, n, dim, rows_per_wg, part);
        SOM_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(cov_mean_kernel, dim3(1), dim3(256), 0, st, part, n > 0 ? G : 0, dim, stats);
    SOM_HIP_TRY(hipGetLastError());
    if (n > 0) {
        hipLaunchKernelGGL((cov_m2_kernel<T, DP>), dim3(G), dim3(COV_NT), 0, st, x, rho, n, dim, rows_per_wg, stats,
                           part);
        SOM_HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL((cov_final_kernel<DP>), dim3(1), dim3(256), 0, st, part, n > 0 ? G : 0, dim, stats, tmp);
    SOM_HIP_TRY(hipGetLastError());
    SOM_HIP_TRY(hipMemcpyAsync(out_host, stats, sizeof(double) * stats_n, hipMemcpyDeviceToHost, st));
    SOM_HIP_TRY(hipStreamSynchronize(st));
    return SOMHIP_OK;
}

}  // namespace

extern "C" int somhip_covariance_chunk(somhip_ctx *C, const void *x, int x_is_f64, const double *rho, int64_t n,
                                       int dim, double *out_host)
{
    if (!C || n < 0 || dim < 1 || dim > 64 || !out_host || (n > 0 && !x)) {
        somhip_set_error("somhip_covariance_chunk: invalid arguments (1 <= dim <= 64)");
        return SOMHIP_ERR_ARG;
    }
    SOM_HIP_TRY(hipSetDevice(C->device));
    const int dp = (dim + 15) / 16 * 16;
#define SOM_COV(D)                                                                                              \
    return x_is_f64 ? cov_launch<double, D>(C, (const double *)x, rho, n, dim, out_host)                       \
                    : cov_launch<float, D>(C, (const float *)x, rho, n, dim, out_host)
    switch (dp) {
        case 16: SOM_COV(16);
        case 32: SOM_COV(32);
        case 48: SOM_COV(48);
        default: SOM_COV(64);
    }
#undef SOM_COV
}
