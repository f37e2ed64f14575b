// Weighted covariance of a chunk of samples (Covariance.compute, decontamination/algo/covariance.py:
// _update_welford_cov_sums :30-84 and _finalize_welford_cov :90-119), the pass that initialises a
// SOM_PCA.
//
// The reference runs Welford's update sample by sample on one CPU thread. Here one chunk takes two
// HBM-bound streaming passes with fixed-order reductions (deterministic):
//   pass 1: W = sum w and sum w x over the valid rows, mean = sum w x / W;
//   pass 2: the corrected two-pass sums on d = x - mean, S = sum w d d^T and c = sum w d, from which
//           the host forms M2 = S - c c^T / W and mean + c / W.
// A row is valid when its weight is finite and > 0 and all its values are finite (the reference's
// rule). Chunks are merged on the host with Chan's update.
//
// Layout: a wave takes 4 rows at a time, lane l holding value (row l/16, column 16b + l%16) of each
// 16-column block b, so one wave load reads 4 consecutive rows of a 16-column block (256 contiguous
// bytes when dim = 16). That is also the operand layout of v_mfma_f64_16x16x4f64 for both A (16 x 4,
// A[i][k] in lane i + 16k) and B (4 x 16, B[k][j] in lane j + 16k), so the 16 x 16 tile
// S_bi,bj += (w d_bi)^T d_bj of 4 rows is one FP64 MFMA; its result sits in lane j + 16q, register
// r, as D[4r + q][j] (rows interleaved across the lane groups, unlike the FP32 16x16 forms).
#include "launch.hpp"

#include <algorithm>
#include <vector>

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int COV_NT = 256;   // threads per workgroup (4 waves)
constexpr int COV_UNROLL = 8; // row quads per wave per step (32 rows, loads in flight together)

template <typename T>
__device__ __forceinline__ bool finite_val(T v)
{
    return __builtin_isfinite(v);
}

// Issues the loads of this lane's values of row `row` (lane column j, blocks 0..NBLK-1) and of the
// row's weight; the tail row and columns beyond dim read a clamped in-bounds address instead.
template <typename T, int NBLK>
__device__ __forceinline__ void load_row(const T *x, const double *rho, int64_t row, int64_t r1, int dim, int j,
                                         T (&v)[NBLK], double &w)
{
    const int64_t rr = row < r1 ? row : r1 - 1;
    w = rho ? rho[rr] : 1.0;
#pragma unroll
    for (int b = 0; b < NBLK; b++) {
        const int col = 16 * b + j;
        v[b] = x[rr * dim + (col < dim ? col : dim - 1)];
    }
}

// The loaded row as FP64 values (0 beyond dim) and its weight, 0 when the row is past r1 or invalid
// (weight not finite or <= 0, or any of its values not finite). Every lane of the wave calls it.
template <typename T, int NBLK>
__device__ __forceinline__ double check_row(const T (&raw)[NBLK], double w, int64_t row, int64_t r1, int dim, int j,
                                            int k, double (&v)[NBLK])
{
    bool fin = true;
#pragma unroll
    for (int b = 0; b < NBLK; b++) {
        const bool in = 16 * b + j < dim;
        fin = fin && (!in || finite_val(raw[b]));
        v[b] = in ? (double)raw[b] : 0.0;
    }
    if (!(row < r1 && __builtin_isfinite(w) && w > 0.0)) w = 0.0;
    const unsigned long long bad = __ballot(!fin);
    if ((bad >> (16 * k)) & 0xFFFFull) w = 0.0;
    return w;
}

// pass 1: per-workgroup partials [sum w, sum w x_col (col < 16*NBLK)]
template <typename T, int NBLK>
__global__ __launch_bounds__(COV_NT) void cov_pass1_kernel(const T *x, const double *rho, int64_t n, int dim,
                                                           int64_t rows_per_wg, double *part)
{
    constexpr int DP = 16 * NBLK;
    __shared__ double red[COV_NT / 64][DP + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, k = lane >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_wg;
    const int64_t r1 = min(n, r0 + rows_per_wg);
    double sw = 0.0, sx[NBLK];
#pragma unroll
    for (int b = 0; b < NBLK; b++) sx[b] = 0.0;
    for (int64_t q0 = r0 + 4 * wave; q0 < r1; q0 += 4 * 4 * COV_UNROLL) {
        T raw[COV_UNROLL][NBLK];
        double v[COV_UNROLL][NBLK], w[COV_UNROLL];
        // all loads of the step first, then the checks (their ballots wait for the data)
#pragma unroll
        for (int u = 0; u < COV_UNROLL; u++) load_row<T, NBLK>(x, rho, q0 + 16 * u + k, r1, dim, j, raw[u], w[u]);
#pragma unroll
        for (int u = 0; u < COV_UNROLL; u++) w[u] = check_row<T, NBLK>(raw[u], w[u], q0 + 16 * u + k, r1, dim, j, k, v[u]);
#pragma unroll
        for (int u = 0; u < COV_UNROLL; u++) {
            if (w[u] > 0.0) {
                sw += w[u];
#pragma unroll
                for (int b = 0; b < NBLK; b++) sx[b] += w[u] * v[u][b];
            }
        }
    }
    // lanes j, j+16, j+32, j+48 hold the four row phases of column j
#pragma unroll
    for (int b = 0; b < NBLK; b++) {
        sx[b] += __shfl_xor(sx[b], 16);
        sx[b] += __shfl_xor(sx[b], 32);
    }
    sw += __shfl_xor(sw, 16);
    sw += __shfl_xor(sw, 32);
    if (lane < 16) {
#pragma unroll
        for (int b = 0; b < NBLK; b++) red[wave][1 + 16 * b + j] = sx[b];
        if (lane == 0) red[wave][0] = sw;
    }
    __syncthreads();
    for (int e = threadIdx.x; e <= dim; e += COV_NT) {
        double s = 0.0;
        for (int wv = 0; wv < COV_NT / 64; wv++) s += red[wv][e];
        part[(int64_t)blockIdx.x * (dim + 1) + e] = s;
    }
}

// pass 2: per-workgroup partials [S tiles (T tiles x 256, D[i][j] at 16 i + j), c (16*NBLK)]
template <typename T, int NBLK>
__global__ __launch_bounds__(COV_NT) void cov_pass2_kernel(const T *x, const double *rho, int64_t n, int dim,
                                                           int64_t rows_per_wg, const double *sums, double *part)
{
    constexpr int DP = 16 * NBLK;
    constexpr int NT = NBLK * (NBLK + 1) / 2;  // upper-triangle 16x16 tiles
    constexpr int OUT = NT * 256 + DP;
    __shared__ double red[COV_NT / 64][OUT];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 15, k = lane >> 4;
    const double W = sums[0];
    double mean[NBLK];
#pragma unroll
    for (int b = 0; b < NBLK; b++) {
        const int col = 16 * b + j;
        mean[b] = (col < dim && W > 0.0) ? sums[1 + col] / W : 0.0;
    }
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_wg;
    const int64_t r1 = min(n, r0 + rows_per_wg);
    f64x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = f64x4{0.0, 0.0, 0.0, 0.0};
    double cs[NBLK];
#pragma unroll
    for (int b = 0; b < NBLK; b++) cs[b] = 0.0;
    for (int64_t q0 = r0 + 4 * wave; q0 < r1; q0 += 4 * 4 * COV_UNROLL) {
        T raw[COV_UNROLL][NBLK];
        double v[COV_UNROLL][NBLK], w[COV_UNROLL];
        // all loads of the step first, then the checks (their ballots wait for the data)
#pragma unroll
        for (int u = 0; u < COV_UNROLL; u++) load_row<T, NBLK>(x, rho, q0 + 16 * u + k, r1, dim, j, raw[u], w[u]);
#pragma unroll
        for (int u = 0; u < COV_UNROLL; u++) w[u] = check_row<T, NBLK>(raw[u], w[u], q0 + 16 * u + k, r1, dim, j, k, v[u]);
#pragma unroll
        for (int u = 0; u < COV_UNROLL; u++) {
            double d[NBLK], wd[NBLK];
#pragma unroll
            for (int b = 0; b < NBLK; b++) {
                d[b] = w[u] > 0.0 ? v[u][b] - mean[b] : 0.0;
                wd[b] = w[u] * d[b];
                cs[b] += wd[b];
            }
            int t = 0;
#pragma unroll
            for (int bi = 0; bi < NBLK; bi++)
#pragma unroll
                for (int bj = bi; bj < NBLK; bj++, t++)
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(wd[bi], d[bj], acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int b = 0; b < NBLK; b++) {
        cs[b] += __shfl_xor(cs[b], 16);
        cs[b] += __shfl_xor(cs[b], 32);
    }
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 4; r++) red[wave][t * 256 + 16 * (4 * r + k) + j] = acc[t][r];
    if (lane < 16)
#pragma unroll
        for (int b = 0; b < NBLK; b++) red[wave][NT * 256 + 16 * b + j] = cs[b];
    __syncthreads();
    for (int e = threadIdx.x; e < OUT; e += COV_NT) {
        double s = 0.0;
        for (int wv = 0; wv < COV_NT / 64; wv++) s += red[wv][e];
        part[(int64_t)blockIdx.x * OUT + e] = s;
    }
}

// out[e] = sum over p of part[p * width + e]: one workgroup per entry, fixed reduction order
__global__ __launch_bounds__(256) void cov_reduce_kernel(const double *part, int nparts, int width, double *out)
{
    __shared__ double red[4];
    const int e = blockIdx.x;
    double s = 0.0;
    for (int p = threadIdx.x; p < nparts; p += 256) s += part[(int64_t)p * width + e];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[e] = (red[0] + red[1]) + (red[2] + red[3]);
}

template <typename T, int NBLK>
int cov_launch(somhip_ctx *C, const T *x, const double *rho, int64_t n, int dim, double *out_host)
{
    constexpr int NT = NBLK * (NBLK + 1) / 2;
    constexpr int OUT = NT * 256 + 16 * NBLK;
    hipStream_t st = C->stream;
    // one round of resident workgroups, each a contiguous run of whole row quads
    static int per_cu = 0;
    if (!per_cu) {
        int a = 0, b = 0;
        SOM_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, cov_pass1_kernel<T, NBLK>, COV_NT, 0));
        SOM_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, cov_pass2_kernel<T, NBLK>, COV_NT, 0));
        per_cu = std::max(1, std::min(a, b));
    }
    const int64_t groups = std::max<int64_t>(1, std::min<int64_t>((int64_t)C->num_cus * per_cu, (n + 1023) / 1024));
    const int64_t rows_per_wg = std::max<int64_t>(4, ((n + groups - 1) / groups + 3) / 4 * 4);
    const int G = (int)std::max<int64_t>(1, (n + rows_per_wg - 1) / rows_per_wg);
    Workspace &Wk = C->ws;
    int rc;
    if ((rc = Wk.part.ensure(sizeof(double) * (size_t)G * std::max(OUT, dim + 1))) ||
        (rc = Wk.maps.ensure(sizeof(double) * (size_t)(dim + 1 + OUT))))
        return rc;
    double *part = (double *)Wk.part.p;
    double *sums = (double *)Wk.maps.p;
    double *tiles = sums + dim + 1;
    hipLaunchKernelGGL((cov_pass1_kernel<T, NBLK>), dim3(G), dim3(COV_NT), 0, st, x, rho, n, dim, rows_per_wg, part);
    SOM_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(cov_reduce_kernel, dim3(dim + 1), dim3(256), 0, st, part, G, dim + 1, sums);
    SOM_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL((cov_pass2_kernel<T, NBLK>), dim3(G), dim3(COV_NT), 0, st, x, rho, n, dim, rows_per_wg, sums,
                       part);
    SOM_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(cov_reduce_kernel, dim3(OUT), dim3(256), 0, st, part, G, OUT, tiles);
    SOM_HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)dim + 1 + OUT);
    SOM_HIP_TRY(hipMemcpyAsync(h.data(), sums, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    SOM_HIP_TRY(hipStreamSynchronize(st));
    // W, mean (the division pass 2 used), then M2 = S - c c^T / W and mean += c / W
    const double W = h[0];
    const double *S = h.data() + dim + 1;
    const double *c = S + NT * 256;
    out_host[0] = W;
    double *mean = out_host + 1, *m2 = out_host + 1 + dim;
    for (int a = 0; a < dim; a++) mean[a] = W > 0.0 ? h[1 + a] / W : 0.0;
    for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) {
            const int lo = std::min(a, b), hi = std::max(a, b);
            const int bi = lo / 16, bj = hi / 16;
            const int t = bi * NBLK - bi * (bi - 1) / 2 + (bj - bi);  // tile (bi, bj) of the upper triangle
            const double s = S[t * 256 + 16 * (lo % 16) + (hi % 16)];
            m2[a * dim + b] = W > 0.0 ? s - c[lo] * c[hi] / W : 0.0;
        }
    if (W > 0.0)
        for (int a = 0; a < dim; a++) mean[a] += c[a] / W;
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
#define SOM_COV(NB)                                                                                             \
    return x_is_f64 ? cov_launch<double, NB>(C, (const double *)x, rho, n, dim, out_host)                      \
                    : cov_launch<float, NB>(C, (const float *)x, rho, n, dim, out_host)
    switch ((dim + 15) / 16) {
        case 1: SOM_COV(1);
        case 2: SOM_COV(2);
        case 3: SOM_COV(3);
        default: SOM_COV(4);
    }
#undef SOM_COV
}
