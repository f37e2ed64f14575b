// Training, activation and error-reduction kernels for gfx950.
//
// Per-BMU sums use exact fixed-point integers: every contribution rho_i * x_ik is scaled by 2^shift_k
// (shift_k chosen per chunk so that the chunk's total cannot overflow int64) and rounded once; the
// sums are then exact integer sums, which do not depend on the order of accumulation. Each workgroup
// accumulates a contiguous sample range into LDS with 64-bit LDS atomics (a slice of neurons at a
// time), writes its slice table, and a reduction kernel adds the workgroup tables (int64) and scales
// the result into the FP64 accumulator. Results are therefore bit-identical for any launch shape,
// batching or chunking into the same calls.
#include "common.hpp"

#include <algorithm>

namespace {

constexpr size_t LDS_SCATTER_MAX = 160 * 1024;  // neuron slice per scatter workgroup (all of gfx950's LDS)
constexpr int SCATTER_NT = 1024;                // threads per scatter workgroup

// ------------------------------------------------------------------------------------------------
// Column-wise max |v| over finite values; out[k] for k < dim is max|x_k|, out[dim] is max rho (after
// the optional float rounding). Stored as FP64 bits with atomicMax (values >= 0 order as their bits),
// so every finite maximum of float32 or float64 data is held exactly.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long bound_bits(double v) { return __builtin_bit_cast(unsigned long long, v); }
__device__ __forceinline__ double bound_value(unsigned long long b) { return __builtin_bit_cast(double, b); }

template <int DP, class T>
__global__ __launch_bounds__(256) void absmax_kernel(const T *__restrict__ x, int dim, const double *__restrict__ rho,
                                                    int rho_f32, int64_t n, unsigned long long *__restrict__ out)
{
    T mx[DP];
#pragma unroll
    for (int k = 0; k < DP; k++) mx[k] = (T)0;
    double mr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (x) {
#pragma unroll
            for (int k = 0; k < DP; k++) {
                if (k < dim) {
                    const T v = fabs(x[i * dim + k]);
                    if (v < (T)__builtin_inf()) mx[k] = fmax(mx[k], v);
                }
            }
        }
        if (rho) {
            double r = rho[i];
            if (rho_f32) r = (double)(float)r;
            if (r > 0.0 && r < 1e300) mr = fmax(mr, r);
        }
    }
    __shared__ unsigned long long red[DP + 1];
    if (threadIdx.x <= DP) red[threadIdx.x] = 0ull;
    __syncthreads();
    if (x) {
#pragma unroll
        for (int k = 0; k < DP; k++)
            if (k < dim) atomicMax(&red[k], bound_bits((double)mx[k]));
    }
    atomicMax(&red[DP], bound_bits(mr));
    __syncthreads();
    if ((int)threadIdx.x < dim && x) atomicMax(&out[threadIdx.x], red[threadIdx.x]);
    if (threadIdx.x == 0) atomicMax(&out[dim], red[DP]);
}

// shift = floor(61 - log2(bx * rb * n)) for a column bound bx and a rho bound rb: every |rho x| * 2^shift
// <= 2^61 / n, so the int64 sum of n terms (each rounded once) stays below 2^62. Where the product
// overflows (float64 data), the exponents are added instead, a bound at most 4 times looser.
__device__ int scale_shift(double bx, double rb, int64_t n)
{
    if (!(bx > 0.0) || !(rb > 0.0)) return 0;
    const double nn = (double)(n > 0 ? n : 1);
    const double t = bx * rb * nn;
    int e;
    if (t < __builtin_inf()) {
        frexp(t, &e);  // t < 2^e
    } else {
        int ex, er, en;
        frexp(bx, &ex);
        frexp(rb, &er);
        frexp(nn, &en);
        e = ex + er + en;
    }
    return min(1000, max(-1000, 61 - e));
}

// Column dim uses the rho bound alone (1 when rho is absent).
__global__ void shifts_kernel(const unsigned long long *__restrict__ bounds, int dim, int has_x, int has_rho,
                              int64_t n, int *__restrict__ shifts)
{
    const int k = threadIdx.x;
    if (k > dim) return;
    const double rb = has_rho ? bound_value(bounds[dim]) : 1.0;
    const double bx = (k == dim) ? 1.0 : (has_x ? bound_value(bounds[k]) : 0.0);
    shifts[k] = scale_shift(bx, rb, n);
}

// parts[b][k] = max over rows r = b, b + RB, b + 2RB, ... of parts[r][k] (b < RB); the other rows are
// only read, so folding in place is safe.
__global__ __launch_bounds__(256) void rowmax_fold_kernel(float *__restrict__ parts, int64_t rows, int dpad, int RB)
{
    const int b = blockIdx.x;
    __shared__ float red[256];
    const int k = threadIdx.x % dpad;
    const int r0 = threadIdx.x / dpad;
    const int rs = blockDim.x / dpad;
    float m = 0.f;
    if (r0 < rs)
        for (int64_t r = b + (int64_t)r0 * RB; r < rows; r += (int64_t)rs * RB) m = fmaxf(m, parts[r * dpad + k]);
    red[threadIdx.x] = m;
    __syncthreads();
    if ((int)threadIdx.x < dpad) {
        float mk = 0.f;
        for (int j = threadIdx.x; j < rs * dpad; j += dpad) mk = fmaxf(mk, red[j]);
        parts[(int64_t)b * dpad + threadIdx.x] = mk;
    }
}

// Same scales from the per-wave max |x_k| rows written by the BMU kernels (rows x dpad floats) and an
// optional max rho (FP64 bits, from absmax_kernel on rho); column dim is the rho column.
__global__ __launch_bounds__(1024) void shifts_parts_kernel(const float *__restrict__ parts, int64_t rows, int dpad,
                                                           int dim, const unsigned long long *__restrict__ rho_bound,
                                                           int64_t n, int *__restrict__ shifts)
{
    __shared__ float red[1024];
    const int k = threadIdx.x % dpad;
    const int r0 = threadIdx.x / dpad;
    const int rs = blockDim.x / dpad;
    float m = 0.f;
    if (r0 < rs)
        for (int64_t r = r0; r < rows; r += rs) m = fmaxf(m, parts[r * dpad + k]);
    red[threadIdx.x] = m;
    __syncthreads();
    if ((int)threadIdx.x <= dim) {
        const double rb = rho_bound ? bound_value(rho_bound[0]) : 1.0;
        double bx = 1.0;
        if ((int)threadIdx.x < dim) {
            float mk = 0.f;
            for (int j = threadIdx.x; j < rs * dpad; j += dpad) mk = fmaxf(mk, red[j]);
            bx = (double)mk;
        }
        shifts[threadIdx.x] = scale_shift(bx, rb, n);
    }
}

// One sample row into registers, zero padded to DP; 16-byte vector loads when XVEC (dim a multiple
// of the vector width, 16-byte aligned rows).
template <int DP, bool XVEC>
__device__ __forceinline__ void load_row(const float *xr, int dim, float (&xv)[DP])
{
    if (XVEC) {
#pragma unroll
        for (int v = 0; v < DP / 4; v++) {
            const f32x4 t = 4 * v < dim ? ((const f32x4 *)xr)[v] : f32x4{0.f, 0.f, 0.f, 0.f};
            xv[4 * v] = t[0]; xv[4 * v + 1] = t[1]; xv[4 * v + 2] = t[2]; xv[4 * v + 3] = t[3];
        }
    } else {
#pragma unroll
        for (int k = 0; k < DP; k++) xv[k] = k < dim ? xr[k] : 0.f;
    }
}
template <int DP, bool XVEC>
__device__ __forceinline__ void load_row(const double *xr, int dim, double (&xv)[DP])
{
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    if (XVEC) {
#pragma unroll
        for (int v = 0; v < DP / 2; v++) {
            const f64x2 t = 2 * v < dim ? ((const f64x2 *)xr)[v] : f64x2{0.0, 0.0};
            xv[2 * v] = t[0]; xv[2 * v + 1] = t[1];
        }
    } else {
#pragma unroll
        for (int k = 0; k < DP; k++) xv[k] = k < dim ? xr[k] : 0.0;
    }
}

// ------------------------------------------------------------------------------------------------
// Fixed-point scatter of per-sample rows into per-neuron integer sums.
//   nv = dim + 1 columns (x != NULL: rho*x_0..rho*x_{dim-1}, rho) or nv = 1 (x == NULL: rho).
//   Samples with bmu < 0 or !(rho > 0) are skipped. rho_f32: round rho to float first (training).
// ------------------------------------------------------------------------------------------------
// Workgroup r adds the samples of its contiguous range into LDS, one neuron slice at a time (the
// slice fills the LDS), and writes each slice to partial[r]. Per slice, every wave sweeps its samples'
// BMU indices (coalesced), queues the ones inside the slice in a per-wave LDS ring, and whenever 64
// are queued adds them with all lanes busy: one sample per lane, its row loaded as 16-byte vectors
// when XVEC (dim % 4 == 0, 16-byte aligned rows), then one 64-bit LDS atomic per column.
template <int DP, bool XVEC, class T>
__global__ __launch_bounds__(SCATTER_NT) void scatter_fixed_kernel(const int32_t *__restrict__ bmu,
                                                                   const T *__restrict__ x, int dim,
                                                                   const double *__restrict__ rho, int rho_f32,
                                                                   int64_t n, int64_t per_wg,
                                                                   const int *__restrict__ shifts, int mn, int nv,
                                                                   int slice_len,
                                                                   unsigned long long *__restrict__ partial)
{
    constexpr int W = 64;
    constexpr int NWV = SCATTER_NT / W;
    constexpr int QCAP = 2 * W;  // ring entries per wave
    extern __shared__ unsigned long long lds[];
    unsigned int *q_idx = (unsigned int *)(lds + (size_t)slice_len * nv);   // [NWV][QCAP] sample - i0
    unsigned short *q_slot = (unsigned short *)(q_idx + NWV * QCAP);       // [NWV][QCAP] neuron - c0
    const int lane = threadIdx.x & (W - 1);
    const int wave = threadIdx.x / W;
    unsigned int *qi = q_idx + wave * QCAP;
    unsigned short *qs = q_slot + wave * QCAP;
    const int64_t i0 = (int64_t)blockIdx.x * per_wg;
    const int64_t i1 = min(n, i0 + per_wg);
    unsigned long long *out = partial + (int64_t)blockIdx.x * mn * nv;

    // adds queue entries [h, h + cnt) (cnt <= 64), one per lane
    auto drain = [&](unsigned h, unsigned cnt) {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if ((unsigned)lane < cnt) {
            const unsigned e = (h + lane) & (QCAP - 1);
            const int64_t i = i0 + qi[e];
            unsigned long long *row = lds + (int64_t)qs[e] * nv;
            double r = rho ? rho[i] : 1.0;
            if (rho_f32) r = (double)(float)r;
            if (r > 0.0 && r < 1e300) {
                if (x) {
                    T xv[DP];
                    load_row<DP, XVEC>(x + i * dim, dim, xv);
#pragma unroll
                    for (int k = 0; k < DP; k++) {
                        // r * x is exact in FP64 for FP32 data (24-bit mantissas; FP64 data rounds it
                        // once), and so is the power-of-two scaling
                        if (k < dim)
                            atomicAdd(&row[k], (unsigned long long)llrint(ldexp(r * (double)xv[k], shifts[k])));
                    }
                }
                atomicAdd(&row[nv - 1], (unsigned long long)llrint(ldexp(r, shifts[nv - 1])));
            }
        }
        __builtin_amdgcn_wave_barrier();
    };

    for (int c0 = 0; c0 < mn; c0 += slice_len) {
        const int cl = min(slice_len, mn - c0);
        for (int e = threadIdx.x; e < cl * nv; e += SCATTER_NT) lds[e] = 0ull;
        __syncthreads();
        unsigned head = 0, tail = 0;  // wave-uniform ring positions
        int64_t ib = i0 + (int64_t)wave * W + lane;
        int cn = ib < i1 ? bmu[ib] : -1;
        for (; ib - lane < i1; ib += SCATTER_NT) {
            const int c = cn;
            const int64_t inext = ib + SCATTER_NT;
            cn = inext < i1 ? bmu[inext] : -1;
            const bool in = c >= c0 && c < c0 + cl;
            const unsigned long long m = __builtin_amdgcn_ballot_w64(in);
            const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            if (in) {
                const unsigned e = (tail + below) & (QCAP - 1);
                qi[e] = (unsigned)(ib - i0);
                qs[e] = (unsigned short)(c - c0);
            }
            tail += (unsigned)__builtin_popcountll(m);
            if (tail - head >= (unsigned)W) {
                drain(head, W);
                head += W;
            }
        }
        while (tail != head) {
            const unsigned cnt = min(tail - head, (unsigned)W);
            drain(head, cnt);
            head += cnt;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < cl * nv; e += SCATTER_NT) out[(int64_t)c0 * nv + e] = lds[e];
        __syncthreads();
    }
}

// Bucketed form for maps that need several neuron slices. Workgroup r first sweeps its range's BMU
// indices once: an LDS histogram per slice (slice = bmu >> log2_slice), a prefix over the slices, and
// every sample's entry ((sample - i0) << 12 | neuron-in-slice) written to bucket[i0 + position], grouped
// by slice. Each slice pass then reads only its own bucket entries (one per thread, the sample row as
// 16-byte vectors when XVEC) instead of sweeping every index of the range again. Positions within a
// bucket depend on atomic order, but the sums are integers, so the tables do not.
constexpr int BUCKET_MAX_SLICES = 1024;
constexpr int BUCKET_LOCAL_BITS = 20;  // samples per workgroup range < 2^20
constexpr int BUCKET_NEURON_BITS = 12; // neurons per slice <= 2^12

template <int DP, bool XVEC, class T>
__global__ __launch_bounds__(SCATTER_NT) void scatter_bucket_kernel(const int32_t *__restrict__ bmu,
                                                                    const T *__restrict__ x, int dim,
                                                                    const double *__restrict__ rho, int rho_f32,
                                                                    int64_t n, int64_t per_wg,
                                                                    const int *__restrict__ shifts, int mn, int nv,
                                                                    int log2_slice, int nslices,
                                                                    unsigned int *__restrict__ bucket,
                                                                    unsigned long long *__restrict__ partial)
{
    extern __shared__ unsigned long long lds[];  // [slice_len][nv]
    __shared__ unsigned int start[BUCKET_MAX_SLICES + 1];
    __shared__ unsigned int cursor[BUCKET_MAX_SLICES];
    const int slice_len = 1 << log2_slice;
    const int64_t i0 = (int64_t)blockIdx.x * per_wg;
    const int64_t i1 = min(n, i0 + per_wg);
    unsigned long long *out = partial + (int64_t)blockIdx.x * mn * nv;
    unsigned int *bk = bucket + i0;

    for (int s = threadIdx.x; s < nslices; s += SCATTER_NT) cursor[s] = 0u;
    __syncthreads();
    for (int64_t i = i0 + threadIdx.x; i < i1; i += SCATTER_NT) {
        const int c = bmu[i];
        if (c >= 0 && c < mn) atomicAdd(&cursor[c >> log2_slice], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned int acc = 0u;
        for (int s = 0; s < nslices; s++) {
            start[s] = acc;
            acc += cursor[s];
            cursor[s] = start[s];
        }
        start[nslices] = acc;
    }
    __syncthreads();
    for (int64_t i = i0 + threadIdx.x; i < i1; i += SCATTER_NT) {
        const int c = bmu[i];
        if (c >= 0 && c < mn) {
            const unsigned int pos = atomicAdd(&cursor[c >> log2_slice], 1u);
            bk[pos] = ((unsigned int)(i - i0) << BUCKET_NEURON_BITS) | (unsigned int)(c & (slice_len - 1));
        }
    }
    __syncthreads();  // the bucket entries written above are read back by other threads of the workgroup

    for (int s = 0; s < nslices; s++) {
        const int c0 = s << log2_slice;
        const int cl = min(slice_len, mn - c0);
        for (int e = threadIdx.x; e < cl * nv; e += SCATTER_NT) lds[e] = 0ull;
        __syncthreads();
        for (unsigned int e = start[s] + threadIdx.x; e < start[s + 1]; e += SCATTER_NT) {
            const unsigned int ent = bk[e];
            const int64_t i = i0 + (ent >> BUCKET_NEURON_BITS);
            unsigned long long *row = lds + (int64_t)(ent & ((1u << BUCKET_NEURON_BITS) - 1u)) * nv;
            double r = rho ? rho[i] : 1.0;
            if (rho_f32) r = (double)(float)r;
            if (!(r > 0.0 && r < 1e300)) continue;
            if (x) {
                T xv[DP];
                load_row<DP, XVEC>(x + i * dim, dim, xv);
#pragma unroll
                for (int k = 0; k < DP; k++) {
                    // r * x is exact in FP64 for FP32 data (24-bit mantissas; FP64 data rounds it once),
                    // and so is the power-of-two scaling
                    if (k < dim) atomicAdd(&row[k], (unsigned long long)llrint(ldexp(r * (double)xv[k], shifts[k])));
                }
            }
            atomicAdd(&row[nv - 1], (unsigned long long)llrint(ldexp(r, shifts[nv - 1])));
        }
        __syncthreads();
        for (int e = threadIdx.x; e < cl * nv; e += SCATTER_NT) out[(int64_t)c0 * nv + e] = lds[e];
        __syncthreads();
    }
}

// Sorted form of the bucketed scatter (dims up to 32, maps whose mn counters fit in the slice table's
// LDS). The first sweep counts every neuron's samples in LDS, an exclusive prefix over the counters gives
// each neuron its range of the workgroup's bucket entries, and a second sweep writes the entries there,
// so the entries of a slice come grouped by neuron. A slice pass then gives each thread a contiguous run
// of entries; consecutive entries mostly share a neuron, so the thread adds their fixed-point
// contributions in registers and issues the row's LDS atomics only when the neuron changes. Active LDS
// atomic lanes, which bound the per-sample form, fall by the mean run length (the partial sums of one
// thread stay below 2^61 as the whole range's do). Entry order within a neuron follows atomic order,
// which the integer sums do not see.
template <int DP, bool XVEC, class T>
__global__ __launch_bounds__(SCATTER_NT) void scatter_sorted_kernel(const int32_t *__restrict__ bmu,
                                                                    const T *__restrict__ x, int dim,
                                                                    const double *__restrict__ rho, int rho_f32,
                                                                    int64_t n, int64_t per_wg,
                                                                    const int *__restrict__ shifts, int mn, int nv,
                                                                    int log2_slice, int nslices,
                                                                    unsigned int *__restrict__ bucket,
                                                                    unsigned long long *__restrict__ partial)
{
    extern __shared__ unsigned long long lds[];  // mn counters during the sweeps, then [slice_len][nv]
    __shared__ unsigned int start[BUCKET_MAX_SLICES + 1];
    __shared__ unsigned int scan[SCATTER_NT];
    unsigned int *cnt = (unsigned int *)lds;
    const int tid = threadIdx.x;
    const int slice_len = 1 << log2_slice;
    const int64_t i0 = (int64_t)blockIdx.x * per_wg;
    const int64_t i1 = min(n, i0 + per_wg);
    unsigned long long *out = partial + (int64_t)blockIdx.x * mn * nv;
    unsigned int *bk = bucket + i0;

    for (int c = tid; c < mn; c += SCATTER_NT) cnt[c] = 0u;
    __syncthreads();
    for (int64_t i = i0 + tid; i < i1; i += SCATTER_NT) {
        const int c = bmu[i];
        if (c >= 0 && c < mn) atomicAdd(&cnt[c], 1u);
    }
    __syncthreads();
    // exclusive prefix over the counters: every thread sums a contiguous chunk, the chunk sums are scanned
    const int per = (mn + SCATTER_NT - 1) / SCATTER_NT;
    const int cb = min(mn, tid * per), ce = min(mn, cb + per);
    unsigned int sum = 0u;
    for (int c = cb; c < ce; c++) sum += cnt[c];
    scan[tid] = sum;
    __syncthreads();
    for (int o = 1; o < SCATTER_NT; o <<= 1) {
        const unsigned int v = tid >= o ? scan[tid - o] : 0u;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
    }
    unsigned int run = scan[tid] - sum;
    for (int c = cb; c < ce; c++) {
        const unsigned int k = cnt[c];
        cnt[c] = run;
        run += k;
    }
    __syncthreads();
    for (int sl = tid; sl < nslices; sl += SCATTER_NT) start[sl] = cnt[sl << log2_slice];
    if (tid == 0) start[nslices] = scan[SCATTER_NT - 1];
    __syncthreads();
    for (int64_t i = i0 + tid; i < i1; i += SCATTER_NT) {
        const int c = bmu[i];
        if (c >= 0 && c < mn) {
            const unsigned int pos = atomicAdd(&cnt[c], 1u);
            bk[pos] = ((unsigned int)(i - i0) << BUCKET_NEURON_BITS) | (unsigned int)(c & (slice_len - 1));
        }
    }
    __syncthreads();  // the entries written above are read back by other threads; the counters are done

    for (int sl = 0; sl < nslices; sl++) {
        const int c0 = sl << log2_slice;
        const int cl = min(slice_len, mn - c0);
        for (int e = tid; e < cl * nv; e += SCATTER_NT) lds[e] = 0ull;
        __syncthreads();
        const unsigned int e0 = start[sl], e1 = start[sl + 1];
        const unsigned int len = (e1 - e0 + SCATTER_NT - 1) / SCATTER_NT;
        const unsigned int eb = min(e1, e0 + (unsigned int)tid * len), ee = min(e1, eb + len);
        long long acc[DP];
        long long acc_r = 0;
        int cur = -1;
        auto flush = [&]() {
            unsigned long long *row = lds + (int64_t)cur * nv;
            if (x) {
#pragma unroll
                for (int k = 0; k < DP; k++)
                    if (k < dim) atomicAdd(&row[k], (unsigned long long)acc[k]);
            }
            atomicAdd(&row[nv - 1], (unsigned long long)acc_r);
        };
        for (unsigned int e = eb; e < ee; e++) {
            const unsigned int ent = bk[e];
            const int nn = (int)(ent & ((1u << BUCKET_NEURON_BITS) - 1u));
            if (nn != cur) {
                if (cur >= 0) flush();
                cur = nn;
#pragma unroll
                for (int k = 0; k < DP; k++) acc[k] = 0;
                acc_r = 0;
            }
            const int64_t i = i0 + (ent >> BUCKET_NEURON_BITS);
            double r = rho ? rho[i] : 1.0;
            if (rho_f32) r = (double)(float)r;
            if (!(r > 0.0 && r < 1e300)) continue;
            if (x) {
                T xv[DP];
                load_row<DP, XVEC>(x + i * dim, dim, xv);
#pragma unroll
                for (int k = 0; k < DP; k++)
                    if (k < dim) acc[k] += llrint(ldexp(r * (double)xv[k], shifts[k]));
            }
            acc_r += llrint(ldexp(r, shifts[nv - 1]));
        }
        if (cur >= 0) flush();
        __syncthreads();
        for (int e = tid; e < cl * nv; e += SCATTER_NT) out[(int64_t)c0 * nv + e] = lds[e];
        __syncthreads();
    }
}

// acc[e] += 2^-shift * sum_b partial[b][e] (integer sum, exact)
__global__ void scatter_reduce_kernel(const unsigned long long *__restrict__ partial, int nwg, int64_t width, int nv,
                                      const int *__restrict__ shifts, double *__restrict__ acc)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= width) return;
    // integer sums are associative: 8 independent partial sums keep 8 loads in flight
    unsigned long long s8[8] = {0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull, 0ull};
    int b = 0;
    for (; b + 8 <= nwg; b += 8)
#pragma unroll
        for (int u = 0; u < 8; u++) s8[u] += partial[(int64_t)(b + u) * width + e];
    for (; b < nwg; b++) s8[0] += partial[(int64_t)b * width + e];
    unsigned long long s = 0ull;
#pragma unroll
    for (int u = 0; u < 8; u++) s += s8[u];
    const int k = (int)(e % nv);
    acc[e] += ldexp((double)(long long)s, -shifts[k]);
}

// ------------------------------------------------------------------------------------------------
// Epoch-mode update: part[p][j][k] = sum_{c in split p} H_jc S_c[k], H_jc = exp(-|t_j-t_c|^2 inv2s2).
// ------------------------------------------------------------------------------------------------
template <int NV, class T>
__global__ __launch_bounds__(256) void update_partial_kernel(const double *__restrict__ S, const T *__restrict__ topo,
                                                             int mn, int nv, double inv_2s2, int chunk,
                                                             double *__restrict__ part)
{
    constexpr int TC = 64;
    __shared__ double s_S[TC * NV];
    __shared__ T s_t[TC * 2];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    const int cbeg = p * chunk, cend = min(mn, cbeg + chunk);
    T tj[2] = {(T)0, (T)0};
    if (j < mn) { tj[0] = topo[2 * j]; tj[1] = topo[2 * j + 1]; }
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] = 0.0;
    for (int c0 = cbeg; c0 < cend; c0 += TC) {
        const int cl = min(TC, cend - c0);
        __syncthreads();
        for (int e = threadIdx.x; e < cl * nv; e += blockDim.x) s_S[(e / nv) * NV + e % nv] = S[(int64_t)c0 * nv + e];
        for (int e = threadIdx.x; e < cl * 2; e += blockDim.x) s_t[e] = topo[2 * c0 + e];
        __syncthreads();
        for (int cc = 0; cc < cl; cc++) {
            const double cnt = s_S[cc * NV + nv - 1];
            if (cnt == 0.0) continue;  // no sample has this BMU: S_c = 0, nothing to add
            const double h = exp(-sqd_topo(tj, &s_t[2 * cc]) * inv_2s2);
#pragma unroll
            for (int k = 0; k < NV; k++)
                if (k < nv) acc[k] = __fma_rn(h, s_S[cc * NV + k], acc[k]);
        }
    }
    if (j < mn) {
#pragma unroll
        for (int k = 0; k < NV; k++)
            if (k < nv) part[((int64_t)p * mn + j) * nv + k] = acc[k];
    }
}

// Separable form of the same update on the reference's grids (neuron a*n + b at x = a (+ 0.5 (b % 2)
// for hexagonal), y = b or b sqrt(3)/2, all FP32): H_jc = exp(-(fl(dx^2) + fl(dy^2)) inv2s2) =
// gx(dx) gy(dy), and dx = (a_j - a_c) + 0.5 (p_j - p_c) with p the column parity (0 on square grids).
// Pass 1, per source column b_c and destination parity p_j:
//   T[b_c][p_j][a_j][k] = sum_{a_c} gx(a_j - a_c + 0.5 (p_j - p_c)) S[a_c n + b_c][k]
// Pass 2, per destination column b_j: num[a_j n + b_j][k] = sum_{b_c} gy(y_j - y_c) T[b_c][p_j][a_j][k],
// then w = num / den as divide_kernel does. m*n*(m+n)*(D+1) products instead of (m n)^2 (D+1).
__global__ __launch_bounds__(256) void update_sep_rows_kernel(const double *__restrict__ S, int m, int n, int nv,
                                                              int hex, double inv_2s2, double *__restrict__ T)
{
    extern __shared__ double sh[];
    double *sS = sh;            // [m][nv] the source column
    double *gx = sh + m * nv;   // [2m - 1] gx(a_j - a_c + delta), a_j - a_c = t - (m - 1)
    const int np = hex ? 2 : 1;
    const int bc = blockIdx.x / np, pj = blockIdx.x % np;
    for (int e = threadIdx.x; e < m * nv; e += blockDim.x) {
        const int a = e / nv, k = e % nv;
        sS[e] = S[((int64_t)a * n + bc) * nv + k];
    }
    const float delta = hex ? 0.5f * (float)(pj - (bc & 1)) : 0.f;
    for (int t = threadIdx.x; t < 2 * m - 1; t += blockDim.x) {
        const float d = (float)(t - (m - 1)) + delta;  // exact: small halves
        const float d2 = __fmul_rn(d, d);
        gx[t] = exp(-(double)d2 * inv_2s2);
    }
    __syncthreads();
    double *out = T + (int64_t)blockIdx.x * m * nv;
    for (int e = threadIdx.x; e < m * nv; e += blockDim.x) {
        const int aj = e / nv, k = e % nv;
        double acc = 0.0;
        for (int ac = 0; ac < m; ac++) acc = __fma_rn(gx[aj - ac + m - 1], sS[ac * nv + k], acc);
        out[e] = acc;
    }
}

template <class R>
__global__ __launch_bounds__(256) void update_sep_cols_kernel(const double *__restrict__ T, const R *__restrict__ topo,
                                                              int m, int n, int dim, int hex, double inv_2s2,
                                                              R *__restrict__ w, R *__restrict__ hist)
{
    extern __shared__ double sh[];
    const int nv = dim + 1;
    double *gy = sh;            // [n]
    double *num = sh + n;       // [m][nv]
    const int bj = blockIdx.x;
    const int np = hex ? 2 : 1;
    const int pj = hex ? (bj & 1) : 0;
    const R yj = topo[2 * bj + 1];  // neuron (0, bj)
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const R tt[2] = {(R)0, yj}, tc[2] = {(R)0, topo[2 * t + 1]};
        gy[t] = exp(-sqd_topo(tt, tc) * inv_2s2);  // fl(dy^2) in the topography's type
    }
    __syncthreads();
    for (int e = threadIdx.x; e < m * nv; e += blockDim.x) {
        double acc = 0.0;
        for (int bc = 0; bc < n; bc++) acc = __fma_rn(gy[bc], T[((int64_t)(bc * np + pj) * m) * nv + e], acc);
        num[e] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < m * dim; e += blockDim.x) {
        const int aj = e / dim, k = e % dim;
        const double den = num[aj * nv + dim];
        const R v = den > SOM_FLOAT_EPS ? (R)(num[aj * nv + k] / den) : (R)0;
        const int64_t o = ((int64_t)aj * n + bj) * dim + k;
        w[o] = v;
        if (hist) hist[o] = v;
    }
}

// w_j = num_j / den_j if den_j > 1e-12 else 0 (som_batch.py:354-359); num/den = fixed-order sums of
// the split partials (or the accumulator itself when nsplit == 1 and part == acc).
template <class R>
__global__ void divide_kernel(const double *__restrict__ part, int nsplit, int mn, int dim, R *__restrict__ w,
                              R *__restrict__ hist)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)mn * dim) return;
    const int j = (int)(e / dim), k = (int)(e % dim);
    const int nv = dim + 1;
    double num = 0.0, den = 0.0;
    for (int p = 0; p < nsplit; p++) {
        num += part[((int64_t)p * mn + j) * nv + k];
        den += part[((int64_t)p * mn + j) * nv + dim];
    }
    const R v = den > SOM_FLOAT_EPS ? (R)(num / den) : (R)0;
    w[e] = v;
    if (hist) hist[e] = v;
}

// ------------------------------------------------------------------------------------------------
// Vector-count mode: part[p][j][k] = sum_{i in split p} rho_i exp(-|t_j - t_bmu_i|^2 / (2 sigma_i^2)) x_ik
// with sigma_i = sigma0 / (1 + 2 (cur + i) / n_vectors) >= 1e-12 (som_batch.py:154-210).
// ------------------------------------------------------------------------------------------------
template <int DP, class T>
__global__ __launch_bounds__(256) void iter_partial_kernel(const int32_t *__restrict__ bmu, const T *__restrict__ x,
                                                           int dim, const double *__restrict__ rho, int64_t n,
                                                           int64_t cur, int64_t nvec, T sigma0,
                                                           const T *__restrict__ topo, int mn, int64_t chunk,
                                                           double *__restrict__ part)
{
    constexpr int TS = 64;
    __shared__ T s_x[TS * DP];
    __shared__ T s_tb[TS * 2];
    __shared__ double s_h0[TS];   // rho_i (rounded to the data type, as the reference casts density to dtype)
    __shared__ double s_inv[TS];  // 1 / (2 sigma_i^2)
    __shared__ int s_ok[TS];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    const int64_t ibeg = (int64_t)p * chunk, iend = min(n, ibeg + chunk);
    T tj[2] = {(T)0, (T)0};
    if (j < mn) { tj[0] = topo[2 * j]; tj[1] = topo[2 * j + 1]; }
    double acc[DP + 1];
#pragma unroll
    for (int k = 0; k <= DP; k++) acc[k] = 0.0;
    for (int64_t i0 = ibeg; i0 < iend; i0 += TS) {
        const int il = (int)min((int64_t)TS, iend - i0);
        __syncthreads();
        for (int e = threadIdx.x; e < il * DP; e += blockDim.x) {
            const int ii = e / DP, k = e % DP;
            s_x[e] = k < dim ? x[(i0 + ii) * dim + k] : (T)0;
        }
        for (int ii = threadIdx.x; ii < il; ii += blockDim.x) {
            const int64_t i = i0 + ii;
            const int c = bmu[i];
            const T r = rho ? (T)rho[i] : (T)1;
            const bool ok = c >= 0 && r > (T)0;
            s_ok[ii] = ok;
            s_tb[2 * ii] = ok ? topo[2 * c] : (T)0;
            s_tb[2 * ii + 1] = ok ? topo[2 * c + 1] : (T)0;
            double sg = (double)sigma0 * (1.0 / (1.0 + 2.0 * (double)(cur + i) / (double)nvec));
            if (sg < SOM_FLOAT_EPS) sg = SOM_FLOAT_EPS;
            s_inv[ii] = 1.0 / (2.0 * sg * sg);
            s_h0[ii] = (double)r;
        }
        __syncthreads();
        for (int ii = 0; ii < il; ii++) {
            if (!s_ok[ii]) continue;
            const double h = s_h0[ii] * exp(-sqd_topo(tj, &s_tb[2 * ii]) * s_inv[ii]);
#pragma unroll
            for (int k = 0; k < DP; k++)
                if (k < dim) acc[k] = __fma_rn(h, (double)s_x[ii * DP + k], acc[k]);
            acc[DP] += h;
        }
    }
    if (j < mn) {
        const int nv = dim + 1;
#pragma unroll
        for (int k = 0; k < DP; k++)
            if (k < dim) part[((int64_t)p * mn + j) * nv + k] = acc[k];
        part[((int64_t)p * mn + j) * nv + dim] = acc[DP];
    }
}

// acc[e] += sum_p part[p][e] (fixed order)
__global__ void sum_parts_kernel(const double *__restrict__ part, int nsplit, int64_t width, double *__restrict__ acc)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= width) return;
    double s = 0.0;
    for (int p = 0; p < nsplit; p++) s += part[(int64_t)p * width + e];
    acc[e] += s;
}

// ------------------------------------------------------------------------------------------------
// Errors (som_abstract.py:815-878): per-workgroup fixed-order sums, then one fixed-order final sum.
// ------------------------------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void errors_partial_kernel(const int32_t *__restrict__ c1, const int32_t *__restrict__ c2,
                                                             const double *__restrict__ d1, const double *__restrict__ rho,
                                                             int64_t n, const T *__restrict__ topo, double penalty,
                                                             int64_t per_wg, double *__restrict__ part)
{
    __shared__ double red[4][3];
    const int64_t i0 = (int64_t)blockIdx.x * per_wg, i1 = min(n, i0 + per_wg);
    double q = 0.0, t = 0.0, s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const double r = rho ? rho[i] : 1.0;
        if (!(r > 0.0)) continue;
        const int a = c1[i], b = c2[i];
        if (a < 0 || b < 0 || a == b) continue;
        q += r * sqrt(d1[i]);
        if (sqd_topo(&topo[2 * a], &topo[2 * b]) > penalty) t += r;
        s += r;
    }
    // fixed-order reduction: butterfly within each wave, then the 4 waves in order
    for (int o = 32; o > 0; o >>= 1) {
        q += __shfl_xor(q, o, 64);
        t += __shfl_xor(t, o, 64);
        s += __shfl_xor(s, o, 64);
    }
    const int wv = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) { red[wv][0] = q; red[wv][1] = t; red[wv][2] = s; }
    __syncthreads();
    if (threadIdx.x < 3)
        part[(int64_t)blockIdx.x * 3 + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

__global__ __launch_bounds__(256) void errors_final_kernel(const double *__restrict__ part, int nparts,
                                                           double *__restrict__ errors)
{
    __shared__ double red[3][256];
    double v[3] = {0.0, 0.0, 0.0};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x)
        for (int k = 0; k < 3; k++) v[k] += part[(int64_t)p * 3 + k];
    for (int k = 0; k < 3; k++) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) errors[threadIdx.x] += red[threadIdx.x][0];
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Host launchers used by capi.hip.
// ------------------------------------------------------------------------------------------------
template <class T>
int somhip_launch_absmax(hipStream_t st, const T *x, int dim, int dpad, const double *rho, int rho_f32, int64_t n,
                         unsigned long long *bounds, int num_cus)
{
    SOM_HIP_TRY(hipMemsetAsync(bounds, 0, sizeof(unsigned long long) * (dim + 1), st));
    if (n <= 0) return SOMHIP_OK;
    const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)num_cus * 8));
    dim3 g((unsigned)blocks), b(256);
    switch (dpad) {
        case 16: hipLaunchKernelGGL((absmax_kernel<16, T>), g, b, 0, st, x, dim, rho, rho_f32, n, bounds); break;
        case 32: hipLaunchKernelGGL((absmax_kernel<32, T>), g, b, 0, st, x, dim, rho, rho_f32, n, bounds); break;
        case 48: hipLaunchKernelGGL((absmax_kernel<48, T>), g, b, 0, st, x, dim, rho, rho_f32, n, bounds); break;
        default: hipLaunchKernelGGL((absmax_kernel<64, T>), g, b, 0, st, x, dim, rho, rho_f32, n, bounds); break;
    }
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

int somhip_launch_shifts(hipStream_t st, const unsigned long long *bounds, int dim, int has_x, int has_rho, int64_t n,
                         int *shifts)
{
    hipLaunchKernelGGL(shifts_kernel, dim3(1), dim3(128), 0, st, bounds, dim, has_x, has_rho, n, shifts);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

int somhip_launch_shifts_parts(hipStream_t st, const float *parts, int64_t rows, int dpad, int dim,
                               const unsigned long long *rho_bound, int64_t n, int *shifts)
{
    // first stage: fold the rows in place into the first RB rows (RB x dpad maxima), then one block
    constexpr int RB = 2048;
    if (rows > RB) {
        hipLaunchKernelGGL(rowmax_fold_kernel, dim3(RB), dim3(256), 0, st, (float *)parts, rows, dpad, RB);
        SOM_HIP_TRY(hipGetLastError());
        rows = RB;
    }
    hipLaunchKernelGGL(shifts_parts_kernel, dim3(1), dim3(1024), 0, st, parts, rows, dpad, dim, rho_bound, n, shifts);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

// Number of scatter workgroups so that the partial table stays within `table_budget` bytes.
int somhip_scatter_nwg(int64_t n, int mn, int nv, int num_cus, size_t table_budget)
{
    int64_t by_budget = (int64_t)(table_budget / ((size_t)mn * nv * 8));
    int64_t by_n = (n + 2047) / 2048;
    int64_t w = std::min<int64_t>(std::min<int64_t>(by_budget, by_n), (int64_t)num_cus);
    return (int)std::max<int64_t>(1, w);
}

template <class T>
int somhip_launch_scatter(hipStream_t st, const int32_t *bmu, const T *x, int dim, const double *rho, int rho_f32,
                          int64_t n, const int *shifts, int mn, int nv, int nwg, unsigned long long *partial,
                          double *acc, size_t lds_limit, unsigned int *bucket)
{
    if (n <= 0) return SOMHIP_OK;
    constexpr size_t queue_bytes = (size_t)(SCATTER_NT / 64) * 128 * (4 + 2);
    const size_t lds_bytes = std::min(LDS_SCATTER_MAX, lds_limit) - queue_bytes;
    int slice_len = std::min((int)(lds_bytes / ((size_t)nv * 8)), 65536);
    if (slice_len < 1) {
        somhip_set_error("LDS too small for one neuron row");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    const bool several_slices = slice_len < mn;
    if (slice_len > mn) slice_len = mn;
    const int64_t per_wg = (n + nwg - 1) / nwg;
    if (per_wg > 0xffffffffll) {
        somhip_set_error("too many samples per scatter workgroup");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    const int dpad = (dim + 15) / 16 * 16;
    constexpr int vec = 16 / (int)sizeof(T);  // values per 16-byte load
    const bool xvec = x && dim % vec == 0 && ((uintptr_t)x & 15) == 0;
    // several slices: the bucketed kernel, with a power-of-two slice that fits beside its static arrays
    int log2_slice = 0, nslices = 0;
    if (bucket && several_slices && per_wg < (1ll << BUCKET_LOCAL_BITS)) {
        const size_t avail = std::min(LDS_SCATTER_MAX, lds_limit) - sizeof(unsigned int) * (2 * BUCKET_MAX_SLICES + 1);
        while (log2_slice < BUCKET_NEURON_BITS && ((size_t)2 << log2_slice) * nv * 8 <= avail) log2_slice++;
        nslices = (mn + (1 << log2_slice) - 1) >> log2_slice;
        if (((size_t)1 << log2_slice) * nv * 8 > avail || nslices > BUCKET_MAX_SLICES) nslices = 0;
    }
    if (nslices > 0) {
        const size_t lds = ((size_t)1 << log2_slice) * nv * 8;
#define SOM_SB(D)                                                                                              \
    do {                                                                                                       \
        auto kern = xvec ? scatter_bucket_kernel<D, true, T> : scatter_bucket_kernel<D, false, T>;             \
        SOM_HIP_TRY(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
        hipLaunchKernelGGL(kern, dim3(nwg), dim3(SCATTER_NT), lds, st, bmu, x, dim, rho, rho_f32, n, per_wg,     \
                           shifts, mn, nv, log2_slice, nslices, bucket, partial);                               \
    } while (0)
#define SOM_SS(D)                                                                                              \
    do {                                                                                                       \
        auto kern = xvec ? scatter_sorted_kernel<D, true, T> : scatter_sorted_kernel<D, false, T>;             \
        SOM_HIP_TRY(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
        hipLaunchKernelGGL(kern, dim3(nwg), dim3(SCATTER_NT), lds, st, bmu, x, dim, rho, rho_f32, n, per_wg,     \
                           shifts, mn, nv, log2_slice, nslices, bucket, partial);                               \
    } while (0)
        // The sorted form needs the mn counters in the table's LDS and registers for the partial sums (at
        // D = 32 only with FP32 rows in vector loads, which leave room for them). It pays when each slice
        // pass reads a small share of the range's rows (100x100/D32, 20 slices: scatter 4.79 -> 3.27 ms
        // per epoch); with few slices the per-sample form, reading in sample order, keeps more DRAM
        // locality than the neuron order (64x64/D16, 4 slices: 0.47 against 0.53 ms).
        const bool sorted = nslices >= 8 && (size_t)mn * sizeof(unsigned int) <= lds &&
                            (!x || dpad <= 16 || (dpad == 32 && xvec && sizeof(T) == sizeof(float)));
        if (sorted) {
            if (!x || dpad <= 16) SOM_SS(16);
            else SOM_SS(32);
        } else if (!x || dpad <= 16) SOM_SB(16);
        else if (dpad == 32) SOM_SB(32);
        else if (dpad == 48) SOM_SB(48);
        else SOM_SB(64);
#undef SOM_SS
#undef SOM_SB
    } else {
        const size_t lds = (size_t)slice_len * nv * 8 + queue_bytes;
#define SOM_SC(D)                                                                                              \
    do {                                                                                                       \
        auto kern = xvec ? scatter_fixed_kernel<D, true, T> : scatter_fixed_kernel<D, false, T>;               \
        SOM_HIP_TRY(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
        hipLaunchKernelGGL(kern, dim3(nwg), dim3(SCATTER_NT), lds, st, bmu, x, dim, rho, rho_f32, n, per_wg,     \
                           shifts, mn, nv, slice_len, partial);                                                \
    } while (0)
        if (!x || dpad <= 16) SOM_SC(16);
        else if (dpad == 32) SOM_SC(32);
        else if (dpad == 48) SOM_SC(48);
        else SOM_SC(64);
#undef SOM_SC
    }
    SOM_HIP_TRY(hipGetLastError());
    const int64_t width = (int64_t)mn * nv;
    hipLaunchKernelGGL(scatter_reduce_kernel, dim3((unsigned)((width + 255) / 256)), dim3(256), 0, st, partial, nwg,
                       width, nv, shifts, acc);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

int somhip_update_nsplit(int mn)
{
    // enough (neuron, split) threads to fill the device: about 256K
    int s = (int)std::max<int64_t>(1, (int64_t)262144 / std::max(mn, 1));
    return std::min(s, std::max(1, mn / 16));
}

template <class T>
int somhip_launch_update(hipStream_t st, const double *S, const T *topo, int mn, int dim, double sigma,
                         double *part, int nsplit, T *w, T *hist)
{
    const int nv = dim + 1;
    const double inv_2s2 = 1.0 / (2.0 * sigma * sigma);
    const int chunk = (mn + nsplit - 1) / nsplit;
    dim3 g((mn + 255) / 256, nsplit), b(256);
    if (nv <= 17) hipLaunchKernelGGL((update_partial_kernel<17, T>), g, b, 0, st, S, topo, mn, nv, inv_2s2, chunk, part);
    else if (nv <= 33) hipLaunchKernelGGL((update_partial_kernel<33, T>), g, b, 0, st, S, topo, mn, nv, inv_2s2, chunk, part);
    else if (nv <= 49) hipLaunchKernelGGL((update_partial_kernel<49, T>), g, b, 0, st, S, topo, mn, nv, inv_2s2, chunk, part);
    else hipLaunchKernelGGL((update_partial_kernel<65, T>), g, b, 0, st, S, topo, mn, nv, inv_2s2, chunk, part);
    SOM_HIP_TRY(hipGetLastError());
    const int64_t tot = (int64_t)mn * dim;
    hipLaunchKernelGGL(divide_kernel<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part, nsplit, mn, dim, w,
                       hist);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

bool somhip_update_separable_fits(int m, int n, int dim)
{
    const size_t nv = (size_t)dim + 1;
    return sizeof(double) * ((size_t)m * nv + 2 * (size_t)m) <= 65536 &&
           sizeof(double) * ((size_t)n + (size_t)m * nv) <= 65536;
}

template <class R>
int somhip_launch_update_separable(hipStream_t st, const double *S, const R *topo, int m, int n, int hex, int dim,
                                   double sigma, double *T, R *w, R *hist)
{
    const int nv = dim + 1;
    const double inv_2s2 = 1.0 / (2.0 * sigma * sigma);
    const int np = hex ? 2 : 1;
    const size_t lds1 = sizeof(double) * ((size_t)m * nv + 2 * m - 1);
    const size_t lds2 = sizeof(double) * ((size_t)n + (size_t)m * nv);
    if (!somhip_update_separable_fits(m, n, dim)) {
        somhip_set_error("separable update: grid too large for one LDS column");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(update_sep_rows_kernel, dim3(n * np), dim3(256), lds1, st, S, m, n, nv, hex, inv_2s2, T);
    SOM_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(update_sep_cols_kernel<R>, dim3(n), dim3(256), lds2, st, T, topo, m, n, dim, hex, inv_2s2, w, hist);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

template <class T>
int somhip_launch_divide(hipStream_t st, const double *acc, int mn, int dim, T *w, T *hist)
{
    const int64_t tot = (int64_t)mn * dim;
    hipLaunchKernelGGL(divide_kernel<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, acc, 1, mn, dim, w, hist);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

int somhip_iter_nsplit(int64_t n, int mn)
{
    int64_t s = std::max<int64_t>(1, (int64_t)262144 / std::max(mn, 1));
    s = std::min<int64_t>(s, std::max<int64_t>(1, n / 64));
    return (int)std::min<int64_t>(s, 4096);
}

template <class T>
int somhip_launch_iter(hipStream_t st, const int32_t *bmu, const T *x, int dim, int dpad, const double *rho,
                       int64_t n, int64_t cur, int64_t nvec, T sigma0, const T *topo, int mn, double *part,
                       int nsplit, double *acc)
{
    if (n <= 0) return SOMHIP_OK;
    const int64_t chunk = (n + nsplit - 1) / nsplit;
    dim3 g((mn + 255) / 256, nsplit), b(256);
#define SOM_IT(D) hipLaunchKernelGGL((iter_partial_kernel<D, T>), g, b, 0, st, bmu, x, dim, rho, n, cur, nvec, sigma0, topo, mn, chunk, part)
    switch (dpad) {
        case 16: SOM_IT(16); break;
        case 32: SOM_IT(32); break;
        case 48: SOM_IT(48); break;
        default: SOM_IT(64); break;
    }
#undef SOM_IT
    SOM_HIP_TRY(hipGetLastError());
    const int64_t width = (int64_t)mn * (dim + 1);
    hipLaunchKernelGGL(sum_parts_kernel, dim3((unsigned)((width + 255) / 256)), dim3(256), 0, st, part, nsplit, width,
                       acc);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

// about 1024 samples (4 per thread) per workgroup: enough workgroups to keep the loads in flight
int somhip_errors_nparts(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(65536, (n + 1023) / 1024)); }

template <class T>
int somhip_launch_errors(hipStream_t st, const int32_t *c1, const int32_t *c2, const double *d1, const double *rho,
                         int64_t n, const T *topo, double penalty, double *part, int nparts, double *errors)
{
    if (n <= 0) return SOMHIP_OK;
    const int64_t per_wg = (n + nparts - 1) / nparts;
    hipLaunchKernelGGL(errors_partial_kernel<T>, dim3(nparts), dim3(256), 0, st, c1, c2, d1, rho, n, topo, penalty, per_wg,
                       part);
    SOM_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(errors_final_kernel, dim3(1), dim3(256), 0, st, part, nparts, errors);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

// FP32 (float32 SOMs) and FP64 (float64 SOMs) forms of the templated launchers
#define SOM_TRAIN_LAUNCHERS(T)                                                                                  \
    template int somhip_launch_absmax<T>(hipStream_t, const T *, int, int, const double *, int, int64_t,        \
                                         unsigned long long *, int);                                            \
    template int somhip_launch_scatter<T>(hipStream_t, const int32_t *, const T *, int, const double *, int,    \
                                          int64_t, const int *, int, int, int, unsigned long long *, double *,  \
                                          size_t, unsigned int *);                                              \
    template int somhip_launch_update<T>(hipStream_t, const double *, const T *, int, int, double, double *,    \
                                         int, T *, T *);                                                        \
    template int somhip_launch_update_separable<T>(hipStream_t, const double *, const T *, int, int, int, int,  \
                                                   double, double *, T *, T *);                                 \
    template int somhip_launch_divide<T>(hipStream_t, const double *, int, int, T *, T *);                      \
    template int somhip_launch_iter<T>(hipStream_t, const int32_t *, const T *, int, int, const double *,       \
                                       int64_t, int64_t, int64_t, T, const T *, int, double *, int, double *);  \
    template int somhip_launch_errors<T>(hipStream_t, const int32_t *, const int32_t *, const double *,         \
                                         const double *, int64_t, const T *, double, double *, int, double *);
SOM_TRAIN_LAUNCHERS(float)
SOM_TRAIN_LAUNCHERS(double)
#undef SOM_TRAIN_LAUNCHERS
