// Best-matching-unit search kernels for gfx950.
//
// Three precisions, one output contract: for every sample i the reference's top-2 BMUs
// (decontamination/algo/som_abstract.py:815-878 update rule = lexicographic (distance, index)),
// written as c1[i], c2[i] (int32, -1 when absent; c1 falls back to `default_index`) and d1[i]
// (FP64 squared distance of c1, NaN when absent).
//
//   EXACT       bmu_exact_kernel: one sample per lane, weights streamed through LDS, the exact
//               distance formula (reference FP32-difference/FP64-sum, or FP64) for every neuron.
//               Also the fallback for samples the MFMA path flags.
//   FP32        bmu_fp32_kernel: |w'|^2 - 2 x'.w' on v_mfma_f32_32x32x2_f32 (x', w' centred on the
//               weight centroid), top-2 on those values (the plain FP32 form).
//   BF16_RERANK bmu_bf16_kernel: the same GEMM on v_mfma_f32_32x32x16_bf16 with a 3-term split
//               (x_h w_h + x_l w_h + x_h w_l). Phase A scans half of a coarse sub-lattice of the
//               map for the top-2 of per-lane block minima; phase B scans every other neuron,
//               recording every neuron inside a provable error window of the running second minimum
//               while it tightens the window; phase C records phase A's neurons under the final window.
//               Then the exact distance formula on the recorded set. Samples whose list overflows,
//               or whose bound is not finite, go to the EXACT fallback.
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace {

constexpr int WAVE = 64;
constexpr int BMU_WAVES = 4;                 // waves per workgroup of the MFMA kernels
constexpr float U32 = 5.9604644775390625e-8f;  // 2^-24

// Pointers in the global address space. The BF16 body reads everything through these, so its accesses
// are global loads and stores also when its pointers come from descriptors in memory (the serving form),
// which the compiler would otherwise treat as generic and access with flat instructions.
template <class T>
using gptr = T __attribute__((address_space(1))) *;

// p itself, or p in the global address space (GC); kernels whose pointers are kernel arguments keep
// them as they are, since the compiler already knows where those point and assigns their registers best
template <bool GC, class T>
__device__ __forceinline__ auto aptr(T *p)
{
    if constexpr (GC) return (gptr<T>)p;
    else return p;
}
__device__ __forceinline__ const f32x4 *as_f32x4(const float *p) { return (const f32x4 *)p; }
__device__ __forceinline__ gptr<const f32x4> as_f32x4(gptr<const float> p) { return (gptr<const f32x4>)p; }

// Each lane needs its partner (lane ^ 32) value: both hold the same sample column.
__device__ __forceinline__ float xchg32(float v) { return __shfl_xor(v, 32, WAVE); }
__device__ __forceinline__ int xchg32i(int v) { return __shfl_xor(v, 32, WAVE); }
__device__ __forceinline__ double xchg32d(double v) { return __shfl_xor(v, 32, WAVE); }

// ------------------------------------------------------------------------------------------------
// EXACT: direct-difference scan. Grid: one thread per listed sample. T = float: the reference's
// FP32-difference formula (MODE REF32) or FP64 differences of the FP32 values (MODE FP64); T = double
// (float64 SOMs): FP64 differences of FP64 data and weights, no |x| maxima (xmax_part must be null).
// ------------------------------------------------------------------------------------------------
template <int DP, int MODE, bool NEED_C2, class T = float>
__global__ __launch_bounds__(256) void bmu_exact_kernel(const T *__restrict__ x, int64_t n, int dim,
                                                        const int64_t *__restrict__ list,
                                                        const unsigned int *__restrict__ list_count,
                                                        const T *__restrict__ w, int mn, int default_index,
                                                        int32_t *__restrict__ c1, int32_t *__restrict__ c2,
                                                        double *__restrict__ d1, float *__restrict__ xmax_part)
{
    constexpr bool F64 = sizeof(T) == sizeof(double);
    constexpr int CH = F64 ? 64 : 128;  // neurons per LDS chunk
    __shared__ T ws[CH * DP];
    const int64_t total = list ? (int64_t)(*list_count) : n;
    float amax[DP];
#pragma unroll
    for (int k = 0; k < DP; k++) amax[k] = 0.f;
    // Grid-stride over samples; every thread of the block takes part in the chunk loads.
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < total; base += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = base + threadIdx.x;
        const bool active = t < total;
        const int64_t i = active ? (list ? list[t] : t) : 0;
        T xr[DP];
#pragma unroll
        for (int k = 0; k < DP; k++) {
            xr[k] = (active && k < dim) ? x[i * dim + k] : (T)0;
            if constexpr (!F64) {
                const float a = fabsf(xr[k]);
                amax[k] = fmaxf(amax[k], a < __builtin_inff() ? a : 0.f);
            }
        }
        double md1 = 1.0e99, md2 = 1.0e99;
        int mi1 = -1, mi2 = -1;
        for (int j0 = 0; j0 < mn; j0 += CH) {
            __syncthreads();
            for (int e = threadIdx.x; e < CH * DP; e += blockDim.x) {
                int jj = j0 + e / DP, k = e % DP;
                ws[e] = (jj < mn && k < dim) ? w[(int64_t)jj * dim + k] : (T)0;
            }
            __syncthreads();
            const int jn = min(CH, mn - j0);
            for (int jj = 0; jj < jn; jj++) {
                double d;  // zero padding adds exact zeros
                if constexpr (F64) d = sqd_f64(&ws[jj * DP], xr, DP);
                else d = sqd_mode<MODE>(&ws[jj * DP], xr, DP);
                if (d != d) continue;
                const int j = j0 + jj;
                if (NEED_C2) {
                    if (md1 > d) {
                        md2 = md1; mi2 = mi1; md1 = d; mi1 = j;
                    } else if (md2 > d) {
                        md2 = d; mi2 = j;
                    }
                } else if (md1 > d) {
                    md1 = d; mi1 = j;
                }
            }
        }
        if (active) {
            c1[i] = mi1 >= 0 ? mi1 : default_index;
            if (c2) c2[i] = mi2;
            if (d1) d1[i] = mi1 >= 0 ? md1 : __builtin_nan("");
        }
    }
    if (!F64 && xmax_part) {
        const int lane = threadIdx.x & (WAVE - 1);
#pragma unroll
        for (int k = 0; k < DP; k++) {
            float v = amax[k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
            amax[k] = v;
        }
        if (lane == 0) {
            const int64_t wid = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
#pragma unroll
            for (int k = 0; k < DP; k++) xmax_part[wid * DP + k] = amax[k];
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Per-sample preparation shared by the MFMA kernels: load this lane's 8 dims (per 16-dim chunk) of
// the sample, centre them, and compute the quantities of the error bound.
// ------------------------------------------------------------------------------------------------
template <int KC>
struct SampleLoad {
    float xc[KC][8];  // centred FP32 values of dims kc*16 + 8*(lane/32) + e
    float ax[KC][8];  // |x| of the same dims (0 when not finite), for the fixed-point scales
};

template <int KC, class XP, class CP>
__device__ __forceinline__ void load_sample(XP x, int64_t n, int dim, int64_t s, int lane, CP centre, SampleLoad<KC> &L)
{
    const int hb = 8 * (lane >> 5);
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = kc * 16 + hb + e;
            float v = (s < n && k < dim) ? x[s * dim + k] : 0.f;
            L.xc[kc][e] = (k < dim) ? __fsub_rn(v, centre[k]) : 0.f;
            const float a = fabsf(v);
            L.ax[kc][e] = a < __builtin_inff() ? a : 0.f;
        }
    }
}

// The same for rows of exactly KC * 16 floats at a 16-byte aligned base: two 16-byte loads per 8 dims,
// no per-dimension tests.
template <int KC, class XP, class CP>
__device__ __forceinline__ void load_sample_rows(XP x, int64_t n, int64_t s, int lane, CP centre, SampleLoad<KC> &L)
{
    const int hb = 8 * (lane >> 5);
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        f32x4 a = {0.f, 0.f, 0.f, 0.f}, b = {0.f, 0.f, 0.f, 0.f};
        if (s < n) {
            const auto row = as_f32x4(x + s * (KC * 16) + kc * 16 + hb);
            a = row[0];
            b = row[1];
        }
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const float v = e < 4 ? a[e] : b[e - 4];
            L.xc[kc][e] = __fsub_rn(v, centre[kc * 16 + hb + e]);
            const float av = fabsf(v);
            L.ax[kc][e] = av < __builtin_inff() ? av : 0.f;
        }
    }
}

// Per-wave max of |x| per dimension: lanes l and l^32 hold different dims, the 32 lanes of a half
// hold the same dims for different samples. Writes DP floats for this wave (plain stores).
template <int KC>
__device__ __forceinline__ void wave_dim_max(float (&amax)[KC][8], int lane, float *__restrict__ part)
{
#pragma unroll
    for (int kc = 0; kc < KC; kc++)
#pragma unroll
        for (int e = 0; e < 8; e++) {
            float v = amax[kc][e];
#pragma unroll
            for (int o = 16; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
            amax[kc][e] = v;
        }
    if ((lane & 31) == 0) {
        const int64_t wid = (int64_t)blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
#pragma unroll
        for (int kc = 0; kc < KC; kc++)
#pragma unroll
            for (int e = 0; e < 8; e++) part[wid * (KC * 16) + kc * 16 + 8 * (lane >> 5) + e] = amax[kc][e];
    }
}

// ------------------------------------------------------------------------------------------------
// FP32 plain form. Wave = G groups of 32 samples; block = 32 neurons (natural order).
// ------------------------------------------------------------------------------------------------
template <int DP, int G>
__global__ __launch_bounds__(BMU_WAVES * WAVE) void bmu_fp32_kernel(
    const float *__restrict__ x, int64_t n, int dim, const float *__restrict__ centre,
    const f32x4 *__restrict__ frag, const f32x4 *__restrict__ norms4, int nblk, int mn, int default_index,
    int32_t *__restrict__ c1, int32_t *__restrict__ c2, double *__restrict__ d1, float *__restrict__ xmax_part)
{
    constexpr int KS = DP / 2;  // MFMA k-steps of 2
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const int64_t sbase = ((int64_t)blockIdx.x * BMU_WAVES + wave) * (G * 32);
    const int col = lane & 31;
    const int khalf = lane >> 5;

    // B operand: b[g][s] = -2 * x'[sample][2s + khalf]
    float b[G][KS];
    float xn[G];  // |x'|^2 partial (this lane's dims), combined below
    float amax[KS];
#pragma unroll
    for (int st = 0; st < KS; st++) amax[st] = 0.f;
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int64_t s = sbase + g * 32 + col;
        float acc = 0.f;
#pragma unroll
        for (int st = 0; st < KS; st++) {
            const int k = 2 * st + khalf;
            const float raw = (s < n && k < dim) ? x[s * dim + k] : 0.f;
            const float a = fabsf(raw);
            amax[st] = fmaxf(amax[st], a < __builtin_inff() ? a : 0.f);
            float v = (k < dim) ? __fsub_rn(raw, centre[k]) : 0.f;
            b[g][st] = -2.f * v;
            acc = fmaf(v, v, acc);
        }
        xn[g] = acc + xchg32(acc);
    }
    if (xmax_part) {
#pragma unroll
        for (int st = 0; st < KS; st++) {
            float v = amax[st];
#pragma unroll
            for (int o = 16; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
            amax[st] = v;
        }
        if ((lane & 31) == 0) {
            const int64_t wid = (int64_t)blockIdx.x * BMU_WAVES + wave;
#pragma unroll
            for (int st = 0; st < KS; st++) xmax_part[wid * DP + 2 * st + khalf] = amax[st];
        }
    }

    float m1[G], m2[G];
    int i1[G], i2[G];
#pragma unroll
    for (int g = 0; g < G; g++) { m1[g] = m2[g] = __builtin_inff(); i1[g] = i2[g] = -1; }

    auto load_block = [&](int blk, f32x4 (&a4)[KS / 4], f32x16 &cin) {
#pragma unroll
        for (int q = 0; q < KS / 4; q++) a4[q] = frag[((int64_t)blk * (KS / 4) + q) * WAVE + lane];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            f32x4 t = norms4[blk * 8 + 2 * q + khalf];
            cin[4 * q + 0] = t[0]; cin[4 * q + 1] = t[1]; cin[4 * q + 2] = t[2]; cin[4 * q + 3] = t[3];
        }
    };
    f32x4 a4[KS / 4];
    f32x16 cin;
    if (nblk > 0) load_block(0, a4, cin);
    for (int blk = 0; blk < nblk; blk++) {
        // prefetch the next block's fragments while this one computes
        f32x4 na4[KS / 4];
        f32x16 ncin;
        if (blk + 1 < nblk) load_block(blk + 1, na4, ncin);
        const int jbase = blk * 32 + 4 * khalf;
#pragma unroll
        for (int g = 0; g < G; g++) {
            f32x16 acc = cin;
#pragma unroll
            for (int st = 0; st < KS; st++)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a4[st / 4][st % 4], b[g][st], acc, 0, 0, 0);
            // in-order top-2 with strict '<' (lowest index wins ties within the lane):
            // values by min/med3, indices by selects
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const float v = acc[r];
                const int j = jbase + 8 * (r >> 2) + (r & 3);
                const bool lt1 = v < m1[g];
                const bool lt2 = v < m2[g];
                const int t2 = lt2 ? j : i2[g];
                i2[g] = lt1 ? i1[g] : t2;
                i1[g] = lt1 ? j : i1[g];
                m2[g] = __builtin_amdgcn_fmed3f(m1[g], m2[g], v);
                m1[g] = fminf(m1[g], v);
            }
        }
        if (blk + 1 < nblk) {
#pragma unroll
            for (int q = 0; q < KS / 4; q++) a4[q] = na4[q];
            cin = ncin;
        }
    }

#pragma unroll
    for (int g = 0; g < G; g++) {
        // merge with the partner lane (lexicographic on (value, index))
        float pm1 = xchg32(m1[g]), pm2 = xchg32(m2[g]);
        int pi1 = xchg32i(i1[g]), pi2 = xchg32i(i2[g]);
        float a1 = m1[g], a2 = m2[g];
        int j1 = i1[g], j2 = i2[g];
        auto lessv = [](float da, int ia, float db, int ib) {
            return da < db || (da == db && (unsigned)ia < (unsigned)ib);
        };
        // insert partner's two candidates
        float cd[2] = {pm1, pm2};
        int ci[2] = {pi1, pi2};
#pragma unroll
        for (int t = 0; t < 2; t++) {
            if (ci[t] < 0) continue;
            if (lessv(cd[t], ci[t], a1, j1)) { a2 = a1; j2 = j1; a1 = cd[t]; j1 = ci[t]; }
            else if (lessv(cd[t], ci[t], a2, j2)) { a2 = cd[t]; j2 = ci[t]; }
        }
        const int64_t s = sbase + g * 32 + col;
        if (lane < 32 && s < n) {
            const bool ok1 = j1 >= 0 && j1 < mn;
            c1[s] = ok1 ? j1 : default_index;
            if (c2) c2[s] = (j2 >= 0 && j2 < mn) ? j2 : -1;
            if (d1) d1[s] = ok1 ? (double)fmaxf(a1 + xn[g], 0.f) : __builtin_nan("");
        }
    }
}

// ------------------------------------------------------------------------------------------------
// BF16 3-term split with candidate window and exact rerank.
//
// A workgroup of NW waves (NW*G*32 samples, one tile) streams the 32-neuron blocks through LDS in
// chunks of CB blocks (A fragments + norms), double-buffered: while the waves compute chunk c from one
// buffer, every thread holds its share of chunk c+1 in registers and writes it to the other buffer
// after its last block of chunk c; one barrier per chunk.
// ------------------------------------------------------------------------------------------------
constexpr int NW_BF16 = 8;      // waves per workgroup
constexpr int G_BF16 = 2;       // 32-sample groups per wave
constexpr int CHUNK_KB = 8;             // (32-neuron block) x (16-wide K slice) units per chunk
constexpr int FRAG_UNITS = CHUNK_KB * 2 * 64;  // 16-byte units of A fragments per chunk
static_assert(NW_BF16 * G_BF16 * 32 == SOM_TILE_ROWS, "tile rows must match the BF16 workgroup");

// Candidate lists. D = 16: 20 per lane; a lane past that sends its sample to the fallback.
// D = 32 (more neurons closer than the window, on large smooth maps): 16 per lane plus a pool of 128
// shared by the 32-sample group for lanes past that, in the same LDS; only a full pool sends a sample
// to the fallback. The pool's spill test costs a few percent at D = 16, where lists rarely fill.
using lds_u16 = unsigned short __attribute__((address_space(3)));
// LDS byte address of an object in LDS
template <class T>
__device__ __forceinline__ unsigned int lds_addr(T *p)
{
    return (unsigned int)(size_t)(T __attribute__((address_space(3))) *)p;
}
constexpr unsigned short CAND_EMPTY = 0xFFFF;  // slots are < 65535 (checked at launch)

template <int KC>
struct Bf16Cand {
    static constexpr int CAPL = KC == 1 ? 20 : 16;  // own list per lane
    static constexpr int POOL = KC == 1 ? 0 : 128;  // shared entries per group
};

template <int KC>
struct Bf16Smem {
    static constexpr int CB = CHUNK_KB / KC;  // blocks per chunk
    static constexpr int CAPL = Bf16Cand<KC>::CAPL;
    u32x4 frag[2][FRAG_UNITS];          // [buf][cb][kc][hi,lo][lane] 16 B
    float norms[2][CB * 32];            // [buf][cb][row]
    unsigned short cand[NW_BF16][G_BF16][CAPL + 1][WAVE];  // last row: writes past a full list
    unsigned int pool[NW_BF16][G_BF16][Bf16Cand<KC>::POOL > 0 ? Bf16Cand<KC>::POOL : 1];  // (lane << 16) | slot
    unsigned short rows[NW_BF16][G_BF16][32];  // tile row of each (wave, group, column), reloaded for the rerank
};

template <int DP, int G, int MODE, bool NEED_C2, bool GC>
__device__ __forceinline__ void bmu_bf16_body(const BmuModelDev &Md, const float *__restrict__ x_in, int64_t n,
                                              int64_t tile_row0, int default_index, int32_t *__restrict__ c1_out,
                                              int32_t *__restrict__ c2_out, double *__restrict__ d1_out,
                                              int64_t *__restrict__ fb_list_out, unsigned int *__restrict__ fb_count,
                                              int64_t fb_tag, Bf16Smem<DP / 16> &sm, float *__restrict__ xmax_part,
                                              const int32_t *__restrict__ prev_c1)
{
    constexpr int KC = DP / 16;
    constexpr int CB = Bf16Smem<KC>::CB;
    constexpr int CAPL = Bf16Cand<KC>::CAPL;
    constexpr int POOL = Bf16Cand<KC>::POOL;
    constexpr int NORM_UNITS = CB * 8;   // 16-byte units of norms per chunk
    constexpr int NT = NW_BF16 * WAVE;   // threads per workgroup
    constexpr int FU = FRAG_UNITS / NT;  // fragment units per thread
    static_assert(FRAG_UNITS % NT == 0 && NORM_UNITS <= NT, "chunk layout");
    const auto centre = aptr<GC>(Md.centre);
    const auto wmax = aptr<GC>(Md.wmax);
    const auto scal = aptr<GC>(Md.scal);
    // loads through a plain vector type: a class type (uint4) cannot be copied from a global-address-space
    // reference in the host compilation pass
    const auto gfrag = aptr<GC>((const u32x4 *)Md.frag);
    const auto gnorm = aptr<GC>((const u32x4 *)Md.norms4);
    const auto perm = aptr<GC>(Md.perm);
    const auto wslot = aptr<GC>(Md.wslot);
    const auto x = aptr<GC>(x_in);
    const auto c1 = aptr<GC>(c1_out);
    const auto c2 = aptr<GC>(c2_out);
    const auto d1 = aptr<GC>(d1_out);
    const auto fb_list = aptr<GC>(fb_list_out);
    const int nblk = Md.nblk, dim = Md.dim;
    // sample rows of exactly DP floats at a 16-byte aligned base are read with vector loads
    const bool rows_vec = dim == DP && ((uintptr_t)x_in & 15) == 0;
    const int tid = threadIdx.x;
    const int lane = tid & (WAVE - 1);
    const int wave = tid / WAVE;
    const int col = lane & 31;
    const int khalf = lane >> 5;
    // rl[g]: the tile row this lane's group g works on. Without earlier BMUs, slot order; with them, the
    // tile's rows counting-sorted by the 8x8 map patch (in Morton order) of their earlier BMU, so a
    // 32-sample group tends to share its nearest neurons and records candidates in fewer blocks. The
    // scratch is the fragment buffer, which the first chunk load overwrites after a barrier.
    constexpr int TR = NW_BF16 * G * 32;  // rows per tile
    constexpr int NBK = 65;               // 64 patches + rows past the end
    int rl[G];
#pragma unroll
    for (int g = 0; g < G; g++) rl[g] = wave * (G * 32) + g * 32 + col;
    if (prev_c1) {
        unsigned int *bcount = (unsigned int *)&sm.frag[0][0];
        unsigned int *bcur = bcount + NBK;
        unsigned short *ord = (unsigned short *)(bcur + NBK);
        for (int b = tid; b < NBK; b += NT) bcount[b] = 0u;
        __syncthreads();
        const int mcol = max(Md.ncol, 1), mrow = max(Md.mn / mcol, 1);
        int key[(TR + NT - 1) / NT];
#pragma unroll
        for (int u = 0; u < (TR + NT - 1) / NT; u++) {
            const int r = tid + u * NT;
            const int64_t row = tile_row0 + r;
            int k = NBK - 1;
            if (r < TR && row < n) {
                const int c = min(max((int)aptr<GC>(prev_c1)[row], 0), Md.mn - 1);
                const int a = c / mcol, b = c - a * mcol;
                const int pa = a * 8 / mrow, pb = b * 8 / mcol;
                k = 0;
#pragma unroll
                for (int bit = 0; bit < 3; bit++) k |= (((pa >> bit) & 1) << (2 * bit + 1)) | (((pb >> bit) & 1) << (2 * bit));
            }
            key[u] = k;
            if (r < TR) atomicAdd(&bcount[k], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            unsigned int acc = 0u;
            for (int b = 0; b < NBK; b++) {
                bcur[b] = acc;
                acc += bcount[b];
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < (TR + NT - 1) / NT; u++) {
            const int r = tid + u * NT;
            if (r < TR) ord[atomicAdd(&bcur[key[u]], 1u)] = (unsigned short)r;
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; g++) rl[g] = ord[rl[g]];
    }
    // kept in LDS rather than in registers across the block loop (both lanes of a column write the same)
#pragma unroll
    for (int g = 0; g < G; g++) sm.rows[wave][g][col] = (unsigned short)rl[g];
    unsigned short(*cw)[CAPL + 1][WAVE] = sm.cand[wave];
    const int nchunks = (nblk + CB - 1) / CB;
    const int64_t frag_units_total = (int64_t)nblk * KC * 2 * 64;
    const int64_t norm_units_total = (int64_t)nblk * 8;

    const float nwmax = scal[0];
    // -inf as a value the compiler cannot see, so v_med3(m1, t, -inf) is not folded back into a minimum
    // that needs canonicalising moves in IEEE mode
    float neg_inf = -__builtin_inff();
    asm volatile("" : "+v"(neg_inf));
    bf16x8 bh[G][KC], bl[G][KC];
    float E[G], K[G];
    float amax[KC][8];
#pragma unroll
    for (int kc = 0; kc < KC; kc++)
#pragma unroll
        for (int e = 0; e < 8; e++) amax[kc][e] = 0.f;
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int64_t s = tile_row0 + rl[g];
        SampleLoad<KC> L;
        if (rows_vec)
            load_sample_rows<KC>(x, n, s, lane, centre, L);
        else
            load_sample<KC>(x, n, dim, s, lane, centre, L);
#pragma unroll
        for (int kc = 0; kc < KC; kc++)
#pragma unroll
            for (int e = 0; e < 8; e++) amax[kc][e] = fmaxf(amax[kc][e], L.ax[kc][e]);
        float sx = 0.f, kk = 0.f;
#pragma unroll
        for (int kc = 0; kc < KC; kc++) {
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const float v = L.xc[kc][e];
                // round-to-nearest-even hi/lo split by the hardware conversion; the weight fragments
                // carry the factor -2
                const __bf16 h = (__bf16)v;
                bh[g][kc][e] = h;
                bl[g][kc][e] = (__bf16)(v - (float)h);
                sx = fmaf(fabsf(v), wmax[kc * 16 + 8 * khalf + e], sx);
                kk = fmaf(v, v, kk);
            }
        }
        sx += xchg32(sx);
        kk += xchg32(kk);
        // Error bound of a_j = |w'_j|^2 - 2 x'.w'_j against its exact value (derivation in the file
        // header of the design notes): split truncation 6.04*2^-16*Sx, MFMA FP32 accumulation
        // (3*DP+6)*2^-24*(nwmax + 2.05*Sx), centring 4*2^-24*(K + nwmax); x1.25 safety.
        const float split = 6.04f * 1.52587890625e-5f * sx;
        const float accum = (3.f * DP + 6.f) * U32 * (nwmax + 2.05f * sx);
        const float cent = 4.f * U32 * (kk + nwmax);
        E[g] = 1.25f * (split + accum + cent) + 1e-30f;
        K[g] = kk * (1.f + (DP + 2) * U32);
    }
    if (xmax_part) wave_dim_max<KC>(amax, lane, xmax_part);

    // chunk streaming ---------------------------------------------------------------------------
    u32x4 pf[FU];
    u32x4 pn = {0u, 0u, 0u, 0u};
    auto load_chunk = [&](int c) {
        const int64_t fb = (int64_t)c * FRAG_UNITS;
#pragma unroll
        for (int u = 0; u < FU; u++) {
            const int64_t idx = fb + tid + (int64_t)u * NT;
            pf[u] = idx < frag_units_total ? gfrag[idx] : u32x4{0u, 0u, 0u, 0u};
        }
        const int64_t nidx = (int64_t)c * NORM_UNITS + tid;
        if (tid < NORM_UNITS && nidx < norm_units_total) pn = gnorm[nidx];
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int u = 0; u < FU; u++) sm.frag[buf][tid + u * NT] = pf[u];
        if (tid < NORM_UNITS) ((u32x4 *)sm.norms[buf])[tid] = pn;
    };
    auto read_block = [&](int buf, int cb, bf16x8 (&ah)[KC], bf16x8 (&al)[KC], f32x16 &cin) {
#pragma unroll
        for (int kc = 0; kc < KC; kc++) {
            ah[kc] = __builtin_bit_cast(bf16x8, sm.frag[buf][((cb * KC + kc) * 2 + 0) * 64 + lane]);
            al[kc] = __builtin_bit_cast(bf16x8, sm.frag[buf][((cb * KC + kc) * 2 + 1) * 64 + lane]);
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const f32x4 t = *(const f32x4 *)&sm.norms[buf][cb * 32 + 8 * q + 4 * khalf];
            cin[4 * q + 0] = t[0]; cin[4 * q + 1] = t[1]; cin[4 * q + 2] = t[2]; cin[4 * q + 3] = t[3];
        }
    };
    auto block_acc = [&](int g, const bf16x8 (&ah)[KC], const bf16x8 (&al)[KC], const f32x16 &cin) {
        f32x16 acc = cin;
#pragma unroll
        for (int kc = 0; kc < KC; kc++) {
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[kc], bh[g][kc], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[kc], bl[g][kc], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[kc], bh[g][kc], acc, 0, 0, 0);
        }
        return acc;
    };
    // Streams chunks [c0, c1) through the LDS buffers; per_block(blk, acc) sees every block's D values,
    // after_chunk() runs once per chunk, after the chunk's closing barrier.
    // The buffer in use alternates across phases too; a phase's last chunk prefetches the first chunk
    // of the next phase (next_c0 >= 0), so only the first phase waits for its first load.
    int buf = 0;
    bool preloaded = false;
    auto stream = [&](int c0, int c1, int next_c0, auto &&per_block, auto &&after_chunk) {
        if (!preloaded) {
            __syncthreads();  // the buffers may still be read by the previous phase
            load_chunk(c0);
            store_chunk(buf);
            __syncthreads();
        }
        for (int c = c0; c < c1; c++) {
            const int nc = c + 1 < c1 ? c + 1 : next_c0;
            if (nc >= 0) load_chunk(nc);
            const int b0 = c * CB;
            const int b1 = min(nblk, b0 + CB);
            for (int blk = b0; blk < b1; blk++) {
                bf16x8 ah[KC], al[KC];
                f32x16 cin;
                read_block(buf, blk - b0, ah, al, cin);
                f32x16 acc[G];
#pragma unroll
                for (int g = 0; g < G; g++) acc[g] = block_acc(g, ah, al, cin);
                per_block(blk, acc);
            }
            if (nc >= 0) store_chunk(buf ^ 1);
            __syncthreads();
            buf ^= 1;
            after_chunk();
        }
        preloaded = next_c0 >= 0;
    };
    auto block_min = [](const f32x16 &v) {
        float t = __builtin_inff();
#pragma unroll
        for (int r = 0; r < 16; r += 2) t = fminf(fminf(t, v[r]), v[r + 1]);
        return t;
    };
    // minima of the lane's four row quads (rows 4i..4i+3 are 4 consecutive neurons) and of the block;
    // chains seeded with +inf, as in block_min, compile to v_min3 without canonicalising the inputs
    auto quad_min = [](const f32x16 &v, float (&q)[4]) {
#pragma unroll
        for (int i = 0; i < 4; i++)
            q[i] = fminf(fminf(fminf(fminf(__builtin_inff(), v[4 * i]), v[4 * i + 1]), v[4 * i + 2]), v[4 * i + 3]);
        return fminf(fminf(fminf(q[0], q[1]), q[2]), q[3]);
    };

    // Own top-2 of this lane's half-block minima. The two lanes of a sample cover disjoint neuron
    // sets, so the merged top-2 are approximate distances of two distinct neurons.
    float m1[G], m2[G], T[G];
#pragma unroll
    for (int g = 0; g < G; g++) m1[g] = m2[g] = __builtin_inff();
    // Threshold from the merged top-2: every neuron whose exact distance can be <= the second-best
    // (best, without NEED_C2) exact distance has D <= T. T only decreases as more blocks are merged.
    auto refresh = [&]() {
#pragma unroll
        for (int g = 0; g < G; g++) {
            const float p1 = xchg32(m1[g]), p2 = xchg32(m2[g]);
            const float A1 = fminf(m1[g], p1);
            const float A2 = fminf(fmaxf(m1[g], p1), fminf(m2[g], p2));
            const float A = NEED_C2 ? A2 : A1;
            const float scale = fmaxf(A + K[g] + E[g], 0.f);
            T[g] = A + 2.f * E[g] + 6.2f * U32 * scale + 8.f * U32 * fabsf(A);
        }
    };
    auto update = [&](int g, float t) {
        // both as v_med3: a NaN t leaves the pair unchanged, as fminf would, without the
        // canonicalising moves fminf needs in IEEE mode
        m2[g] = __builtin_amdgcn_fmed3f(m1[g], m2[g], t);
        m1[g] = __builtin_amdgcn_fmed3f(m1[g], t, neg_inf);
    };
    // Without a pool (D = 16), lp is the LDS byte address of this lane's next list entry, held at the
    // spare last row once the list is full (a store then costs one add and one min); the spare row
    // starts as CAND_EMPTY, which no slot takes, so a list that overflowed is told apart from one that
    // filled exactly. With a pool (D = 32), cnt counts the lane's candidates (two fewer registers).
    unsigned int lp[G], lcap[G];
    int cnt[G], pool_n[G];  // pool_n: wave-uniform fill of the group's shared pool
    bool lost[G];           // a candidate of this lane found both its list and the pool full (D = 32)
#pragma unroll
    for (int g = 0; g < G; g++) {
        if constexpr (POOL == 0) {
            lp[g] = lds_addr(&cw[g][0][lane]);
            lcap[g] = lds_addr(&cw[g][CAPL][lane]);
            cw[g][CAPL][lane] = CAND_EMPTY;
        } else {
            cnt[g] = 0;
        }
        pool_n[g] = 0;
        lost[g] = false;
    }
    // Appends every row of this lane's half-block with D <= T; the pre-check skips the per-value
    // compares when no lane of the wave has a hit in the block, and a second one per row quad skips
    // the quads without a hit (hits in a block usually fall in one quad).
    // Rows of a quad with a hit cost one wave compare and one scalar test each when none of their lanes
    // hits. Past CAPL a lane's candidates go to its group's shared pool when there is one (slots handed
    // out in lane order from the wave's ballot), else the lane is lost to the fallback. Only the lanes
    // with a hit store (an exec-masked store); candidates past the list land in the spare last row.
    auto record = [&](int blk, int g, const f32x16 &v, const float (&q)[4], float t) {
        if (__builtin_amdgcn_ballot_w64(t <= T[g]) != 0ull) {
            const int sbase_blk = blk * 32 + 4 * khalf;
            unsigned long long bq[4];
#pragma unroll
            for (int i = 0; i < 4; i++) bq[i] = __builtin_amdgcn_ballot_w64(q[i] <= T[g]);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (bq[i] == 0ull) continue;  // uniform: no hit in this quad
#pragma unroll
                for (int r = 4 * i; r < 4 * i + 4; r++) {
                    if (__builtin_expect(__builtin_amdgcn_ballot_w64(v[r] <= T[g]) != 0ull, 0)) {  // uniform, rare
                        const bool hit = v[r] <= T[g];
                        const unsigned short slot = (unsigned short)(sbase_blk + 8 * (r >> 2) + (r & 3));
                        if constexpr (POOL > 0) {
                            const bool spill = hit && cnt[g] >= CAPL;
                            const unsigned long long sp = __builtin_amdgcn_ballot_w64(spill);
                            if (__builtin_expect(sp != 0ull, 0)) {  // uniform, rarer still
                                const int off = pool_n[g] + (int)__builtin_amdgcn_mbcnt_hi(
                                                    (unsigned)(sp >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)sp, 0u));
                                if (spill && off < POOL) sm.pool[wave][g][off] = ((unsigned)lane << 16) | slot;
                                lost[g] = lost[g] || (spill && off >= POOL);
                                pool_n[g] += __builtin_popcountll(sp);
                            }
                        }
                        if (hit) {
                            if constexpr (POOL == 0) {
                                *(lds_u16 *)(size_t)lp[g] = slot;
                                // orders the address update after the store, so it is made in place (no
                                // move); the empty statement emits no instruction
                                asm volatile("" : "+v"(lp[g]) : : "memory");
                                lp[g] = min(lp[g] + (unsigned)sizeof(cw[g][0]), lcap[g]);
                            } else {
                                cw[g][min(cnt[g], CAPL)][lane] = slot;
                                cnt[g] += 1;
                            }
                        }
                    }
                }
            }
        }
    };
    auto nothing = [] {};
    const int nchunks_a = (Md.nblk_a + CB - 1) / CB;

    // ---------------- phase A: the first half of the coarse section, top-2 only ----------------
    const int nchunks_a1 = max(1, nchunks_a / 2);
    stream(0, nchunks_a1, nchunks_a1 < nchunks ? nchunks_a1 : 0, [&](int, f32x16 (&acc)[G]) {
#pragma unroll
        for (int g = 0; g < G; g++) update(g, block_min(acc[g]));
    }, nothing);
    refresh();

    // ---------------- phase B: every other block, recording under a running threshold ----------------
    // The rest of the coarse section comes first, then the other blocks. A block is recorded against
    // the threshold in force before its own minima are merged; the threshold in force at any time is
    // >= the final one, so the recorded set contains every neuron with D <= final T.
    if (nchunks_a1 < nchunks) {
        stream(nchunks_a1, nchunks, 0, [&](int blk, f32x16 (&acc)[G]) {
#pragma unroll
            for (int g = 0; g < G; g++) {
                float q[4];
                const float t = quad_min(acc[g], q);
                record(blk, g, acc[g], q, t);
                update(g, t);
            }
        }, refresh);
    }

    // ---------------- phase C: phase A's chunks again, recording under the final threshold ----------------
    stream(0, nchunks_a1, -1, [&](int blk, f32x16 (&acc)[G]) {
#pragma unroll
        for (int g = 0; g < G; g++) {
            float q[4];
            const float t = quad_min(acc[g], q);
            record(blk, g, acc[g], q, t);
        }
    }, nothing);

    // ---------------- exact rerank of the candidate set ----------------
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int64_t s = tile_row0 + sm.rows[wave][g][col];
        const bool valid_s = s < n;
        Top2 tp;
        tp.init();
        int mycnt;
        bool ovf;
        if constexpr (POOL == 0) {
            mycnt = (int)((lp[g] - lds_addr(&cw[g][0][lane])) / (unsigned)sizeof(cw[g][0]));
            ovf = mycnt == CAPL && cw[g][CAPL][lane] != CAND_EMPTY;
        } else {
            mycnt = min(cnt[g], CAPL);
            ovf = lost[g];
        }
        // The two lanes of a sample share its candidates: the pair's lists, the khalf 0 lane's first,
        // are split in two halves, one per lane, so a wave runs as many candidate distances as its most
        // loaded pair's mean, not its longest list. Top-2 insertion is order-independent.
        const int pcnt = xchg32i(mycnt);
        const bool povf = xchg32i(ovf ? 1 : 0) != 0;
        if (valid_s && !ovf && !povf) {
            float xr[DP];
#pragma unroll
            for (int k = 0; k < DP; k++) xr[k] = 0.f;
            if (rows_vec) {
                const auto row = as_f32x4(x + s * DP);
#pragma unroll
                for (int q = 0; q < DP / 4; q++) {
                    const f32x4 t = row[q];
                    xr[4 * q] = t[0]; xr[4 * q + 1] = t[1]; xr[4 * q + 2] = t[2]; xr[4 * q + 3] = t[3];
                }
            } else {
#pragma unroll
                for (int k = 0; k < DP; k++)
                    if (k < dim) xr[k] = x[s * dim + k];
            }
            const int nlo = khalf == 0 ? mycnt : pcnt;  // entries of the khalf 0 lane's list
            const int half = (mycnt + pcnt + 1) / 2;
            const int eb = khalf == 0 ? 0 : half, ee = khalf == 0 ? half : mycnt + pcnt;
            // candidate rows come from the slot-ordered copy of W (no index indirection before the
            // load); the next candidate's row is loaded while this one is evaluated
            auto load_row = [&](int e, float (&wr)[DP], int &j) {
                const bool lo = e < nlo;
                const int slot = cw[g][lo ? e : e - nlo][lo ? col : col + 32];
                j = perm[slot];
                const auto row = as_f32x4(wslot + (int64_t)slot * DP);
#pragma unroll
                for (int q = 0; q < DP / 4; q++) {
                    const f32x4 t = row[q];
                    wr[4 * q] = t[0]; wr[4 * q + 1] = t[1]; wr[4 * q + 2] = t[2]; wr[4 * q + 3] = t[3];
                }
            };
            if constexpr (DP <= 16) {
                float wn[DP];
                int jn = -1;
                if (eb < ee) load_row(eb, wn, jn);
                for (int e = eb; e < ee; e++) {
                    float wr[DP];
#pragma unroll
                    for (int k = 0; k < DP; k++) wr[k] = wn[k];
                    const int j = jn;
                    if (e + 1 < ee) load_row(e + 1, wn, jn);
                    if (j >= 0) tp.push(sqd_mode<MODE>(wr, xr, DP), j);  // zero-padded dims add exact zeros
                }
            } else {
                // wider rows: one row in registers at a time, so the rerank does not set the kernel's
                // register peak (which would halve the waves per SIMD)
                for (int e = eb; e < ee; e++) {
                    float wr[DP];
                    int j;
                    load_row(e, wr, j);
                    // keeps the FP64 arithmetic from holding the sample converted to double across the
                    // loop (64 more registers); the empty statement emits no instruction
#pragma unroll
                    for (int k = 0; k < DP; k++) asm volatile("" : "+v"(xr[k]));
                    if (j >= 0) tp.push(sqd_mode<MODE>(wr, xr, DP), j);
                }
            }
            // the group's spilled candidates: each entry is reranked by the lane that recorded it
            const int np = POOL > 0 ? min(pool_n[g], POOL) : 0;
            for (int e = 0; e < np; e++) {
                const unsigned int ent = sm.pool[wave][g][e];
                if ((int)(ent >> 16) == lane) {
                    float wr[DP];
                    const int slot = (int)(ent & 0xFFFFu);
                    const int j = perm[slot];
                    const auto row = as_f32x4(wslot + (int64_t)slot * DP);
#pragma unroll
                    for (int q = 0; q < DP / 4; q++) {
                        const f32x4 t = row[q];
                        wr[4 * q] = t[0]; wr[4 * q + 1] = t[1]; wr[4 * q + 2] = t[2]; wr[4 * q + 3] = t[3];
                    }
                    if constexpr (DP > 16) {
#pragma unroll
                        for (int k = 0; k < DP; k++) asm volatile("" : "+v"(xr[k]));
                    }
                    if (j >= 0) tp.push(sqd_mode<MODE>(wr, xr, DP), j);
                }
            }
        }
        // merge with the partner lane
        const double pd1 = xchg32d(tp.d1), pd2 = xchg32d(tp.d2);
        const int pi1 = xchg32i(tp.i1), pi2 = xchg32i(tp.i2);
        if (pi1 >= 0) tp.push(pd1, pi1);
        if (pi2 >= 0) tp.push(pd2, pi2);
        const bool bad = ovf || povf || !(T[g] < __builtin_inff()) || (T[g] != T[g]);
        if (lane < 32 && valid_s) {
            if (bad) {
                const unsigned int slot = atomicAdd(fb_count, 1u);
                fb_list[slot] = fb_tag | s;
            } else {
                c1[s] = tp.i1 >= 0 ? tp.i1 : default_index;
                if (c2) c2[s] = tp.i2;
                if (d1) d1[s] = tp.i1 >= 0 ? tp.d1 : __builtin_nan("");
            }
        }
    }
}

// 4 waves per SIMD up to D = 32; at D = 64 the sample fragments and four K slices per block need more
// registers, 2 waves per SIMD
template <int DP, int G, int MODE, bool NEED_C2>
__global__ __launch_bounds__(NW_BF16 * WAVE, DP <= 32 ? 4 : 2) void bmu_bf16_kernel(BmuModelDev Md, const float *__restrict__ x,
                                                                 int64_t n, int default_index,
                                                                 int32_t *__restrict__ c1, int32_t *__restrict__ c2,
                                                                 double *__restrict__ d1,
                                                                 int64_t *__restrict__ fb_list,
                                                                 unsigned int *__restrict__ fb_count,
                                                                 float *__restrict__ xmax_part,
                                                                 const int32_t *__restrict__ prev_c1)
{
    __shared__ Bf16Smem<DP / 16> sm;
    bmu_bf16_body<DP, G, MODE, NEED_C2, false>(Md, x, n, (int64_t)blockIdx.x * SOM_TILE_ROWS, default_index, c1, c2,
                                               d1, fb_list, fb_count, 0, sm, xmax_part, prev_c1);
}

// Serving form: one workgroup per tile; tiles of different requests and models share one launch.
template <int DP, int G, int MODE, bool NEED_C2>
__global__ __launch_bounds__(NW_BF16 * WAVE, 4) void bmu_bf16_tiles_kernel(const BmuModelDev *__restrict__ models,
                                                                       const BmuTile *__restrict__ tiles,
                                                                       int64_t *__restrict__ fb_list,
                                                                       unsigned int *__restrict__ fb_count)
{
    __shared__ Bf16Smem<DP / 16> sm;
    const BmuTile T = tiles[blockIdx.x];
    if (T.n <= 0) return;  // uniform over the workgroup
    const BmuModelDev Md = models[T.model];
    // the descriptors' pointers are generic to the compiler: global_* accesses instead of flat ones
    bmu_bf16_body<DP, G, MODE, NEED_C2, true>(Md, T.x, T.n, T.row0, T.default_index, T.c1, T.c2, T.d1, fb_list,
                                              fb_count, (int64_t)blockIdx.x << FB_TILE_SHIFT, sm, nullptr, nullptr);
}

// Exact top-2 of one sample by one workgroup of FB_THREADS threads: thread t scans the permuted
// slots t, t + FB_THREADS, ... of the slot-ordered weight copy (4 rows in flight per thread), then
// the top-2 lists are merged (butterfly in each wave, then the waves through LDS), lexicographic on
// (distance, neuron index), so the scan order does not matter. Uniform over the workgroup.
constexpr int FB_THREADS = 256;

struct Top2Slot {
    double d1, d2;
    int i1, i2;
};

template <int DP, int MODE>
__device__ __forceinline__ Top2 group_exact_top2(const float *__restrict__ xrow, int dim, const BmuModelDev &Md,
                                                 Top2Slot *red)
{
    constexpr int U = 4;
    const int tid = threadIdx.x;
    float xr[DP];
#pragma unroll
    for (int k = 0; k < DP; k++) xr[k] = k < dim ? ((gptr<const float>)xrow)[k] : 0.f;
    Top2 tp;
    tp.init();
    const int nslots = Md.nblk * 32;
    const gptr<const f32x4> rows = (gptr<const f32x4>)Md.wslot;
    const gptr<const int> perm = (gptr<const int>)Md.perm;
    for (int p0 = tid; p0 < nslots; p0 += U * FB_THREADS) {
        float wr[U][DP];
        int jj[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int p = p0 + u * FB_THREADS;
            jj[u] = -1;
            if (p < nslots) jj[u] = perm[p];
#pragma unroll
            for (int q = 0; q < DP / 4; q++) {
                f32x4 t = f32x4{0.f, 0.f, 0.f, 0.f};
                if (p < nslots) t = rows[(int64_t)p * (DP / 4) + q];
                wr[u][4 * q] = t[0]; wr[u][4 * q + 1] = t[1]; wr[u][4 * q + 2] = t[2]; wr[u][4 * q + 3] = t[3];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (jj[u] >= 0) tp.push(sqd_mode<MODE>(wr[u], xr, DP), jj[u]);  // zero-padded dims add zeros
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const double pd1 = __shfl_xor(tp.d1, o, WAVE), pd2 = __shfl_xor(tp.d2, o, WAVE);
        const int pi1 = __shfl_xor(tp.i1, o, WAVE), pi2 = __shfl_xor(tp.i2, o, WAVE);
        if (pi1 >= 0) tp.push(pd1, pi1);
        if (pi2 >= 0) tp.push(pd2, pi2);
    }
    if ((tid & (WAVE - 1)) == 0) red[tid / WAVE] = Top2Slot{tp.d1, tp.d2, tp.i1, tp.i2};
    __syncthreads();
    if (tid == 0)
        for (int v = 1; v < FB_THREADS / WAVE; v++) {
            if (red[v].i1 >= 0) tp.push(red[v].d1, red[v].i1);
            if (red[v].i2 >= 0) tp.push(red[v].d2, red[v].i2);
        }
    __syncthreads();  // red is reused by the next sample
    return tp;        // complete in thread 0
}

// Fallback of the plain BF16 form: one workgroup per listed sample, grid-stride over the device-side
// count.
template <int DP, int MODE>
__global__ __launch_bounds__(FB_THREADS) void bmu_exact_group_kernel(BmuModelDev Md, const float *__restrict__ x,
                                                                     const int64_t *__restrict__ list,
                                                                     const unsigned int *__restrict__ list_count,
                                                                     int default_index, int32_t *__restrict__ c1,
                                                                     int32_t *__restrict__ c2,
                                                                     double *__restrict__ d1)
{
    __shared__ Top2Slot red[FB_THREADS / WAVE];
    const int64_t total = (int64_t)(*list_count);
    for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
        const int64_t i = list[t];
        const Top2 tp = group_exact_top2<DP, MODE>(x + i * Md.dim, Md.dim, Md, red);
        if (threadIdx.x == 0) {
            c1[i] = tp.i1 >= 0 ? tp.i1 : default_index;
            if (c2) c2[i] = tp.i2;
            if (d1) d1[i] = tp.i1 >= 0 ? tp.d1 : __builtin_nan("");
        }
    }
}

// Fallback of the serving form: one workgroup per (tile, row) entry of the list (entries of one launch
// may belong to different models).
template <int DP, int MODE>
__global__ __launch_bounds__(FB_THREADS) void bmu_exact_tiles_kernel(const BmuModelDev *__restrict__ models,
                                                                     const BmuTile *__restrict__ tiles,
                                                                     const int64_t *__restrict__ list,
                                                                     const unsigned int *__restrict__ list_count)
{
    __shared__ Top2Slot red[FB_THREADS / WAVE];
    const int64_t total = (int64_t)(*list_count);
    for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
        const int64_t e = list[t];
        const BmuTile T = tiles[e >> FB_TILE_SHIFT];
        const int64_t i = e & ((1ll << FB_TILE_SHIFT) - 1);
        const BmuModelDev Md = models[T.model];
        const Top2 tp = group_exact_top2<DP, MODE>(T.x + i * Md.dim, Md.dim, Md, red);
        if (threadIdx.x == 0) {
            ((gptr<int32_t>)T.c1)[i] = tp.i1 >= 0 ? tp.i1 : T.default_index;
            if (T.c2) ((gptr<int32_t>)T.c2)[i] = tp.i2;
            if (T.d1) ((gptr<double>)T.d1)[i] = tp.i1 >= 0 ? tp.d1 : __builtin_nan("");
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Preparation of the derived weight data.
// ------------------------------------------------------------------------------------------------
__global__ void prep_centre_kernel(const float *__restrict__ w, int mn, int dim, int dpad, float *__restrict__ centre,
                                   float *__restrict__ wmax, float *__restrict__ scal)
{
    // one block per padded dimension; fixed-order FP64 sum. Also clears the maxima that
    // prep_rows_kernel raises (same stream, next launch).
    const int k = blockIdx.x;
    if (threadIdx.x == 0) wmax[k] = 0.f;
    if (k == 0 && threadIdx.x < 4) scal[threadIdx.x] = 0.f;
    __shared__ double red[256];
    double acc = 0.0;
    if (k < dim)
        for (int j = threadIdx.x; j < mn; j += blockDim.x) acc += (double)w[(int64_t)j * dim + k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) centre[k] = k < dim ? (float)(red[0] / (double)mn) : 0.f;
}

// One thread per slot: centred row, |w'|^2 (FP32 fma in dimension order), the bf16 hi/lo fragments
// and the raw row of the permuted order, the FP32 fragments of the natural order; 16-byte stores.
// The per-dimension max |w'| and the max |w'|^2 are reduced over the wave before one atomic each.
template <int DP>
__global__ __launch_bounds__(256) void prep_rows_kernel(const float *__restrict__ w, int mn, int dim, int nblk,
                                                        int pblk, const float *__restrict__ centre,
                                                        const int *__restrict__ perm, uint4 *__restrict__ frag_bf,
                                                        f32x4 *__restrict__ frag_f32, f32x4 *__restrict__ wslot,
                                                        float *__restrict__ norms_perm, float *__restrict__ norms_nat,
                                                        unsigned int *__restrict__ wmax_bits,
                                                        unsigned int *__restrict__ nwmax_bits)
{
    constexpr int KC = DP / 16;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;  // slot (permuted) == neuron slot (natural)
    const int lane = threadIdx.x & (WAVE - 1);
    const int blk = p >> 5, row = p & 31;
    float am[DP];
#pragma unroll
    for (int d = 0; d < DP; d++) am[d] = 0.f;
    float nmax = 0.f;
    // permuted slot -> bf16 fragments, raw row and permuted norms
    if (p < pblk * 32) {
        const int j = perm[p];
        float raw[DP], v[DP];
        float nrm = 0.f;
#pragma unroll
        for (int d = 0; d < DP; d++) {
            raw[d] = (j >= 0 && d < dim) ? w[(int64_t)j * dim + d] : 0.f;
            v[d] = (j >= 0 && d < dim) ? __fsub_rn(raw[d], centre[d]) : 0.f;
            nrm = __fmaf_rn(v[d], v[d], nrm);
        }
#pragma unroll
        for (int q = 0; q < DP / 4; q++)
            wslot[(int64_t)p * (DP / 4) + q] = f32x4{raw[4 * q], raw[4 * q + 1], raw[4 * q + 2], raw[4 * q + 3]};
#pragma unroll
        for (int kc = 0; kc < KC; kc++)
#pragma unroll
            for (int half = 0; half < 2; half++) {
                unsigned int hw[4], lw[4];
#pragma unroll
                for (int e = 0; e < 8; e += 2) {
                    unsigned int hp = 0, lp = 0;
#pragma unroll
                    for (int s = 0; s < 2; s++) {
                        const float x = v[kc * 16 + half * 8 + e + s];
                        const unsigned short h = f32_to_bf16_bits(x);
                        const unsigned short l = f32_to_bf16_bits(__fsub_rn(x, bf16_bits_to_f32(h)));
                        // the fragments hold -2 w'_hi and -2 w'_lo (exact in bf16: an exponent step), so
                        // the MFMAs give -2 x'.w' from the sample's unscaled split
                        const unsigned short h2 = f32_to_bf16_bits(-2.f * bf16_bits_to_f32(h));
                        const unsigned short l2 = f32_to_bf16_bits(-2.f * bf16_bits_to_f32(l));
                        hp |= (unsigned int)h2 << (16 * s);
                        lp |= (unsigned int)l2 << (16 * s);
                    }
                    hw[e / 2] = hp;
                    lw[e / 2] = lp;
                }
                const int64_t u = (((int64_t)blk * KC + kc) * 2) * 64 + row + 32 * half;
                frag_bf[u] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
                frag_bf[u + 64] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
            }
        norms_perm[p] = j >= 0 ? nrm : __builtin_inff();
        if (j >= 0) {
#pragma unroll
            for (int d = 0; d < DP; d++) am[d] = fabsf(v[d]);
            nmax = nrm;
        }
    }
#pragma unroll
    for (int d = 0; d < DP; d++) {
        float m = am[d];
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, WAVE));
        if (lane == 0 && m > 0.f) atomicMax(&wmax_bits[d], __builtin_bit_cast(unsigned int, m));
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) nmax = fmaxf(nmax, __shfl_xor(nmax, o, WAVE));
    if (lane == 0 && nmax > 0.f) atomicMax(nwmax_bits, __builtin_bit_cast(unsigned int, nmax));
    // natural slot -> FP32 fragments and natural norms
    if (p < nblk * 32) {
        const int j = p < mn ? p : -1;
        float v[DP];
        float nrm = 0.f;
#pragma unroll
        for (int d = 0; d < DP; d++) {
            v[d] = (j >= 0 && d < dim) ? __fsub_rn(w[(int64_t)j * dim + d], centre[d]) : 0.f;
            nrm = __fmaf_rn(v[d], v[d], nrm);
        }
        // MFMA 32x32x2 operand order: K step st = d/2 goes to lane row + 32*(d&1), element st&3
#pragma unroll
        for (int q = 0; q < DP / 8; q++) {
            const int64_t u = (((int64_t)blk * (DP / 8) + q) * 64 + row);
            frag_f32[u] = f32x4{v[8 * q], v[8 * q + 2], v[8 * q + 4], v[8 * q + 6]};
            frag_f32[u + 32] = f32x4{v[8 * q + 1], v[8 * q + 3], v[8 * q + 5], v[8 * q + 7]};
        }
        norms_nat[p] = j >= 0 ? nrm : __builtin_inff();
    }
}

// Launch helpers -----------------------------------------------------------------------------------

int64_t exact_blocks(int64_t grid_items)
{
    int64_t blocks = (grid_items + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 65535 * 8) blocks = 65535 * 8;
    return blocks;
}

template <int DP>
void launch_exact(int mode, bool need_c2, hipStream_t st, const float *x, int64_t n, int dim, const int64_t *list,
                  const unsigned int *list_count, int64_t grid_items, const float *w, int mn, int default_index,
                  int32_t *c1, int32_t *c2, double *d1, float *xmax_part)
{
    const int64_t blocks = exact_blocks(grid_items);
    dim3 grid((unsigned)blocks), block(256);
#define SOM_EX(M, C)                                                                                            \
    hipLaunchKernelGGL((bmu_exact_kernel<DP, M, C>), grid, block, 0, st, x, n, dim, list, list_count, w, mn,  \
                       default_index, c1, c2, d1, xmax_part)
    if (mode == SOMHIP_RERANK_FP64) {
        if (need_c2) SOM_EX(SOMHIP_RERANK_FP64, true); else SOM_EX(SOMHIP_RERANK_FP64, false);
    } else {
        if (need_c2) SOM_EX(SOMHIP_RERANK_REF32, true); else SOM_EX(SOMHIP_RERANK_REF32, false);
    }
#undef SOM_EX
}

template <int DP>
void launch_exact64(bool need_c2, hipStream_t st, const double *x, int64_t n, int dim, const double *w, int mn,
                    int default_index, int32_t *c1, int32_t *c2, double *d1)
{
    dim3 grid((unsigned)exact_blocks(n)), block(256);
    if (need_c2)
        hipLaunchKernelGGL((bmu_exact_kernel<DP, SOMHIP_RERANK_FP64, true, double>), grid, block, 0, st, x, n, dim,
                           nullptr, nullptr, w, mn, default_index, c1, c2, d1, nullptr);
    else
        hipLaunchKernelGGL((bmu_exact_kernel<DP, SOMHIP_RERANK_FP64, false, double>), grid, block, 0, st, x, n, dim,
                           nullptr, nullptr, w, mn, default_index, c1, c2, d1, nullptr);
}

template <int DP>
void launch_exact_group(int mode, hipStream_t st, const BmuModelDev &Md, const float *x, const int64_t *list,
                        const unsigned int *list_count, int64_t blocks, int default_index, int32_t *c1, int32_t *c2,
                        double *d1)
{
    if (mode == SOMHIP_RERANK_FP64)
        hipLaunchKernelGGL((bmu_exact_group_kernel<DP, SOMHIP_RERANK_FP64>), dim3((unsigned)blocks), dim3(FB_THREADS),
                           0, st, Md, x, list, list_count, default_index, c1, c2, d1);
    else
        hipLaunchKernelGGL((bmu_exact_group_kernel<DP, SOMHIP_RERANK_REF32>), dim3((unsigned)blocks), dim3(FB_THREADS),
                           0, st, Md, x, list, list_count, default_index, c1, c2, d1);
}

constexpr int G_FP32 = 2;
// 32-sample groups per wave of the FP32 kernel: one at D = 64, where two would spill
constexpr int fp32_groups(int dpad) { return dpad <= 32 ? G_FP32 : 1; }

__global__ void zero_u32_kernel(unsigned int *p) { *p = 0u; }

template <int DP>
void launch_fp32(hipStream_t st, somhip_model *M, const float *x, int64_t n, int default_index, int32_t *c1,
                 int32_t *c2, double *d1, float *xmax_part)
{
    constexpr int G = fp32_groups(DP);
    const int64_t per_block = (int64_t)BMU_WAVES * G * 32;
    const int64_t blocks = (n + per_block - 1) / per_block;
    hipLaunchKernelGGL((bmu_fp32_kernel<DP, G>), dim3((unsigned)blocks), dim3(BMU_WAVES * WAVE), 0, st, x, n,
                       M->dim, M->centre, (const f32x4 *)M->frag_f32, (const f32x4 *)M->norms_nat, M->nblk, M->mn,
                       default_index, c1, c2, d1, xmax_part);
}

template <int DP>
void launch_bf16(int mode, bool need_c2, hipStream_t st, somhip_model *M, const float *x, int64_t n,
                 int default_index, int32_t *c1, int32_t *c2, double *d1, int64_t *fb_list, unsigned int *fb_count,
                 float *xmax_part, const int32_t *prev_c1)
{
    const int64_t blocks = (n + SOM_TILE_ROWS - 1) / SOM_TILE_ROWS;
    dim3 grid((unsigned)blocks), block(NW_BF16 * WAVE);
#define SOM_BF(MO, C)                                                                                           \
    hipLaunchKernelGGL((bmu_bf16_kernel<DP, G_BF16, MO, C>), grid, block, 0, st, somhip_model_dev(M), x, n, default_index, c1, \
                       c2, d1, fb_list, fb_count, xmax_part, prev_c1)
    if (mode == SOMHIP_RERANK_FP64) {
        if (need_c2) SOM_BF(SOMHIP_RERANK_FP64, true); else SOM_BF(SOMHIP_RERANK_FP64, false);
    } else {
        if (need_c2) SOM_BF(SOMHIP_RERANK_REF32, true); else SOM_BF(SOMHIP_RERANK_REF32, false);
    }
#undef SOM_BF
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// Internal entry points.
// ------------------------------------------------------------------------------------------------
BmuModelDev somhip_model_dev(const somhip_model *M)
{
    BmuModelDev d;
    d.frag = (const bf16x8 *)M->frag_bf;
    d.norms4 = (const f32x4 *)M->norms_perm;
    d.perm = M->perm;
    d.w = M->w;
    d.centre = M->centre;
    d.wmax = M->wmax;
    d.scal = M->scal;
    d.wslot = M->wslot;
    d.nblk = M->pblk;
    d.nblk_a = M->pblk_a;
    d.mn = M->mn;
    d.dim = M->dim;
    d.ncol = M->n;
    return d;
}

int somhip_launch_zero_u32(hipStream_t st, unsigned int *p)
{
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, st, p);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

int somhip_launch_bmu_tiles(hipStream_t st, int dpad, int arithmetic, int need_c2, const BmuModelDev *models,
                            const BmuTile *tiles, int grid, int64_t *fb_list, unsigned int *fb_count)
{
    if (grid <= 0) return SOMHIP_OK;
    if (dpad != 16 && dpad != 32) {
        somhip_set_error("serving supports dim <= 32");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    dim3 g((unsigned)grid), b(NW_BF16 * WAVE);
    const int fb_grid = std::max(1, std::min(grid, 1024));
#define SOM_TL(D, MO, C)                                                                                  \
    do {                                                                                                  \
        hipLaunchKernelGGL((bmu_bf16_tiles_kernel<D, G_BF16, MO, C>), g, b, 0, st, models, tiles, fb_list, \
                           fb_count);                                                                    \
        hipLaunchKernelGGL((bmu_exact_tiles_kernel<D, MO>), dim3(fb_grid), dim3(256), 0, st, models,       \
                           tiles, fb_list, fb_count);                                                    \
    } while (0)
#define SOM_TL2(D)                                                                                        \
    do {                                                                                                  \
        if (arithmetic == SOMHIP_RERANK_FP64) {                                                           \
            if (need_c2) SOM_TL(D, SOMHIP_RERANK_FP64, true); else SOM_TL(D, SOMHIP_RERANK_FP64, false);  \
        } else {                                                                                          \
            if (need_c2) SOM_TL(D, SOMHIP_RERANK_REF32, true); else SOM_TL(D, SOMHIP_RERANK_REF32, false); \
        }                                                                                                 \
    } while (0)
    if (dpad == 16) SOM_TL2(16);
    else SOM_TL2(32);
#undef SOM_TL2
#undef SOM_TL
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}
// ------------------------------------------------------------------------------------------------
int somhip_prepare(somhip_model *M)
{
    if (M->prepared) return SOMHIP_OK;
    hipStream_t st = M->ctx->stream;
    TimerScope ts(M->ctx, TK_PREP);
    if (M->dpad != 16 && M->dpad != 32 && M->dpad != 64) {
        somhip_set_error("MFMA precisions support dim <= 64");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    hipLaunchKernelGGL(prep_centre_kernel, dim3(M->dpad), dim3(256), 0, st, M->w, M->mn, M->dim, M->dpad, M->centre,
                       M->wmax, M->scal);
    const int slots = std::max(M->nblk, M->pblk) * 32;
#define SOM_PREP(D)                                                                                           \
    hipLaunchKernelGGL(prep_rows_kernel<D>, dim3((slots + 255) / 256), dim3(256), 0, st, M->w, M->mn, M->dim,  \
                       M->nblk, M->pblk, M->centre, M->perm, (uint4 *)M->frag_bf, (f32x4 *)M->frag_f32,       \
                       (f32x4 *)M->wslot, M->norms_perm, M->norms_nat, (unsigned int *)M->wmax,               \
                       (unsigned int *)M->scal)
    if (M->dpad == 16) SOM_PREP(16);
    else if (M->dpad == 32) SOM_PREP(32);
    else SOM_PREP(64);
#undef SOM_PREP
    SOM_HIP_TRY(hipGetLastError());
    M->prepared = true;
    return SOMHIP_OK;
}

// Workspace for the BF16 fallback list (grow-only, per context).
static int fb_workspace(somhip_ctx *C, int64_t n, int64_t **list, unsigned int **count)
{
    size_t need = sizeof(int64_t) * (size_t)(n > 0 ? n : 1) + 256;
    if (C->scratch_bytes < need) {
        if (C->scratch) SOM_HIP_TRY(hipFree(C->scratch));
        C->scratch = nullptr;
        C->scratch_bytes = 0;
        SOM_HIP_TRY(hipMalloc(&C->scratch, need));
        C->scratch_bytes = need;
    }
    *count = (unsigned int *)C->scratch;
    *list = (int64_t *)((char *)C->scratch + 256);
    return SOMHIP_OK;
}

int64_t somhip_bmu_xmax_rows(int precision, int dpad, int64_t n)
{
    if (n <= 0) return 0;
    if (precision == SOMHIP_PREC_EXACT) return exact_blocks(n) * 4;
    if (precision == SOMHIP_PREC_FP32) {
        const int64_t per_block = (int64_t)BMU_WAVES * fp32_groups(dpad) * 32;
        return (n + per_block - 1) / per_block * BMU_WAVES;
    }
    return (n + SOM_TILE_ROWS - 1) / SOM_TILE_ROWS * NW_BF16;
}

int somhip_bmu_precision(const somhip_model *M, int precision)
{
    // Candidate slots are 16-bit in LDS (0xFFFF marks an empty entry). BF16_RERANK promises the exact
    // winners of its rerank arithmetic, which the exact kernel computes directly, so maps with more
    // slots run that kernel.
    if (precision == SOMHIP_PREC_BF16_RERANK && (int64_t)M->pblk * 32 >= 65536) return SOMHIP_PREC_EXACT;
    return precision;
}

int somhip_launch_bmu(somhip_model *M, const float *x, int64_t n, int precision, int arithmetic, int need_c2,
                      int default_index, int32_t *c1, int32_t *c2, double *d1, float *xmax_part, int64_t *fb_list_ext,
                      unsigned int *fb_count_ext, const int32_t *prev_c1)
{
    if (n <= 0) return SOMHIP_OK;
    precision = somhip_bmu_precision(M, precision);
    hipStream_t st = M->ctx->stream;
    const int dp = M->dpad;
    if (precision == SOMHIP_PREC_EXACT) {
        TimerScope ts(M->ctx, TK_BMU_EXACT);
        switch (dp) {
            case 16: launch_exact<16>(arithmetic, need_c2, st, x, n, M->dim, nullptr, nullptr, n, M->w, M->mn, default_index, c1, c2, d1, xmax_part); break;
            case 32: launch_exact<32>(arithmetic, need_c2, st, x, n, M->dim, nullptr, nullptr, n, M->w, M->mn, default_index, c1, c2, d1, xmax_part); break;
            case 48: launch_exact<48>(arithmetic, need_c2, st, x, n, M->dim, nullptr, nullptr, n, M->w, M->mn, default_index, c1, c2, d1, xmax_part); break;
            case 64: launch_exact<64>(arithmetic, need_c2, st, x, n, M->dim, nullptr, nullptr, n, M->w, M->mn, default_index, c1, c2, d1, xmax_part); break;
            default: somhip_set_error("dim > 64 is not supported"); return SOMHIP_ERR_UNSUPPORTED;
        }
        SOM_HIP_TRY(hipGetLastError());
        return SOMHIP_OK;
    }
    if (dp > 64) {
        somhip_set_error("MFMA precisions support dim <= 64");
        return SOMHIP_ERR_UNSUPPORTED;
    }
    int rc = somhip_prepare(M);
    if (rc) return rc;
    if (precision == SOMHIP_PREC_FP32) {
        TimerScope ts(M->ctx, TK_BMU_FP32);
        if (dp == 16) launch_fp32<16>(st, M, x, n, default_index, c1, c2, d1, xmax_part);
        else if (dp == 32) launch_fp32<32>(st, M, x, n, default_index, c1, c2, d1, xmax_part);
        else launch_fp32<64>(st, M, x, n, default_index, c1, c2, d1, xmax_part);
        SOM_HIP_TRY(hipGetLastError());
        return SOMHIP_OK;
    }
    if (precision != SOMHIP_PREC_BF16_RERANK) {
        somhip_set_error("unknown precision");
        return SOMHIP_ERR_ARG;
    }
    int64_t *fb_list = fb_list_ext;
    unsigned int *fb_count = fb_count_ext;
    if (fb_list) {
        // caller's workspace: cleared by a kernel, so the call holds no memset node under graph capture
        if ((rc = somhip_launch_zero_u32(st, fb_count))) return rc;
    } else {
        rc = fb_workspace(M->ctx, n, &fb_list, &fb_count);
        if (rc) return rc;
        SOM_HIP_TRY(hipMemsetAsync(fb_count, 0, sizeof(unsigned int), st));
    }
    {
        TimerScope ts(M->ctx, TK_BMU_BF16);
        if (dp == 16) launch_bf16<16>(arithmetic, need_c2, st, M, x, n, default_index, c1, c2, d1, fb_list, fb_count, xmax_part, prev_c1);
        else if (dp == 32) launch_bf16<32>(arithmetic, need_c2, st, M, x, n, default_index, c1, c2, d1, fb_list, fb_count, xmax_part, prev_c1);
        else launch_bf16<64>(arithmetic, need_c2, st, M, x, n, default_index, c1, c2, d1, fb_list, fb_count, xmax_part, prev_c1);
        SOM_HIP_TRY(hipGetLastError());
    }
    TimerScope tf(M->ctx, TK_BMU_FALLBACK);
    // Fallback over the flagged samples: one workgroup per sample, a fixed grid that reads the
    // device-side count (an empty list costs one launch).
    const int64_t fb_blocks = std::max<int64_t>(1, std::min<int64_t>(n, (int64_t)M->ctx->num_cus * 8));
    const BmuModelDev Md = somhip_model_dev(M);
    if (dp == 16) launch_exact_group<16>(arithmetic, st, Md, x, fb_list, fb_count, fb_blocks, default_index, c1, c2, d1);
    else if (dp == 32) launch_exact_group<32>(arithmetic, st, Md, x, fb_list, fb_count, fb_blocks, default_index, c1, c2, d1);
    else launch_exact_group<64>(arithmetic, st, Md, x, fb_list, fb_count, fb_blocks, default_index, c1, c2, d1);
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}

int somhip_launch_bmu_f64(somhip_model *M, const double *x, int64_t n, int need_c2, int default_index, int32_t *c1,
                          int32_t *c2, double *d1)
{
    if (n <= 0) return SOMHIP_OK;
    hipStream_t st = M->ctx->stream;
    TimerScope ts(M->ctx, TK_BMU_EXACT);
    const bool c2n = need_c2 != 0;
    switch ((M->dim + 15) / 16 * 16) {
        case 16: launch_exact64<16>(c2n, st, x, n, M->dim, M->w64, M->mn, default_index, c1, c2, d1); break;
        case 32: launch_exact64<32>(c2n, st, x, n, M->dim, M->w64, M->mn, default_index, c1, c2, d1); break;
        case 48: launch_exact64<48>(c2n, st, x, n, M->dim, M->w64, M->mn, default_index, c1, c2, d1); break;
        case 64: launch_exact64<64>(c2n, st, x, n, M->dim, M->w64, M->mn, default_index, c1, c2, d1); break;
        default: somhip_set_error("dim > 64 is not supported"); return SOMHIP_ERR_UNSUPPORTED;
    }
    SOM_HIP_TRY(hipGetLastError());
    return SOMHIP_OK;
}
