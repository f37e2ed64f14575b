// Galaxy number density map from the footprint winners (Decontamination_SOM.compute_gndm and
// _compute_gndm, decontamination/decontamination/decontamination_som.py:600-667).
//
// The reference runs two more BMU searches over the footprint (one activation map weighted by the
// coverage-corrected density, one unweighted) after the one that produced the winners, divides
// the two maps on the host and gathers gnd[winners]. The BMUs of all three are the same (same
// weights, same data), so here the two maps are exact fixed-point scatters of the winners, and
// the division and the gather are two small kernels: no distance is computed at all.
#include "launch.hpp"

namespace {

// catalog / footprint maps as the reference hands them over (get_activation_map casts its FP64
// accumulator to the SOM dtype, som_abstract.py:666), and gnd = catalog / footprint in FP32 where
// footprint > 0, NaN elsewhere (np.divide(..., where=footprint > 0), decontamination_som.py:613-618)
__global__ void gndm_maps_kernel(const double *cat64, const double *foot64, int mn, float *cat, float *foot,
                                 float *gnd)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= mn) return;
    const float c = (float)cat64[j];
    const float f = (float)foot64[j];
    cat[j] = c;
    foot[j] = f;
    gnd[j] = f > 0.0f ? __fdiv_rn(c, f) : __builtin_nanf("");
}

// gndm[i] = gnd[winners[i]] (decontamination_som.py:624); an index outside [0, mn) gives NaN
__global__ void gndm_gather_kernel(const int32_t *winners, int64_t n, const float *gnd, int mn, float *gndm)
{
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i0 + 3 < n) {
        const int4 w = *(const int4 *)(winners + i0);
        float4 o;
        o.x = (unsigned)w.x < (unsigned)mn ? gnd[w.x] : __builtin_nanf("");
        o.y = (unsigned)w.y < (unsigned)mn ? gnd[w.y] : __builtin_nanf("");
        o.z = (unsigned)w.z < (unsigned)mn ? gnd[w.z] : __builtin_nanf("");
        o.w = (unsigned)w.w < (unsigned)mn ? gnd[w.w] : __builtin_nanf("");
        *(float4 *)(gndm + i0) = o;
    } else {
        for (int64_t i = i0; i < n; i++) {
            const int32_t w = winners[i];
            gndm[i] = (unsigned)w < (unsigned)mn ? gnd[w] : __builtin_nanf("");
        }
    }
}

}  // namespace

extern "C" int somhip_gndm(somhip_ctx *C, const int32_t *winners, const double *rho, int64_t n, int mn,
                           float *catalog_map, float *footprint_map, float *gnd, float *gndm)
{
    if (!C || n < 0 || mn <= 0 || (n > 0 && (!winners || !rho || !gndm)) || !catalog_map || !footprint_map || !gnd) {
        somhip_set_error("somhip_gndm: invalid arguments");
        return SOMHIP_ERR_ARG;
    }
    if (n > 0 && (((uintptr_t)winners & 15) || ((uintptr_t)gndm & 15))) {
        somhip_set_error("somhip_gndm: winners and gndm must be 16-byte aligned");
        return SOMHIP_ERR_ARG;
    }
    SOM_HIP_TRY(hipSetDevice(C->device));
    hipStream_t st = C->stream;
    Workspace &W = C->ws;
    const int nwg = somhip_scatter_nwg(std::max<int64_t>(n, 1), mn, 1, C->num_cus, SCATTER_TABLE_BUDGET);
    int rc;
    if ((rc = W.maps.ensure(sizeof(double) * 2 * (size_t)mn)) || (rc = W.bounds.ensure(256)) ||
        (rc = W.shifts.ensure(256)) || (rc = W.table.ensure((size_t)nwg * mn * 8)))
        return rc;
    double *cat64 = (double *)W.maps.p, *foot64 = cat64 + mn;
    SOM_HIP_TRY(hipMemsetAsync(cat64, 0, sizeof(double) * 2 * (size_t)mn, st));
    // the two activation maps (the reference's _count_bmus_kernel rule: rows with rho > 0 add rho)
    for (int pass = 0; pass < 2 && n > 0; pass++) {
        const double *r = pass == 0 ? rho : nullptr;
        if ((rc = somhip_launch_absmax(st, NO_ROWS, 0, 16, r, 0, r ? n : 0, (unsigned long long *)W.bounds.p, C->num_cus)) ||
            (rc = somhip_launch_shifts(st, (const unsigned long long *)W.bounds.p, 0, 0, r != nullptr, n, (int *)W.shifts.p)) ||
            (rc = somhip_launch_scatter(st, winners, NO_ROWS, 0, r, 0, n, (const int *)W.shifts.p, mn, 1, nwg,
                                        (unsigned long long *)W.table.p, pass == 0 ? cat64 : foot64,
                                        C->lds_per_block)))
            return rc;
    }
    hipLaunchKernelGGL(gndm_maps_kernel, dim3((mn + 255) / 256), dim3(256), 0, st, cat64, foot64, mn, catalog_map,
                       footprint_map, gnd);
    SOM_HIP_TRY(hipGetLastError());
    if (n > 0) {
        const int64_t groups = (n + 1023) / 1024;
        hipLaunchKernelGGL(gndm_gather_kernel, dim3((unsigned)groups), dim3(256), 0, st, winners, n, gnd, mn, gndm);
        SOM_HIP_TRY(hipGetLastError());
    }
    return SOMHIP_OK;
}
