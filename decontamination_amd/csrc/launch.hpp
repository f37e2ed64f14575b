// Launchers of the training / reduction kernels (k_train.hip), shared by the entry points.
#pragma once

#include "common.hpp"

// bytes of per-workgroup partial tables: enough for one table per CU up to 100x100 neurons x 33 values
// (2.6 MB each), so large maps keep every CU busy; a small share of the 288 GB of HBM
constexpr size_t SCATTER_TABLE_BUDGET = (size_t)1 << 30;

// The launchers taking sample rows, weights or topography are instantiated for float (float32 SOMs)
// and double (float64 SOMs); T is that type. NO_ROWS: the density column alone (activation maps).
constexpr const float *NO_ROWS = nullptr;
template <class T>
int somhip_launch_absmax(hipStream_t st, const T *x, int dim, int dpad, const double *rho, int rho_f32, int64_t n,
                         unsigned long long *bounds, int num_cus);
int somhip_launch_shifts(hipStream_t st, const unsigned long long *bounds, int dim, int has_x, int has_rho, int64_t n,
                         int *shifts);
int somhip_launch_shifts_parts(hipStream_t st, const float *parts, int64_t rows, int dpad, int dim,
                               const unsigned long long *rho_bound, int64_t n, int *shifts);
int somhip_scatter_nwg(int64_t n, int mn, int nv, int num_cus, size_t table_budget);
// bucket (nullable): n uint32 of workspace; with it, maps that need several neuron slices use the
// bucketed kernel (each slice pass reads only its own samples' entries)
template <class T>
int somhip_launch_scatter(hipStream_t st, const int32_t *bmu, const T *x, int dim, const double *rho, int rho_f32,
                          int64_t n, const int *shifts, int mn, int nv, int nwg, unsigned long long *partial,
                          double *acc, size_t lds_limit, unsigned int *bucket = nullptr);
int somhip_update_nsplit(int mn);
template <class T>
int somhip_launch_update(hipStream_t st, const double *S, const T *topo, int mn, int dim, double sigma,
                         double *part, int nsplit, T *w, T *hist);
bool somhip_update_separable_fits(int m, int n, int dim);
template <class R>
int somhip_launch_update_separable(hipStream_t st, const double *S, const R *topo, int m, int n, int hex, int dim,
                                   double sigma, double *T, R *w, R *hist);
template <class T>
int somhip_launch_divide(hipStream_t st, const double *acc, int mn, int dim, T *w, T *hist);
int somhip_iter_nsplit(int64_t n, int mn);
template <class T>
int somhip_launch_iter(hipStream_t st, const int32_t *bmu, const T *x, int dim, int dpad, const double *rho,
                       int64_t n, int64_t cur, int64_t nvec, T sigma0, const T *topo, int mn, double *part,
                       int nsplit, double *acc);
int somhip_errors_nparts(int64_t n);
template <class T>
int somhip_launch_errors(hipStream_t st, const int32_t *c1, const int32_t *c2, const double *d1, const double *rho,
                         int64_t n, const T *topo, double penalty, double *part, int nparts, double *errors);
