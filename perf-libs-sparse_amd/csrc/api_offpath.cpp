// api_offpath.cpp — the entry points of the reference interface that lie
// outside the sparse x dense execute path: complex-valued objects and
// operations, sparse vectors and their BLAS-1 operations, sparse + sparse,
// triangular solves, element-wise products, SDDMM and supernodal objects.
// Each is defined so that a program written against the
// reference header links unchanged; a call records an error on the handle it
// names (and as the library's last error), sets any object it would create to
// NULL and returns PERFLIBS_STATUS_EXECUTION_FAILURE.
#include <cstdio>
#include <string>

#include "plsparse.hpp"

namespace {

perflibs_status_t off_path(const char *name, perflibs_spmat_top_t *A) {
  return pls::exec_fail(A, std::string(name) +
                               " is not provided: libplsparse_amd implements the real-valued "
                               "sparse x dense path (SpMV, SpMM) of perflibs_sparse.h");
}

}  // namespace

extern "C" {

// complex-valued matrix objects and supernodal creation
// (reference include/perflibs_sparse.h:167-264)
perflibs_status_t
perflibs_spmat_create_csr_c(perflibs_spmat_t *A, perflibs_int_t, perflibs_int_t,
                            const perflibs_int_t *, const perflibs_int_t *,
                            const perflibs_singlecomplex_t *, perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_csr_c", nullptr);
}
perflibs_status_t
perflibs_spmat_create_csr_z(perflibs_spmat_t *A, perflibs_int_t, perflibs_int_t,
                            const perflibs_int_t *, const perflibs_int_t *,
                            const perflibs_doublecomplex_t *, perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_csr_z", nullptr);
}
perflibs_status_t
perflibs_spmat_create_csc_c(perflibs_spmat_t *A, perflibs_int_t, perflibs_int_t,
                            const perflibs_int_t *, const perflibs_int_t *,
                            const perflibs_singlecomplex_t *, perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_csc_c", nullptr);
}
perflibs_status_t
perflibs_spmat_create_csc_z(perflibs_spmat_t *A, perflibs_int_t, perflibs_int_t,
                            const perflibs_int_t *, const perflibs_int_t *,
                            const perflibs_doublecomplex_t *, perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_csc_z", nullptr);
}
perflibs_status_t
perflibs_spmat_create_coo_c(perflibs_spmat_t *A, perflibs_int_t, perflibs_int_t,
                            perflibs_int_t, const perflibs_int_t *,
                            const perflibs_int_t *,
                            const perflibs_singlecomplex_t *, perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_coo_c", nullptr);
}
perflibs_status_t
perflibs_spmat_create_coo_z(perflibs_spmat_t *A, perflibs_int_t, perflibs_int_t,
                            perflibs_int_t, const perflibs_int_t *,
                            const perflibs_int_t *,
                            const perflibs_doublecomplex_t *, perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_coo_z", nullptr);
}
perflibs_status_t
perflibs_spmat_create_dense_c(perflibs_spmat_t *A, enum perflibs_dense_layout,
                              perflibs_int_t, perflibs_int_t, perflibs_int_t,
                              const perflibs_singlecomplex_t *,
                              perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_dense_c", nullptr);
}
perflibs_status_t
perflibs_spmat_create_dense_z(perflibs_spmat_t *A, enum perflibs_dense_layout,
                              perflibs_int_t, perflibs_int_t, perflibs_int_t,
                              const perflibs_doublecomplex_t *,
                              perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_dense_z", nullptr);
}
perflibs_status_t
perflibs_spmat_create_bsr_c(perflibs_spmat_t *A, enum perflibs_dense_layout,
                            perflibs_int_t, perflibs_int_t, perflibs_int_t,
                            const perflibs_int_t *, const perflibs_int_t *,
                            const perflibs_singlecomplex_t *, perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_bsr_c", nullptr);
}
perflibs_status_t
perflibs_spmat_create_bsr_z(perflibs_spmat_t *A, enum perflibs_dense_layout,
                            perflibs_int_t, perflibs_int_t, perflibs_int_t,
                            const perflibs_int_t *, const perflibs_int_t *,
                            const perflibs_doublecomplex_t *, perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_bsr_z", nullptr);
}
perflibs_status_t
perflibs_spmat_create_supernodal_s(perflibs_spmat_t *A, perflibs_int_t,
                                   perflibs_int_t, perflibs_int_t,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *, const float *,
                                   perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_supernodal_s", nullptr);
}
perflibs_status_t
perflibs_spmat_create_supernodal_d(perflibs_spmat_t *A, perflibs_int_t,
                                   perflibs_int_t, perflibs_int_t,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *, const double *,
                                   perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_supernodal_d", nullptr);
}
perflibs_status_t
perflibs_spmat_create_supernodal_c(perflibs_spmat_t *A, perflibs_int_t,
                                   perflibs_int_t, perflibs_int_t,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_singlecomplex_t *,
                                   perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_supernodal_c", nullptr);
}
perflibs_status_t
perflibs_spmat_create_supernodal_z(perflibs_spmat_t *A, perflibs_int_t,
                                   perflibs_int_t, perflibs_int_t,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_int_t *,
                                   const perflibs_doublecomplex_t *,
                                   perflibs_int_t) {
  if (A) *A = nullptr;
  return off_path("perflibs_spmat_create_supernodal_z", nullptr);
}
// sparse vector objects
// (reference include/perflibs_sparse.h:286-376)
perflibs_status_t
perflibs_spvec_create_s(perflibs_spvec_t *x, perflibs_int_t, perflibs_int_t,
                        perflibs_int_t, const perflibs_int_t *, const float *,
                        perflibs_int_t) {
  if (x) *x = nullptr;
  return off_path("perflibs_spvec_create_s", nullptr);
}
perflibs_status_t
perflibs_spvec_create_d(perflibs_spvec_t *x, perflibs_int_t, perflibs_int_t,
                        perflibs_int_t, const perflibs_int_t *, const double *,
                        perflibs_int_t) {
  if (x) *x = nullptr;
  return off_path("perflibs_spvec_create_d", nullptr);
}
perflibs_status_t
perflibs_spvec_create_c(perflibs_spvec_t *x, perflibs_int_t, perflibs_int_t,
                        perflibs_int_t, const perflibs_int_t *,
                        const perflibs_singlecomplex_t *, perflibs_int_t) {
  if (x) *x = nullptr;
  return off_path("perflibs_spvec_create_c", nullptr);
}
perflibs_status_t
perflibs_spvec_create_z(perflibs_spvec_t *x, perflibs_int_t, perflibs_int_t,
                        perflibs_int_t, const perflibs_int_t *,
                        const perflibs_doublecomplex_t *, perflibs_int_t) {
  if (x) *x = nullptr;
  return off_path("perflibs_spvec_create_z", nullptr);
}
perflibs_status_t
perflibs_spvec_query(perflibs_spvec_t, perflibs_int_t *, perflibs_int_t *,
                     perflibs_int_t *) {
  return off_path("perflibs_spvec_query", nullptr);
}
perflibs_status_t perflibs_spvec_destroy(perflibs_spvec_t) {
  return off_path("perflibs_spvec_destroy", nullptr);
}
perflibs_status_t
perflibs_spvec_export_s(perflibs_spvec_t, perflibs_int_t *, perflibs_int_t *,
                        perflibs_int_t *, perflibs_int_t *, float *) {
  return off_path("perflibs_spvec_export_s", nullptr);
}
perflibs_status_t
perflibs_spvec_export_d(perflibs_spvec_t, perflibs_int_t *, perflibs_int_t *,
                        perflibs_int_t *, perflibs_int_t *, double *) {
  return off_path("perflibs_spvec_export_d", nullptr);
}
perflibs_status_t
perflibs_spvec_export_c(perflibs_spvec_t, perflibs_int_t *, perflibs_int_t *,
                        perflibs_int_t *, perflibs_int_t *,
                        perflibs_singlecomplex_t *) {
  return off_path("perflibs_spvec_export_c", nullptr);
}
perflibs_status_t
perflibs_spvec_export_z(perflibs_spvec_t, perflibs_int_t *, perflibs_int_t *,
                        perflibs_int_t *, perflibs_int_t *,
                        perflibs_doublecomplex_t *) {
  return off_path("perflibs_spvec_export_z", nullptr);
}
perflibs_status_t
perflibs_spvec_gather_s(const float *, perflibs_int_t, perflibs_int_t,
                        perflibs_spvec_t *x_s, perflibs_int_t) {
  if (x_s) *x_s = nullptr;
  return off_path("perflibs_spvec_gather_s", nullptr);
}
perflibs_status_t
perflibs_spvec_gather_d(const double *, perflibs_int_t, perflibs_int_t,
                        perflibs_spvec_t *x_s, perflibs_int_t) {
  if (x_s) *x_s = nullptr;
  return off_path("perflibs_spvec_gather_d", nullptr);
}
perflibs_status_t
perflibs_spvec_gather_c(const perflibs_singlecomplex_t *, perflibs_int_t,
                        perflibs_int_t, perflibs_spvec_t *x_s, perflibs_int_t) {
  if (x_s) *x_s = nullptr;
  return off_path("perflibs_spvec_gather_c", nullptr);
}
perflibs_status_t
perflibs_spvec_gather_z(const perflibs_doublecomplex_t *, perflibs_int_t,
                        perflibs_int_t, perflibs_spvec_t *x_s, perflibs_int_t) {
  if (x_s) *x_s = nullptr;
  return off_path("perflibs_spvec_gather_z", nullptr);
}
perflibs_status_t perflibs_spvec_scatter_s(perflibs_spvec_t, float *) {
  return off_path("perflibs_spvec_scatter_s", nullptr);
}
perflibs_status_t perflibs_spvec_scatter_d(perflibs_spvec_t, double *) {
  return off_path("perflibs_spvec_scatter_d", nullptr);
}
perflibs_status_t
perflibs_spvec_scatter_c(perflibs_spvec_t, perflibs_singlecomplex_t *) {
  return off_path("perflibs_spvec_scatter_c", nullptr);
}
perflibs_status_t
perflibs_spvec_scatter_z(perflibs_spvec_t, perflibs_doublecomplex_t *) {
  return off_path("perflibs_spvec_scatter_z", nullptr);
}
perflibs_status_t
perflibs_spvec_update_s(perflibs_spvec_t, perflibs_int_t,
                        const perflibs_int_t *, const float *) {
  return off_path("perflibs_spvec_update_s", nullptr);
}
perflibs_status_t
perflibs_spvec_update_d(perflibs_spvec_t, perflibs_int_t,
                        const perflibs_int_t *, const double *) {
  return off_path("perflibs_spvec_update_d", nullptr);
}
perflibs_status_t
perflibs_spvec_update_c(perflibs_spvec_t, perflibs_int_t,
                        const perflibs_int_t *,
                        const perflibs_singlecomplex_t *) {
  return off_path("perflibs_spvec_update_c", nullptr);
}
perflibs_status_t
perflibs_spvec_update_z(perflibs_spvec_t, perflibs_int_t,
                        const perflibs_int_t *,
                        const perflibs_doublecomplex_t *) {
  return off_path("perflibs_spvec_update_z", nullptr);
}
// sparse vector x dense vector operations
// (reference include/perflibs_sparse.h:381-686)
perflibs_status_t
perflibs_spdot_exec_s(perflibs_spvec_t, const float *, float *) {
  return off_path("perflibs_spdot_exec_s", nullptr);
}
perflibs_status_t
perflibs_spdot_exec_d(perflibs_spvec_t, const double *, double *) {
  return off_path("perflibs_spdot_exec_d", nullptr);
}
perflibs_status_t
perflibs_spdotu_exec_c(perflibs_spvec_t, const perflibs_singlecomplex_t *,
                       perflibs_singlecomplex_t *) {
  return off_path("perflibs_spdotu_exec_c", nullptr);
}
perflibs_status_t
perflibs_spdotu_exec_z(perflibs_spvec_t, const perflibs_doublecomplex_t *,
                       perflibs_doublecomplex_t *) {
  return off_path("perflibs_spdotu_exec_z", nullptr);
}
perflibs_status_t
perflibs_spdotc_exec_c(perflibs_spvec_t, const perflibs_singlecomplex_t *,
                       perflibs_singlecomplex_t *) {
  return off_path("perflibs_spdotc_exec_c", nullptr);
}
perflibs_status_t
perflibs_spdotc_exec_z(perflibs_spvec_t, const perflibs_doublecomplex_t *,
                       perflibs_doublecomplex_t *) {
  return off_path("perflibs_spdotc_exec_z", nullptr);
}
perflibs_status_t
perflibs_spaxpby_exec_s(const float, perflibs_spvec_t, const float, float *) {
  return off_path("perflibs_spaxpby_exec_s", nullptr);
}
perflibs_status_t
perflibs_spaxpby_exec_d(const double, perflibs_spvec_t, const double,
                        double *) {
  return off_path("perflibs_spaxpby_exec_d", nullptr);
}
perflibs_status_t
perflibs_spaxpby_exec_c(const perflibs_singlecomplex_t, perflibs_spvec_t,
                        const perflibs_singlecomplex_t,
                        perflibs_singlecomplex_t *) {
  return off_path("perflibs_spaxpby_exec_c", nullptr);
}
perflibs_status_t
perflibs_spaxpby_exec_z(const perflibs_doublecomplex_t, perflibs_spvec_t,
                        const perflibs_doublecomplex_t,
                        perflibs_doublecomplex_t *) {
  return off_path("perflibs_spaxpby_exec_z", nullptr);
}
perflibs_status_t
perflibs_spwaxpby_exec_s(const float, perflibs_spvec_t, const float,
                         const float *, float *) {
  return off_path("perflibs_spwaxpby_exec_s", nullptr);
}
perflibs_status_t
perflibs_spwaxpby_exec_d(const double, perflibs_spvec_t, const double,
                         const double *, double *) {
  return off_path("perflibs_spwaxpby_exec_d", nullptr);
}
perflibs_status_t
perflibs_spwaxpby_exec_c(const perflibs_singlecomplex_t, perflibs_spvec_t,
                         const perflibs_singlecomplex_t,
                         const perflibs_singlecomplex_t *,
                         perflibs_singlecomplex_t *) {
  return off_path("perflibs_spwaxpby_exec_c", nullptr);
}
perflibs_status_t
perflibs_spwaxpby_exec_z(const perflibs_doublecomplex_t, perflibs_spvec_t,
                         const perflibs_doublecomplex_t,
                         const perflibs_doublecomplex_t *,
                         perflibs_doublecomplex_t *) {
  return off_path("perflibs_spwaxpby_exec_z", nullptr);
}
perflibs_status_t
perflibs_sprot_exec_s(perflibs_spvec_t, float *, float, float) {
  return off_path("perflibs_sprot_exec_s", nullptr);
}
perflibs_status_t
perflibs_sprot_exec_d(perflibs_spvec_t, double *, double, double) {
  return off_path("perflibs_sprot_exec_d", nullptr);
}
perflibs_status_t
perflibs_sprot_exec_c(perflibs_spvec_t, perflibs_singlecomplex_t *, float,
                      perflibs_singlecomplex_t) {
  return off_path("perflibs_sprot_exec_c", nullptr);
}
perflibs_status_t
perflibs_sprot_exec_cs(perflibs_spvec_t, perflibs_singlecomplex_t *, float,
                       float) {
  return off_path("perflibs_sprot_exec_cs", nullptr);
}
perflibs_status_t
perflibs_sprot_exec_z(perflibs_spvec_t, perflibs_doublecomplex_t *, double,
                      perflibs_doublecomplex_t) {
  return off_path("perflibs_sprot_exec_z", nullptr);
}
perflibs_status_t
perflibs_sprot_exec_zd(perflibs_spvec_t, perflibs_doublecomplex_t *, double,
                       double) {
  return off_path("perflibs_sprot_exec_zd", nullptr);
}
// optimize calls of the other operations
// (reference include/perflibs_sparse.h:438-457)
perflibs_status_t
perflibs_spadd_optimize(enum perflibs_sparse_hint_value,
                        enum perflibs_sparse_hint_value,
                        enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                        enum perflibs_sparse_hint_value, perflibs_spmat_t,
                        perflibs_spmat_t) {
  return off_path("perflibs_spadd_optimize", A);
}
perflibs_status_t
perflibs_spsm_optimize(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                       perflibs_spmat_t, enum perflibs_sparse_hint_value,
                       perflibs_spmat_t) {
  return off_path("perflibs_spsm_optimize", A);
}
perflibs_status_t perflibs_spsv_optimize(perflibs_spmat_t A) {
  return off_path("perflibs_spsv_optimize", A);
}
perflibs_status_t
perflibs_spelmm_optimize(enum perflibs_sparse_hint_value,
                         enum perflibs_sparse_hint_value,
                         enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                         perflibs_spmat_t, enum perflibs_sparse_hint_value,
                         perflibs_spmat_t) {
  return off_path("perflibs_spelmm_optimize", A);
}
perflibs_status_t
perflibs_sddmm_optimize(enum perflibs_sparse_hint_value,
                        enum perflibs_sparse_hint_value,
                        enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                        perflibs_spmat_t, enum perflibs_sparse_hint_value,
                        perflibs_spmat_t) {
  return off_path("perflibs_sddmm_optimize", A);
}
// complex-valued SpMV and SpMM
// (reference include/perflibs_sparse.h:470-499)
perflibs_status_t
perflibs_spmv_exec_c(enum perflibs_sparse_hint_value, perflibs_singlecomplex_t,
                     perflibs_spmat_t A, const perflibs_singlecomplex_t *,
                     perflibs_singlecomplex_t, perflibs_singlecomplex_t *) {
  return off_path("perflibs_spmv_exec_c", A);
}
perflibs_status_t
perflibs_spmv_exec_z(enum perflibs_sparse_hint_value, perflibs_doublecomplex_t,
                     perflibs_spmat_t A, const perflibs_doublecomplex_t *,
                     perflibs_doublecomplex_t, perflibs_doublecomplex_t *) {
  return off_path("perflibs_spmv_exec_z", A);
}
perflibs_status_t
perflibs_spmm_exec_c(enum perflibs_sparse_hint_value,
                     enum perflibs_sparse_hint_value, perflibs_singlecomplex_t,
                     perflibs_spmat_t A, perflibs_spmat_t,
                     perflibs_singlecomplex_t, perflibs_spmat_t) {
  return off_path("perflibs_spmm_exec_c", A);
}
perflibs_status_t
perflibs_spmm_exec_z(enum perflibs_sparse_hint_value,
                     enum perflibs_sparse_hint_value, perflibs_doublecomplex_t,
                     perflibs_spmat_t A, perflibs_spmat_t,
                     perflibs_doublecomplex_t, perflibs_spmat_t) {
  return off_path("perflibs_spmm_exec_z", A);
}
// sparse + sparse, triangular solves, element-wise product, SDDMM
// (reference include/perflibs_sparse.h:506-742)
perflibs_status_t
perflibs_spadd_exec_s(enum perflibs_sparse_hint_value,
                      enum perflibs_sparse_hint_value, float,
                      perflibs_spmat_t A, float, perflibs_spmat_t,
                      perflibs_spmat_t) {
  return off_path("perflibs_spadd_exec_s", A);
}
perflibs_status_t
perflibs_spadd_exec_d(enum perflibs_sparse_hint_value,
                      enum perflibs_sparse_hint_value, double,
                      perflibs_spmat_t A, double, perflibs_spmat_t,
                      perflibs_spmat_t) {
  return off_path("perflibs_spadd_exec_d", A);
}
perflibs_status_t
perflibs_spadd_exec_c(enum perflibs_sparse_hint_value,
                      enum perflibs_sparse_hint_value, perflibs_singlecomplex_t,
                      perflibs_spmat_t A, perflibs_singlecomplex_t,
                      perflibs_spmat_t, perflibs_spmat_t) {
  return off_path("perflibs_spadd_exec_c", A);
}
perflibs_status_t
perflibs_spadd_exec_z(enum perflibs_sparse_hint_value,
                      enum perflibs_sparse_hint_value, perflibs_doublecomplex_t,
                      perflibs_spmat_t A, perflibs_doublecomplex_t,
                      perflibs_spmat_t, perflibs_spmat_t) {
  return off_path("perflibs_spadd_exec_z", A);
}
perflibs_status_t
perflibs_spsm_exec_s(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                     perflibs_spmat_t, float, perflibs_spmat_t) {
  return off_path("perflibs_spsm_exec_s", A);
}
perflibs_status_t
perflibs_spsm_exec_d(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                     perflibs_spmat_t, double, perflibs_spmat_t) {
  return off_path("perflibs_spsm_exec_d", A);
}
perflibs_status_t
perflibs_spsm_exec_c(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                     perflibs_spmat_t, perflibs_singlecomplex_t,
                     perflibs_spmat_t) {
  return off_path("perflibs_spsm_exec_c", A);
}
perflibs_status_t
perflibs_spsm_exec_z(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                     perflibs_spmat_t, perflibs_doublecomplex_t,
                     perflibs_spmat_t) {
  return off_path("perflibs_spsm_exec_z", A);
}
perflibs_status_t
perflibs_spsv_exec_s(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                     float *, float, const float *) {
  return off_path("perflibs_spsv_exec_s", A);
}
perflibs_status_t
perflibs_spsv_exec_d(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                     double *, double, const double *) {
  return off_path("perflibs_spsv_exec_d", A);
}
perflibs_status_t
perflibs_spsv_exec_c(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                     perflibs_singlecomplex_t *, perflibs_singlecomplex_t,
                     const perflibs_singlecomplex_t *) {
  return off_path("perflibs_spsv_exec_c", A);
}
perflibs_status_t
perflibs_spsv_exec_z(enum perflibs_sparse_hint_value, perflibs_spmat_t A,
                     perflibs_doublecomplex_t *, perflibs_doublecomplex_t,
                     const perflibs_doublecomplex_t *) {
  return off_path("perflibs_spsv_exec_z", A);
}
perflibs_status_t
perflibs_spelmm_exec_s(enum perflibs_sparse_hint_value,
                       enum perflibs_sparse_hint_value, float,
                       perflibs_spmat_t A, perflibs_spmat_t, float,
                       perflibs_spmat_t) {
  return off_path("perflibs_spelmm_exec_s", A);
}
perflibs_status_t
perflibs_spelmm_exec_d(enum perflibs_sparse_hint_value,
                       enum perflibs_sparse_hint_value, double,
                       perflibs_spmat_t A, perflibs_spmat_t, double,
                       perflibs_spmat_t) {
  return off_path("perflibs_spelmm_exec_d", A);
}
perflibs_status_t
perflibs_spelmm_exec_c(enum perflibs_sparse_hint_value,
                       enum perflibs_sparse_hint_value,
                       perflibs_singlecomplex_t, perflibs_spmat_t A,
                       perflibs_spmat_t, perflibs_singlecomplex_t,
                       perflibs_spmat_t) {
  return off_path("perflibs_spelmm_exec_c", A);
}
perflibs_status_t
perflibs_spelmm_exec_z(enum perflibs_sparse_hint_value,
                       enum perflibs_sparse_hint_value,
                       perflibs_doublecomplex_t, perflibs_spmat_t A,
                       perflibs_spmat_t, perflibs_doublecomplex_t,
                       perflibs_spmat_t) {
  return off_path("perflibs_spelmm_exec_z", A);
}
perflibs_status_t
perflibs_sddmm_exec_s(enum perflibs_sparse_hint_value,
                      enum perflibs_sparse_hint_value, float,
                      perflibs_spmat_t A, perflibs_spmat_t, float,
                      perflibs_spmat_t) {
  return off_path("perflibs_sddmm_exec_s", A);
}
perflibs_status_t
perflibs_sddmm_exec_d(enum perflibs_sparse_hint_value,
                      enum perflibs_sparse_hint_value, double,
                      perflibs_spmat_t A, perflibs_spmat_t, double,
                      perflibs_spmat_t) {
  return off_path("perflibs_sddmm_exec_d", A);
}
perflibs_status_t
perflibs_sddmm_exec_c(enum perflibs_sparse_hint_value,
                      enum perflibs_sparse_hint_value, perflibs_singlecomplex_t,
                      perflibs_spmat_t A, perflibs_spmat_t,
                      perflibs_singlecomplex_t, perflibs_spmat_t) {
  return off_path("perflibs_sddmm_exec_c", A);
}
perflibs_status_t
perflibs_sddmm_exec_z(enum perflibs_sparse_hint_value,
                      enum perflibs_sparse_hint_value, perflibs_doublecomplex_t,
                      perflibs_spmat_t A, perflibs_spmat_t,
                      perflibs_doublecomplex_t, perflibs_spmat_t) {
  return off_path("perflibs_sddmm_exec_z", A);
}
// complex-valued update and export; CSC and COO export
// (reference include/perflibs_sparse.h:571-667)
perflibs_status_t
perflibs_spmat_update_c(perflibs_spmat_t A, perflibs_int_t,
                        const perflibs_int_t *, const perflibs_int_t *,
                        const perflibs_singlecomplex_t *) {
  return off_path("perflibs_spmat_update_c", A);
}
perflibs_status_t
perflibs_spmat_update_z(perflibs_spmat_t A, perflibs_int_t,
                        const perflibs_int_t *, const perflibs_int_t *,
                        const perflibs_doublecomplex_t *) {
  return off_path("perflibs_spmat_update_z", A);
}
perflibs_status_t
perflibs_spmat_export_csr_c(perflibs_const_spmat_t, perflibs_int_t,
                            perflibs_int_t *, perflibs_int_t *,
                            perflibs_int_t **, perflibs_int_t **,
                            perflibs_singlecomplex_t **) {
  return off_path("perflibs_spmat_export_csr_c", nullptr);
}
perflibs_status_t
perflibs_spmat_export_csr_z(perflibs_const_spmat_t, perflibs_int_t,
                            perflibs_int_t *, perflibs_int_t *,
                            perflibs_int_t **, perflibs_int_t **,
                            perflibs_doublecomplex_t **) {
  return off_path("perflibs_spmat_export_csr_z", nullptr);
}
perflibs_status_t
perflibs_spmat_export_csc_c(perflibs_const_spmat_t, perflibs_int_t,
                            perflibs_int_t *, perflibs_int_t *,
                            perflibs_int_t **, perflibs_int_t **,
                            perflibs_singlecomplex_t **) {
  return off_path("perflibs_spmat_export_csc_c", nullptr);
}
perflibs_status_t
perflibs_spmat_export_csc_z(perflibs_const_spmat_t, perflibs_int_t,
                            perflibs_int_t *, perflibs_int_t *,
                            perflibs_int_t **, perflibs_int_t **,
                            perflibs_doublecomplex_t **) {
  return off_path("perflibs_spmat_export_csc_z", nullptr);
}
perflibs_status_t
perflibs_spmat_export_coo_c(perflibs_const_spmat_t, perflibs_int_t *,
                            perflibs_int_t *, perflibs_int_t *,
                            perflibs_int_t **, perflibs_int_t **,
                            perflibs_singlecomplex_t **) {
  return off_path("perflibs_spmat_export_coo_c", nullptr);
}
perflibs_status_t
perflibs_spmat_export_coo_z(perflibs_const_spmat_t, perflibs_int_t *,
                            perflibs_int_t *, perflibs_int_t *,
                            perflibs_int_t **, perflibs_int_t **,
                            perflibs_doublecomplex_t **) {
  return off_path("perflibs_spmat_export_coo_z", nullptr);
}
perflibs_status_t
perflibs_spmat_export_dense_c(perflibs_const_spmat_t,
                              enum perflibs_dense_layout, perflibs_int_t *,
                              perflibs_int_t *, perflibs_singlecomplex_t **) {
  return off_path("perflibs_spmat_export_dense_c", nullptr);
}
perflibs_status_t
perflibs_spmat_export_dense_z(perflibs_const_spmat_t,
                              enum perflibs_dense_layout, perflibs_int_t *,
                              perflibs_int_t *, perflibs_doublecomplex_t **) {
  return off_path("perflibs_spmat_export_dense_z", nullptr);
}
perflibs_status_t
perflibs_spmat_export_bsr_c(perflibs_const_spmat_t, enum perflibs_dense_layout,
                            perflibs_int_t, perflibs_int_t *, perflibs_int_t *,
                            perflibs_int_t *, perflibs_int_t **,
                            perflibs_int_t **, perflibs_singlecomplex_t **) {
  return off_path("perflibs_spmat_export_bsr_c", nullptr);
}
perflibs_status_t
perflibs_spmat_export_bsr_z(perflibs_const_spmat_t, enum perflibs_dense_layout,
                            perflibs_int_t, perflibs_int_t *, perflibs_int_t *,
                            perflibs_int_t *, perflibs_int_t **,
                            perflibs_int_t **, perflibs_doublecomplex_t **) {
  return off_path("perflibs_spmat_export_bsr_z", nullptr);
}
// norms, in-place scaling and transposition
// (reference include/perflibs_sparse.h:690-766)
perflibs_status_t
perflibs_spnorm_exec_c(perflibs_spmat_t A, enum perflibs_sparse_norm, float *) {
  return off_path("perflibs_spnorm_exec_c", A);
}
perflibs_status_t
perflibs_spnorm_exec_z(perflibs_spmat_t A, enum perflibs_sparse_norm,
                       double *) {
  return off_path("perflibs_spnorm_exec_z", A);
}
perflibs_status_t
perflibs_spscale_exec_c(perflibs_singlecomplex_t, perflibs_spmat_t A) {
  return off_path("perflibs_spscale_exec_c", A);
}
perflibs_status_t
perflibs_spscale_exec_z(perflibs_doublecomplex_t, perflibs_spmat_t A) {
  return off_path("perflibs_spscale_exec_z", A);
}
perflibs_status_t
perflibs_sptranspose_exec_c(enum perflibs_sparse_hint_value,
                            perflibs_spmat_t A) {
  return off_path("perflibs_sptranspose_exec_c", A);
}
perflibs_status_t
perflibs_sptranspose_exec_z(enum perflibs_sparse_hint_value,
                            perflibs_spmat_t A) {
  return off_path("perflibs_sptranspose_exec_z", A);
}
// errors of sparse vector objects
// (reference include/perflibs_sparse.h:770-770)
void perflibs_spvec_print_err(perflibs_spvec_t) {
  // no sparse vector object can exist here
  std::fprintf(stderr, "perflibs_spvec_print_err: sparse vector objects are not provided by "
                       "libplsparse_amd\n");
}

}  // extern "C"
