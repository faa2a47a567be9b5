/*
 * oracle.h -- CPU oracle for the device-load path.  TEST INFRASTRUCTURE ONLY:
 * used by tests/, __graft_entry__.smoke() and the CPU-baseline leg of
 * bench.py as the checker; never linked into or called by the product.
 *
 * A plain-C FP64 restatement of Xyce 6.2's device loads, statement by
 * statement, with array-of-structs instances and sequential `+=` through
 * per-instance CSR slot offsets exactly as the reference's loops do.  Each
 * model file cites the reference lines it follows.  The call sequence
 * mirrors the product ABI (include/xyce_amd.h) with host pointers:
 *   or_group_create -> or_group_add_model -> or_group_add_instances ->
 *   or_group_setup -> or_group_counts -> or_group_register_lids ->
 *   or_group_jac_pairs -> or_group_register_csr -> or_group_finalize ->
 *   or_update_state / or_update_secondary_state / or_load_vectors /
 *   or_load_matrices.
 */
#ifndef XA_ORACLE_H
#define XA_ORACLE_H
#include <stdint.h>
#include "../include/xyce_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_group_s* or_group;

int  or_group_create(int model_type, const xa_device_options* opt, or_group* out);
int  or_group_destroy(or_group g);
int  or_group_add_model(or_group g, int level, int n_params, const char* const* names,
                        const double* values, int* model_index);
int  or_group_add_instances(or_group g, int n, const int* model_index, int n_params,
                            const char* const* names, const double* const* values);
int  or_group_request_lead_current(or_group g, const int32_t* flags);
int  or_group_setup(or_group g);
int  or_group_counts(or_group g, int* n_ext, int* n_int, int* n_sta, int* n_sto);
int  or_group_register_lids(or_group g, const int32_t* node_lids, const int32_t* sta_lids,
                            const int32_t* sto_lids);
int  or_group_jac_pairs(or_group g, int64_t* count, int32_t* rows, int32_t* cols);
int  or_group_register_csr(or_group g, int n_rows, const int32_t* row_ptr,
                           const int32_t* col_idx);
int  or_group_finalize(or_group g);
int  or_update_state(or_group g, const xa_solver_state* ss, const xa_vectors* v);
int  or_update_secondary_state(or_group g, const xa_solver_state* ss, const xa_vectors* v);
int  or_load_vectors(or_group g, const xa_solver_state* ss, const xa_vectors* v);
int  or_load_matrices(or_group g, const xa_solver_state* ss, const xa_vectors* v);
int  or_read_intermediate(or_group g, const char* name, double* out);
const char* or_last_error(void);
/* OpenMP threads for the updateState loops (instances are independent there;
 * 1 = the serial order).  The loads that add into shared rows stay serial. */
void or_set_threads(int n);
double or_math(int which, double a, double b);
void or_default_options(xa_device_options* o);
/* 1: loads accumulate |contribution| (error scale of each entry). */
void or_set_magnitude_mode(int on);
/* exp/log/pow results of the evaluation moved by up to `ulps` ulps (0: off) */
void or_set_libm_perturbation(int ulps, uint64_t seed);
int  or_group_save_state(or_group g);
int  or_group_restore_state(or_group g);

#ifdef __cplusplus
}
#endif
#endif
