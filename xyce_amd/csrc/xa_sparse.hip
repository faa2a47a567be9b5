// xa_sparse.hip -- ILU(0) preconditioner for the sparse Newton systems of a
// netlist circuit's transient (host code over rocSPARSE).
//
// The reference solves each Newton system of a transient with KLU (or
// AztecOO GMRES + an ILU preconditioner for large circuits,
// N_LAS_AztecOOSolver.C / N_LAS_IfpackPrecond.C).  Here the Krylov solve
// runs on the device (flexible GMRES in xyce_amd.krylov, products by
// xa_csr_spmv) and the preconditioner is an incomplete LU with the matrix's
// own pattern, factored and applied on the device: rocSPARSE's csrilu0 and
// level-scheduled triangular solves (L unit lower, U upper).  The pattern
// must hold every diagonal; a voltage source's branch row, whose diagonal is
// structurally zero in the Jacobian, gets its pivot from the elimination of
// its node column (the node precedes the branch in the LID order).
#include <rocsparse/rocsparse.h>
#include <algorithm>
#include <memory>
#include <string>
#include <vector>
#include "xa_internal.h"

namespace xa {

struct Ilu0 {
  Ctx* ctx = nullptr;
  rocsparse_handle h = nullptr;
  rocsparse_mat_descr dM = nullptr, dL = nullptr, dU = nullptr;
  rocsparse_mat_info info = nullptr;
  int m = 0, nnz = 0;
  int32_t* rp = nullptr;
  int32_t* ci = nullptr;
  double* lu = nullptr;    // the factors, in the pattern's slots
  double* tmp = nullptr;   // L^-1 b
  void* buf = nullptr;
  bool analysed = false;
  ~Ilu0() {
    if (info) rocsparse_destroy_mat_info(info);
    if (dM) rocsparse_destroy_mat_descr(dM);
    if (dL) rocsparse_destroy_mat_descr(dL);
    if (dU) rocsparse_destroy_mat_descr(dU);
    if (h) rocsparse_destroy_handle(h);
    if (rp) (void)hipFree(rp);
    if (ci) (void)hipFree(ci);
    if (lu) (void)hipFree(lu);
    if (tmp) (void)hipFree(tmp);
    if (buf) (void)hipFree(buf);
  }
};

static int sp_check(rocsparse_status s, const char* what) {
  if (s == rocsparse_status_success) return XA_OK;
  return fail(XA_ERR_HIP, std::string("rocSPARSE ") + what + " failed (status " +
                              std::to_string((int)s) + ")");
}
#define XA_SP(call) do { int _r = sp_check((call), #call); if (_r != XA_OK) return _r; } while (0)

}  // namespace xa

using namespace xa;

struct xa_ilu0_s { Ilu0 p; };

extern "C" {

int xa_ilu0_create(xa_ctx ctx, int n, int nnz, const int32_t* row_ptr, const int32_t* col_idx,
                   xa_ilu0* out) {
  Ctx* cx = reinterpret_cast<Ctx*>(ctx);
  if (!cx || !row_ptr || !col_idx || !out || n <= 0 || nnz <= 0)
    return fail(XA_ERR_ARG, "ilu0: null argument or empty matrix");
  if (row_ptr[0] != 0 || row_ptr[n] != nnz) return fail(XA_ERR_ARG, "ilu0: row_ptr does not span nnz");
  for (int r = 0; r < n; ++r) {
    bool diag = false;
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
      if (col_idx[k] < 0 || col_idx[k] >= n) return fail(XA_ERR_ARG, "ilu0: column out of range");
      if (k > row_ptr[r] && col_idx[k] <= col_idx[k - 1])
        return fail(XA_ERR_ARG, "ilu0: columns must be sorted and unique in each row");
      diag |= col_idx[k] == r;
    }
    if (!diag) return fail(XA_ERR_ARG, "ilu0: row " + std::to_string(r) + " has no diagonal slot");
  }
  std::unique_ptr<xa_ilu0_s> hh(new xa_ilu0_s());
  Ilu0& p = hh->p;
  p.ctx = cx;
  p.m = n;
  p.nnz = nnz;
  XA_HIP(hipSetDevice(cx->device));
  XA_HIP(hipMalloc(&p.rp, sizeof(int32_t) * (n + 1)));
  XA_HIP(hipMalloc(&p.ci, sizeof(int32_t) * nnz));
  XA_HIP(hipMalloc(&p.lu, sizeof(double) * nnz));
  XA_HIP(hipMalloc(&p.tmp, sizeof(double) * n));
  XA_HIP(hipMemcpy(p.rp, row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
  XA_HIP(hipMemcpy(p.ci, col_idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
  XA_SP(rocsparse_create_handle(&p.h));
  XA_SP(rocsparse_create_mat_descr(&p.dM));
  XA_SP(rocsparse_create_mat_descr(&p.dL));
  XA_SP(rocsparse_set_mat_fill_mode(p.dL, rocsparse_fill_mode_lower));
  XA_SP(rocsparse_set_mat_diag_type(p.dL, rocsparse_diag_type_unit));
  XA_SP(rocsparse_create_mat_descr(&p.dU));
  XA_SP(rocsparse_set_mat_fill_mode(p.dU, rocsparse_fill_mode_upper));
  XA_SP(rocsparse_set_mat_diag_type(p.dU, rocsparse_diag_type_non_unit));
  XA_SP(rocsparse_create_mat_info(&p.info));
  size_t sM = 0, sL = 0, sU = 0;
  XA_SP(rocsparse_dcsrilu0_buffer_size(p.h, n, nnz, p.dM, p.lu, p.rp, p.ci, p.info, &sM));
  XA_SP(rocsparse_dcsrsv_buffer_size(p.h, rocsparse_operation_none, n, nnz, p.dL, p.lu, p.rp,
                                     p.ci, p.info, &sL));
  XA_SP(rocsparse_dcsrsv_buffer_size(p.h, rocsparse_operation_none, n, nnz, p.dU, p.lu, p.rp,
                                     p.ci, p.info, &sU));
  XA_HIP(hipMalloc(&p.buf, std::max(sM, std::max(sL, sU)) + 256));
  *out = hh.release();
  return XA_OK;
}

int xa_ilu0_destroy(xa_ilu0 h) {
  delete h;
  return XA_OK;
}

int xa_ilu0_factor(xa_ilu0 hh, const double* values, int* zero_pivot) {
  if (!hh || !values) return fail(XA_ERR_ARG, "ilu0 factor: null argument");
  Ilu0& p = hh->p;
  XA_HIP(hipSetDevice(p.ctx->device));
  XA_SP(rocsparse_set_stream(p.h, p.ctx->stream));
  XA_HIP(hipMemcpyAsync(p.lu, values, sizeof(double) * p.nnz, hipMemcpyDeviceToDevice,
                        p.ctx->stream));
  if (!p.analysed) {
    XA_SP(rocsparse_dcsrilu0_analysis(p.h, p.m, p.nnz, p.dM, p.lu, p.rp, p.ci, p.info,
                                      rocsparse_analysis_policy_reuse,
                                      rocsparse_solve_policy_auto, p.buf));
    XA_SP(rocsparse_dcsrsv_analysis(p.h, rocsparse_operation_none, p.m, p.nnz, p.dL, p.lu, p.rp,
                                    p.ci, p.info, rocsparse_analysis_policy_reuse,
                                    rocsparse_solve_policy_auto, p.buf));
    XA_SP(rocsparse_dcsrsv_analysis(p.h, rocsparse_operation_none, p.m, p.nnz, p.dU, p.lu, p.rp,
                                    p.ci, p.info, rocsparse_analysis_policy_reuse,
                                    rocsparse_solve_policy_auto, p.buf));
    p.analysed = true;
  }
  XA_SP(rocsparse_dcsrilu0(p.h, p.m, p.nnz, p.dM, p.lu, p.rp, p.ci, p.info,
                           rocsparse_solve_policy_auto, p.buf));
  rocsparse_int pos = -1;
  const rocsparse_status zs = rocsparse_csrilu0_zero_pivot(p.h, p.info, &pos);
  if (zs != rocsparse_status_success && zs != rocsparse_status_zero_pivot)
    return sp_check(zs, "csrilu0_zero_pivot");
  if (zero_pivot) *zero_pivot = zs == rocsparse_status_zero_pivot ? (int)pos : -1;
  return XA_OK;
}

int xa_ilu0_solve(xa_ilu0 hh, const double* b, double* x) {
  if (!hh || !b || !x) return fail(XA_ERR_ARG, "ilu0 solve: null argument");
  Ilu0& p = hh->p;
  if (!p.analysed) return fail(XA_ERR_STATE, "ilu0 solve before a factor");
  XA_HIP(hipSetDevice(p.ctx->device));
  XA_SP(rocsparse_set_stream(p.h, p.ctx->stream));
  const double one = 1.0;
  XA_SP(rocsparse_dcsrsv_solve(p.h, rocsparse_operation_none, p.m, p.nnz, &one, p.dL, p.lu, p.rp,
                               p.ci, p.info, b, p.tmp, rocsparse_solve_policy_auto, p.buf));
  XA_SP(rocsparse_dcsrsv_solve(p.h, rocsparse_operation_none, p.m, p.nnz, &one, p.dU, p.lu, p.rp,
                               p.ci, p.info, p.tmp, x, rocsparse_solve_policy_auto, p.buf));
  return XA_OK;
}

}  // extern "C"
