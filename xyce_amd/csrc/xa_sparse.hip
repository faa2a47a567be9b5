// xa_sparse.hip -- ILU(0) preconditioner for the sparse Newton systems of a
// netlist circuit's transient (host code over rocSPARSE).
//
// The reference solves each Newton system of a transient with KLU (or
// AztecOO GMRES + an ILU preconditioner for large circuits,
// N_LAS_AztecOOSolver.C / N_LAS_IfpackPrecond.C).  Here the Krylov solve
// runs on the device (flexible GMRES in xyce_amd.krylov, products by
// xa_csr_spmv) and the preconditioner is an incomplete LU with the matrix's
// own pattern, factored on the device by rocSPARSE's csrilu0 and applied by
// level-scheduled triangular solves of our own (the levels come from the
// pattern at set-up; a large level is one launch, a run of small levels is
// walked by one workgroup with a barrier per level) or by rocSPARSE's csrsv.  The pattern
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
  // level schedules of the triangular solves (rows grouped so that a row's
  // L (U) entries refer to rows of earlier levels only), host offsets
  int32_t* diag = nullptr;           // slot of each row's diagonal
  int32_t* lrows = nullptr;          // rows by L level
  int32_t* urows = nullptr;          // rows by U level
  std::vector<int> lptr, uptr;
  int32_t* d_lptr = nullptr;         // the level offsets on the device
  int32_t* d_uptr = nullptr;
  int mode = 1;                      // 1 level kernels launched (default; a graph
                                     // replay, 2, measured no faster), 0 rocSPARSE csrsv
  double* bin = nullptr;             // the graph's fixed input and output
  double* xout = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
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
    if (diag) (void)hipFree(diag);
    if (d_lptr) (void)hipFree(d_lptr);
    if (d_uptr) (void)hipFree(d_uptr);
    if (lrows) (void)hipFree(lrows);
    if (urows) (void)hipFree(urows);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (bin) (void)hipFree(bin);
    if (xout) (void)hipFree(xout);
  }
};

// forward substitution with the unit lower factor over one level's rows
__global__ void k_lsolve_level(const int32_t* __restrict__ rows, int cnt,
                               const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                               const int32_t* __restrict__ diag, const double* __restrict__ lu,
                               const double* __restrict__ b, double* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int r = rows[t];
  double s = b[r];
  for (int k = rp[r]; k < diag[r]; ++k) s -= lu[k] * y[ci[k]];
  y[r] = s;
}

// back substitution with the upper factor over one level's rows
__global__ void k_usolve_level(const int32_t* __restrict__ rows, int cnt,
                               const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                               const int32_t* __restrict__ diag, const double* __restrict__ lu,
                               const double* __restrict__ y, double* __restrict__ x) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int r = rows[t];
  double s = y[r];
  for (int k = diag[r] + 1; k < rp[r + 1]; ++k) s -= lu[k] * x[ci[k]];
  x[r] = s / lu[diag[r]];
}

// A run of consecutive small levels (each at most kChain rows) in one
// workgroup: the block walks the levels in order with a barrier between
// them, which makes the previous level's results visible to the whole block
// -- one launch for the run instead of one per level.  Every thread takes
// every barrier (the loop bounds are uniform).
constexpr int kChain = 1024;

__global__ void __launch_bounds__(kChain) k_lsolve_chain(
    const int32_t* __restrict__ rows, const int32_t* __restrict__ lp, int l0, int l1,
    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const int32_t* __restrict__ diag, const double* __restrict__ lu,
    const double* __restrict__ b, double* y) {
  for (int l = l0; l < l1; ++l) {
    for (int t = lp[l] + (int)threadIdx.x; t < lp[l + 1]; t += kChain) {
      const int r = rows[t];
      double s = b[r];
      for (int k = rp[r]; k < diag[r]; ++k) s -= lu[k] * y[ci[k]];
      y[r] = s;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kChain) k_usolve_chain(
    const int32_t* __restrict__ rows, const int32_t* __restrict__ up, int l0, int l1,
    const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const int32_t* __restrict__ diag, const double* __restrict__ lu,
    const double* __restrict__ y, double* x) {
  for (int l = l0; l < l1; ++l) {
    for (int t = up[l] + (int)threadIdx.x; t < up[l + 1]; t += kChain) {
      const int r = rows[t];
      double s = y[r];
      for (int k = diag[r] + 1; k < rp[r + 1]; ++k) s -= lu[k] * x[ci[k]];
      x[r] = s / lu[diag[r]];
    }
    __syncthreads();
  }
}

static void launch_levels(const Ilu0& p, hipStream_t st, const double* b, double* x) {
  // large levels one launch each (many blocks); runs of small ones chained
  auto sweep = [&](const std::vector<int>& ptr, bool lower) {
    const int nl = (int)ptr.size() - 1;
    int l = 0;
    while (l < nl) {
      const int cnt = ptr[l + 1] - ptr[l];
      if (cnt > kChain) {
        if (lower)
          hipLaunchKernelGGL(k_lsolve_level, dim3((cnt + 255) / 256), dim3(256), 0, st,
                             p.lrows + ptr[l], cnt, p.rp, p.ci, p.diag, p.lu, b, p.tmp);
        else
          hipLaunchKernelGGL(k_usolve_level, dim3((cnt + 255) / 256), dim3(256), 0, st,
                             p.urows + ptr[l], cnt, p.rp, p.ci, p.diag, p.lu, p.tmp, x);
        ++l;
        continue;
      }
      int l1 = l;
      while (l1 < nl && ptr[l1 + 1] - ptr[l1] <= kChain) ++l1;
      if (lower)
        hipLaunchKernelGGL(k_lsolve_chain, dim3(1), dim3(kChain), 0, st, p.lrows, p.d_lptr, l,
                           l1, p.rp, p.ci, p.diag, p.lu, b, p.tmp);
      else
        hipLaunchKernelGGL(k_usolve_chain, dim3(1), dim3(kChain), 0, st, p.urows, p.d_uptr, l,
                           l1, p.rp, p.ci, p.diag, p.lu, p.tmp, x);
      l = l1;
    }
  };
  sweep(p.lptr, true);
  sweep(p.uptr, false);
}

// rows bucketed by level: level[i] = 1 + the largest level among the rows
// its strictly lower (upper) entries name
static void levels(int n, const int32_t* rp, const int32_t* ci, const std::vector<int>& dg,
                   bool lower, std::vector<int32_t>& rows, std::vector<int>& ptr) {
  std::vector<int> lev(n, 0);
  int nl = 0;
  for (int t = 0; t < n; ++t) {
    const int i = lower ? t : n - 1 - t;
    int m = 0;
    const int k0 = lower ? rp[i] : dg[i] + 1, k1 = lower ? dg[i] : rp[i + 1];
    for (int k = k0; k < k1; ++k) m = std::max(m, lev[ci[k]] + 1);
    lev[i] = m;
    nl = std::max(nl, m + 1);
  }
  ptr.assign(nl + 1, 0);
  for (int i = 0; i < n; ++i) ptr[lev[i] + 1]++;
  for (int l = 0; l < nl; ++l) ptr[l + 1] += ptr[l];
  rows.assign(n, 0);
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int t = 0; t < n; ++t) {     // increasing row order inside a level
    rows[fill[lev[t]]++] = t;
  }
}

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
  std::vector<int> dg(n);
  for (int r = 0; r < n; ++r)
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      if (col_idx[k] == r) dg[r] = k;
  std::vector<int32_t> lr, ur;
  levels(n, row_ptr, col_idx, dg, true, lr, p.lptr);
  levels(n, row_ptr, col_idx, dg, false, ur, p.uptr);
  XA_HIP(hipMalloc(&p.diag, sizeof(int32_t) * n));
  XA_HIP(hipMalloc(&p.lrows, sizeof(int32_t) * n));
  XA_HIP(hipMalloc(&p.urows, sizeof(int32_t) * n));
  XA_HIP(hipMemcpy(p.diag, dg.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  XA_HIP(hipMemcpy(p.lrows, lr.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  XA_HIP(hipMemcpy(p.urows, ur.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
  XA_HIP(hipMalloc(&p.d_lptr, sizeof(int32_t) * p.lptr.size()));
  XA_HIP(hipMalloc(&p.d_uptr, sizeof(int32_t) * p.uptr.size()));
  XA_HIP(hipMemcpy(p.d_lptr, p.lptr.data(), sizeof(int32_t) * p.lptr.size(),
                   hipMemcpyHostToDevice));
  XA_HIP(hipMemcpy(p.d_uptr, p.uptr.data(), sizeof(int32_t) * p.uptr.size(),
                   hipMemcpyHostToDevice));
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
  if (p.mode == 1) {
    launch_levels(p, p.ctx->stream, b, x);
    XA_HIP(hipGetLastError());
    return XA_OK;
  }
  if (p.mode == 2) {
    // the level launches captured once (on a private stream, kernels only)
    // on fixed buffers, replayed on the context's stream between two copies
    if (!p.exec) {
      XA_HIP(hipMalloc(&p.bin, sizeof(double) * p.m));
      XA_HIP(hipMalloc(&p.xout, sizeof(double) * p.m));
      hipStream_t cs;
      XA_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        launch_levels(p, cs, p.bin, p.xout);
        e = hipStreamEndCapture(cs, &p.graph);
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0);
      (void)hipStreamDestroy(cs);
      if (e != hipSuccess) return hip_check(e, "ilu0 solve graph capture");
    }
    hipStream_t st = p.ctx->stream;
    XA_HIP(hipMemcpyAsync(p.bin, b, sizeof(double) * p.m, hipMemcpyDeviceToDevice, st));
    XA_HIP(hipGraphLaunch(p.exec, st));
    XA_HIP(hipMemcpyAsync(x, p.xout, sizeof(double) * p.m, hipMemcpyDeviceToDevice, st));
    return XA_OK;
  }
  XA_SP(rocsparse_set_stream(p.h, p.ctx->stream));
  const double one = 1.0;
  XA_SP(rocsparse_dcsrsv_solve(p.h, rocsparse_operation_none, p.m, p.nnz, &one, p.dL, p.lu, p.rp,
                               p.ci, p.info, b, p.tmp, rocsparse_solve_policy_auto, p.buf));
  XA_SP(rocsparse_dcsrsv_solve(p.h, rocsparse_operation_none, p.m, p.nnz, &one, p.dU, p.lu, p.rp,
                               p.ci, p.info, p.tmp, x, rocsparse_solve_policy_auto, p.buf));
  return XA_OK;
}

// solve method (2 level kernels as a replayed HIP graph, 1 level kernels
// launched, 0 rocSPARSE csrsv) and the level
// counts of the two triangular solves
int xa_ilu0_config(xa_ilu0 hh, int mode, int* l_levels, int* u_levels) {
  if (!hh) return fail(XA_ERR_ARG, "ilu0 config: null handle");
  if (mode == 0 || mode == 1 || mode == 2) hh->p.mode = mode;
  else if (mode != -1) return fail(XA_ERR_ARG, "ilu0 config: mode must be 0, 1, 2 or -1 (keep)");
  if (l_levels) *l_levels = (int)hh->p.lptr.size() - 1;
  if (u_levels) *u_levels = (int)hh->p.uptr.size() - 1;
  return XA_OK;
}

}  // extern "C"
