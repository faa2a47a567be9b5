// xa_batch.hip -- Newton update of a batch of independent circuits.
//
// A Monte Carlo or .STEP batch places many small circuits side by side in
// one system: circuit c owns the unknowns [c_start[c], c_start[c+1]) and no
// matrix entry couples two circuits.  After a device load of the whole batch,
// this kernel forms each circuit's Newton system the way the reference's
// backward-Euler integrator forms it (N_TIA_BackwardDifferentiation15.C:
// 215-265, with one linear capacitor per node added to Q) and solves it:
//
//   q      = Q(x) + c x
//   rhs    = -((q - q_prev) / h + F) + dFdxdVp + dQdxdVp / h
//   J      = dF/dx + dQ/dx / h + diag(c / h)
//
// or, at a DC operating point, rhs = -F + dFdxdVp and J = dF/dx.  Rows of
// held nodes (a supply, an initial condition) are identity rows with zero
// right-hand side.  One workgroup factors one circuit's dense matrix in LDS
// by LU with partial pivoting (LAPACK getrf order of operations, column by
// column), solves, updates x and reports the circuit's largest update.
#include <cmath>
#include "xa_internal.h"

namespace xa {

struct BatchArgs {
  int n_circ;
  const int32_t* c_start;
  const int32_t* row_ptr;
  const int32_t* col_idx;
  const double* dFdx;
  const double* dQdx;
  const double* f;
  const double* q;
  const double* dFdxdVp;
  const double* dQdxdVp;
  const double* q_prev;
  const double* c_lin;
  const uint8_t* held;
  double h;
  int dc;
  int max_n;
  double* x;
  double* q_now;
  double* dx_max;
};

constexpr int kBatchThreads = 256;

__global__ void __launch_bounds__(kBatchThreads) k_batch_newton(BatchArgs b) {
  extern __shared__ double sm[];
  __shared__ int s_piv;
  __shared__ double s_red[kBatchThreads / 64];
  __shared__ int s_redi[kBatchThreads / 64];
  const int c = blockIdx.x;
  const int tid = threadIdx.x;
  const int r0 = b.c_start[c], n = b.c_start[c + 1] - r0;
  if (n <= 0 || n > b.max_n) {   // outside the LDS this launch was sized for
    if (tid == 0) b.dx_max[c] = NAN;
    return;
  }
  double* A = sm;            // n x n, row-major
  double* rhs = sm + n * n;  // n
  for (int idx = tid; idx < n * n; idx += kBatchThreads) A[idx] = 0.0;
  __syncthreads();
  for (int i = tid; i < n; i += kBatchThreads) {
    const int gi = r0 + i;
    const double qn = b.q[gi] + b.c_lin[gi] * b.x[gi];
    b.q_now[gi] = qn;
    if (b.held[gi]) {
      A[i * n + i] = 1.0;
      rhs[i] = 0.0;
      continue;
    }
    for (int p = b.row_ptr[gi]; p < b.row_ptr[gi + 1]; ++p) {
      const int gj = b.col_idx[p], j = gj - r0;
      // held columns multiply a zero update; entries outside the circuit
      // (none in a valid batch) are left out
      if (j < 0 || j >= n || b.held[gj]) continue;
      A[i * n + j] += b.dc ? b.dFdx[p] : b.dFdx[p] + b.dQdx[p] / b.h;
    }
    if (b.dc) {
      rhs[i] = -b.f[gi] + b.dFdxdVp[gi];
    } else {
      A[i * n + i] += b.c_lin[gi] / b.h;
      rhs[i] = -((qn - b.q_prev[gi]) / b.h + b.f[gi]) + b.dFdxdVp[gi] + b.dQdxdVp[gi] / b.h;
    }
  }
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  for (int k = 0; k < n; ++k) {
    // pivot: largest |A[i][k]|, i >= k (first such i on ties)
    double best = -1.0;
    int bi = n;
    for (int i = k + tid; i < n; i += kBatchThreads) {
      const double v = fabs(A[i * n + k]);
      if (v > best) { best = v; bi = i; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_down(best, off);
      const int oi = __shfl_down(bi, off);
      if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    if (lane == 0) { s_red[wave] = best; s_redi[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      double bv = s_red[0];
      int bj = s_redi[0];
      for (int w = 1; w < kBatchThreads / 64; ++w)
        if (s_red[w] > bv || (s_red[w] == bv && s_redi[w] < bj)) { bv = s_red[w]; bj = s_redi[w]; }
      s_piv = bj;
    }
    __syncthreads();
    const int p = s_piv;
    if (p != k && p < n) {
      for (int j = tid; j <= n; j += kBatchThreads) {
        double* a = (j < n) ? &A[k * n + j] : &rhs[k];
        double* bb = (j < n) ? &A[p * n + j] : &rhs[p];
        const double t = *a;
        *a = *bb;
        *bb = t;
      }
    }
    __syncthreads();
    const double akk = A[k * n + k];
    for (int i = k + 1 + tid; i < n; i += kBatchThreads) A[i * n + k] /= akk;
    __syncthreads();
    const int m = n - k - 1;
    for (int idx = tid; idx < m * (m + 1); idx += kBatchThreads) {
      const int i = k + 1 + idx / (m + 1), jj = idx % (m + 1);
      if (jj < m) A[i * n + k + 1 + jj] -= A[i * n + k] * A[k * n + k + 1 + jj];
      else rhs[i] -= A[i * n + k] * rhs[k];
    }
    __syncthreads();
  }
  // back substitution, column by column
  for (int k = n - 1; k >= 0; --k) {
    if (tid == 0) rhs[k] /= A[k * n + k];
    __syncthreads();
    for (int i = tid; i < k; i += kBatchThreads) rhs[i] -= A[i * n + k] * rhs[k];
    __syncthreads();
  }
  double mx = 0.0;
  for (int i = tid; i < n; i += kBatchThreads) {
    const int gi = r0 + i;
    if (!b.held[gi]) {
      b.x[gi] += rhs[i];
      mx = fmax(mx, fabs(rhs[i]));
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off));
  if (lane == 0) s_red[wave] = mx;
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < kBatchThreads / 64; ++w) mx = fmax(mx, s_red[w]);
    b.dx_max[c] = mx;
  }
}

}  // namespace xa

using namespace xa;

extern "C" int xa_batch_newton_step(xa_ctx ctx, int n_circ, int max_unknowns,
                                    const int32_t* c_start, const int32_t* row_ptr,
                                    const int32_t* col_idx, const double* dFdx,
                                    const double* dQdx, const double* f, const double* q,
                                    const double* dFdxdVp, const double* dQdxdVp,
                                    const double* q_prev, const double* c_lin,
                                    const uint8_t* held, double h, int dc, double* x,
                                    double* q_now, double* dx_max) {
  Ctx* cx = reinterpret_cast<Ctx*>(ctx);
  if (!cx) return fail(XA_ERR_ARG, "null context");
  if (n_circ <= 0) return XA_OK;
  if (max_unknowns <= 0) return fail(XA_ERR_ARG, "batch: max_unknowns must be positive");
  const size_t lds = ((size_t)max_unknowns * max_unknowns + max_unknowns) * sizeof(double);
  if (lds > kMaxLds - 1024)
    return fail(XA_ERR_ARG, "batch: a circuit's dense system does not fit in LDS");
  if (!dc && !(h > 0.0)) return fail(XA_ERR_ARG, "batch: transient step needs h > 0");
  BatchArgs a{n_circ, c_start, row_ptr, col_idx, dFdx, dQdx, f, q, dFdxdVp, dQdxdVp,
              q_prev, c_lin, held, h, dc, max_unknowns, x, q_now, dx_max};
  if (lds > 65536)
    XA_HIP(hipFuncSetAttribute((const void*)k_batch_newton,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_batch_newton, dim3(n_circ), dim3(kBatchThreads), lds, cx->stream, a);
  XA_HIP(hipGetLastError());
  return XA_OK;
}
