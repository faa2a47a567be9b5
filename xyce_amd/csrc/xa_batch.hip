// xa_batch.hip -- Newton update of a batch of independent circuits.
//
// A Monte Carlo or .STEP batch places many small circuits side by side in
// one system: circuit c owns the unknowns [c_start[c], c_start[c+1]) and no
// matrix entry couples two circuits.  After a device load of the whole batch,
// this kernel forms each circuit's Newton system the way the reference's
// integrators form it, with one linear capacitor per node added to Q, and
// solves it.  Backward Euler (N_TIA_BackwardDifferentiation15.C:215-265, and
// the one-step integrator at order 1):
//
//   q      = Q(x) + c x
//   rhs    = -((q - q_prev) / h + F) + dFdxdVp + dQdxdVp / h
//   J      = dF/dx + dQ/dx / h + diag(c / h)
//
// trapezoid, the one-step integrator at order 2 (N_TIA_OneStep.C:272-335 and
// 463-500; F_prev is the previous step's F, its qHistory[2]):
//
//   rhs    = -((q - q_prev) / h + F / 2 + F_prev / 2) + dFdxdVp / 2 + dQdxdVp / h
//   J      = dF/dx / 2 + dQ/dx / h + diag(c / h)
//
// or, at a DC operating point, rhs = -F + dFdxdVp and J = dF/dx.  Rows of
// held nodes (a supply, an initial condition) are identity rows with zero
// right-hand side and their columns are left out.
//
// The matrix is banded once the circuit's unknowns are taken in the caller's
// elimination order (pos[g]: position of unknown g within its circuit; for a
// ring, a reverse Cuthill-McKee order gives two bands either side).  One
// wavefront per circuit assembles the band in LDS and factors it by banded LU
// with partial pivoting in LAPACK dgbtf2's order of operations (kl extra
// rows for the pivoting fill), forward-eliminates the right-hand side
// alongside, back-substitutes, updates x and reports the circuit's largest
// update.
#include <cmath>
#include "xa_internal.h"

namespace xa {

struct BatchArgs {
  int n_circ;
  const int32_t* c_start;
  const int32_t* pos;
  const int32_t* row_ptr;
  const int32_t* col_idx;
  const double* dFdx;
  const double* dQdx;
  const double* f;
  const double* q;
  const double* dFdxdVp;
  const double* dQdxdVp;
  const double* q_prev;
  const double* f_prev;
  const double* c_lin;
  const uint8_t* held;
  double h;
  int method;   // XA_STEP_DCOP, XA_STEP_BE or XA_STEP_TRAP
  int max_n;
  int kl, ku;
  double* x;
  double* q_now;
  double* dx_max;
};

constexpr int kBatchThreads = 64;   // one wavefront per circuit

// Orders the wavefront's LDS accesses across lanes where no barrier stands
// between two phases: a workgroup is one wavefront, whose LDS operations run
// in program order; this keeps the compiler to that order.
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Largest of two magnitudes, NaN when either is NaN (fmax would drop a NaN
// and let a circuit whose update is not finite pass as converged).
__device__ __forceinline__ double nan_max(double a, double b) {
  return (a != a || b != b) ? (double)NAN : (a > b ? a : b);
}

__global__ void __launch_bounds__(kBatchThreads) k_batch_newton(BatchArgs b) {
  extern __shared__ double sm[];
  __shared__ int s_bad;
  const int c = blockIdx.x;
  const int tid = threadIdx.x;
  const int r0 = b.c_start[c], n = b.c_start[c + 1] - r0;
  if (n <= 0 || n > b.max_n) {   // outside the LDS this launch was sized for
    if (tid == 0) b.dx_max[c] = NAN;
    return;
  }
  const int kl = b.kl, ku = b.ku, kv = kl + ku, L = 2 * kl + ku + 1;
  double* AB = sm;             // L x n band: A(i, j) at AB[(kv + i - j) * n + j]
  double* rhs = sm + L * n;    // n, in elimination order
  for (int idx = tid; idx < L * n + n; idx += kBatchThreads) sm[idx] = 0.0;
  if (tid == 0) s_bad = 0;
  __syncthreads();
  for (int i = tid; i < n; i += kBatchThreads) {
    const int gi = r0 + i;
    const int ip = b.pos ? b.pos[gi] : i;
    const double qn = b.q[gi] + b.c_lin[gi] * b.x[gi];
    b.q_now[gi] = qn;
    if (ip < 0 || ip >= n) { s_bad = 1; continue; }
    if (b.held[gi]) {
      AB[kv * n + ip] = 1.0;
      continue;
    }
    for (int p = b.row_ptr[gi]; p < b.row_ptr[gi + 1]; ++p) {
      const int gj = b.col_idx[p], j = gj - r0;
      // held columns multiply a zero update; entries outside the circuit
      // (none in a valid batch) are left out
      if (j < 0 || j >= n || b.held[gj]) continue;
      const int jp = b.pos ? b.pos[gj] : j;
      if (jp < 0 || jp >= n || ip - jp > kl || jp - ip > ku) { s_bad = 1; continue; }
      AB[(kv + ip - jp) * n + jp] +=
          b.method == XA_STEP_DCOP ? b.dFdx[p]
          : b.method == XA_STEP_TRAP ? 0.5 * b.dFdx[p] + b.dQdx[p] / b.h
                                     : b.dFdx[p] + b.dQdx[p] / b.h;
    }
    if (b.method == XA_STEP_DCOP) {
      rhs[ip] = -b.f[gi] + b.dFdxdVp[gi];
    } else if (b.method == XA_STEP_TRAP) {
      AB[kv * n + ip] += b.c_lin[gi] / b.h;
      rhs[ip] = -((qn - b.q_prev[gi]) / b.h + 0.5 * b.f[gi] + 0.5 * b.f_prev[gi]) +
                0.5 * b.dFdxdVp[gi] + b.dQdxdVp[gi] / b.h;
    } else {
      AB[kv * n + ip] += b.c_lin[gi] / b.h;
      rhs[ip] = -((qn - b.q_prev[gi]) / b.h + b.f[gi]) + b.dFdxdVp[gi] + b.dQdxdVp[gi] / b.h;
    }
  }
  __syncthreads();
  if (s_bad) {   // an entry outside the band or a bad position: reported, not solved
    if (tid == 0) b.dx_max[c] = NAN;
    return;
  }
  // Every lane reads the pivot column and the pivots itself (LDS broadcast
  // reads), so the only barriers are after the row interchange and after the
  // update; a lane recomputes its row's multiplier a / pivot, the same
  // division the stored multiplier would hold (the multipliers are not kept:
  // the right-hand side is eliminated alongside).
  int ju = 0;   // last column touched by the row interchanges so far
  for (int j = 0; j < n; ++j) {
    const int km = min(kl, n - 1 - j);
    int jp = 0;
    double best = fabs(AB[kv * n + j]);
    for (int i = 1; i <= km; ++i) {
      const double v = fabs(AB[(kv + i) * n + j]);
      if (v > best) { best = v; jp = i; }
    }
    // A zero pivot means a singular Jacobian (a floating node, a cut-off
    // branch): the circuit gets no update and reports NaN.  Every lane read
    // the same column, so every lane leaves here together.
    if (AB[(kv + jp) * n + j] == 0.0) {
      if (tid == 0) b.dx_max[c] = NAN;
      return;
    }
    ju = max(ju, min(j + ku + jp, n - 1));
    if (jp != 0) {
      wave_fence();   // every lane has read the pivot column before the swap
      for (int col = j + tid; col <= ju; col += kBatchThreads) {
        double* a = &AB[(kv + j - col) * n + col];
        double* e = &AB[(kv + j + jp - col) * n + col];
        const double t = *a;
        *a = *e;
        *e = t;
      }
      if (tid == 0) {
        const double t = rhs[j];
        rhs[j] = rhs[j + jp];
        rhs[j + jp] = t;
      }
      __syncthreads();
    }
    if (km > 0) {
      const double piv = AB[kv * n + j];
      // rank-1 update of the km x (ju - j) block and of the right-hand side
      const int w = ju - j + 1;
      for (int idx = tid; idx < km * w; idx += kBatchThreads) {
        const int i = 1 + idx / w, d = idx % w;
        const double l = AB[(kv + i) * n + j] / piv;
        if (d == 0) rhs[j + i] -= l * rhs[j];
        else AB[(kv + i - d) * n + j + d] -= l * AB[(kv - d) * n + j + d];
      }
      __syncthreads();
    }
  }
  // back substitution with U (bandwidth kl + ku): every lane forms x_j, the
  // lanes update the rows above, lane 0 keeps x_j
  for (int j = n - 1; j >= 0; --j) {
    const double xj = rhs[j] / AB[kv * n + j];
    for (int i = max(0, j - kv) + tid; i < j; i += kBatchThreads)
      rhs[i] -= AB[(kv + i - j) * n + j] * xj;
    wave_fence();   // every lane has read rhs[j]
    if (tid == 0) rhs[j] = xj;
    __syncthreads();
  }
  double mx = 0.0;
  for (int i = tid; i < n; i += kBatchThreads) {
    const int gi = r0 + i;
    if (!b.held[gi]) {
      const double d = rhs[b.pos ? b.pos[gi] : i];
      b.x[gi] += d;
      mx = nan_max(mx, fabs(d));
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = nan_max(mx, __shfl_down(mx, off));
  if (tid == 0) b.dx_max[c] = mx;
}

}  // namespace xa

using namespace xa;

extern "C" int xa_batch_newton_step_ex(xa_ctx ctx, int n_circ, int max_unknowns, int kl, int ku,
                                       const int32_t* c_start, const int32_t* pos,
                                       const int32_t* row_ptr, const int32_t* col_idx,
                                       const double* dFdx, const double* dQdx, const double* f,
                                       const double* q, const double* dFdxdVp,
                                       const double* dQdxdVp, const double* q_prev,
                                       const double* f_prev, const double* c_lin,
                                       const uint8_t* held, double h, int method, double* x,
                                       double* q_now, double* dx_max) {
  Ctx* cx = reinterpret_cast<Ctx*>(ctx);
  if (!cx) return fail(XA_ERR_ARG, "null context");
  if (n_circ <= 0) return XA_OK;
  if (max_unknowns <= 0 || kl < 0 || ku < 0)
    return fail(XA_ERR_ARG, "batch: max_unknowns must be positive and kl, ku >= 0");
  if (method != XA_STEP_DCOP && method != XA_STEP_BE && method != XA_STEP_TRAP)
    return fail(XA_ERR_ARG, "batch: method must be XA_STEP_DCOP, XA_STEP_BE or XA_STEP_TRAP");
  const size_t lds = ((size_t)(2 * kl + ku + 2) * (size_t)max_unknowns) * sizeof(double);
  if (lds > kMaxLds - 1024)
    return fail(XA_ERR_ARG, "batch: a circuit's band does not fit in LDS");
  if (method != XA_STEP_DCOP && !(h > 0.0))
    return fail(XA_ERR_ARG, "batch: transient step needs h > 0");
  if (method == XA_STEP_TRAP && !f_prev)
    return fail(XA_ERR_ARG, "batch: the trapezoid step needs the previous step's F");
  BatchArgs a{n_circ, c_start, pos, row_ptr, col_idx, dFdx, dQdx, f, q, dFdxdVp, dQdxdVp,
              q_prev, f_prev, c_lin, held, h, method, max_unknowns, kl, ku, x, q_now, dx_max};
  if (lds > 65536)
    XA_HIP(hipFuncSetAttribute((const void*)k_batch_newton,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_batch_newton, dim3(n_circ), dim3(kBatchThreads), lds, cx->stream, a);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

extern "C" int xa_batch_newton_step(xa_ctx ctx, int n_circ, int max_unknowns, int kl, int ku,
                                    const int32_t* c_start, const int32_t* pos,
                                    const int32_t* row_ptr, const int32_t* col_idx,
                                    const double* dFdx, const double* dQdx, const double* f,
                                    const double* q, const double* dFdxdVp,
                                    const double* dQdxdVp, const double* q_prev,
                                    const double* c_lin, const uint8_t* held, double h, int dc,
                                    double* x, double* q_now, double* dx_max) {
  return xa_batch_newton_step_ex(ctx, n_circ, max_unknowns, kl, ku, c_start, pos, row_ptr,
                                 col_idx, dFdx, dQdx, f, q, dFdxdVp, dQdxdVp, q_prev, nullptr,
                                 c_lin, held, h, dc ? XA_STEP_DCOP : XA_STEP_BE, x, q_now,
                                 dx_max);
}
