// xa_hb.hip -- harmonic balance on the device: the permuted transforms, the
// batched-sample HB load, and the block-Jacobi preconditioner solve in FP32
// MFMA LU with FP64 iterative refinement.
//
// Layouts (the reference's HB block vectors):
//   time domain      xt[s * n + j]: sample s (fast time s T / N) of unknown j,
//                    one block per sample (N_LOA_HBLoader.C:565-600);
//   frequency domain xf[j * 2N + 2p (+1)]: (Re, Im) of unknown j at position
//                    p, harmonic k = p for p <= H and k = p - N above (the
//                    permuted layout of N_LAS_BlockSystemHelpers.C:512-609),
//                    N = 2H + 1 samples (N_ANP_HB.C:1470).
//
// The transforms are direct sums over the samples with an FP64 table of
// cos / sin (2 pi r / N), one thread per output (unknown, harmonic) or
// (sample, unknown): for the sample counts of HB (a few hundred) this is a
// few microseconds and exact to the table's rounding, where the reference
// calls FFTW (N_UTL_FFTW_Interface.C:66-130) -- the same sums in another
// order.
#include <cmath>
#include "xa_internal.h"

namespace xa {

int lu_check(int batch, int m, int ld);
int lu_factor(Ctx* cx, int batch, int m, int ld, float* a, int32_t* ipiv, int32_t* info);
int lu_solve(Ctx* cx, int batch, int nrhs, int m, int ld, const float* lu, const int32_t* ipiv,
             float* b);
int lu_prepare(Ctx* cx, int batch, int m, int ld, const float* lu, const int32_t* ipiv,
               int32_t* perm, float* linv, float* uinv);
int lu_solve_prepared(Ctx* cx, int batch, int nrhs, int m, int ld, const float* lu,
                      const int32_t* perm, const float* linv, const float* uinv, float* b);

struct HBPlan {
  Ctx* ctx = nullptr;
  int N = 0, H = 0, n = 0;
  double period = 0.0;
  double* d_cos = nullptr;   // cos(2 pi r / N), r = 0..N-1
  double* d_sin = nullptr;
};

struct HBPrecond {
  HBPlan* plan = nullptr;
  int nnz = 0, m = 0, ld = 0, nh = 0;   // m = 2n real-equivalent rows; nh = H + 1 factors
  int32_t* d_row_ptr = nullptr;
  int32_t* d_col_idx = nullptr;
  double* d_G = nullptr;     // sample averages of dF/dx and dQ/dx (the CSR pattern)
  double* d_C = nullptr;
  float* d_lu = nullptr;     // nh blocks ld x ld
  int32_t* d_ipiv = nullptr;
  int32_t* d_info = nullptr;
  int32_t* d_perm = nullptr;  // the exchanges as one permutation, per block
  float* d_linv = nullptr;    // inverses of the 64x64 diagonal blocks of L and U
  float* d_uinv = nullptr;
  double* d_b = nullptr;     // N systems x m: right-hand sides, solutions, residuals
  double* d_x = nullptr;
  double* d_r = nullptr;
  float* d_rhs = nullptr;    // nh x 2 x ld FP32 right-hand sides of the factors
  double* d_norm = nullptr;  // N x 2: max |r|, max |b|
  bool factored = false;
};

__global__ void k_hb_zero(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// xt[s][j] = Re_0 + 2 sum_{k=1..H} (Re_k cos(2 pi k s / N) - Im_k sin(...)),
// the k-th term added in order (computePermutedIFT with the HC2R packing and
// the 1/N and N scalings that cancel, N_UTL_FFTW_Interface.C:101-130).
__global__ void k_hb_ift(const double* __restrict__ xf, double* __restrict__ xt,
                         const double* __restrict__ cs, const double* __restrict__ sn, int N,
                         int H, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = blockIdx.y;
  if (j >= n) return;
  const double* f = xf + (size_t)j * 2 * N;
  double acc = f[0];
  int r = 0;
  for (int k = 1; k <= H; ++k) {
    r += s;
    if (r >= N) r -= N;   // (k s) mod N
    acc += 2.0 * (f[2 * k] * cs[r] - f[2 * k + 1] * sn[r]);
  }
  xt[(size_t)s * n + j] = acc;
}

// Harmonic k (0..H) of every unknown: r_k = sum_s x_s cos(2 pi k s / N),
// i_k = -sum_s x_s sin(...), stored / N at k and, conjugated, at N - k
// (computePermutedDFT, N_LAS_BlockSystemHelpers.C:512-562).
__global__ void k_hb_dft(const double* __restrict__ xt, double* __restrict__ xf,
                         const double* __restrict__ cs, const double* __restrict__ sn, int N,
                         int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (j >= n) return;
  double re = 0.0, im = 0.0;
  int r = 0;
  for (int s = 0; s < N; ++s) {
    const double x = xt[(size_t)s * n + j];
    re += x * cs[r];
    im -= x * sn[r];
    r += k;
    if (r >= N) r -= N;
  }
  double* f = xf + (size_t)j * 2 * N;
  if (k == 0) {
    f[0] = re / N;
    f[1] = 0.0;
  } else {
    f[2 * k] = re / N;
    f[2 * k + 1] = im / N;
    f[2 * (N - k)] = re / N;
    f[2 * (N - k) + 1] = -im / N;
  }
}

// F_f += j w k Q_f (N_LOA_HBLoader.C:733-775): harmonic k = 1..H gains
// (-Im Q_k k w, Re Q_k k w) and harmonic -k (-Im Q_k k w, -Re Q_k k w).
__global__ void k_hb_jwq(double* __restrict__ ff, const double* __restrict__ qf, int N, int H,
                         int n, double w) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y + 1;
  if (j >= n || k > H) return;
  double* f = ff + (size_t)j * 2 * N;
  const double* q = qf + (size_t)j * 2 * N;
  const double qre = q[2 * k], qim = q[2 * k + 1];
  f[2 * k] += -qim * k * w;
  f[2 * (N - k)] += -qim * k * w;
  f[2 * k + 1] += qre * k * w;
  f[2 * (N - k) + 1] += -qre * k * w;
}

// G = (sum_s dF/dx_s) * (1/N), C likewise (N_LAS_HBBlockJacobiPrecond.C:
// 298-311: the stored per-sample matrices added, then scaled).
__global__ void k_hb_average(const double* __restrict__ dfdx, const double* __restrict__ dqdx,
                             double* __restrict__ G, double* __restrict__ C, int N, int nnz) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  double g = 0.0, c = 0.0;
  for (int s = 0; s < N; ++s) {
    g += dfdx[(size_t)s * nnz + e];
    c += dqdx[(size_t)s * nnz + e];
  }
  G[e] = g * (1.0 / N);
  C[e] = c * (1.0 / N);
}

// Row i of harmonic k's real-equivalent block [G, -k w C; k w C, G]
// (N_LAS_HBBlockJacobiPrecond.C:325-370) in FP32, and the identity beyond
// row m: a workgroup clears the row, then its threads write the nonzeros.
__global__ void __launch_bounds__(256)
k_hb_assemble(float* __restrict__ lu, const int32_t* __restrict__ row_ptr,
              const int32_t* __restrict__ col_idx, const double* __restrict__ G,
              const double* __restrict__ C, int n, int ld, double w) {
  const int i = blockIdx.x, k = blockIdx.y;
  float* row = lu + ((size_t)k * ld + i) * ld;
  for (int c = threadIdx.x; c < ld; c += blockDim.x) row[c] = 0.0f;
  __syncthreads();
  const int m = 2 * n;
  if (i >= m) {
    if (threadIdx.x == 0) row[i] = 1.0f;
    return;
  }
  const int base = i < n ? i : i - n;
  const double cplx = k * w;
  for (int e = row_ptr[base] + threadIdx.x; e < row_ptr[base + 1]; e += blockDim.x) {
    const int j = col_idx[e];
    if (i < n) {
      row[j] = (float)G[e];
      if (k) row[n + j] = (float)(-(C[e] * cplx));
    } else {
      if (k) row[j] = (float)(C[e] * cplx);
      row[n + j] = (float)G[e];
    }
  }
}

// harmonic of system position p, and the factor that serves it
__device__ __forceinline__ int hb_harmonic(int p, int N, int H) { return p <= H ? p : p - N; }

// b[p][i] from the permuted layout: rows i < n the real parts, n + i the
// imaginary parts of position p
__global__ void k_hb_gather(const double* __restrict__ bf, double* __restrict__ b, int N, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= 2 * n) return;
  const int j = i < n ? i : i - n;
  b[(size_t)p * 2 * n + i] = bf[(size_t)j * 2 * N + 2 * p + (i < n ? 0 : 1)];
}

__global__ void k_hb_scatter(const double* __restrict__ x, double* __restrict__ xf, int N, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= 2 * n) return;
  const int j = i < n ? i : i - n;
  xf[(size_t)j * 2 * N + 2 * p + (i < n ? 0 : 1)] = x[(size_t)p * 2 * n + i];
}

// r = b - K_k x for every system (FP64, from the CSR averages): row i < n is
// G x_re - k w C x_im, row n + i is k w C x_re + G x_im.
__global__ void k_hb_residual(const int32_t* __restrict__ row_ptr,
                              const int32_t* __restrict__ col_idx, const double* __restrict__ G,
                              const double* __restrict__ C, const double* __restrict__ b,
                              const double* __restrict__ x, double* __restrict__ r, int N, int H,
                              int n, double w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= 2 * n) return;
  const double cplx = hb_harmonic(p, N, H) * w;
  const double* xs = x + (size_t)p * 2 * n;
  const int base = i < n ? i : i - n;
  double gx = 0.0, cx = 0.0;
  for (int e = row_ptr[base]; e < row_ptr[base + 1]; ++e) {
    const int j = col_idx[e];
    if (i < n) {
      gx += G[e] * xs[j];
      cx += C[e] * xs[n + j];
    } else {
      gx += G[e] * xs[n + j];
      cx += C[e] * xs[j];
    }
  }
  const double kx = i < n ? gx - cplx * cx : gx + cplx * cx;
  r[(size_t)p * 2 * n + i] = b[(size_t)p * 2 * n + i] - kx;
}

// r of system p, in FP32, into its factor's right-hand side: position p of
// harmonic k >= 0 is slot 0 of factor k; harmonic -k is slot 1 of factor k
// with the imaginary half negated (K_-k = P K_k P, P = diag(I, -I)).
__global__ void k_hb_to_rhs(const double* __restrict__ r, float* __restrict__ rhs, int N, int H,
                            int n, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= ld) return;
  const int k = hb_harmonic(p, N, H);
  const int f = k < 0 ? -k : k, slot = k < 0 ? 1 : 0;
  float v = 0.0f;
  if (i < 2 * n) {
    const double d = r[(size_t)p * 2 * n + i];
    v = (float)(k < 0 && i >= n ? -d : d);
  }
  rhs[((size_t)f * 2 + slot) * ld + i] = v;
}

// x += the correction (FP32 solve result, undoing P for negative harmonics)
__global__ void k_hb_from_rhs(const float* __restrict__ rhs, double* __restrict__ x, int N, int H,
                              int n, int ld) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= 2 * n) return;
  const int k = hb_harmonic(p, N, H);
  const int f = k < 0 ? -k : k, slot = k < 0 ? 1 : 0;
  const double d = (double)rhs[((size_t)f * 2 + slot) * ld + i];
  x[(size_t)p * 2 * n + i] += (k < 0 && i >= n) ? -d : d;
}

// max |r| and max |b| of every system (one workgroup per system)
__global__ void __launch_bounds__(256)
k_hb_norms(const double* __restrict__ r, const double* __restrict__ b, double* __restrict__ out,
           int n) {
  const int p = blockIdx.x;
  __shared__ double sr[256], sb[256];
  double mr = 0.0, mb = 0.0;
  for (int i = threadIdx.x; i < 2 * n; i += 256) {
    const double a = fabs(r[(size_t)p * 2 * n + i]), c = fabs(b[(size_t)p * 2 * n + i]);
    mr = (a > mr || a != a) ? a : mr;
    mb = c > mb ? c : mb;
  }
  sr[threadIdx.x] = mr;
  sb[threadIdx.x] = mb;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double a = sr[threadIdx.x + s], c = sb[threadIdx.x + s];
      if (a > sr[threadIdx.x] || a != a) sr[threadIdx.x] = a;
      if (c > sb[threadIdx.x]) sb[threadIdx.x] = c;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * p] = sr[0];
    out[2 * p + 1] = sb[0];
  }
}

static dim3 hb_grid(int count, int y) { return dim3((count + 255) / 256, y); }

int hb_ift(HBPlan* h, const double* xf, double* xt) {
  hipLaunchKernelGGL(k_hb_ift, hb_grid(h->n, h->N), dim3(256), 0, h->ctx->stream, xf, xt,
                     h->d_cos, h->d_sin, h->N, h->H, h->n);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

int hb_dft(HBPlan* h, const double* xt, double* xf) {
  hipLaunchKernelGGL(k_hb_dft, hb_grid(h->n, h->H + 1), dim3(256), 0, h->ctx->stream, xt, xf,
                     h->d_cos, h->d_sin, h->N, h->n);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

int hb_jwq(HBPlan* h, double* ff, const double* qf) {
  if (h->H == 0) return XA_OK;
  hipLaunchKernelGGL(k_hb_jwq, hb_grid(h->n, h->H), dim3(256), 0, h->ctx->stream, ff, qf, h->N,
                     h->H, h->n, 2.0 * M_PI / h->period);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

}  // namespace xa

using namespace xa;

struct xa_hb_plan_s : HBPlan {};
struct xa_hb_precond_s : HBPrecond {};

extern "C" int xa_hb_plan_create(xa_ctx ctx, int n_samples, int n_unknowns, double period,
                                 xa_hb_plan* out) {
  Ctx* cx = reinterpret_cast<Ctx*>(ctx);
  if (!out) return fail(XA_ERR_ARG, "hb: null output handle");
  *out = nullptr;
  if (!cx) return fail(XA_ERR_ARG, "null context");
  if (n_samples < 1 || n_samples % 2 == 0)
    return fail(XA_ERR_ARG, "hb: the sample count must be odd (2H + 1)");
  if (n_unknowns < 1 || !(period > 0.0)) return fail(XA_ERR_ARG, "hb: unknowns >= 1, period > 0");
  XA_HIP(hipSetDevice(cx->device));
  auto* h = new xa_hb_plan_s();
  h->ctx = cx;
  h->N = n_samples;
  h->H = (n_samples - 1) / 2;
  h->n = n_unknowns;
  h->period = period;
  std::vector<double> c(n_samples), s(n_samples);
  for (int r = 0; r < n_samples; ++r) {
    const double a = 2.0 * M_PI * r / n_samples;
    c[r] = std::cos(a);
    s[r] = std::sin(a);
  }
  hipError_t e = hipMalloc((void**)&h->d_cos, sizeof(double) * n_samples);
  if (e == hipSuccess) e = hipMalloc((void**)&h->d_sin, sizeof(double) * n_samples);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_cos, c.data(), sizeof(double) * n_samples, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_sin, s.data(), sizeof(double) * n_samples, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(h->d_cos);
    (void)hipFree(h->d_sin);
    delete h;
    return hip_check(e, "hb plan");
  }
  *out = h;
  return XA_OK;
}

extern "C" int xa_hb_plan_destroy(xa_hb_plan p) {
  if (!p) return XA_OK;
  (void)hipFree(p->d_cos);
  (void)hipFree(p->d_sin);
  delete p;
  return XA_OK;
}

extern "C" int xa_hb_ift(xa_hb_plan p, const double* xf, double* xt) {
  if (!p || !xf || !xt) return fail(XA_ERR_ARG, "hb ift: null argument");
  return hb_ift(p, xf, xt);
}

extern "C" int xa_hb_dft(xa_hb_plan p, const double* xt, double* xf) {
  if (!p || !xt || !xf) return fail(XA_ERR_ARG, "hb dft: null argument");
  return hb_dft(p, xt, xf);
}

extern "C" int xa_hb_add_jwq(xa_hb_plan p, double* ff, const double* qf) {
  if (!p || !ff || !qf) return fail(XA_ERR_ARG, "hb jwq: null argument");
  return hb_jwq(p, ff, qf);
}

// y = A x for a CSR matrix (device arrays): the Jacobian-vector products of
// the Newton-Krylov solves (for HB the time-domain dF/dx and dQ/dx applied to
// a direction, N_LOA_HBLoader.C:350-456).  One wavefront per row: lane l sums
// the row's products l, l+64, ... in slot order, then the 64 partial sums are
// added in a fixed tree, so the result does not depend on the launch -- and a
// supply's KCL row (one entry per transistor on the rail) is spread over the
// lanes instead of one thread.
__global__ void __launch_bounds__(256) k_csr_spmv(int n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ ci,
                                                  const double* __restrict__ a,
                                                  const double* __restrict__ x,
                                                  double* __restrict__ y) {
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n) return;
  double s = 0.0;
  for (int k = rp[r] + lane; k < rp[r + 1]; k += 64) s += a[k] * x[ci[k]];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) y[r] = s;
}

// y = A x for a matrix of few, long rows (a supply's KCL row): one
// 1024-thread block per row, thread-strided partial sums added in a fixed
// tree in LDS.
__global__ void __launch_bounds__(1024) k_csr_spmv_long(const int32_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci,
                                                        const double* __restrict__ a,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y) {
  __shared__ double part[1024];
  const int r = blockIdx.x;
  double s = 0.0;
  for (int k = rp[r] + (int)threadIdx.x; k < rp[r + 1]; k += 1024) s += a[k] * x[ci[k]];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) y[r] = part[0];
}

extern "C" int xa_csr_spmv_long(xa_ctx ctx, int n_rows, const int32_t* row_ptr,
                                const int32_t* col_idx, const double* values, const double* x,
                                double* y) {
  Ctx* cx = reinterpret_cast<Ctx*>(ctx);
  if (!cx || !row_ptr || !col_idx || !values || !x || !y || n_rows < 0)
    return fail(XA_ERR_ARG, "csr spmv (long rows): null argument or negative size");
  if (n_rows == 0) return XA_OK;
  hipLaunchKernelGGL(k_csr_spmv_long, dim3(n_rows), dim3(1024), 0, cx->stream, row_ptr, col_idx,
                     values, x, y);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

extern "C" int xa_csr_spmv(xa_ctx ctx, int n_rows, const int32_t* row_ptr, const int32_t* col_idx,
                           const double* values, const double* x, double* y) {
  Ctx* cx = reinterpret_cast<Ctx*>(ctx);
  if (!cx || !row_ptr || !col_idx || !values || !x || !y || n_rows < 0)
    return fail(XA_ERR_ARG, "csr spmv: null argument or negative size");
  if (n_rows == 0) return XA_OK;
  hipLaunchKernelGGL(k_csr_spmv, dim3((n_rows + 3) / 4), dim3(256), 0, cx->stream, n_rows,
                     row_ptr, col_idx, values, x, y);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

extern "C" int xa_hb_load(xa_hb_plan p, xa_group g, const xa_solver_state* ss, const double* xf,
                          const xa_vectors* vt, int64_t nnz_t, double* ff, double* qf,
                          double* dfdxdvp_f, double* dqdxdvp_f) {
  if (!p || !g || !ss || !xf || !vt) return fail(XA_ERR_ARG, "hb load: null argument");
  if (!vt->nextSol) return fail(XA_ERR_ARG, "hb load: the time-domain nextSol is required");
  if (nnz_t < 0) return fail(XA_ERR_ARG, "hb load: negative nnz");
  int rc;
  hipStream_t st = p->ctx->stream;
  // the solution vector the loads read: the caller's device buffer, which
  // this call fills (xa_vectors carries it as read-only for the loads)
  if ((rc = hb_ift(p, xf, const_cast<double*>(vt->nextSol))) != XA_OK) return rc;
  // the loads add: every time-domain output starts from zero (the reference
  // zeroes appQ, appF, ... before each sample's load, N_LOA_HBLoader.C:670-675)
  const int64_t nt = (int64_t)p->N * p->n;
  double* vecs[4] = {vt->f, vt->q, vt->dFdxdVp, vt->dQdxdVp};
  for (double* v : vecs)
    if (v) hipLaunchKernelGGL(k_hb_zero, dim3(grid_for(nt)), dim3(kBlock), 0, st, v, nt);
  double* mats[2] = {vt->dFdx, vt->dQdx};
  for (double* m : mats)
    if (m && nnz_t > 0)
      hipLaunchKernelGGL(k_hb_zero, dim3(std::min<int64_t>(grid_for(nnz_t), 65536)), dim3(kBlock),
                         0, st, m, nnz_t);
  XA_HIP(hipGetLastError());
  if ((rc = xa_load_all(g, ss, vt)) != XA_OK) return rc;
  if (ff && vt->f && (rc = hb_dft(p, vt->f, ff)) != XA_OK) return rc;
  if (qf && vt->q && (rc = hb_dft(p, vt->q, qf)) != XA_OK) return rc;
  if (dfdxdvp_f && vt->dFdxdVp && (rc = hb_dft(p, vt->dFdxdVp, dfdxdvp_f)) != XA_OK) return rc;
  if (dqdxdvp_f && vt->dQdxdVp && (rc = hb_dft(p, vt->dQdxdVp, dqdxdvp_f)) != XA_OK) return rc;
  if (ff && qf && (rc = hb_jwq(p, ff, qf)) != XA_OK) return rc;
  if (dfdxdvp_f && dqdxdvp_f && (rc = hb_jwq(p, dfdxdvp_f, dqdxdvp_f)) != XA_OK) return rc;
  return XA_OK;
}

extern "C" int xa_hb_precond_destroy(xa_hb_precond pc) {
  if (!pc) return XA_OK;
  void* bufs[] = {pc->d_row_ptr, pc->d_col_idx, pc->d_G, pc->d_C, pc->d_lu, pc->d_ipiv,
                  pc->d_info, pc->d_perm, pc->d_linv, pc->d_uinv, pc->d_b, pc->d_x, pc->d_r,
                  pc->d_rhs, pc->d_norm};
  for (void* b : bufs) (void)hipFree(b);
  delete pc;
  return XA_OK;
}

extern "C" int xa_hb_precond_create(xa_hb_plan p, int nnz, const int32_t* row_ptr,
                                    const int32_t* col_idx, xa_hb_precond* out) {
  if (!out) return fail(XA_ERR_ARG, "hb precond: null output handle");
  *out = nullptr;
  if (!p || !row_ptr || (nnz > 0 && !col_idx)) return fail(XA_ERR_ARG, "hb precond: null argument");
  const int n = p->n, m = 2 * n, ld = (m + 63) / 64 * 64;
  int rc = lu_check(1, m, ld);
  if (rc != XA_OK) return rc;
  if (nnz < 0 || row_ptr[0] != 0 || row_ptr[n] != nnz)
    return fail(XA_ERR_ARG, "hb precond: row_ptr must run from 0 to nnz");
  for (int i = 0; i < n; ++i)
    if (row_ptr[i + 1] < row_ptr[i]) return fail(XA_ERR_ARG, "hb precond: row_ptr decreases");
  for (int e = 0; e < nnz; ++e)
    if (col_idx[e] < 0 || col_idx[e] >= n) return fail(XA_ERR_ARG, "hb precond: column out of range");
  auto* pc = new xa_hb_precond_s();
  pc->plan = p;
  pc->nnz = nnz;
  pc->m = m;
  pc->ld = ld;
  pc->nh = p->H + 1;
  const size_t N = p->N;
  hipError_t e = hipSuccess;
  auto alloc = [&](void** ptr, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(ptr, std::max<size_t>(bytes, 8));
  };
  alloc((void**)&pc->d_row_ptr, sizeof(int32_t) * (n + 1));
  alloc((void**)&pc->d_col_idx, sizeof(int32_t) * nnz);
  alloc((void**)&pc->d_G, sizeof(double) * nnz);
  alloc((void**)&pc->d_C, sizeof(double) * nnz);
  alloc((void**)&pc->d_lu, sizeof(float) * pc->nh * (size_t)ld * ld);
  alloc((void**)&pc->d_ipiv, sizeof(int32_t) * pc->nh * (size_t)ld);
  alloc((void**)&pc->d_info, sizeof(int32_t) * pc->nh);
  alloc((void**)&pc->d_perm, sizeof(int32_t) * pc->nh * (size_t)ld);
  alloc((void**)&pc->d_linv, sizeof(float) * pc->nh * (size_t)ld * 64);
  alloc((void**)&pc->d_uinv, sizeof(float) * pc->nh * (size_t)ld * 64);
  alloc((void**)&pc->d_b, sizeof(double) * N * m);
  alloc((void**)&pc->d_x, sizeof(double) * N * m);
  alloc((void**)&pc->d_r, sizeof(double) * N * m);
  alloc((void**)&pc->d_rhs, sizeof(float) * pc->nh * 2 * (size_t)ld);
  alloc((void**)&pc->d_norm, sizeof(double) * 2 * N);
  if (e == hipSuccess)
    e = hipMemcpy(pc->d_row_ptr, row_ptr, sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz > 0)
    e = hipMemcpy(pc->d_col_idx, col_idx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    xa_hb_precond_destroy(pc);
    return hip_check(e, "hb precond");
  }
  *out = pc;
  return XA_OK;
}

extern "C" int xa_hb_precond_factor(xa_hb_precond pc, const double* dFdx_t, const double* dQdx_t,
                                    int32_t* info) {
  if (!pc || !dFdx_t || !dQdx_t) return fail(XA_ERR_ARG, "hb factor: null argument");
  HBPlan* h = pc->plan;
  hipStream_t st = h->ctx->stream;
  const double w = 2.0 * M_PI / h->period;
  if (pc->nnz > 0)
    hipLaunchKernelGGL(k_hb_average, hb_grid(pc->nnz, 1), dim3(256), 0, st, dFdx_t, dQdx_t,
                       pc->d_G, pc->d_C, h->N, pc->nnz);
  hipLaunchKernelGGL(k_hb_assemble, dim3(pc->ld, pc->nh), dim3(256), 0, st, pc->d_lu,
                     pc->d_row_ptr, pc->d_col_idx, pc->d_G, pc->d_C, h->n, pc->ld, w);
  XA_HIP(hipGetLastError());
  int rc = lu_factor(h->ctx, pc->nh, pc->m, pc->ld, pc->d_lu, pc->d_ipiv, pc->d_info);
  if (rc != XA_OK) return rc;
  rc = lu_prepare(h->ctx, pc->nh, pc->m, pc->ld, pc->d_lu, pc->d_ipiv, pc->d_perm, pc->d_linv,
                  pc->d_uinv);
  if (rc != XA_OK) return rc;
  if (info) {
    XA_HIP(hipMemcpyAsync(info, pc->d_info, sizeof(int32_t) * pc->nh, hipMemcpyDeviceToHost, st));
    XA_HIP(hipStreamSynchronize(st));
  }
  pc->factored = true;
  return XA_OK;
}

extern "C" int xa_hb_precond_solve(xa_hb_precond pc, const double* bf, double* xf,
                                   int max_refine, double tol, double* rel_fp32,
                                   double* rel_final, int32_t* iterations) {
  if (!pc || !bf || !xf) return fail(XA_ERR_ARG, "hb solve: null argument");
  if (!pc->factored) return fail(XA_ERR_STATE, "hb solve: factor first");
  if (max_refine < 0) return fail(XA_ERR_ARG, "hb solve: max_refine must be >= 0");
  HBPlan* h = pc->plan;
  hipStream_t st = h->ctx->stream;
  const int N = h->N, H = h->H, n = h->n, ld = pc->ld;
  const double w = 2.0 * M_PI / h->period;
  const size_t sys = (size_t)N * pc->m;
  std::vector<double> norms(2 * N), rel(N);
  // r = b - K x in FP64 and every system's max |r| / max |b|
  auto residual = [&](bool with_norms) -> int {
    hipLaunchKernelGGL(k_hb_residual, hb_grid(pc->m, N), dim3(256), 0, st, pc->d_row_ptr,
                       pc->d_col_idx, pc->d_G, pc->d_C, pc->d_b, pc->d_x, pc->d_r, N, H, n, w);
    XA_HIP(hipGetLastError());
    if (!with_norms) return XA_OK;
    hipLaunchKernelGGL(k_hb_norms, dim3(N), dim3(256), 0, st, pc->d_r, pc->d_b, pc->d_norm, n);
    XA_HIP(hipGetLastError());
    XA_HIP(hipMemcpyAsync(norms.data(), pc->d_norm, sizeof(double) * 2 * N, hipMemcpyDeviceToHost,
                          st));
    XA_HIP(hipStreamSynchronize(st));
    for (int q = 0; q < N; ++q)
      rel[q] = norms[2 * q + 1] > 0.0 ? norms[2 * q] / norms[2 * q + 1] : norms[2 * q];
    return XA_OK;
  };
  // x += K^-1 r through the FP32 factors
  auto correct = [&]() -> int {
    hipLaunchKernelGGL(k_hb_to_rhs, hb_grid(ld, N), dim3(256), 0, st, pc->d_r, pc->d_rhs, N, H, n,
                       ld);
    XA_HIP(hipGetLastError());
    int rc = lu_solve_prepared(h->ctx, pc->nh, 2, pc->m, ld, pc->d_lu, pc->d_perm, pc->d_linv,
                               pc->d_uinv, pc->d_rhs);
    if (rc != XA_OK) return rc;
    hipLaunchKernelGGL(k_hb_from_rhs, hb_grid(pc->m, N), dim3(256), 0, st, pc->d_rhs, pc->d_x, N,
                       H, n, ld);
    XA_HIP(hipGetLastError());
    return XA_OK;
  };
  hipLaunchKernelGGL(k_hb_gather, hb_grid(pc->m, N), dim3(256), 0, st, bf, pc->d_b, N, n);
  hipLaunchKernelGGL(k_hb_zero, dim3(grid_for((int64_t)sys)), dim3(kBlock), 0, st, pc->d_x,
                     (int64_t)sys);
  XA_HIP(hipGetLastError());
  int rc;
  if ((rc = residual(false)) != XA_OK) return rc;   // r = b
  if ((rc = correct()) != XA_OK) return rc;         // the FP32 solve
  if ((rc = residual(true)) != XA_OK) return rc;
  std::vector<int32_t> its(N, 0);
  // a system is done when its residual meets tol or stops falling (halving
  // no more: the FP64 residual's own rounding floor, or a factor whose
  // condition number times the FP32 precision is near one)
  std::vector<char> done(N, 0);
  if (rel_fp32) for (int q = 0; q < N; ++q) rel_fp32[q] = rel[q];
  for (int q = 0; q < N; ++q) done[q] = rel[q] <= tol;
  for (int it = 1; it <= max_refine; ++it) {
    bool all = true;
    for (int q = 0; q < N; ++q) all = all && done[q];
    if (all) break;
    if ((rc = correct()) != XA_OK) return rc;
    std::vector<double> before = rel;
    if ((rc = residual(true)) != XA_OK) return rc;
    for (int q = 0; q < N; ++q) {
      if (done[q]) continue;
      its[q] = it;   // refinements this system took
      if (rel[q] <= tol || !(rel[q] < 0.5 * before[q])) done[q] = 1;
    }
  }
  if (rel_final) for (int q = 0; q < N; ++q) rel_final[q] = rel[q];
  if (iterations) for (int q = 0; q < N; ++q) iterations[q] = its[q];
  hipLaunchKernelGGL(k_hb_scatter, hb_grid(pc->m, N), dim3(256), 0, st, pc->d_x, xf, N, n);
  XA_HIP(hipGetLastError());
  return XA_OK;
}
