// xa_lu.hip -- batched dense LU in FP32 on the matrix cores, and its solve.
//
// The mixed-precision linear solves of the batched forms (the harmonic-
// balance block-Jacobi blocks, N_LAS_HBBlockJacobiPrecond.C:283-370, which
// the reference factors with KLU in FP64) factor in FP32 and refine in FP64
// around it; this file is the FP32 half.  No reference counterpart as code:
// it is LAPACK's right-looking blocked LU with partial pivoting (dgetrf:
// dgetf2 on a panel, dlaswp, dtrsm, dgemm), laid out for CDNA4:
//
//   k_lu_panel      one workgroup per matrix factors the ld - k0 rows of a
//                   16-column panel held in registers (rows spread over 1024
//                   threads, three per thread); per column a block-wide argmax
//                   (first largest magnitude, as idamax), the interchange
//                   recorded, the scaling and the rank-1 update;
//   k_lu_swap_trsm  applies the panel's 16 row exchanges to every other
//                   column and forms U12 = L11^-1 A12 (one thread a column);
//   k_lu_update     A22 -= L21 U12 in 64x64 tiles with
//                   v_mfma_f32_16x16x4f32, the tiles spread over the GPU.
//
// Matrices are row-major, ld x ld with ld a multiple of 64 (at most 3072);
// rows and columns [m, ld) must hold the identity (a padded block factors as
// the matrix and an identity beside it).  ipiv[i] is the 0-based row that row
// i was exchanged with at step i; info = first zero pivot's column + 1.
#include <cmath>
#include "xa_internal.h"

namespace xa {

constexpr int kLuNB = 16;              // panel width
constexpr int kPanelThreads = 1024;
constexpr int kPanelQ = 3;             // panel rows per thread
constexpr int kLuMaxLd = kPanelThreads * kPanelQ;
constexpr int kLuTile = 64;            // trailing-update tile
constexpr int kLuOuter = 64;           // columns per trailing update

typedef float floatx4 __attribute__((ext_vector_type(4)));

// a fill kernel, not hipMemsetAsync: the factorization may be captured in a
// HIP graph
__global__ void k_lu_zero_i32(int32_t* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Shared state of one panel factorization.  The panel rows stay in the
// registers of the threads that loaded them; the row interchanges are kept as
// a permutation (s_pos: logical position of each loaded row, s_row: the
// loaded row at each logical position) and applied when the panel is written
// back, so no register moves between rows.
struct PanelShared {
  float piv[kLuNB];                  // the pivot row
  float val[kPanelThreads / 64];     // per-wavefront argmax: magnitude,
  int pos[kPanelThreads / 64];       // logical position,
  int row[kPanelThreads / 64];       // loaded row
  int ipiv[kLuNB];                   // panel-relative interchanges (LAPACK form)
  int s_pos[kLuMaxLd];
  int s_row[kLuMaxLd];
};

// better candidate: larger magnitude, then the earlier logical position
// (idamax takes the first largest)
__device__ __forceinline__ bool panel_better(float v, int pos, float bv, int bpos) {
  return v > bv || (v == bv && pos < bpos);
}

// Column C of the panel (a template, so that every register of the panel
// rows is addressed by a constant): argmax over the rows not yet pivots,
// the interchange recorded, then scaling and the rank-1 update of those rows.
template <int C>
__device__ __forceinline__ void panel_column(float (&p)[kPanelQ][kLuNB], unsigned& done,
                                             PanelShared& sh, int t, int R, int& first_zero) {
  const int lane = t & 63, wave = t >> 6;
  float best = -1.0f;
  int bpos = 0x7fffffff, brow = -1;
#pragma unroll
  for (int q = 0; q < kPanelQ; ++q) {
    const int r = t + q * kPanelThreads;
    if (r < R && !(done >> q & 1u)) {
      const float v = fabsf(p[q][C]);
      const int pos = sh.s_pos[r];
      if (panel_better(v, pos, best, bpos)) { best = v; bpos = pos; brow = r; }
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const float ov = __shfl_xor(best, off);
    const int op = __shfl_xor(bpos, off), orow = __shfl_xor(brow, off);
    if (panel_better(ov, op, best, bpos)) { best = ov; bpos = op; brow = orow; }
  }
  if (lane == 0) { sh.val[wave] = best; sh.pos[wave] = bpos; sh.row[wave] = brow; }
  __syncthreads();
  best = sh.val[0];
  bpos = sh.pos[0];
  brow = sh.row[0];
  for (int w = 1; w < kPanelThreads / 64; ++w)
    if (panel_better(sh.val[w], sh.pos[w], best, bpos)) {
      best = sh.val[w]; bpos = sh.pos[w]; brow = sh.row[w];
    }
  if (brow < 0) {   // no candidate (every one NaN): the row at position C
    brow = sh.s_row[C];
    bpos = C;
  }
  const int tp = brow % kPanelThreads, qp = brow / kPanelThreads;
  if (t == tp) {   // read by selects: an indexed read would put the rows in scratch
    float v[kLuNB];
#pragma unroll
    for (int cc = 0; cc < kLuNB; ++cc) {
      v[cc] = p[0][cc];
#pragma unroll
      for (int q = 1; q < kPanelQ; ++q) v[cc] = qp == q ? p[q][cc] : v[cc];
      sh.piv[cc] = v[cc];
    }
    done |= 1u << qp;
  }
  if (t == 0) {   // interchange positions C and bpos
    const int other = sh.s_row[C];
    sh.s_row[C] = brow;
    sh.s_row[bpos] = other;
    sh.s_pos[brow] = C;
    sh.s_pos[other] = bpos;
    sh.ipiv[C] = bpos;
  }
  __syncthreads();
  const float d = sh.piv[C];
  if (d == 0.0f) {
    if (!first_zero) first_zero = C + 1;
  } else {
#pragma unroll
    for (int q = 0; q < kPanelQ; ++q) {
      const int r = t + q * kPanelThreads;
      if (r < R && !(done >> q & 1u)) {
        const float l = p[q][C] / d;
        p[q][C] = l;
#pragma unroll
        for (int cc = C + 1; cc < kLuNB; ++cc) p[q][cc] -= l * sh.piv[cc];
      }
    }
  }
}

template <int C>
struct PanelColumns {
  static __device__ __forceinline__ void run(float (&p)[kPanelQ][kLuNB], unsigned& done,
                                             PanelShared& sh, int t, int R, int& first_zero) {
    panel_column<C>(p, done, sh, t, R, first_zero);
    PanelColumns<C + 1>::run(p, done, sh, t, R, first_zero);
  }
};
template <>
struct PanelColumns<kLuNB> {
  static __device__ __forceinline__ void run(float (&)[kPanelQ][kLuNB], unsigned&, PanelShared&,
                                             int, int, int&) {}
};

// four waves per SIMD: 128 registers for the three panel rows of a thread
__global__ void __launch_bounds__(kPanelThreads, 4)
k_lu_panel(float* __restrict__ a, int32_t* __restrict__ ipiv, int32_t* __restrict__ info, int ld,
           int k0) {
  const int bt = blockIdx.x;
  float* A = a + (size_t)bt * ld * ld;
  const int t = threadIdx.x;
  const int R = ld - k0;   // panel rows
  __shared__ PanelShared sh;
  float p[kPanelQ][kLuNB];
#pragma unroll
  for (int q = 0; q < kPanelQ; ++q) {
    const int r = t + q * kPanelThreads;
    if (r < R) {
      sh.s_pos[r] = r;
      sh.s_row[r] = r;
    }
#pragma unroll
    for (int c = 0; c < kLuNB; ++c) p[q][c] = r < R ? A[(size_t)(k0 + r) * ld + k0 + c] : 0.0f;
  }
  __syncthreads();
  unsigned done = 0;
  int first_zero = 0;
  PanelColumns<0>::run(p, done, sh, t, R, first_zero);
  // rows to their positions after the interchanges
#pragma unroll
  for (int q = 0; q < kPanelQ; ++q) {
    const int r = t + q * kPanelThreads;
    if (r < R) {
      const int pos = sh.s_pos[r];
#pragma unroll
      for (int c = 0; c < kLuNB; ++c) A[(size_t)(k0 + pos) * ld + k0 + c] = p[q][c];
    }
  }
  if (t < kLuNB) ipiv[(size_t)bt * ld + k0 + t] = k0 + sh.ipiv[t];
  if (t == 0 && first_zero && info[bt] == 0) info[bt] = k0 + first_zero;
}

// The row exchanges ipiv[k0, k0 + NB) on the columns [jlo, jhi) outside the
// block's own columns [k0, k0 + NB), then, on those right of it,
// U12 = L11^-1 A12 (L11 the block's unit lower triangle, staged in LDS).
// One thread a column.
template <int NB>
__global__ void __launch_bounds__(256)
k_lu_swap_trsm(float* __restrict__ a, const int32_t* __restrict__ ipiv, int ld, int k0, int jlo,
               int jhi) {
  const int bt = blockIdx.y;
  float* A = a + (size_t)bt * ld * ld;
  __shared__ float sL[NB][NB + 1];
  __shared__ int sp[NB];
  const int t = threadIdx.x;
  for (int x = t; x < NB * NB; x += 256) {
    const int r = x / NB, c = x % NB;
    sL[r][c] = A[(size_t)(k0 + r) * ld + k0 + c];
  }
  if (t < NB) sp[t] = ipiv[(size_t)bt * ld + k0 + t];
  __syncthreads();
  const int j = jlo + blockIdx.x * 256 + t;
  if (j >= jhi || (j >= k0 && j < k0 + NB)) return;
  for (int c = 0; c < NB; ++c) {
    const int r2 = sp[c];
    if (r2 != k0 + c) {
      float* x1 = &A[(size_t)(k0 + c) * ld + j];
      float* x2 = &A[(size_t)r2 * ld + j];
      const float tmp = *x1;
      *x1 = *x2;
      *x2 = tmp;
    }
  }
  if (j < k0) return;   // left of the block: the exchanges only
  float u[NB];
#pragma unroll
  for (int c = 0; c < NB; ++c) u[c] = A[(size_t)(k0 + c) * ld + j];
#pragma unroll
  for (int c = 1; c < NB; ++c) {
#pragma unroll
    for (int cc = 0; cc < c; ++cc) u[c] -= sL[c][cc] * u[cc];
  }
#pragma unroll
  for (int c = 0; c < NB; ++c) A[(size_t)(k0 + c) * ld + j] = u[c];
}

// A22 -= L21 U12 on rows [k0 + K, ld) and columns [k0 + K, jhi), L21 the
// columns [k0, k0 + K) below the block and U12 its rows right of it.  A 64x64
// tile per workgroup, a 32x32 quarter per wavefront as 2x2 blocks of the
// 16x16x4 FP32 MFMA: lane l feeds A(i = l % 16, k = l / 16),
// B(k = l / 16, j = l % 16) and holds D(i = 4 (l / 16) + v, j = l % 16),
// v = 0..3.  The extents are multiples of 16, so a 16x16 block lies wholly
// inside or outside.
template <int K>
__global__ void __launch_bounds__(256)
k_lu_update(float* __restrict__ a, int ld, int k0, int jhi, int ntj) {
  const int bt = blockIdx.y;
  const int s0 = k0 + K;
  const int ti = blockIdx.x / ntj, tj = blockIdx.x % ntj;
  const int i0 = s0 + ti * kLuTile, j0 = s0 + tj * kLuTile;
  float* A = a + (size_t)bt * ld * ld;
  __shared__ float Ls[kLuTile][K + 1];   // -L21 rows of the tile
  __shared__ float Us[K][kLuTile];       // U12 columns of the tile
  const int t = threadIdx.x;
  for (int x = t; x < kLuTile * K; x += 256) {
    const int r = x / K, c = x % K;
    Ls[r][c] = i0 + r < ld ? -A[(size_t)(i0 + r) * ld + k0 + c] : 0.0f;
  }
  for (int x = t; x < K * kLuTile; x += 256) {
    const int r = x / kLuTile, c = x % kLuTile;
    Us[r][c] = j0 + c < jhi ? A[(size_t)(k0 + r) * ld + j0 + c] : 0.0f;
  }
  __syncthreads();
  const int w = t >> 6, lane = t & 63;
  const int wr = 32 * (w >> 1), wc = 32 * (w & 1);
  if (i0 + wr >= ld || j0 + wc >= jhi) return;
  const int li = lane & 15, lk = lane >> 4;
  // blocks inside the updated range (uniform over the wavefront)
  const bool in_r[2] = {true, i0 + wr + 16 < ld}, in_c[2] = {true, j0 + wc + 16 < jhi};
  floatx4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        acc[bi][bj][v] = in_r[bi] && in_c[bj]
            ? A[(size_t)(i0 + wr + 16 * bi + 4 * lk + v) * ld + j0 + wc + 16 * bj + li] : 0.0f;
#pragma unroll
  for (int kk = 0; kk < K; kk += 4) {
    float av[2], bv[2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) av[bi] = Ls[wr + 16 * bi + li][kk + lk];
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) bv[bj] = Us[kk + lk][wc + 16 * bj + li];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[bi], bv[bj], acc[bi][bj], 0, 0, 0);
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
      if (in_r[bi] && in_c[bj]) {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          A[(size_t)(i0 + wr + 16 * bi + 4 * lk + v) * ld + j0 + wc + 16 * bj + li] =
              acc[bi][bj][v];
      }
}

// x = A^-1 b for up to kSolveRhs right-hand sides of one matrix per
// workgroup (matrix blockIdx.y, right-hand sides kSolveRhs * blockIdx.x on):
// the row exchanges, then forward (unit L) and backward (U) substitution in
// 64-row blocks -- the block's diagonal part solved by one wavefront from
// LDS, the rows beyond it updated by all 1024 threads, every row of the
// factor read once for all the right-hand sides.
constexpr int kSolveRhs = 2;
constexpr int kSolveThreads = 1024;

// The 64-column segment of a factor row (16-byte aligned: ld and the block
// start are multiples of 64) against the block's entries of each
// right-hand side, in 16-byte loads.
__device__ __forceinline__ void row_dot(const float* __restrict__ row,
                                        const float (*sx)[kLuMaxLd], int b0,
                                        float (&s)[kSolveRhs]) {
  const float4* r4 = reinterpret_cast<const float4*>(row);
  float4 a[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) a[k] = r4[k];
#pragma unroll
  for (int r = 0; r < kSolveRhs; ++r) {
    const float4* x4 = reinterpret_cast<const float4*>(&sx[r][b0]);
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const float4 x = x4[k];
      acc += a[k].x * x.x;
      acc += a[k].y * x.y;
      acc += a[k].z * x.z;
      acc += a[k].w * x.w;
    }
    s[r] = acc;
  }
}
__global__ void __launch_bounds__(kSolveThreads)
k_lu_solve(const float* __restrict__ lu, const int32_t* __restrict__ ipiv, float* __restrict__ b,
           int ld, int mr, int nrhs) {
  const int bt = blockIdx.y, r0 = blockIdx.x * kSolveRhs;
  const int nr = nrhs - r0 < kSolveRhs ? nrhs - r0 : kSolveRhs;
  const float* A = lu + (size_t)bt * ld * ld;
  __shared__ __attribute__((aligned(16))) float sx[kSolveRhs][kLuMaxLd];
  __shared__ float sd[64][65];
  const int t = threadIdx.x;
  for (int r = 0; r < nr; ++r) {
    const float* x = b + ((size_t)bt * nrhs + r0 + r) * ld;
    for (int i = t; i < ld; i += kSolveThreads) sx[r][i] = x[i];
  }
  __syncthreads();
  if (t < nr) {
    for (int i = 0; i < mr; ++i) {
      const int p = ipiv[(size_t)bt * ld + i];
      if (p != i) {
        const float tmp = sx[t][i];
        sx[t][i] = sx[t][p];
        sx[t][p] = tmp;
      }
    }
  }
  __syncthreads();
  for (int b0 = 0; b0 < mr; b0 += 64) {
    for (int xi = t; xi < 64 * 64; xi += kSolveThreads)
      sd[xi / 64][xi % 64] = A[(size_t)(b0 + xi / 64) * ld + b0 + xi % 64];
    __syncthreads();
    if (t < 64) {
      float xv[kSolveRhs];
      for (int r = 0; r < kSolveRhs; ++r) xv[r] = r < nr ? sx[r][b0 + t] : 0.0f;
      for (int i = 0; i < 64; ++i) {
        const float l = sd[t][i];
#pragma unroll
        for (int r = 0; r < kSolveRhs; ++r) {
          const float xi = __shfl(xv[r], i);
          if (t > i) xv[r] -= l * xi;
        }
      }
      for (int r = 0; r < nr; ++r) sx[r][b0 + t] = xv[r];
    }
    __syncthreads();
    for (int i = b0 + 64 + t; i < mr; i += kSolveThreads) {
      float s[kSolveRhs];
      row_dot(A + (size_t)i * ld + b0, sx, b0, s);
      for (int r = 0; r < nr; ++r) sx[r][i] -= s[r];
    }
    __syncthreads();
  }
  for (int b0 = mr - 64; b0 >= 0; b0 -= 64) {
    for (int xi = t; xi < 64 * 64; xi += kSolveThreads)
      sd[xi / 64][xi % 64] = A[(size_t)(b0 + xi / 64) * ld + b0 + xi % 64];
    __syncthreads();
    if (t < 64) {
      float xv[kSolveRhs];
      for (int r = 0; r < kSolveRhs; ++r) xv[r] = r < nr ? sx[r][b0 + t] : 0.0f;
      for (int i = 63; i >= 0; --i) {
        const float u = sd[t][i], d = sd[i][i];
#pragma unroll
        for (int r = 0; r < kSolveRhs; ++r) {
          if (t == i) xv[r] /= d;
          const float xi = __shfl(xv[r], i);
          if (t < i) xv[r] -= u * xi;
        }
      }
      for (int r = 0; r < nr; ++r) sx[r][b0 + t] = xv[r];
    }
    __syncthreads();
    for (int i = t; i < b0; i += kSolveThreads) {
      float s[kSolveRhs];
      row_dot(A + (size_t)i * ld + b0, sx, b0, s);
      for (int r = 0; r < nr; ++r) sx[r][i] -= s[r];
    }
    __syncthreads();
  }
  for (int r = 0; r < nr; ++r) {
    float* x = b + ((size_t)bt * nrhs + r0 + r) * ld;
    for (int i = t; i < ld; i += kSolveThreads) x[i] = sx[r][i];
  }
}

// ---------------------------------------------------------------------------
// Solve support for repeated solves with one factor (the refinement steps):
// the row exchanges composed into one permutation, and the inverses of the
// 64x64 diagonal blocks of L (unit lower) and U, so that a solve's block
// steps are small products instead of 64 dependent substitution steps.

// perm[i] = the original row that ends at position i after the exchanges
__global__ void __launch_bounds__(256)
k_lu_perm(const int32_t* __restrict__ ipiv, int32_t* __restrict__ perm, int ld, int mr) {
  const int bt = blockIdx.x;
  __shared__ int32_t idx[kLuMaxLd];
  for (int i = threadIdx.x; i < ld; i += 256) idx[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < mr; ++i) {
      const int p = ipiv[(size_t)bt * ld + i];
      const int tmp = idx[i];
      idx[i] = idx[p];
      idx[p] = tmp;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ld; i += 256) perm[(size_t)bt * ld + i] = idx[i];
}

// Inverse of diagonal block blockIdx.x of matrix blockIdx.y: of L (unit
// lower) for blockIdx.z = 0, of U for 1; thread c forms column c by
// substitution against the block staged in LDS.  inv: [batch][blocks][64][64].
__global__ void __launch_bounds__(64)
k_lu_diag_inverse(const float* __restrict__ lu, float* __restrict__ linv, float* __restrict__ uinv,
                  int ld, int nblk) {
  const int blk = blockIdx.x, bt = blockIdx.y, upper = blockIdx.z;
  const float* A = lu + (size_t)bt * ld * ld;
  __shared__ float sb[64][65];
  const int c = threadIdx.x, b0 = blk * 64;
  for (int x = c; x < 64 * 64; x += 64) sb[x / 64][x % 64] = A[(size_t)(b0 + x / 64) * ld + b0 + x % 64];
  __syncthreads();
  float col[64];
  float* out = (upper ? uinv : linv) + ((size_t)bt * nblk + blk) * 64 * 64;
  if (!upper) {
#pragma unroll
    for (int i = 0; i < 64; ++i) {
      float v = i == c ? 1.0f : 0.0f;
#pragma unroll
      for (int k = 0; k < i; ++k) v -= sb[i][k] * col[k];
      col[i] = i < c ? 0.0f : v;
    }
  } else {
#pragma unroll
    for (int i = 63; i >= 0; --i) {
      float v = i == c ? 1.0f : 0.0f;
#pragma unroll
      for (int k = i + 1; k < 64; ++k) v -= sb[i][k] * col[k];
      col[i] = i > c ? 0.0f : v / sb[i][i];
    }
  }
#pragma unroll
  for (int i = 0; i < 64; ++i) out[i * 64 + c] = col[i];
}

// x_blk = inv x_blk for every right-hand side: 1024 threads, sixteen lanes a
// row of the 64x64 inverse, four columns a lane, summed over the sixteen.
__device__ __forceinline__ void block_apply(const float* __restrict__ inv,
                                            float (*sx)[kLuMaxLd], int b0, int nr) {
  const int t = threadIdx.x, row = t >> 4, part = t & 15;
  float s[kSolveRhs];
  const float4 a = reinterpret_cast<const float4*>(inv + row * 64)[part];
#pragma unroll
  for (int r = 0; r < kSolveRhs; ++r) {
    const float4 x = reinterpret_cast<const float4*>(&sx[r][b0])[part];
    float acc = a.x * x.x;
    acc += a.y * x.y;
    acc += a.z * x.z;
    acc += a.w * x.w;
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 16);
    s[r] = acc;
  }
  __syncthreads();   // every lane has read the block before it is overwritten
  if (part == 0)
    for (int r = 0; r < nr; ++r) sx[r][b0 + row] = s[r];
  __syncthreads();
}

// The solve with the composed permutation and the inverted diagonal blocks.
__global__ void __launch_bounds__(kSolveThreads)
k_lu_solve_inv(const float* __restrict__ lu, const int32_t* __restrict__ perm,
               const float* __restrict__ linv, const float* __restrict__ uinv,
               float* __restrict__ b, int ld, int mr, int nrhs) {
  const int bt = blockIdx.y, r0 = blockIdx.x * kSolveRhs;
  const int nr = nrhs - r0 < kSolveRhs ? nrhs - r0 : kSolveRhs;
  const int nblk = mr / 64;
  const float* A = lu + (size_t)bt * ld * ld;
  __shared__ __attribute__((aligned(16))) float sx[kSolveRhs][kLuMaxLd];
  const int t = threadIdx.x;
  for (int r = 0; r < kSolveRhs; ++r) {
    const float* x = b + ((size_t)bt * nrhs + r0 + (r < nr ? r : 0)) * ld;
    for (int i = t; i < ld; i += kSolveThreads)
      sx[r][i] = r < nr ? x[perm[(size_t)bt * ld + i]] : 0.0f;
  }
  __syncthreads();
  for (int blk = 0; blk < nblk; ++blk) {
    const int b0 = blk * 64;
    block_apply(linv + ((size_t)bt * nblk + blk) * 64 * 64, sx, b0, kSolveRhs);
    for (int i = b0 + 64 + t; i < mr; i += kSolveThreads) {
      float s[kSolveRhs];
      row_dot(A + (size_t)i * ld + b0, sx, b0, s);
      for (int r = 0; r < kSolveRhs; ++r) sx[r][i] -= s[r];
    }
    __syncthreads();
  }
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int b0 = blk * 64;
    block_apply(uinv + ((size_t)bt * nblk + blk) * 64 * 64, sx, b0, kSolveRhs);
    for (int i = t; i < b0; i += kSolveThreads) {
      float s[kSolveRhs];
      row_dot(A + (size_t)i * ld + b0, sx, b0, s);
      for (int r = 0; r < kSolveRhs; ++r) sx[r][i] -= s[r];
    }
    __syncthreads();
  }
  for (int r = 0; r < nr; ++r) {
    float* x = b + ((size_t)bt * nrhs + r0 + r) * ld;
    for (int i = t; i < ld; i += kSolveThreads) x[i] = sx[r][i];
  }
}

int lu_check(int batch, int m, int ld) {
  if (batch < 0 || m < 0 || ld < m || ld % 64 != 0 || ld > kLuMaxLd)
    return fail(XA_ERR_ARG, "lu: need 0 <= m <= ld, ld a multiple of 64 and at most " +
                                std::to_string(kLuMaxLd));
  return XA_OK;
}

// rows factored: m rounded up to 64 (the identity beyond m factors trivially)
static int lu_rows(int m) { return (m + 63) / 64 * 64; }

int lu_factor(Ctx* cx, int batch, int m, int ld, float* a, int32_t* ipiv, int32_t* info) {
  int rc = lu_check(batch, m, ld);
  if (rc != XA_OK || batch == 0 || m == 0) return rc;
  const int mr = lu_rows(m);
  hipStream_t st = cx->stream;
  hipLaunchKernelGGL(k_lu_zero_i32, dim3(grid_for(batch)), dim3(kBlock), 0, st, info, (int64_t)batch);
  // Blocks of kLuOuter columns, each factored as four 16-column panels with
  // their exchanges, U12 and updates kept inside the block; then the block's
  // exchanges on the other columns, its U12 and one trailing update of depth
  // kLuOuter (a quarter of the trailing-matrix traffic of 16-deep updates).
  auto update = [&](auto kernel, int K, int k0, int jhi) {
    const int rows = ld - k0 - K, cols = jhi - k0 - K;
    if (rows <= 0 || cols <= 0) return;
    const int nti = (rows + kLuTile - 1) / kLuTile, ntj = (cols + kLuTile - 1) / kLuTile;
    hipLaunchKernelGGL(kernel, dim3(nti * ntj, batch), dim3(256), 0, st, a, ld, k0, jhi, ntj);
  };
  for (int K0 = 0; K0 < mr; K0 += kLuOuter) {
    for (int c0 = K0; c0 < K0 + kLuOuter; c0 += kLuNB) {
      hipLaunchKernelGGL(k_lu_panel, dim3(batch), dim3(kPanelThreads), 0, st, a, ipiv, info, ld,
                         c0);
      hipLaunchKernelGGL(k_lu_swap_trsm<kLuNB>, dim3(1, batch), dim3(256), 0, st, a, ipiv, ld, c0,
                         K0, K0 + kLuOuter);
      update(k_lu_update<kLuNB>, kLuNB, c0, K0 + kLuOuter);
    }
    hipLaunchKernelGGL(k_lu_swap_trsm<kLuOuter>, dim3((ld + 255) / 256, batch), dim3(256), 0, st,
                       a, ipiv, ld, K0, 0, ld);
    update(k_lu_update<kLuOuter>, kLuOuter, K0, ld);
  }
  XA_HIP(hipGetLastError());
  return XA_OK;
}

int lu_solve(Ctx* cx, int batch, int nrhs, int m, int ld, const float* lu, const int32_t* ipiv,
             float* b) {
  int rc = lu_check(batch, m, ld);
  if (rc != XA_OK || batch == 0 || m == 0 || nrhs <= 0) return rc;
  hipLaunchKernelGGL(k_lu_solve, dim3((nrhs + kSolveRhs - 1) / kSolveRhs, batch),
                     dim3(kSolveThreads), 0, cx->stream, lu, ipiv, b, ld, lu_rows(m), nrhs);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

// perm: batch x ld; linv, uinv: batch x (mr / 64) x 64 x 64 floats
int lu_prepare(Ctx* cx, int batch, int m, int ld, const float* lu, const int32_t* ipiv,
               int32_t* perm, float* linv, float* uinv) {
  int rc = lu_check(batch, m, ld);
  if (rc != XA_OK || batch == 0 || m == 0) return rc;
  const int mr = lu_rows(m), nblk = mr / 64;
  hipLaunchKernelGGL(k_lu_perm, dim3(batch), dim3(256), 0, cx->stream, ipiv, perm, ld, mr);
  hipLaunchKernelGGL(k_lu_diag_inverse, dim3(nblk, batch, 2), dim3(64), 0, cx->stream, lu, linv,
                     uinv, ld, nblk);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

int lu_solve_prepared(Ctx* cx, int batch, int nrhs, int m, int ld, const float* lu,
                      const int32_t* perm, const float* linv, const float* uinv, float* b) {
  int rc = lu_check(batch, m, ld);
  if (rc != XA_OK || batch == 0 || m == 0 || nrhs <= 0) return rc;
  hipLaunchKernelGGL(k_lu_solve_inv, dim3((nrhs + kSolveRhs - 1) / kSolveRhs, batch),
                     dim3(kSolveThreads), 0, cx->stream, lu, perm, linv, uinv, b, ld, lu_rows(m),
                     nrhs);
  XA_HIP(hipGetLastError());
  return XA_OK;
}

}  // namespace xa

using namespace xa;

extern "C" int xa_lu_f32_factor(xa_ctx ctx, int batch, int m, int ld, float* a, int32_t* ipiv,
                                int32_t* info) {
  Ctx* cx = reinterpret_cast<Ctx*>(ctx);
  if (!cx) return fail(XA_ERR_ARG, "null context");
  if (batch > 0 && m > 0 && (!a || !ipiv || !info)) return fail(XA_ERR_ARG, "lu: null array");
  return lu_factor(cx, batch, m, ld, a, ipiv, info);
}

extern "C" int xa_lu_f32_solve(xa_ctx ctx, int batch, int nrhs, int m, int ld, const float* lu,
                               const int32_t* ipiv, float* b) {
  Ctx* cx = reinterpret_cast<Ctx*>(ctx);
  if (!cx) return fail(XA_ERR_ARG, "null context");
  if (batch > 0 && m > 0 && nrhs > 0 && (!lu || !ipiv || !b))
    return fail(XA_ERR_ARG, "lu: null array");
  return lu_solve(cx, batch, nrhs, m, ld, lu, ipiv, b);
}
