// xa_runtime.hip -- C ABI implementation: contexts, instance groups, the
// setup sequence (parameters -> LIDs -> CSR slots -> scatter plans) and the
// fixed-order segmented scatter that assembles F, Q, limiter vectors and the
// dF/dx, dQ/dx CSR values without atomics.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <initializer_list>
#include <mutex>
#include <vector>
#include "xa_internal.h"
#include "xa_kernel.h"

namespace xa {

static thread_local std::string g_err;

// Scratch of the allocation-free load (xa_load_all_noalloc): the stamp value
// buffers and supply-rail partial sums of one call, in a file-scope device
// array that every GPU the library runs on holds until the process ends.
// Host code never passes `scratch` itself to a kernel (on the host it names a
// stand-in, not device memory): each device's address comes from
// hipGetSymbolAddress, taken once per device at context creation, and only
// that address is used.
constexpr size_t kScratchDoubles = (size_t)4 << 20;   // 32 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices];
static std::mutex g_scratch_mu;

// Resolves (once) the scratch address on `device`, which must be current.
static int resolve_scratch(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(XA_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_check(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return XA_OK;
}

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_check(hipError_t e, const char* what) {
  return fail(XA_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static const ModelImpl* impl_for(int type) {
  switch (type) {
    case XA_DIODE: return model_diode();
    case XA_RESISTOR: return model_resistor();
    case XA_CAPACITOR: return model_capacitor();
    case XA_VSOURCE: return model_vsrc();
    case XA_MOSFET1: return model_mosfet1();
    case XA_BJT: return model_bjt();
    case XA_BSIM4: return model_bsim4();
    default: return nullptr;
  }
}

static std::string upper(const char* s) {
  std::string r(s);
  for (auto& c : r) c = (char)toupper((unsigned char)c);
  return r;
}

static const ParamDef* find_param(const ParamDef* tab, int n, const std::string& name) {
  for (int k = 0; k < n; ++k)
    if (name == tab[k].name) return &tab[k];
  return nullptr;
}

template <typename T>
static int upload(T** dst, const std::vector<T>& src) {
  if (*dst) { (void)hipFree(*dst); *dst = nullptr; }
  if (src.empty()) return XA_OK;
  XA_HIP(hipMalloc((void**)dst, src.size() * sizeof(T)));
  XA_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return XA_OK;
}

// ---------------------------------------------------------------------------
// Fixed-order segmented scatter.  One thread per output target; the target's
// contributions are added in plan order starting from the value already in
// the output, which reproduces the sequential `*ptr += value` of the
// reference bit for bit.
// ---------------------------------------------------------------------------
// Segments are read eight sources at a time, all loads issued before the
// adds, which stay in segment order.  Segments longer than kLongSegment (a
// supply rail's diagonal collects every device on it) are left to
// k_scatter_long.
constexpr int kLongSegment = 64;
// Segments longer than kRailSegment (a supply rail of a large netlist) are
// not added one by one: an in-order FP64 sum is one dependent add (about 16
// cycles) per term, which would hold one wavefront for most of a pass.  They
// are summed in fixed chunks of kRailChunk contributions, each chunk in
// order; lane l of one wavefront adds chunk sums l, l + 64, ... in order;
// the 64 lane sums combine in a fixed pairwise tree and the total is added
// onto the output.  Deterministic (the same bits on every run), and within
// (kRailChunk + chunks / 64 + 7) rounding steps of the contributions'
// magnitude sum of the one-by-one order.
constexpr int64_t kRailSegment = 1024;
constexpr int kRailChunk = 256;

// One thread per target: the target's terms added onto the output in order.
template <typename I>
__device__ __forceinline__ void add_short(int64_t t, const int32_t* __restrict__ target, int32_t t0,
                                          bool contig, const I* __restrict__ seg,
                                          const I* __restrict__ src,
                                          const double* __restrict__ vals,
                                          double* __restrict__ out) {
  const I s1 = seg[t + 1];
  I j = seg[t];
  if (s1 - j > kLongSegment) return;   // k_scatter_long adds this one
  const int64_t o = contig ? t0 + t : target[t];
  double acc = out[o];
  constexpr int U = 8;
  for (; j + U <= s1; j += U) {
    I ix[U];
    double v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) ix[u] = src[j + u];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = vals[ix[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u];
  }
  for (; j < s1; ++j) acc += vals[src[j]];
  out[o] = acc;
}

template <typename I, bool CONTIG>
__global__ void __launch_bounds__(256)
k_scatter_add(int64_t nt, const int32_t* __restrict__ target, int32_t t0,
              const I* __restrict__ seg, const I* __restrict__ src,
              const double* __restrict__ vals, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  add_short<I>(t, target, t0, CONTIG, seg, src, vals, out);
}

// In-order sums of long runs of terms by one wavefront.  A round of ROUND
// terms is gathered (ROUND / 64 per lane), staged in LDS, and every lane adds
// the staged terms in order, each read as an LDS broadcast; the next round's
// gathers are in flight meanwhile.  The sum is the one-by-one sum (the same
// on every lane; lane 0 writes it).

// Orders the wavefront's LDS accesses on either side (one-wave blocks: no
// hardware barrier is needed, only the compiler's and the memory model's).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Term indices of the round at base (lanes past the end repeat the last
// term; the adds below stop at the end).
template <typename I, bool GATHER, int PER>
__device__ __forceinline__ void round_index(I (&ix)[PER], I base, I end,
                                            const I* __restrict__ src, int lane) {
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    I j = base + r * 64 + lane;
    j = j < end ? j : end - 1;
    ix[r] = GATHER ? src[j] : j;
  }
}

// acc + t[s0] + t[s0 + 1] + ... + t[s1 - 1] in order, t[j] = vals[src[j]]
// (GATHER) or vals[j].  The next round's terms and the index of the round
// after are loaded while a round is added.  stage: ROUND doubles of LDS.
template <typename I, bool GATHER, int ROUND>
__device__ __forceinline__ double ordered_sum(double acc, I s0, I s1, const I* __restrict__ src,
                                              const double* __restrict__ vals, double* stage) {
  constexpr int PER = ROUND / 64;
  if (s1 <= s0) return acc;
  const int lane = (int)threadIdx.x & 63;   // one wavefront per sum
  I ix[PER];
  double v[PER];
  round_index<I, GATHER, PER>(ix, s0, s1, src, lane);
#pragma unroll
  for (int r = 0; r < PER; ++r) v[r] = vals[ix[r]];
  if (s0 + ROUND < s1) round_index<I, GATHER, PER>(ix, s0 + ROUND, s1, src, lane);
  for (I base = s0; base < s1; base += ROUND) {
    wave_sync();   // the previous round's reads of stage are done
#pragma unroll
    for (int r = 0; r < PER; ++r) stage[r * 64 + lane] = v[r];
    wave_sync();
    if (base + ROUND < s1) {
#pragma unroll
      for (int r = 0; r < PER; ++r) v[r] = vals[ix[r]];
      if (base + 2 * ROUND < s1) round_index<I, GATHER, PER>(ix, base + 2 * ROUND, s1, src, lane);
    }
    const int cnt = (s1 - base) < ROUND ? (int)(s1 - base) : ROUND;
    int u = 0;
    for (; u + 64 <= cnt; u += 64) {
#pragma unroll
      for (int k = 0; k < 64; ++k) acc += stage[u + k];
    }
    for (; u < cnt; ++u) acc += stage[u];
  }
  return acc;
}

// One wavefront per long segment, added onto the output in segment order.
template <typename I, bool CONTIG>
__global__ void __launch_bounds__(64)
k_scatter_long(const int32_t* __restrict__ long_t, const int32_t* __restrict__ target, int32_t t0,
               const I* __restrict__ seg, const I* __restrict__ src,
               const double* __restrict__ vals, double* __restrict__ out) {
  constexpr int ROUND = (int)kRailSegment;   // a long segment is one round
  __shared__ double stage[ROUND];
  const int32_t t = long_t[blockIdx.x];
  const int64_t o = CONTIG ? (int64_t)t0 + t : target[t];
  const double acc = ordered_sum<I, true, ROUND>(out[o], seg[t], seg[t + 1], src, vals, stage);
  if (threadIdx.x == 0) out[o] = acc;
}

// The segmented plans of several kinds of one call (32-bit indices, distinct
// outputs) in one launch each for the short and the long segments: block b
// belongs to the kind whose block range holds it and does for that kind's
// plan what k_scatter_add / k_scatter_long do, so every sum is the same.
struct SegKind {
  int64_t nt;
  const int32_t* target;   // null: contiguous targets from t0
  int32_t t0;
  const int32_t* seg;
  const int32_t* src;
  const double* vals;
  double* out;
  const int32_t* long_t;
  int64_t b0, lb0;          // first short / long block of this kind
};
struct SegSet {
  SegKind k[NK];
  int n;
};

// the kind whose short-block (LONG false) or long-target range holds x
template <bool LONG>
__device__ __forceinline__ int seg_kind(const SegSet& s, int64_t x) {
  int k = 0;
  while (k + 1 < s.n && x >= (LONG ? s.k[k + 1].lb0 : s.k[k + 1].b0)) ++k;
  return k;
}

// The short and the long segments of the set in one launch: the first
// short_blocks blocks add the short segments (one thread per target), the
// rest give each of their four wavefronts one long target.
constexpr int kLongPerBlock = 4;
__global__ void __launch_bounds__(256) k_scatter_set(SegSet s, int64_t short_blocks,
                                                     int64_t n_long) {
  constexpr int ROUND = (int)kRailSegment;
  __shared__ double stage[kLongPerBlock][ROUND];
  if ((int64_t)blockIdx.x < short_blocks) {
    const SegKind& d = s.k[seg_kind<false>(s, blockIdx.x)];
    const int64_t t = ((int64_t)blockIdx.x - d.b0) * blockDim.x + threadIdx.x;
    if (t >= d.nt) return;
    add_short<int32_t>(t, d.target, d.t0, d.target == nullptr, d.seg, d.src, d.vals, d.out);
    return;
  }
  const int wave = (int)threadIdx.x >> 6;
  const int64_t L = ((int64_t)blockIdx.x - short_blocks) * kLongPerBlock + wave;
  if (L >= n_long) return;   // the whole wavefront
  const SegKind& d = s.k[seg_kind<true>(s, L)];
  const int32_t t = d.long_t[L - d.lb0];
  const int64_t o = d.target ? (int64_t)d.target[t] : (int64_t)d.t0 + t;
  const double acc = ordered_sum<int32_t, true, ROUND>(d.out[o], d.seg[t], d.seg[t + 1], d.src,
                                                        d.vals, stage[wave]);
  if ((threadIdx.x & 63) == 0) d.out[o] = acc;
}

// One wavefront per rail chunk of at most kRailChunk contributions, summed
// in order (from -0.0, the exact identity of IEEE addition, so the sum is
// the first term plus the rest).
template <typename I>
__global__ void __launch_bounds__(64)
k_rail_partial(const I* __restrict__ j0s, const I* __restrict__ j1s, const I* __restrict__ src,
               const double* __restrict__ vals, double* __restrict__ partial) {
  __shared__ double stage[kRailChunk];
  const int64_t c = blockIdx.x;
  const double acc = ordered_sum<I, true, kRailChunk>(-0.0, j0s[c], j1s[c], src, vals, stage);
  if (threadIdx.x == 0) partial[c] = acc;
}

// One wavefront per rail: lane l adds chunk sums l, l + 64, ... in order,
// then a fixed pairwise tree over the lanes; the total goes onto the output.
template <bool CONTIG>
__global__ void __launch_bounds__(64)
k_rail_sum(const int32_t* __restrict__ rail_t, const int32_t* __restrict__ rail_c,
           const int32_t* __restrict__ target, int32_t t0, const double* __restrict__ partial,
           double* __restrict__ out) {
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const int32_t t = rail_t[r];
  const int32_t c0 = rail_c[r], c1 = rail_c[r + 1];
  double s = -0.0;   // the identity of IEEE addition
  constexpr int R = 4;
  int32_t c = c0 + lane;
  for (; c + 64 * (R - 1) < c1; c += 64 * R) {
    double v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = partial[c + 64 * k];
#pragma unroll
    for (int k = 0; k < R; ++k) s += v[k];
  }
  for (; c < c1; c += 64) s += partial[c];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) {
    const int64_t o = CONTIG ? (int64_t)t0 + t : target[t];
    out[o] = out[o] + s;
  }
}

// Pattern plan (see Plan::pattern): one thread per instance.  Instance i's
// targets lie at t0 + i * P + offset; a block stages the output range from
// its first instance's first target to its last instance's last target in
// LDS (instance rows padded to P + 1 doubles against bank conflicts), every
// thread adds its instance's entries in entry order (the plan order of a
// private target) from the entry-major value rows, eight rows' loads issued
// before their adds, and the block writes the range back (slots of the range
// that no entry feeds are written back unchanged).  Same bits as the
// segmented scatter, with coalesced reads and writes and no index arrays.
__global__ void __launch_bounds__(256)
k_scatter_pattern(int64_t n, int P, int last, int E, int32_t t0, const int32_t* __restrict__ poff,
                  const double* __restrict__ vals, double* __restrict__ out) {
  extern __shared__ double blk[];   // blockDim.x * (P + 1)
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
  const int m = (int)((n - i0) < (int64_t)blockDim.x ? (n - i0) : (int64_t)blockDim.x);
  const int W = P + 1;
  const int span = (m - 1) * P + last + 1;
  const float inv_p = 1.0f / (float)P;
  double* o = out + t0 + i0 * P;
  for (int k = threadIdx.x; k < span; k += blockDim.x) blk[staged_slot(k, P, inv_p)] = o[k];
  __syncthreads();
  const int il = threadIdx.x;
  if (il < m) {
    const double* vi = vals + i0 + il;
    double* row = blk + il * W;
    int e = 0;
    constexpr int U = 8;   // value rows in flight per thread
    for (; e + U <= E; e += U) {
      double v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) v[u] = vi[(int64_t)(e + u) * n];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int off = poff[e + u];
        if (off >= 0) row[off] += v[u];
      }
    }
    for (; e < E; ++e) {
      const int off = poff[e];
      if (off >= 0) row[off] += vi[(int64_t)e * n];
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < span; k += blockDim.x) o[k] = blk[staged_slot(k, P, inv_p)];
}

template <typename I, bool CONTIG>
__global__ void __launch_bounds__(256)
k_scatter_assign(int64_t nt, const int32_t* __restrict__ target, int32_t t0,
                 const I* __restrict__ seg, const I* __restrict__ src,
                 const double* __restrict__ vals, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nt) return;
  const int64_t o = CONTIG ? t0 + t : target[t];
  out[o] = vals[src[seg[t + 1] - 1]];
}

template <typename I, bool CONTIG>
static void launch_scatter(const Plan& p, bool assign, hipStream_t st, double* out) {
  const I* seg = static_cast<const I*>(p.d_seg);
  const I* src = static_cast<const I*>(p.d_src);
  if (assign)
    hipLaunchKernelGGL((k_scatter_assign<I, CONTIG>), dim3(grid_for(p.n_targets)), dim3(kBlock), 0,
                       st, p.n_targets, p.d_target, p.t0, seg, src, p.d_vals, out);
  else if (p.pattern) {
    const int64_t n = p.n_targets / p.pS;
    const size_t lds = sizeof(double) * (size_t)p.pblock * (size_t)(p.pP + 1);
    hipLaunchKernelGGL(k_scatter_pattern, dim3(grid_for(n, p.pblock)), dim3(p.pblock), lds, st, n,
                       p.pP, p.plast, p.E, p.t0, p.d_pslot, p.d_vals, out);
  } else {
    hipLaunchKernelGGL((k_scatter_add<I, CONTIG>), dim3(grid_for(p.n_targets)), dim3(kBlock), 0,
                       st, p.n_targets, p.d_target, p.t0, seg, src, p.d_vals, out);
    if (p.n_long > 0)
      hipLaunchKernelGGL((k_scatter_long<I, CONTIG>), dim3((unsigned)p.n_long), dim3(64), 0, st,
                         p.d_long, p.d_target, p.t0, seg, src, p.d_vals, out);
    if (p.n_rails > 0) {
      hipLaunchKernelGGL((k_rail_partial<I>), dim3((unsigned)p.n_chunks), dim3(64), 0, st,
                         static_cast<const I*>(p.d_chunk_j0),
                         static_cast<const I*>(p.d_chunk_j1), src, p.d_vals, p.d_partial);
      hipLaunchKernelGGL((k_rail_sum<CONTIG>), dim3((unsigned)p.n_rails), dim3(64), 0, st,
                         p.d_rail_t, p.d_rail_c, p.d_target, p.t0, p.d_partial, out);
    }
  }
}

static double* kind_out(const xa_vectors& v, int k) {
  switch (k) {
    case K_F: return v.f;
    case K_Q: return v.q;
    case K_FLIM: return v.dFdxdVp;
    case K_QLIM: return v.dQdxdVp;
    case K_DFDX: return v.dFdx;
    case K_DQDX: return v.dQdx;
    case K_LEADF: return v.storeLeadF;
    case K_LEADQ: return v.storeLeadQ;
  }
  return nullptr;
}

static int run_scatter(Group& g, const xa_solver_state& ss, const xa_vectors& v,
                       int k) {
  Plan& p = g.plan[k];
  double* out = kind_out(v, k);
  if (!out || p.n_targets == 0) return XA_OK;
  if (g.impl->skip_kind && g.impl->skip_kind(g, ss, k)) return XA_OK;
  const bool assign = kind_is_assign(k);
  const bool contig = p.d_target == nullptr;
  if (p.idx32) {
    if (contig) launch_scatter<int32_t, true>(p, assign, g.ctx->stream, out);
    else launch_scatter<int32_t, false>(p, assign, g.ctx->stream, out);
  } else {
    if (contig) launch_scatter<int64_t, true>(p, assign, g.ctx->stream, out);
    else launch_scatter<int64_t, false>(p, assign, g.ctx->stream, out);
  }
  XA_HIP(hipGetLastError());
  return XA_OK;
}

void fill_kargs(Group& g, const xa_solver_state& ss, const xa_vectors& v,
                bool fused, KArgs& a) {
  memset(&a, 0, sizeof(a));
  a.n = g.n;
  a.dfield = g.d_dfield;
  a.ifield = g.d_ifield;
  a.mcard = g.d_mcard;
  a.mstride = g.mcard_stride;
  a.n_models = (int)g.models.size();
  for (int w = 0; w < 3; ++w) a.ucol[w] = g.ucol[w];
  a.blk_cards = g.d_blk_cards;
  a.card_slot = g.d_card_slot;
  a.nodes = g.d_nodes;
  a.sto = g.d_sto;
  a.sta = g.d_sta;
  a.state = g.d_state;
  a.fail = g.d_fail;
  a.ss = ss;
  a.opt = g.opt;
  a.v = v;
  a.kmask = 0;
  if (fused) {
    a.fblk = g.d_fblk;
    a.frec = g.d_frec;
    a.fsrc = g.d_fsrc;
    a.flblk = g.d_flblk;
    a.flong = g.d_flong;
  }
  for (int k = 0; k < NK; ++k) {
    const Plan& p = g.plan[k];
    a.vals[k] = p.d_vals;
    a.E[k] = p.E;
    double* o = kind_out(v, k);
    a.out[k] = o;
    bool on = p.E > 0;
    // fused mode stages into LDS and writes out by the block plan: needs an
    // output, a live target and the kind not skipped on this call
    if (fused) {
      on = on && o != nullptr && p.n_targets > 0 &&
           !(g.impl->skip_kind && g.impl->skip_kind(g, ss, k));
      a.loff[k] = p.loff;
    }
    if (on) a.kmask |= 1u << k;
  }
}

int launch_model_kernel(Group& g, const KArgs& a, bool fused, const void* k_stored,
                        const void* k_fused) {
  void* args[] = {(void*)&a};
  if (!fused) {
    XA_HIP(hipLaunchKernel(k_stored, dim3(grid_for(g.n, g.split_block)), dim3(g.split_block), args,
                           0, g.ctx->stream));
    return XA_OK;
  }
  const int B = g.fused_B;
  if (g.fused_lds > 65536)
    XA_HIP(hipFuncSetAttribute(k_fused, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)g.fused_lds));
  XA_HIP(hipLaunchKernel(k_fused, dim3(grid_for(g.n, B)), dim3(B), args, g.fused_lds,
                         g.ctx->stream));
  return XA_OK;
}

int launch_pattern_kernel(Group& g, KArgs& a, const void* kernel) {
  const int B = g.pattern_B;
  int base = 0;
  for (int k = 0; k < NK; ++k) {
    const Plan& p = g.plan[k];
    a.poff[k] = p.d_pslot;
    a.pbase[k] = base;
    a.pP[k] = p.pP;
    a.plast[k] = p.plast;
    a.pt0[k] = p.t0;
    if (p.pattern) base += B * (p.pP + 1);
  }
  const size_t lds = sizeof(double) * (size_t)base;
  if (lds > 65536)
    XA_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  void* args[] = {(void*)&a};
  XA_HIP(hipLaunchKernel(kernel, dim3(grid_for(g.n, B)), dim3(B), args, lds, g.ctx->stream));
  return XA_OK;
}

}  // namespace xa

using namespace xa;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" {

const char* xa_last_error(void) { return g_err.c_str(); }
int xa_abi_version(void) { return XA_ABI_VERSION; }

void xa_default_options(xa_device_options* o) {
  memset(o, 0, sizeof(*o));
  o->gmin = 1.0e-12;
  o->tnom = kRefTemp;
  o->temp = kRefTemp;
  o->defl = 1.0e-4;
  o->defw = 1.0e-4;
  o->defad = 0.0;
  o->defas = 0.0;
  o->excessPhaseScalar1 = 1.0;
  o->excessPhaseScalar2 = 1.0;
  o->voltageLimiterFlag = 1;
  o->newMeyerFlag = 0;
  o->newExcessPhase = 0;
  o->lambertWFlag = 0;
}

void xa_default_solver_state(xa_solver_state* s) {
  memset(s, 0, sizeof(*s));
  s->gainScale = 1.0;
  s->nltermScale = 1.0;
}

int xa_ctx_create(int device_ordinal, xa_ctx* out) {
  if (!out) return fail(XA_ERR_ARG, "xa_ctx_create: null out");
  int ndev = 0;
  XA_HIP(hipGetDeviceCount(&ndev));
  if (device_ordinal < 0 || device_ordinal >= ndev)
    return fail(XA_ERR_ARG, "xa_ctx_create: no such device");
  XA_HIP(hipSetDevice(device_ordinal));
  double* scr = nullptr;
  int rc = resolve_scratch(device_ordinal, &scr);
  if (rc != XA_OK) return rc;
  Ctx* c = new Ctx();
  c->device = device_ordinal;
  c->scratch = scr;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return hip_check(e, "hipStreamCreate"); }
  c->own_stream = true;
  *out = reinterpret_cast<xa_ctx>(c);
  return XA_OK;
}

int xa_ctx_set_stream(xa_ctx ctx, void* stream) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(XA_ERR_ARG, "null ctx");
  if (c->own_stream && c->stream) XA_HIP(hipStreamDestroy(c->stream));
  c->stream = reinterpret_cast<hipStream_t>(stream);
  c->own_stream = false;
  return XA_OK;
}

int xa_ctx_destroy(xa_ctx ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return XA_OK;
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return XA_OK;
}

int xa_ctx_sync(xa_ctx ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(XA_ERR_ARG, "null ctx");
  XA_HIP(hipStreamSynchronize(c->stream));
  return XA_OK;
}

void* xa_ctx_stream(xa_ctx ctx) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  return c ? (void*)c->stream : nullptr;
}

int xa_device_alloc(xa_ctx ctx, size_t bytes, void** out) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c || !out) return fail(XA_ERR_ARG, "xa_device_alloc: null argument");
  *out = nullptr;
  if (bytes == 0) return XA_OK;
  XA_HIP(hipSetDevice(c->device));
  XA_HIP(hipMalloc(out, bytes));
  return XA_OK;
}

int xa_device_free(xa_ctx ctx, void* ptr) {
  Ctx* c = reinterpret_cast<Ctx*>(ctx);
  if (!c) return fail(XA_ERR_ARG, "xa_device_free: null context");
  if (!ptr) return XA_OK;
  XA_HIP(hipStreamSynchronize(c->stream));
  XA_HIP(hipFree(ptr));
  return XA_OK;
}

static int copy_sync(Ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (!c || ((!dst || !src) && bytes)) return fail(XA_ERR_ARG, "xa_copy: null argument");
  if (bytes == 0) return XA_OK;
  XA_HIP(hipMemcpyAsync(dst, src, bytes, kind, c->stream));
  XA_HIP(hipStreamSynchronize(c->stream));
  return XA_OK;
}

int xa_copy_to_device(xa_ctx ctx, void* dst, const void* src, size_t bytes) {
  return copy_sync(reinterpret_cast<Ctx*>(ctx), dst, src, bytes, hipMemcpyHostToDevice);
}

int xa_copy_to_host(xa_ctx ctx, void* dst, const void* src, size_t bytes) {
  return copy_sync(reinterpret_cast<Ctx*>(ctx), dst, src, bytes, hipMemcpyDeviceToHost);
}

// ---------------------------------------------------------------------------
// group setup
// ---------------------------------------------------------------------------
int xa_group_create(xa_ctx ctx, int model_type, const xa_device_options* opt,
                    xa_group* out) {
  const ModelImpl* impl = impl_for(model_type);
  if (!impl) return fail(XA_ERR_ARG, "xa_group_create: unknown model type");
  if (!out) return fail(XA_ERR_ARG, "xa_group_create: null out");
  Group* g = new Group();
  g->ctx = reinterpret_cast<Ctx*>(ctx);
  g->impl = impl;
  if (opt) g->opt = *opt; else xa_default_options(&g->opt);
  g->ival.assign(impl->n_islots, {});
  g->igiven.assign(impl->n_islots, 0);
  for (int k = 0; k < NK; ++k) g->ent[k] = impl->max_entries[k];
  *out = reinterpret_cast<xa_group>(g);
  return XA_OK;
}

static void free_flat_plan(Group& g);

static void free_plan(Plan& p) {
  (void)hipFree(p.d_target); (void)hipFree(p.d_seg); (void)hipFree(p.d_src);
  (void)hipFree(p.d_long); (void)hipFree(p.d_rail_t); (void)hipFree(p.d_rail_c);
  (void)hipFree(p.d_chunk_j0); (void)hipFree(p.d_chunk_j1); (void)hipFree(p.d_partial);
  (void)hipFree(p.d_vals); (void)hipFree(p.d_pslot);
  p = Plan();
}

int xa_group_destroy(xa_group gh) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return XA_OK;
  (void)hipFree(g->d_dfield); (void)hipFree(g->d_ifield); (void)hipFree(g->d_mcard);
  (void)hipFree(g->d_blk_cards); (void)hipFree(g->d_card_slot);
  (void)hipFree(g->d_nodes); (void)hipFree(g->d_sto); (void)hipFree(g->d_sta);
  (void)hipFree(g->d_state); (void)hipFree(g->d_fail);
  for (int k = 0; k < NK; ++k) free_plan(g->plan[k]);
  free_flat_plan(*g);
  for (auto& v : g->ev) for (auto& e : v) (void)hipEventDestroy(e);
  delete g;
  return XA_OK;
}

int xa_group_add_model(xa_group gh, int level, int n_params, const char* const* names,
                       const double* values, int* model_index) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase != PH_CREATED) return fail(XA_ERR_STATE, "add_model after setup");
  const ModelImpl* m = g->impl;
  Card c;
  c.level = level;
  c.v.assign(m->n_mslots, 0.0);
  c.given.assign(m->n_mslots, 0);
  // defaults: first table row of each slot carries the default
  std::vector<char> seen(m->n_mslots, 0);
  for (int k = 0; k < m->n_mparams; ++k) {
    const ParamDef& d = m->mparams[k];
    if (!seen[d.index]) { c.v[d.index] = d.def; seen[d.index] = 1; }
  }
  for (int p = 0; p < n_params; ++p) {
    std::string nm = upper(names[p]);
    const ParamDef* d = find_param(m->mparams, m->n_mparams, nm);
    if (!d) return fail(XA_ERR_ARG, std::string(m->name) + ": unknown model parameter " + nm);
    c.v[d->index] = values[p] + (d->is_temp ? kCtoK : 0.0);
    c.given[d->index] = 1;
  }
  g->models.push_back(std::move(c));
  if (model_index) *model_index = (int)g->models.size() - 1;
  return XA_OK;
}

int xa_group_add_instances(xa_group gh, int n, const int* model_index, int n_params,
                           const char* const* names, const double* const* values) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase != PH_CREATED) return fail(XA_ERR_STATE, "add_instances after setup");
  if (n < 0) return fail(XA_ERR_ARG, "negative instance count");
  const ModelImpl* m = g->impl;
  const int64_t n0 = g->n;
  // validate names first
  std::vector<const ParamDef*> defs(n_params);
  for (int p = 0; p < n_params; ++p) {
    defs[p] = find_param(m->iparams, m->n_iparams, upper(names[p]));
    if (!defs[p]) return fail(XA_ERR_ARG, std::string(m->name) + ": unknown instance parameter " + names[p]);
  }
  for (int i = 0; i < n; ++i) {
    const int mi = model_index ? model_index[i] : 0;
    if (mi < 0 || mi >= (int)g->models.size())
      return fail(XA_ERR_ARG, "instance refers to a model index not added");
  }
  if (n0 > 0) {
    // a later batch must give the same parameter set (given flags are per group)
    for (int s = 0; s < m->n_islots; ++s) {
      bool in_batch = false;
      for (int p = 0; p < n_params; ++p) in_batch |= defs[p]->index == s;
      if (in_batch != (bool)g->igiven[s])
        return fail(XA_ERR_ARG, "instance batches must give the same parameters");
    }
  }
  g->n = n0 + n;
  for (int i = 0; i < n; ++i) g->inst_model.push_back(model_index ? model_index[i] : 0);
  std::vector<char> seen(m->n_islots, 0);
  for (int k = 0; k < m->n_iparams; ++k) {
    const ParamDef& d = m->iparams[k];
    if (seen[d.index]) continue;
    seen[d.index] = 1;
    g->ival[d.index].resize(g->n, d.def);
  }
  for (int p = 0; p < n_params; ++p) {
    const ParamDef* d = defs[p];
    g->igiven[d->index] = 1;
    for (int i = 0; i < n; ++i)
      g->ival[d->index][n0 + i] = values[p][i] + (d->is_temp ? kCtoK : 0.0);
  }
  g->lead.resize(g->n, 0);
  return XA_OK;
}

int xa_group_request_lead_current(xa_group gh, const int32_t* flags) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase != PH_CREATED) return fail(XA_ERR_STATE, "lead current request after setup");
  g->lead.assign(g->n, 0);
  if (flags) for (int64_t i = 0; i < g->n; ++i) g->lead[i] = flags[i] != 0;
  return XA_OK;
}

int xa_group_setup(xa_group gh) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase != PH_CREATED) return fail(XA_ERR_STATE, "setup called twice");
  if (g->models.empty()) return fail(XA_ERR_STATE, "setup without a model card");
  g->n_ext.assign(g->n, 0); g->n_int.assign(g->n, 0);
  g->n_sta.assign(g->n, 0); g->n_sto.assign(g->n, 0);
  int rc = g->impl->setup(*g);
  if (rc != XA_OK) return rc;
  g->phase = PH_SETUP;
  return XA_OK;
}

int xa_group_num_instances(xa_group gh) {
  Group* g = reinterpret_cast<Group*>(gh);
  return g ? (int)g->n : -1;
}

int xa_group_counts(xa_group gh, int* n_ext, int* n_int, int* n_sta, int* n_sto) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase < PH_SETUP) return fail(XA_ERR_STATE, "counts before setup");
  for (int64_t i = 0; i < g->n; ++i) {
    if (n_ext) n_ext[i] = g->n_ext[i];
    if (n_int) n_int[i] = g->n_int[i];
    if (n_sta) n_sta[i] = g->n_sta[i];
    if (n_sto) n_sto[i] = g->n_sto[i];
  }
  return XA_OK;
}

int xa_group_register_lids(xa_group gh, const int32_t* node_lids, const int32_t* sta_lids,
                           const int32_t* sto_lids) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase != PH_SETUP) return fail(XA_ERR_STATE, "register_lids out of order");
  const ModelImpl* m = g->impl;
  const int64_t n = g->n;
  g->nodes.assign((size_t)m->max_nodes * n, -1);
  g->sto_lid.assign((size_t)std::max(m->max_sto, 1) * n, -1);
  g->sta_lid.assign((size_t)std::max(m->max_sta, 1) * n, -1);
  std::vector<int32_t> full(m->max_nodes);
  int64_t off = 0, osta = 0, osto = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int nreg = g->n_ext[i] + g->n_int[i];
    m->resolve_nodes(*g, i, node_lids + off, full.data());
    for (int j = 0; j < m->max_nodes; ++j) g->nodes[j * n + i] = full[j];
    off += nreg;
    for (int j = 0; j < g->n_sta[i]; ++j) {
      if (j < m->max_sta) g->sta_lid[j * n + i] = sta_lids ? sta_lids[osta + j] : -1;
    }
    osta += g->n_sta[i];
    for (int j = 0; j < g->n_sto[i]; ++j) {
      if (j < m->max_sto) g->sto_lid[j * n + i] = sto_lids ? sto_lids[osto + j] : -1;
    }
    osto += g->n_sto[i];
  }
  g->phase = PH_LIDS;
  return XA_OK;
}

int xa_group_jac_pairs(xa_group gh, int64_t* count, int32_t* rows, int32_t* cols) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase < PH_LIDS) return fail(XA_ERR_STATE, "jac_pairs before register_lids");
  const int64_t n = g->n;
  int64_t c = 0;
  std::vector<Stamp> st;
  for (int k : {K_DFDX, K_DQDX}) {
    for (int64_t i = 0; i < n; ++i) {
      g->impl->stamps(*g, i, k, st);
      for (const Stamp& s : st) {
        if (s.row < 0 || s.col < 0) continue;   // entry not loaded by this instance
        const int32_t r = g->nodes[s.row * n + i], q = g->nodes[s.col * n + i];
        if (r < 0 || q < 0) continue;
        if (rows) { rows[c] = r; cols[c] = q; }
        ++c;
      }
    }
  }
  if (count) *count = c;
  return XA_OK;
}

int xa_group_register_csr(xa_group gh, int n_rows, const int32_t* row_ptr,
                          const int32_t* col_idx) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase != PH_LIDS) return fail(XA_ERR_STATE, "register_csr out of order");
  const ModelImpl* m = g->impl;
  const int64_t n = g->n;
  std::vector<Stamp> st;
  for (int k = 0; k < NK; ++k) {
    const int E = g->ent[k];
    g->tgt[k].assign((size_t)E * n, -1);
    for (int64_t i = 0; i < n; ++i) {
      g->impl->stamps(*g, i, k, st);
      if ((int)st.size() > E) return fail(XA_ERR_STATE, "stamp list exceeds the group's entries");
      for (size_t e = 0; e < st.size(); ++e) {
        int32_t t = -1;
        if (st[e].row < 0 || (kind_is_matrix(k) && st[e].col < 0)) {
          // entry not loaded by this instance: no target
        } else if (kind_is_assign(k)) {
          t = (st[e].row < m->max_sto) ? g->sto_lid[st[e].row * n + i] : -1;
        } else if (!kind_is_matrix(k)) {
          t = g->nodes[st[e].row * n + i];
        } else {
          const int32_t r = g->nodes[st[e].row * n + i], q = g->nodes[st[e].col * n + i];
          if (r >= 0 && q >= 0) {
            if (r >= n_rows) return fail(XA_ERR_ARG, "stamp row outside the CSR structure");
            const int32_t* b = col_idx + row_ptr[r];
            const int32_t* en = col_idx + row_ptr[r + 1];
            const int32_t* f = std::lower_bound(b, en, q);
            if (f == en || *f != q) {
              // unsorted row: fall back to a linear search
              f = std::find(b, en, q);
              if (f == en) return fail(XA_ERR_ARG, "stamp (row,col) missing from CSR structure");
            }
            t = (int32_t)(f - col_idx);
          }
        }
        g->tgt[k][i * E + e] = t;
      }
    }
  }
  g->phase = PH_CSR;
  return XA_OK;
}

// Block plan of the fused load.  The fused kernel runs blocks of B
// instances; when every target of a kind receives contributions from the
// instances of one block only, the block can add them itself, in the global
// plan's order, from LDS.  One kind's plan on the host: its targets grouped by
// block (ascending within a block), a source becoming the LDS slot
// (loff + e) * B + lane; loff is set by the caller before (the kind's first
// LDS row).  The kinds' plans are then joined into the group's flat plan.
struct BlockHost {
  std::vector<int32_t> bblk, btgt, bseg;
  std::vector<uint16_t> bsrc;
};

static void build_block_plan(Group& g, Plan& p, const std::vector<int32_t>& target,
                             const std::vector<int64_t>& seg, const std::vector<int64_t>& src,
                             int B, BlockHost& h) {
  h = BlockHost();
  const int64_t n = g.n, nt = (int64_t)target.size();
  p.block_private = B > 0;
  if (!p.block_private || nt == 0) return;
  const int64_t nb = (n + B - 1) / B;
  std::vector<int32_t> tblock(nt);
  for (int64_t t = 0; t < nt && p.block_private; ++t) {
    const int64_t b = (src[seg[t]] % n) / B;
    for (int64_t j = seg[t]; j < seg[t + 1]; ++j)
      if ((src[j] % n) / B != b) { p.block_private = false; break; }
    tblock[t] = (int32_t)b;
  }
  if (!p.block_private) return;
  h.bblk.assign(nb + 1, 0);
  for (int64_t t = 0; t < nt; ++t) h.bblk[tblock[t] + 1]++;
  for (int64_t b = 0; b < nb; ++b) h.bblk[b + 1] += h.bblk[b];
  std::vector<int32_t> pos(h.bblk.begin(), h.bblk.end() - 1);
  std::vector<int64_t> order(nt);
  for (int64_t t = 0; t < nt; ++t) order[pos[tblock[t]]++] = t;   // stable per block
  h.btgt.resize(nt);
  h.bseg.assign(nt + 1, 0);
  h.bsrc.reserve(src.size());
  for (int64_t u = 0; u < nt; ++u) {
    const int64_t t = order[u];
    h.btgt[u] = target[t];
    for (int64_t j = seg[t]; j < seg[t + 1]; ++j) {
      const int64_t inst = src[j] % n, e = src[j] / n;
      h.bsrc.push_back((uint16_t)((p.loff + e) * B + inst % B));
    }
    h.bseg[u + 1] = (int32_t)h.bsrc.size();
  }
}

static void free_flat_plan(Group& g) {
  (void)hipFree(g.d_fblk); (void)hipFree(g.d_frec); (void)hipFree(g.d_fsrc);
  (void)hipFree(g.d_flblk); (void)hipFree(g.d_flong);
  g.d_fblk = g.d_flblk = nullptr;
  g.d_frec = nullptr;
  g.d_fsrc = nullptr;
  g.d_flong = nullptr;
}

// Joins the kinds' block plans (blocks of B) into the flat plan: per block,
// kind by kind, every target's record and slots, and the long targets' list.
// Leaves no flat plan (and the fused load off) when a kind is not
// block-private or a target index or the slot count does not fit the records.
static int build_flat_plan(Group& g, int B, const BlockHost (&h)[NK]) {
  free_flat_plan(g);
  if (B <= 0 || g.n == 0) return XA_OK;
  for (int k = 0; k < NK; ++k)
    if (!g.plan[k].block_private) return XA_OK;
  const int64_t nb = (g.n + B - 1) / B;
  int64_t n_rec = 0, n_src = 0;
  for (int k = 0; k < NK; ++k) {
    n_rec += (int64_t)h[k].btgt.size();
    n_src += (int64_t)h[k].bsrc.size();
    for (int32_t t : h[k].btgt)
      if ((uint32_t)t > kTargetMask) {
        for (int kk = 0; kk < NK; ++kk) g.plan[kk].block_private = false;
        return XA_OK;
      }
  }
  if (n_rec + 1 > INT32_MAX || n_src > INT32_MAX) {
    for (int kk = 0; kk < NK; ++kk) g.plan[kk].block_private = false;
    return XA_OK;
  }
  std::vector<int32_t> fblk(nb + 1, 0), flblk(nb + 1, 0);
  std::vector<FRec> frec;
  std::vector<uint16_t> fsrc;
  std::vector<FLong> flong;
  frec.reserve(n_rec + 1);
  fsrc.reserve(n_src);
  for (int64_t b = 0; b < nb; ++b) {
    fblk[b] = (int32_t)frec.size();
    flblk[b] = (int32_t)flong.size();
    for (int k = 0; k < NK; ++k) {
      const BlockHost& hk = h[k];
      if (hk.btgt.empty()) continue;
      for (int32_t t = hk.bblk[b]; t < hk.bblk[b + 1]; ++t) {
        const int32_t tk = (int32_t)((uint32_t)hk.btgt[t] | ((uint32_t)k << kKindShift));
        const int32_t j0 = (int32_t)fsrc.size();
        fsrc.insert(fsrc.end(), hk.bsrc.begin() + hk.bseg[t], hk.bsrc.begin() + hk.bseg[t + 1]);
        const int32_t j1 = (int32_t)fsrc.size();
        frec.push_back({tk, j0});
        if (!kind_is_assign(k) && j1 - j0 > kWaveSegment) flong.push_back({tk, j0, j1, 0});
      }
    }
  }
  fblk[nb] = (int32_t)frec.size();
  flblk[nb] = (int32_t)flong.size();
  frec.push_back({0, (int32_t)fsrc.size()});
  int rc;
  if ((rc = upload(&g.d_fblk, fblk)) != XA_OK) return rc;
  if ((rc = upload(&g.d_frec, frec)) != XA_OK) return rc;
  if ((rc = upload(&g.d_fsrc, fsrc)) != XA_OK) return rc;
  if ((rc = upload(&g.d_flblk, flblk)) != XA_OK) return rc;
  return upload(&g.d_flong, flong);
}

// A pattern plan applies when each of the n instances feeds S targets at the
// same offsets from its own base, the bases P apart in instance order
// (target[i * S + s] = target[0] + i * P + off[s], off[S-1] < P), and every
// instance's slot s is fed by the same entries.  Sources are e * n + i.
static int detect_pattern(const Group& g, Plan& p, const std::vector<int32_t>& target,
                          const std::vector<int64_t>& seg, const std::vector<int64_t>& src) {
  p.pattern = false;
  const int64_t n = g.n, nt = (int64_t)target.size();
  if (n == 0 || nt == 0 || nt % n != 0) return XA_OK;
  const int S = (int)(nt / n), E = p.E;
  const int64_t last = (int64_t)target[S - 1] - target[0];
  const int64_t P = n > 1 ? (int64_t)target[S] - target[0] : last + 1;
  if (last >= P || P > 127) return XA_OK;
  for (int64_t i = 0; i < n; ++i)
    for (int s = 0; s < S; ++s)
      if ((int64_t)target[i * S + s] != (int64_t)target[0] + i * P + (target[s] - target[0]))
        return XA_OK;
  std::vector<int32_t> poff((size_t)E, -1), slot_of((size_t)E, -1);
  for (int s = 0; s < S; ++s)   // instance 0 defines the pattern
    for (int64_t j = seg[s]; j < seg[s + 1]; ++j) {
      if (src[j] % n != 0) return XA_OK;
      slot_of[src[j] / n] = s;
      poff[src[j] / n] = target[s] - target[0];
    }
  // every instance: slot s holds exactly the entries of slot s, in entry
  // order (the plan's order within a private target)
  std::vector<int32_t> cnt((size_t)S, 0);
  for (int e = 0; e < E; ++e) if (slot_of[e] >= 0) cnt[slot_of[e]]++;
  for (int64_t i = 0; i < n; ++i)
    for (int s = 0; s < S; ++s) {
      const int64_t t = i * S + s;
      if (seg[t + 1] - seg[t] != cnt[s]) return XA_OK;
      int64_t prev = -1;
      for (int64_t j = seg[t]; j < seg[t + 1]; ++j) {
        const int64_t e = src[j] / n;
        if (src[j] % n != i || slot_of[e] != s || e <= prev) return XA_OK;
        prev = e;
      }
    }
  int B = 256;   // LDS of the staged ranges within 64 KB
  while (B > 64 && (size_t)B * (P + 1) * sizeof(double) > 65536) B /= 2;
  if ((size_t)B * (P + 1) * sizeof(double) > 65536) return XA_OK;
  int rc;
  if ((rc = upload(&p.d_pslot, poff)) != XA_OK) return rc;
  p.pattern = true;
  p.pS = S;
  p.pP = (int)P;
  p.plast = (int)last;
  p.pblock = B;
  return XA_OK;
}

// Counting sort of (target, instance, entry) triples by target; stable, so
// contributions of one target stay in instance order then stamp order.
static int build_plan(Group& g, int k, BlockHost& bh) {
  Plan& p = g.plan[k];
  const int E = g.ent[k];
  p.E = E;
  const std::vector<int32_t>& tg = g.tgt[k];
  if (E == 0 || g.n == 0) return XA_OK;
  int32_t tmax = -1;
  for (int32_t t : tg) tmax = std::max(tmax, t);
  if (tmax < 0) {
    XA_HIP(hipMalloc((void**)&p.d_vals, sizeof(double) * (size_t)E * g.n));
    return XA_OK;
  }
  std::vector<int64_t> cnt((size_t)tmax + 2, 0);
  for (int64_t idx = 0; idx < (int64_t)tg.size(); ++idx) {
    const int32_t t = tg[idx];
    if (t < 0) continue;
    cnt[t + 1]++;
  }
  std::vector<int32_t> target;
  std::vector<int64_t> seg;
  target.reserve(tg.size());
  seg.reserve(tg.size() + 1);
  seg.push_back(0);
  std::vector<int64_t> start((size_t)tmax + 1, -1);
  int64_t acc = 0;
  for (int32_t t = 0; t <= tmax; ++t) {
    if (cnt[t + 1] == 0) continue;
    start[t] = acc;
    acc += cnt[t + 1];
    target.push_back(t);
    seg.push_back(acc);
  }
  // sources are slots of the entry-major value buffer: e * n + instance
  const int64_t n = g.n;
  std::vector<int64_t> src(acc);
  std::vector<int64_t> fillp(start);
  for (int64_t idx = 0; idx < (int64_t)tg.size(); ++idx) {
    const int32_t t = tg[idx];
    if (t < 0) continue;
    src[fillp[t]++] = (idx % E) * n + idx / E;
  }
  p.n_targets = (int64_t)target.size();
  p.n_src = acc;
  int rc;
  if ((rc = detect_pattern(g, p, target, seg, src)) != XA_OK) return rc;
  std::vector<int32_t> long_t, rail_t, rail_c(1, 0);
  std::vector<int64_t> chunk_j0, chunk_j1;
  for (int64_t t = 0; t < p.n_targets; ++t) {
    const int64_t len = seg[t + 1] - seg[t];
    if (len > kRailSegment) {
      rail_t.push_back((int32_t)t);
      for (int64_t j = seg[t]; j < seg[t + 1]; j += kRailChunk) {
        chunk_j0.push_back(j);
        chunk_j1.push_back(std::min<int64_t>(j + kRailChunk, seg[t + 1]));
      }
      rail_c.push_back((int32_t)chunk_j0.size());
    } else if (len > kLongSegment) {
      long_t.push_back((int32_t)t);
    }
  }
  p.n_long = (int64_t)long_t.size();
  p.n_rails = (int64_t)rail_t.size();
  p.n_chunks = (int64_t)chunk_j0.size();
  if (p.n_chunks > ((int64_t)1 << 25))   // one 64-lane launch slot per chunk
    return fail(XA_ERR_ARG, "plan: supply-rail contributions exceed the chunked launch");
  if ((rc = upload(&p.d_long, long_t)) != XA_OK) return rc;
  if ((rc = upload(&p.d_rail_t, rail_t)) != XA_OK) return rc;
  if (p.n_rails > 0) {
    if ((rc = upload(&p.d_rail_c, rail_c)) != XA_OK) return rc;
    XA_HIP(hipMalloc((void**)&p.d_partial, sizeof(double) * (size_t)p.n_chunks));
  }
  // contiguous targets need no target array; 32-bit indices when they fit
  bool contig = true;
  for (size_t t = 0; t < target.size() && contig; ++t) contig = target[t] == target[0] + (int32_t)t;
  p.t0 = target.empty() ? 0 : target[0];
  if (!contig && (rc = upload(&p.d_target, target)) != XA_OK) return rc;
  p.idx32 = acc < ((int64_t)1 << 31) && (int64_t)E * n < ((int64_t)1 << 31);
  if (p.idx32) {
    std::vector<int32_t> s32(seg.begin(), seg.end()), r32(src.begin(), src.end());
    int32_t* ds = nullptr;
    int32_t* dr = nullptr;
    if ((rc = upload(&ds, s32)) != XA_OK) return rc;
    if ((rc = upload(&dr, r32)) != XA_OK) return rc;
    p.d_seg = ds;
    p.d_src = dr;
    std::vector<int32_t> a0(chunk_j0.begin(), chunk_j0.end()), a1(chunk_j1.begin(), chunk_j1.end());
    int32_t* c0 = nullptr;
    int32_t* c1 = nullptr;
    if ((rc = upload(&c0, a0)) != XA_OK) return rc;
    if ((rc = upload(&c1, a1)) != XA_OK) return rc;
    p.d_chunk_j0 = c0;
    p.d_chunk_j1 = c1;
  } else {
    int64_t* ds = nullptr;
    int64_t* dr = nullptr;
    if ((rc = upload(&ds, seg)) != XA_OK) return rc;
    if ((rc = upload(&dr, src)) != XA_OK) return rc;
    p.d_seg = ds;
    p.d_src = dr;
    int64_t* c0 = nullptr;
    int64_t* c1 = nullptr;
    if ((rc = upload(&c0, chunk_j0)) != XA_OK) return rc;
    if ((rc = upload(&c1, chunk_j1)) != XA_OK) return rc;
    p.d_chunk_j0 = c0;
    p.d_chunk_j1 = c1;
  }
  build_block_plan(g, p, target, seg, src, g.impl->fused_block, bh);
  XA_HIP(hipMalloc((void**)&p.d_vals, sizeof(double) * (size_t)E * g.n));
  XA_HIP(hipMemset(p.d_vals, 0, sizeof(double) * (size_t)E * g.n));
  return XA_OK;
}

// (Re)builds the per-block card tables of the stored-mode update for blocks
// of B instances.
static int build_card_tables(Group& g, int B) {
  if (g.card_block == B) return XA_OK;
  const int64_t n = g.n, nb = (n + B - 1) / B;
  // per block: the number of cards, then the cards (kBlockCards + 1 ints)
  std::vector<int32_t> cards((size_t)nb * (kBlockCards + 1), 0);
  std::vector<uint8_t> slot((size_t)n, 0);
  bool ok = true;
  for (int64_t b = 0; b < nb && ok; ++b) {
    int used = 0;
    int32_t* tab = &cards[(size_t)b * (kBlockCards + 1) + 1];
    for (int64_t i = b * B; i < std::min<int64_t>(n, (b + 1) * B) && ok; ++i) {
      const int32_t m = g.inst_model[i];
      int k = 0;
      while (k < used && tab[k] != m) ++k;
      if (k == used) {
        if (used == kBlockCards) { ok = false; break; }
        tab[used++] = m;
      }
      slot[i] = (uint8_t)k;
    }
    tab[-1] = used;
  }
  int rc;
  if (ok) {
    if ((rc = upload(&g.d_blk_cards, cards)) != XA_OK) return rc;
    if ((rc = upload(&g.d_card_slot, slot)) != XA_OK) return rc;
  }
  g.card_tables = ok;
  g.card_block = B;
  return XA_OK;
}

// The fused load's geometry for blocks of B instances: its LDS and whether
// every kind's block plan (built for B) exists.
static bool fused_fits(const Group& g, int B) {
  const size_t lds = (size_t)g.fused_rows * (size_t)B * sizeof(double);
  return B > 0 && B <= kBlock && lds <= kMaxLds && (int64_t)g.fused_rows * B <= 65536;
}
static void set_fused_block(Group& g, int B) {
  g.fused_B = B;
  g.fused_lds = (size_t)g.fused_rows * (size_t)std::max(B, 0) * sizeof(double);
  g.fused_ok = fused_fits(g, B) && !g.split_only;
  for (int k = 0; k < NK; ++k) g.fused_ok = g.fused_ok && g.plan[k].block_private;
}

int xa_group_finalize(xa_group gh) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase != PH_CSR) return fail(XA_ERR_STATE, "finalize out of order");
  XA_HIP(hipSetDevice(g->ctx->device));
  int rc;
  if ((rc = upload(&g->d_dfield, g->dfield)) != XA_OK) return rc;
  if ((rc = upload(&g->d_ifield, g->ifield)) != XA_OK) return rc;
  if ((rc = upload(&g->d_mcard, g->mcard)) != XA_OK) return rc;
  if ((rc = upload(&g->d_nodes, g->nodes)) != XA_OK) return rc;
  if ((rc = upload(&g->d_sto, g->sto_lid)) != XA_OK) return rc;
  if ((rc = upload(&g->d_sta, g->sta_lid)) != XA_OK) return rc;
  if ((rc = upload(&g->d_state, g->state_init)) != XA_OK) return rc;
  XA_HIP(hipMalloc((void**)&g->d_fail, sizeof(int32_t)));
  XA_HIP(hipMemset(g->d_fail, 0, sizeof(int32_t)));
  // LDS rows of the fused load: the kinds with a live target, in kind order
  int rows = 0;
  for (int k = 0; k < NK; ++k) {
    bool live = false;
    for (int32_t t : g->tgt[k]) if (t >= 0) { live = true; break; }
    g->plan[k].loff = rows;
    if (live) rows += g->ent[k];
  }
  {
    BlockHost bh[NK];
    for (int k = 0; k < NK; ++k)
      if ((rc = build_plan(*g, k, bh[k])) != XA_OK) return rc;
    if ((rc = build_flat_plan(*g, g->impl->fused_block, bh)) != XA_OK) return rc;
  }
  g->fused_rows = rows;
  set_fused_block(*g, g->impl->fused_block);
  g->pattern_ok = g->impl->launch_pattern != nullptr && g->n > 0 && !g->split_only;
  g->pattern_rows = 0;
  for (int k = 0; k < NK; ++k) {
    const Plan& p = g->plan[k];
    if (p.n_targets == 0) continue;
    if (!p.pattern || kind_is_assign(k)) g->pattern_ok = false;
    else g->pattern_rows += p.pP + 1;
  }
  if ((rc = build_card_tables(*g, g->split_block)) != XA_OK) return rc;
  g->phase = PH_FINAL;
  return XA_OK;
}

// ---------------------------------------------------------------------------
// loads
// ---------------------------------------------------------------------------
static int ready(Group* g) {
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase != PH_FINAL) return fail(XA_ERR_STATE, "load before finalize");
  return XA_OK;
}

static void tic(Group* g, int which) {
  if (!g->timing) return;
  std::vector<hipEvent_t>& v = g->ev[which];
  const size_t k = 2 * g->n_ev[which];
  while (v.size() < k + 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { g->timing = false; return; }
    v.push_back(e);
  }
  (void)hipEventRecord(v[k], g->ctx->stream);
}
static void toc(Group* g, int which) {
  if (!g->timing) return;
  (void)hipEventRecord(g->ev[which][2 * g->n_ev[which] + 1], g->ctx->stream);
  g->n_ev[which]++;
}

int xa_update_state(xa_group gh, const xa_solver_state* ss, const xa_vectors* v) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (!ss || !v) return fail(XA_ERR_ARG, "null solver state / vectors");
  if (g->n == 0) return XA_OK;
  tic(g, 0);
  rc = g->impl->launch_update(*g, *ss, *v, false);
  toc(g, 0);
  return rc;
}

int xa_update_secondary_state(xa_group gh, const xa_solver_state* ss, const xa_vectors* v) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (!ss || !v) return fail(XA_ERR_ARG, "null solver state / vectors");
  if (g->n == 0 || !g->impl->launch_secondary) return XA_OK;
  return g->impl->launch_secondary(*g, *ss, *v);
}

// The scatters of one call.  When every kind on the call writes its own
// output vector, the segmented plans with 32-bit indices go in one launch for
// their short and long segments (k_scatter_set), then their rails;
// the other kinds (pattern plans, lead assignments, 64-bit plans) follow one
// by one.  Different kinds write different vectors, so the order among them
// does not change any bit.
static int run_scatters(Group& g, const xa_solver_state& ss, const xa_vectors& v,
                        std::initializer_list<int> kinds) {
  std::vector<int> live;
  for (int k : kinds) {
    const Plan& p = g.plan[k];
    if (!kind_out(v, k) || p.n_targets == 0) continue;
    if (g.impl->skip_kind && g.impl->skip_kind(g, ss, k)) continue;
    live.push_back(k);
  }
  bool distinct = true;
  for (size_t a = 0; a < live.size(); ++a)
    for (size_t b = a + 1; b < live.size(); ++b)
      distinct = distinct && kind_out(v, live[a]) != kind_out(v, live[b]);
  SegSet s{};
  int64_t blocks = 0, lblocks = 0;
  std::vector<int> rest;
  for (int k : live) {
    const Plan& p = g.plan[k];
    if (!distinct || kind_is_assign(k) || p.pattern || !p.idx32) {
      rest.push_back(k);
      continue;
    }
    SegKind& d = s.k[s.n++];
    d.nt = p.n_targets;
    d.target = p.d_target;
    d.t0 = p.t0;
    d.seg = static_cast<const int32_t*>(p.d_seg);
    d.src = static_cast<const int32_t*>(p.d_src);
    d.vals = p.d_vals;
    d.out = kind_out(v, k);
    d.long_t = p.d_long;
    d.b0 = blocks;
    d.lb0 = lblocks;
    blocks += grid_for(p.n_targets);
    lblocks += p.n_long;
  }
  hipStream_t st = g.ctx->stream;
  if (blocks + lblocks > 0)
    hipLaunchKernelGGL(k_scatter_set,
                       dim3((unsigned)(blocks + (lblocks + kLongPerBlock - 1) / kLongPerBlock)),
                       dim3(kBlock), 0, st, s, blocks, lblocks);
  for (int k : live) {
    const Plan& p = g.plan[k];
    if (!distinct || kind_is_assign(k) || p.pattern || !p.idx32 || p.n_rails == 0) continue;
    const int32_t* src = static_cast<const int32_t*>(p.d_src);
    hipLaunchKernelGGL((k_rail_partial<int32_t>), dim3((unsigned)p.n_chunks), dim3(64), 0, st,
                       static_cast<const int32_t*>(p.d_chunk_j0),
                       static_cast<const int32_t*>(p.d_chunk_j1), src, p.d_vals, p.d_partial);
    if (p.d_target == nullptr)
      hipLaunchKernelGGL((k_rail_sum<true>), dim3((unsigned)p.n_rails), dim3(64), 0, st,
                         p.d_rail_t, p.d_rail_c, p.d_target, p.t0, p.d_partial, kind_out(v, k));
    else
      hipLaunchKernelGGL((k_rail_sum<false>), dim3((unsigned)p.n_rails), dim3(64), 0, st,
                         p.d_rail_t, p.d_rail_c, p.d_target, p.t0, p.d_partial, kind_out(v, k));
  }
  XA_HIP(hipGetLastError());
  int rc;
  for (int k : rest)
    if ((rc = run_scatter(g, ss, v, k)) != XA_OK) return rc;
  return XA_OK;
}

int xa_load_vectors(xa_group gh, const xa_solver_state* ss, const xa_vectors* v) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (!ss || !v) return fail(XA_ERR_ARG, "null solver state / vectors");
  tic(g, 1);
  if ((rc = run_scatters(*g, *ss, *v, {K_F, K_Q, K_FLIM, K_QLIM, K_LEADF, K_LEADQ})) != XA_OK)
    return rc;
  toc(g, 1);
  return XA_OK;
}

int xa_load_matrices(xa_group gh, const xa_solver_state* ss, const xa_vectors* v) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (!ss || !v) return fail(XA_ERR_ARG, "null solver state / vectors");
  tic(g, 2);
  if ((rc = run_scatters(*g, *ss, *v, {K_DFDX, K_DQDX})) != XA_OK) return rc;
  toc(g, 2);
  return XA_OK;
}

// The fused write-out adds every kind's targets concurrently: it needs the
// kinds on the call to write distinct vectors.
static bool fused_outputs_distinct(const Group& g, const xa_solver_state& ss,
                                   const xa_vectors& v) {
  const double* seen[NK];
  int n = 0;
  for (int k = 0; k < NK; ++k) {
    const double* o = kind_out(v, k);
    if (!o || g.plan[k].n_targets == 0) continue;
    if (g.impl->skip_kind && g.impl->skip_kind(g, ss, k)) continue;
    for (int q = 0; q < n; ++q)
      if (seen[q] == o) return false;
    seen[n++] = o;
  }
  return true;
}

// The route xa_load_all takes (timed when the group's timing is on).
static int load_all(Group* g, const xa_solver_state* ss, const xa_vectors* v) {
  int rc;
  xa_group gh = reinterpret_cast<xa_group>(g);
  // the fused kernel needs every target block-private and no secondary pass;
  // it is taken for launch-bound sizes where the model's fused kernel is the
  // faster (fused_small) and at scale where fused_at_scale says so; otherwise
  // the four calls run in sequence (same results)
  const int64_t blocks = (g->n + g->fused_B - 1) / std::max(g->fused_B, 1);
  const bool fused_policy =
      g->impl->fused_at_scale || (g->impl->fused_small && blocks <= kSmallLaunchBlocks);
  if (g->route == 3 && g->pattern_ok && g->pattern_B > 0) {
    tic(g, 0);
    rc = g->impl->launch_pattern(*g, *ss, *v);
    toc(g, 0);
    return rc;
  }
  if (g->fused_ok && !g->impl->launch_secondary &&
      (g->route == 2 || (g->route == 0 && fused_policy)) &&
      fused_outputs_distinct(*g, *ss, *v)) {
    tic(g, 0);
    rc = g->impl->launch_update(*g, *ss, *v, true);
    toc(g, 0);
    return rc;
  }
  if ((rc = xa_update_state(gh, ss, v)) != XA_OK) return rc;
  if ((rc = xa_update_secondary_state(gh, ss, v)) != XA_OK) return rc;
  if (g->timing) {   // the phases timed one by one
    if ((rc = xa_load_vectors(gh, ss, v)) != XA_OK) return rc;
    return xa_load_matrices(gh, ss, v);
  }
  // the vector and the matrix kinds write different vectors: their scatters
  // share one launch (a small group's load is launch-latency bound)
  return run_scatters(*g, *ss, *v, {K_F, K_Q, K_FLIM, K_QLIM, K_LEADF, K_LEADQ, K_DFDX, K_DQDX});
}

int xa_load_all(xa_group gh, const xa_solver_state* ss, const xa_vectors* v) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (!ss || !v) return fail(XA_ERR_ARG, "null solver state / vectors");
  if (g->n == 0) return XA_OK;
  return load_all(g, ss, v);
}

__global__ void k_zero(double* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// xa_load_all with the call's scratch -- every kind's stamp value buffer
// (E doubles per instance) and supply-rail partial sums -- laid out in the
// device's `scratch` array instead of the group's own buffers.  The plans'
// pointers are swapped for the call and restored before returning (the
// kernels take them by value at launch).  The used range is cleared first by
// a fill kernel, so no entry starts from an earlier call's contents.
int xa_load_all_noalloc(xa_group gh, const xa_solver_state* ss, const xa_vectors* v) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (!ss || !v) return fail(XA_ERR_ARG, "null solver state / vectors");
  if (g->n == 0) return XA_OK;
  double* base = g->ctx->scratch;
  if (!base) return fail(XA_ERR_STATE, "no-alloc load: scratch address not resolved");
  size_t need = 0;
  for (int k = 0; k < NK; ++k) {
    const Plan& p = g->plan[k];
    if (p.d_vals) need += (size_t)p.E * (size_t)g->n;
    if (p.d_partial) need += (size_t)p.n_chunks;
  }
  if (need > kScratchDoubles)
    return fail(XA_ERR_ARG, "no-alloc load: the group needs " + std::to_string(need) +
                                " doubles of scratch; the maximum is " +
                                std::to_string(kScratchDoubles) +
                                " (use xa_load_all for larger groups)");
  double* saved_vals[NK];
  double* saved_part[NK];
  size_t off = 0;
  for (int k = 0; k < NK; ++k) {
    Plan& p = g->plan[k];
    saved_vals[k] = p.d_vals;
    saved_part[k] = p.d_partial;
    if (p.d_vals) { p.d_vals = base + off; off += (size_t)p.E * (size_t)g->n; }
    if (p.d_partial) { p.d_partial = base + off; off += (size_t)p.n_chunks; }
  }
  if (off > 0) {
    const unsigned blocks = (unsigned)std::min<size_t>((off + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(kBlock), 0, g->ctx->stream, base,
                       (int64_t)off);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = hip_check(e, "no-alloc load: clearing scratch");
  }
  if (rc == XA_OK) {
    const bool timing = g->timing;
    g->timing = false;   // timing records events, which are created on demand
    rc = load_all(g, ss, v);
    g->timing = timing;
  }
  for (int k = 0; k < NK; ++k) {
    g->plan[k].d_vals = saved_vals[k];
    g->plan[k].d_partial = saved_part[k];
  }
  return rc;
}

// ---------------------------------------------------------------------------
// autotuning of the block shapes
// ---------------------------------------------------------------------------
// Host copy of one kind's global plan: targets, segment starts, sources.
static int download_plan(const Plan& p, std::vector<int32_t>& target, std::vector<int64_t>& seg,
                         std::vector<int64_t>& src) {
  const int64_t nt = p.n_targets;
  target.assign((size_t)nt, 0);
  seg.assign((size_t)nt + 1, 0);
  src.assign((size_t)p.n_src, 0);
  if (nt == 0) return XA_OK;
  if (p.d_target)
    XA_HIP(hipMemcpy(target.data(), p.d_target, sizeof(int32_t) * nt, hipMemcpyDeviceToHost));
  else
    for (int64_t t = 0; t < nt; ++t) target[t] = p.t0 + (int32_t)t;
  if (p.idx32) {
    std::vector<int32_t> s32((size_t)nt + 1), r32((size_t)p.n_src);
    XA_HIP(hipMemcpy(s32.data(), p.d_seg, sizeof(int32_t) * (nt + 1), hipMemcpyDeviceToHost));
    if (p.n_src)
      XA_HIP(hipMemcpy(r32.data(), p.d_src, sizeof(int32_t) * p.n_src, hipMemcpyDeviceToHost));
    seg.assign(s32.begin(), s32.end());
    src.assign(r32.begin(), r32.end());
  } else {
    XA_HIP(hipMemcpy(seg.data(), p.d_seg, sizeof(int64_t) * (nt + 1), hipMemcpyDeviceToHost));
    if (p.n_src)
      XA_HIP(hipMemcpy(src.data(), p.d_src, sizeof(int64_t) * p.n_src, hipMemcpyDeviceToHost));
  }
  return XA_OK;
}

// Block plans of every kind for blocks of B instances (one kind's host copy
// at a time), then the fused geometry for B.
static int build_block_plans(Group& g, int B) {
  std::vector<int32_t> target;
  std::vector<int64_t> seg, src;
  BlockHost bh[NK];
  int rc;
  for (int k = 0; k < NK; ++k) {
    Plan& p = g.plan[k];
    if (p.E == 0) continue;
    if ((rc = download_plan(p, target, seg, src)) != XA_OK) return rc;
    build_block_plan(g, p, target, seg, src, B, bh[k]);
  }
  if ((rc = build_flat_plan(g, B, bh)) != XA_OK) return rc;
  set_fused_block(g, B);
  return XA_OK;
}

// One load pass as xa_load_all would run it on the fused or the split path.
static int load_pass(Group& g, const xa_solver_state& ss, const xa_vectors& v, int path) {
  int rc;
  if (path == 3) return g.impl->launch_pattern(g, ss, v);
  if (path == 2) return g.impl->launch_update(g, ss, v, true);
  if ((rc = g.impl->launch_update(g, ss, v, false)) != XA_OK) return rc;
  if (g.impl->launch_secondary && (rc = g.impl->launch_secondary(g, ss, v)) != XA_OK) return rc;
  // the same merged launch as xa_load_all's split route
  return run_scatters(g, ss, v, {K_F, K_Q, K_FLIM, K_QLIM, K_LEADF, K_LEADQ, K_DFDX, K_DQDX});
}

// Mean milliseconds of one pass over `reps` passes after one untimed pass.
static int time_passes(Group& g, const xa_solver_state& ss, const xa_vectors& v, int path,
                       int reps, float* ms) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  XA_HIP(hipEventCreate(&e0));
  XA_HIP(hipEventCreate(&e1));
  int rc = load_pass(g, ss, v, path);
  if (rc == XA_OK) (void)hipEventRecord(e0, g.ctx->stream);
  for (int r = 0; r < reps && rc == XA_OK; ++r) rc = load_pass(g, ss, v, path);
  if (rc == XA_OK) {
    (void)hipEventRecord(e1, g.ctx->stream);
    hipError_t e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e != hipSuccess) rc = hip_check(e, "autotune timing");
    else *ms /= (float)reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// The search: the four-call path with update-kernel blocks of 64, 128 and
// 256 threads, then the fused load with blocks of 64, 128 and 256 instances
// where it exists; the fastest wins (the split block is kept for the four
// calls in any case).
static int tune(Group& g, const xa_solver_state& ss, const xa_vectors& v, int reps, int* route,
                int* block) {
  const int candidates[] = {64, 128, 256};
  int rc;
  float best_split = 0.0f;
  int split_B = g.split_block;
  for (int c = 0; c < 3; ++c) {
    g.split_block = candidates[c];
    if ((rc = build_card_tables(g, g.split_block)) != XA_OK) return rc;
    float ms = 0.0f;
    if ((rc = time_passes(g, ss, v, 1, reps, &ms)) != XA_OK) return rc;
    if (c == 0 || ms < best_split) { best_split = ms; split_B = candidates[c]; }
  }
  g.split_block = split_B;
  if ((rc = build_card_tables(g, g.split_block)) != XA_OK) return rc;
  float best = best_split;
  int best_route = 1, best_B = split_B;
  if (!g.impl->launch_secondary && g.impl->fused_block > 0) {
    int built = g.fused_B;
    for (int B : candidates) {
      if (!fused_fits(g, B)) continue;
      if ((rc = build_block_plans(g, B)) != XA_OK) return rc;
      built = B;
      if (!g.fused_ok) continue;
      float ms = 0.0f;
      if ((rc = time_passes(g, ss, v, 2, reps, &ms)) != XA_OK) return rc;
      if (ms < best) { best = ms; best_route = 2; best_B = B; }
    }
    const int keep = best_route == 2 ? best_B : g.impl->fused_block;
    if (built != keep && (rc = build_block_plans(g, keep)) != XA_OK) return rc;
  }
  if (g.pattern_ok) {   // the pattern-fused load: instances per block
    const int keep_B = g.pattern_B;
    for (int B : candidates) {
      if ((size_t)B * g.pattern_rows * sizeof(double) > kMaxLds - 16 * 1024) continue;
      g.pattern_B = B;
      float ms = 0.0f;
      if ((rc = time_passes(g, ss, v, 3, reps, &ms)) != XA_OK) return rc;
      if (ms < best) { best = ms; best_route = 3; best_B = B; }
    }
    g.pattern_B = best_route == 3 ? best_B : keep_B;
  }
  g.route = best_route;
  if (route) *route = best_route;
  if (block) *block = best_B;
  return XA_OK;
}

int xa_group_set_path(xa_group gh, int path, int block) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (block != 64 && block != 128 && block != 256)
    return fail(XA_ERR_ARG, "set_path: block must be 64, 128 or 256");
  switch (path) {
    case 1:
      g->split_block = block;
      if ((rc = build_card_tables(*g, block)) != XA_OK) return rc;
      break;
    case 2:
      if (g->impl->launch_secondary || g->impl->fused_block <= 0 || !fused_fits(*g, block))
        return fail(XA_ERR_UNSUPPORTED, "set_path: no fused load of that block size");
      if ((rc = build_block_plans(*g, block)) != XA_OK) return rc;
      if (!g->fused_ok) return fail(XA_ERR_UNSUPPORTED, "set_path: targets shared across blocks");
      break;
    case 3:
      if (!g->pattern_ok || (size_t)block * g->pattern_rows * sizeof(double) > kMaxLds - 16 * 1024)
        return fail(XA_ERR_UNSUPPORTED, "set_path: no pattern-fused load of that block size");
      g->pattern_B = block;
      break;
    default:
      return fail(XA_ERR_ARG, "set_path: path must be 1, 2 or 3");
  }
  g->route = path;
  return XA_OK;
}

int xa_group_set_persistent(xa_group gh, int blocks) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (blocks == 0) {
    g->persist_blocks = 0;
    return XA_OK;
  }
  if (!g->impl->persistent_capacity)
    return fail(XA_ERR_UNSUPPORTED, "set_persistent: no persistent update for this model");
  XA_HIP(hipSetDevice(g->ctx->device));
  g->persist_blocks = blocks > 0 ? blocks : g->impl->persistent_capacity(*g);
  if (g->persist_blocks <= 0) return fail(XA_ERR_DEVICE, "set_persistent: no resident capacity");
  return XA_OK;
}

int xa_group_autotune(xa_group gh, const xa_solver_state* ss, const xa_vectors* v, int reps,
                      int* route, int* block) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (!ss || !v) return fail(XA_ERR_ARG, "null solver state / vectors");
  if (g->n == 0) {
    if (route) *route = 1;
    if (block) *block = g->split_block;
    return XA_OK;
  }
  // the tuning loads must not advance the instances' own state (the values
  // the reference keeps in its Instance objects) or the failure count
  const size_t state_bytes = sizeof(double) * (size_t)g->n_state * (size_t)g->n;
  double* saved = nullptr;
  int32_t fail0 = 0;
  if (state_bytes) XA_HIP(hipMalloc((void**)&saved, state_bytes));
  hipError_t e = hipSuccess;
  if (state_bytes)
    e = hipMemcpyAsync(saved, g->d_state, state_bytes, hipMemcpyDeviceToDevice, g->ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&fail0, g->d_fail, sizeof(int32_t), hipMemcpyDeviceToHost, g->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
  rc = e == hipSuccess ? tune(*g, *ss, *v, std::max(reps, 1), route, block)
                       : hip_check(e, "autotune: saving the instance state");
  e = hipSuccess;
  if (state_bytes)
    e = hipMemcpyAsync(g->d_state, saved, state_bytes, hipMemcpyDeviceToDevice, g->ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(g->d_fail, &fail0, sizeof(int32_t), hipMemcpyHostToDevice, g->ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->ctx->stream);
  (void)hipFree(saved);
  if (rc == XA_OK && e != hipSuccess) rc = hip_check(e, "autotune: restoring the instance state");
  return rc;
}

int xa_group_param_layout(xa_group gh, int* per_instance_columns, int* card_columns) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (g->phase < PH_SETUP) return fail(XA_ERR_STATE, "param_layout before setup");
  const int card = g->impl->card_columns ? g->impl->card_columns(*g) : 0;
  if (per_instance_columns) *per_instance_columns = g->n_dfield - card;
  if (card_columns) *card_columns = card;
  return XA_OK;
}

int xa_group_plan_info(xa_group gh, int kind, int64_t* n_targets, int64_t* n_sources,
                       int* pattern_slots) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  if (kind < 0 || kind >= XA_NKINDS) return fail(XA_ERR_ARG, "plan_info: unknown kind");
  const Plan& p = g->plan[kind];
  if (n_targets) *n_targets = p.n_targets;
  if (n_sources) *n_sources = p.n_src;
  if (pattern_slots) *pattern_slots = p.pattern ? p.pS : 0;
  return XA_OK;
}

int xa_group_failures(xa_group gh, int64_t* n_failed) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  int32_t h = 0;
  XA_HIP(hipMemcpyAsync(&h, g->d_fail, sizeof(int32_t), hipMemcpyDeviceToHost, g->ctx->stream));
  XA_HIP(hipStreamSynchronize(g->ctx->stream));
  if (n_failed) *n_failed = h;
  return XA_OK;
}

int xa_group_read_intermediate(xa_group gh, const char* name, double* out) {
  Group* g = reinterpret_cast<Group*>(gh);
  int rc = ready(g);
  if (rc != XA_OK) return rc;
  for (int s = 0; s < g->impl->n_state; ++s) {
    if (strcmp(g->impl->state_names[s], name) == 0) {
      XA_HIP(hipStreamSynchronize(g->ctx->stream));
      XA_HIP(hipMemcpy(out, g->d_state + (size_t)s * g->n, sizeof(double) * g->n,
                       hipMemcpyDeviceToHost));
      return XA_OK;
    }
  }
  return fail(XA_ERR_ARG, std::string("unknown intermediate ") + name);
}

int xa_group_traffic_bytes(xa_group gh, double* bu, double* bv, double* bm) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  if (!g->impl->traffic) return fail(XA_ERR_UNSUPPORTED, "no traffic model");
  if (bu) *bu = g->impl->traffic(*g, 0) * (double)g->n;
  if (bv) *bv = g->impl->traffic(*g, 1) * (double)g->n;
  if (bm) *bm = g->impl->traffic(*g, 2) * (double)g->n;
  return XA_OK;
}

int xa_group_enable_timing(xa_group gh, int on) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  g->timing = on != 0;
  for (int w = 0; w < 3; ++w) g->n_ev[w] = 0;
  return XA_OK;
}

int xa_group_last_times(xa_group gh, float* mu, float* mv, float* mm) {
  Group* g = reinterpret_cast<Group*>(gh);
  if (!g) return fail(XA_ERR_ARG, "null group");
  float* outs[3] = {mu, mv, mm};
  for (int w = 0; w < 3; ++w) {
    double tot = 0.0;
    for (size_t k = 0; k < g->n_ev[w]; ++k) {
      float ms = 0.f;
      XA_HIP(hipEventSynchronize(g->ev[w][2 * k + 1]));
      XA_HIP(hipEventElapsedTime(&ms, g->ev[w][2 * k], g->ev[w][2 * k + 1]));
      tot += ms;
    }
    if (outs[w]) *outs[w] = g->n_ev[w] ? (float)(tot / g->n_ev[w]) : 0.f;
  }
  return XA_OK;
}

}  // extern "C"
