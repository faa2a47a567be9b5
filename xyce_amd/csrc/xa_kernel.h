// xa_kernel.h -- device-side argument block and stamp sink shared by the
// per-model evaluation kernels.
#pragma once
#include "xa_internal.h"
#include "xa_devsupport.h"

namespace xa {

// Everything a per-instance kernel needs, passed by value.
struct KArgs {
  int64_t n;
  const double* __restrict__ dfield;   // [field][inst]
  const int32_t* __restrict__ ifield;  // [field][inst]
  const double* __restrict__ mcard;    // [model][mstride]
  int mstride;
  int n_models;                        // cards in mcard
  uint64_t ucol[3];                    // card-uniform per-instance columns (Group::ucol)
  const int32_t* blk_cards;            // [block][1 + kBlockCards]: count, models to stage
  const uint8_t* card_slot;            // [inst] slot of its model in its block's table
  const int32_t* __restrict__ nodes;   // [local node][inst]
  const int32_t* __restrict__ sto;     // [local sto][inst]
  const int32_t* __restrict__ sta;     // [local sta][inst]
  double* __restrict__ state;          // [state][inst]
  int32_t* fail;                       // failure counter (one int)
  xa_solver_state ss;
  xa_device_options opt;
  xa_vectors v;
  double* vals[NK];                    // value buffers, entry-major [e][inst]
  int E[NK];
  double* out[NK];                     // fused mode: output arrays
  unsigned kmask;                      // kinds to emit
  // fused mode: LDS row of each kind and the flat block plan
  int loff[NK];
  const int32_t* fblk;
  const FRec* frec;
  const uint16_t* fsrc;
  const int32_t* flblk;
  const FLong* flong;
  // pattern-fused mode: each kind's staged output range (see pattern_stage)
  const int32_t* poff[NK];             // offset of each entry in its instance's span
  int pbase[NK];                       // first LDS double of the kind's range
  int pP[NK];                          // span of one instance
  int plast[NK];                       // offset of its last target
  int32_t pt0[NK];                     // first target of instance 0
};

// LDS of the fused load: row (loff[k] + e) holds entry e of kind k for every
// lane of the block.
extern __shared__ double xa_lds[];

// Stamp sink: an instance hands its contributions over in stamp order.
// MODE 0 (stored) keeps them in the entry-major value buffers for the
// fixed-order scatter (one put across a wavefront is 64 consecutive slots);
// MODE 1 (fused) keeps them in LDS rows for the block's ordered write-out
// (fused_writeout); MODE 2 (pattern-fused) adds each one at once into the
// instance's staged output span in LDS (pattern_stage), which in stamp order
// is the plan order of a private target.
template <int MODE>
struct Sink {
  const KArgs& a;
  int64_t i;
  __device__ Sink(const KArgs& args, int64_t inst) : a(args), i(inst) {}
  __device__ void put(int k, int e, double val) const {
    if (!(a.kmask & (1u << k))) return;
    if (MODE == 1) {
      xa_lds[(a.loff[k] + e) * (int)blockDim.x + (int)threadIdx.x] = val;
    } else if (MODE == 2) {
      const int off = a.poff[k][e];
      if (off >= 0) xa_lds[a.pbase[k] + (int)threadIdx.x * (a.pP[k] + 1) + off] += val;
    } else {
      a.vals[k][(int64_t)e * a.n + i] = val;
    }
  }
};

// LDS slot of element k of a staged range: row k / P padded to P + 1.  The
// quotient comes from a float reciprocal and one correction, exact for
// k < 2^16 and P < 128 (the estimate is within one of the quotient).
__device__ __forceinline__ int staged_slot(int k, int P, float inv_p) {
  int q = (int)((float)k * inv_p);
  int r = k - q * P;
  if (r < 0) { --q; r += P; }
  if (r >= P) { ++q; r -= P; }
  return q * (P + 1) + r;
}

// Pattern-fused staging of the block's outputs: for every kind on this call,
// the output range from the block's first instance's first target to its
// last instance's last target, instance rows padded to P + 1 doubles (load:
// outputs into LDS; otherwise LDS back into the outputs, slots no entry feeds
// unchanged).  Every thread of the block must call it.
__device__ __forceinline__ void pattern_stage(const KArgs& a, bool load) {
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
  const int m = (int)((a.n - i0) < (int64_t)blockDim.x ? (a.n - i0) : (int64_t)blockDim.x);
  for (int k = 0; k < NK; ++k) {
    if (!(a.kmask & (1u << k))) continue;
    const int P = a.pP[k];
    const int span = (m - 1) * P + a.plast[k] + 1;
    const float inv_p = 1.0f / (float)P;
    double* o = a.out[k] + a.pt0[k] + i0 * P;
    double* base = xa_lds + a.pbase[k];
    for (int x = threadIdx.x; x < span; x += blockDim.x) {
      if (load) base[staged_slot(x, P, inv_p)] = o[x];
      else o[x] = base[staged_slot(x, P, inv_p)];
    }
  }
}

// Fused write-out: after every lane of the block has put its values, the
// block adds them into its targets in the plan's order -- instance then stamp
// order, the reference's `+=` order -- so the result equals the scatter's bit
// for bit.  The block's targets of every kind form one list (the flat block
// plan), shared out over all threads, kUnroll targets per thread at a time:
// the records of a batch load together, then every target's first kInline
// source slots and its output, so a batch costs two memory round trips
// whatever the number of kinds.  Targets of more than kWaveSegment terms are
// summed by a whole wavefront: its lanes gather 64 terms at a time from the
// LDS rows, then every lane adds them in segment order from register
// broadcasts -- the same additions in the same order as one thread would
// make.  Every thread of the block must call it (it synchronises).
constexpr int kUnroll = 8;
constexpr int kInline = 8;

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
  return __hiloint2double(hi, lo);
}

// a.out[k] for a per-lane kind (a select chain, no indexed private array)
__device__ __forceinline__ double* kind_out_ptr(const KArgs& a, int k) {
  double* p = a.out[0];
#pragma unroll
  for (int kk = 1; kk < NK; ++kk) p = k == kk ? a.out[kk] : p;
  return p;
}

__device__ __forceinline__ bool kind_assigns(int k) { return k == K_LEADF || k == K_LEADQ; }

__device__ __forceinline__ void fused_writeout(const KArgs& a) {
  __syncthreads();
  const int b = blockIdx.x;
  const int32_t B = (int32_t)blockDim.x;
  const int32_t tb = a.fblk[b], te = a.fblk[b + 1];
  const FRec* rec = a.frec;
  const uint16_t* src = a.fsrc;
  for (int32_t x0 = tb + (int32_t)threadIdx.x; x0 < te; x0 += kUnroll * B) {
    int32_t tk[kUnroll], j0[kUnroll], j1[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int32_t t = x0 + u * B;
      tk[u] = j0[u] = j1[u] = 0;
      if (t < te) {
        const FRec r = rec[t];
        tk[u] = r.tk;
        j0[u] = r.j0;
        j1[u] = rec[t + 1].j0;
      }
    }
    // a target this pass adds: in the batch, of a kind on this call, and a
    // lead assignment or at most kWaveSegment terms
    bool mine[kUnroll];
    int kind[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      kind[u] = (int)((uint32_t)tk[u] >> kKindShift);
      mine[u] = x0 + u * B < te && (a.kmask >> kind[u] & 1u) &&
                (kind_assigns(kind[u]) || j1[u] - j0[u] <= kWaveSegment);
    }
    uint16_t s[kUnroll][kInline];
    double acc[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (!mine[u]) continue;
      if (kind_assigns(kind[u])) {
        s[u][0] = src[j1[u] - 1];
      } else {
#pragma unroll
        for (int w = 0; w < kInline; ++w)
          if (w < j1[u] - j0[u]) s[u][w] = src[j0[u] + w];
        acc[u] = kind_out_ptr(a, kind[u])[(uint32_t)tk[u] & kTargetMask];
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (!mine[u]) continue;
      double* o = kind_out_ptr(a, kind[u]) + ((uint32_t)tk[u] & kTargetMask);
      if (kind_assigns(kind[u])) {
        *o = xa_lds[s[u][0]];
      } else {
        double v = acc[u];
#pragma unroll
        for (int w = 0; w < kInline; ++w)
          if (w < j1[u] - j0[u]) v += xa_lds[s[u][w]];
        for (int32_t j = j0[u] + kInline; j < j1[u]; ++j) v += xa_lds[src[j]];
        *o = v;
      }
    }
  }
  // long targets: one per wavefront at a time
  const int lane = (int)threadIdx.x & 63;
  const int wave = (int)threadIdx.x >> 6, nwaves = (B + 63) >> 6;
  const int32_t le = a.flblk[b + 1];
  for (int32_t L = a.flblk[b] + wave; L < le; L += nwaves) {
    const FLong r = a.flong[L];
    const int k = (int)((uint32_t)r.tk >> kKindShift);
    if (!(a.kmask >> k & 1u)) continue;
    double* o = kind_out_ptr(a, k) + ((uint32_t)r.tk & kTargetMask);
    double acc = *o;
    for (int32_t base = r.j0; base < r.j1; base += 64) {
      const int32_t j = base + lane;
      const double x = j < r.j1 ? xa_lds[src[j]] : 0.0;
      const int cnt = (r.j1 - base) < 64 ? (int)(r.j1 - base) : 64;
#pragma unroll
      for (int q = 0; q < 64; ++q)
        if (q < cnt) acc += readlane_f64(x, q);
    }
    if (lane == 0) *o = acc;
  }
}

XA_HD double dfld(const KArgs& a, int f, int64_t i) { return a.dfield[f * a.n + i]; }
XA_HD int32_t ifld(const KArgs& a, int f, int64_t i) { return a.ifield[f * a.n + i]; }
XA_HD int32_t node(const KArgs& a, int j, int64_t i) { return a.nodes[j * a.n + i]; }
XA_HD int32_t stolid(const KArgs& a, int j, int64_t i) { return a.sto[j * a.n + i]; }
XA_HD int32_t stalid(const KArgs& a, int j, int64_t i) { return a.sta[j * a.n + i]; }
XA_HD double& st(const KArgs& a, int s, int64_t i) { return a.state[s * a.n + i]; }
// solution value of a node, 0 for ground
XA_HD double solv(const double* x, int32_t lid) { return lid >= 0 ? x[lid] : 0.0; }

void fill_kargs(Group& g, const xa_solver_state& ss, const xa_vectors& v,
                bool fused, KArgs& a);

// Launch a model's k_update<false> (stored) or k_update<true> (fused, block
// size and LDS from the group) on the context stream.
int launch_model_kernel(Group& g, const KArgs& a, bool fused, const void* k_stored,
                        const void* k_fused);

// Pattern-fused launch: fills the pattern fields of `a` (staging layout for
// blocks of g.pattern_B instances) and launches `kernel` with that LDS.
int launch_pattern_kernel(Group& g, KArgs& a, const void* kernel);

}  // namespace xa
