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
  const int32_t* direct[NK];           // fused mode: target per [e][inst]
  double* out[NK];                     // fused mode: output arrays
  unsigned kmask;                      // kinds to emit
};

// Stamp sink: an instance hands its contributions over in stamp order.
// Stored mode keeps them for the fixed-order scatter; fused mode adds them
// straight into outputs whose targets belong to this instance alone.  Both
// buffers are entry-major, so one put across a wavefront touches 64
// consecutive slots.
template <bool FUSED>
struct Sink {
  const KArgs& a;
  int64_t i;
  XA_HD Sink(const KArgs& args, int64_t inst) : a(args), i(inst) {}
  XA_HD void put(int k, int e, double val) const {
    if (!(a.kmask & (1u << k))) return;
    const int64_t idx = (int64_t)e * a.n + i;
    if (FUSED) {
      const int32_t t = a.direct[k][idx];
      if (t >= 0) {
        if (k == K_LEADF || k == K_LEADQ) a.out[k][t] = val;
        else a.out[k][t] += val;
      }
    } else {
      a.vals[k][idx] = val;
    }
  }
};

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

}  // namespace xa
