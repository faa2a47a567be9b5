// model_vsrc.hip -- independent voltage source ("v", level 1): DC, SIN, EXP
// and PULSE waveforms, one branch unknown per source, so that a netlist's
// supplies and drive inputs load on the device with its transistors.
//
// Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_Vsrc.C
//   parameter table :74-210; Instance ctor :300-380 (the DC value always,
//   the transient waveform when TRANSIENTSOURCETYPE is given or HB is set
//   up); registerLIDs :470-500; jacStamp / registerJacLIDs :560-600;
//   Master::updateState :941-990; loadDAEVectors :1000-1020 (the build
//   does not define SEPARATE_F_AND_B, so the source value enters F);
//   loadDAEMatrices :1030-1047.
// Waveforms: src/DeviceModelPKG/Core/src/N_DEV_SourceData.C -- SinData
//   :380-470, ExpData :650-730, PulseData :900-1010, ConstData :2060-2075;
//   the waveform codes are Src_index, N_DEV_SourceData.h:61-72.
// Waveform parameters that are not given take their defaults from the
// solver state (startingTimeStep, finalTime) at every call; a NaN in an
// instance column marks that instance's value as not given.  The source
// scaling option (DeviceOptions scale_src) is taken as its default, off.
#include <cmath>
#include <cstring>
#include <vector>

#include "xa_internal.h"
#include "xa_kernel.h"

namespace xa {
namespace vsrc {

static const ParamDef kModelParams[] = {{"_unused", 0, 0.0, 0}};
enum ISlot {
  I_DCV0, I_V0, I_V1, I_V2, I_TD, I_TR, I_TF, I_PW, I_PER, I_VA, I_FREQ, I_THETA, I_PHASE,
  I_TD1, I_TAU1, I_TD2, I_TAU2, I_TRAN, NI
};
static const ParamDef kInstParams[] = {
  {"DCV0", I_DCV0, 0.0, 0}, {"V0", I_V0, 0.0, 0}, {"V1", I_V1, 0.0, 0},
  {"V2", I_V2, 0.0, 0}, {"TD", I_TD, 0.0, 0}, {"TR", I_TR, 0.0, 0}, {"TF", I_TF, 0.0, 0},
  {"PW", I_PW, 0.0, 0}, {"PER", I_PER, 0.0, 0}, {"VA", I_VA, 0.0, 0},
  {"FREQ", I_FREQ, 0.0, 0}, {"THETA", I_THETA, 0.0, 0}, {"PHASE", I_PHASE, 0.0, 0},
  {"TD1", I_TD1, 0.0, 0}, {"TAU1", I_TAU1, 0.0, 0}, {"TD2", I_TD2, 0.0, 0},
  {"TAU2", I_TAU2, 0.0, 0}, {"TRANSIENTSOURCETYPE", I_TRAN, -1.0, 0},
};
// Src_index (N_DEV_SourceData.h:61-72); -1: no transient waveform
enum { SRC_NONE = -1, SRC_SIN = 0, SRC_EXP = 1, SRC_PULSE = 2, SRC_DC = 6 };
struct DevCard { double unused; };
// per-instance columns: the 17 waveform values (slot order), then
enum IField { F_TRAN, F_GIVEN, NIF };   // waveform code, given bits per slot
constexpr int ND = I_TRAN;              // value columns I_DCV0 .. I_TAU2
enum Node { N_POS, N_NEG, N_BRA, NNODE };

static int setup(Group& g) {
  const int64_t n = g.n;
  for (const Card& c : g.models)
    if (c.level != 1) return fail(XA_ERR_UNSUPPORTED, "vsource: level must be 1");
  g.n_dfield = ND;
  g.n_ifield = NIF;
  g.dfield.assign((size_t)ND * n, 0.0);
  g.ifield.assign((size_t)NIF * n, 0);
  g.n_state = 0;
  g.state_init.clear();
  for (int64_t i = 0; i < n; ++i) {
    int32_t given = 0;
    for (int k = 0; k < NI; ++k) {
      const double x = g.ival[k][i];
      const bool gv = g.igiven[k] && x == x;
      if (gv) given |= 1 << k;
      if (k < ND) g.dfield[(size_t)k * n + i] = gv ? x : kInstParams[k].def;
    }
    const int tran = (given & (1 << I_TRAN)) ? (int)g.ival[I_TRAN][i] : SRC_NONE;
    if (tran != SRC_NONE && tran != SRC_SIN && tran != SRC_EXP && tran != SRC_PULSE &&
        tran != SRC_DC)
      return fail(XA_ERR_UNSUPPORTED, "vsource: waveform other than SIN, EXP, PULSE or DC");
    g.ifield[F_TRAN * n + i] = tran;
    g.ifield[F_GIVEN * n + i] = given;
    g.n_ext[i] = 2;
    g.n_int[i] = 1;   // the branch current
    g.n_sta[i] = 0;
    g.n_sto[i] = 0;
  }
  g.mcard_stride = sizeof(DevCard) / sizeof(double);
  g.mcard.assign(g.models.size() * g.mcard_stride, 0.0);
  return XA_OK;
}

static void resolve_nodes(const Group&, int64_t, const int32_t* reg, int32_t* full) {
  full[N_POS] = reg[0];
  full[N_NEG] = reg[1];
  full[N_BRA] = reg[2];
}

static const Stamp kF[] = {{N_POS, 0}, {N_NEG, 0}, {N_BRA, 0}, {N_BRA, 0}};
static const Stamp kDF[] = {{N_POS, N_BRA}, {N_NEG, N_BRA}, {N_BRA, N_POS}, {N_BRA, N_NEG}};

static void stamps(const Group&, int64_t, int kind, std::vector<Stamp>& out) {
  out.clear();
  if (kind == K_F) out.assign(kF, kF + 4);
  else if (kind == K_DFDX) out.assign(kDF, kDF + 4);
}

// SourceData::updateSource of instance i's waveform at the solver's time.
__device__ __forceinline__ double waveform(const KArgs& a, int64_t i, int tran, int32_t gv) {
  auto P = [&](int k) { return dfld(a, k, i); };
  auto G = [&](int k) { return (gv >> k) & 1; };
  const double tstep = a.ss.startingTimeStep, tstop = a.ss.finalTime;
  double time = a.ss.currTime;
  if (tran == SRC_SIN) {
    const double FREQ = G(I_FREQ) ? P(I_FREQ) : 1.0 / tstop;
    const double TD = G(I_TD) ? P(I_TD) : 0.0;
    const double THETA = G(I_THETA) ? P(I_THETA) : 0.0;
    const double mpi = M_PI;
    time -= TD;
    if (time <= 0) return P(I_V0) + P(I_VA) * sin(2.0 * mpi * (P(I_PHASE) / 360));
    return P(I_V0) + P(I_VA) * sin(2.0 * mpi * (FREQ * time + P(I_PHASE) / 360)) *
                         exp(-(time * THETA));
  }
  if (tran == SRC_EXP) {
    const double TD1 = G(I_TD1) ? P(I_TD1) : 0.0;
    const double TAU1 = G(I_TAU1) ? P(I_TAU1) : tstep;
    const double TD2 = G(I_TD2) ? P(I_TD2) : TD1 + tstep;
    const double TAU2 = G(I_TAU2) ? P(I_TAU2) : tstep;
    const double V1 = P(I_V1), V2 = P(I_V2);
    if (time <= TD1) return V1;
    if (time <= TD2) return V1 + (V2 - V1) * (1 - exp(-(time - TD1) / TAU1));
    return V1 + (V2 - V1) * (1 - exp(-(time - TD1) / TAU1)) +
           (V1 - V2) * (1 - exp(-(time - TD2) / TAU2));
  }
  if (tran == SRC_PULSE) {
    const double TD = G(I_TD) ? P(I_TD) : 0.0;
    const double TR = G(I_TR) ? P(I_TR) : tstep;
    const double TF = G(I_TF) ? P(I_TF) : tstep;
    const double PW = G(I_PW) ? P(I_PW) : tstop;
    const double PER = G(I_PER) ? P(I_PER) : tstop;
    const double V1 = P(I_V1), V2 = P(I_V2), bpTol = a.ss.bpTol;
    time -= TD;
    if (time > PER && PER != 0.0) {
      const double basetime = PER * floor(time / PER);
      time -= basetime;
    }
    if (time <= 0 || (time > (TR + PW + TF) && (fabs(time - (TR + PW + TF)) > bpTol)))
      return V1;
    if ((time > TR && fabs(time - TR) > bpTol) &&
        (time < (TR + PW) || fabs(time - (TR + PW)) < bpTol))
      return V2;
    if (time > 0 && (time < TR || fabs(time - TR) < bpTol))
      return TR != 0.0 ? V1 + (V2 - V1) * (time) / TR : V1;
    return TF != 0.0 ? V2 + (V1 - V2) * (time - (TR + PW)) / TF : V2;
  }
  return P(I_DCV0);   // ConstData
}

template <int MODE>
__device__ __forceinline__ void eval(const KArgs& a, int64_t i) {
  if (i >= a.n) return;
  // Master::updateState
  const int tran = ifld(a, F_TRAN, i);
  const bool use_tran =
      (a.ss.HBspecified || a.ss.tranopFlag || a.ss.transientFlag) && tran != SRC_NONE;
  const double source = use_tran ? waveform(a, i, tran, ifld(a, F_GIVEN, i)) : dfld(a, I_DCV0, i);
  const double v_pos = solv(a.v.nextSol, node(a, N_POS, i));
  const double v_neg = solv(a.v.nextSol, node(a, N_NEG, i));
  const double srcCurrent = solv(a.v.nextSol, node(a, N_BRA, i));
  const double srcDrop = (v_pos - v_neg);
  const double srcBC = source;
  // Master::loadDAEVectors / loadDAEMatrices
  Sink<MODE> s(a, i);
  s.put(K_F, 0, srcCurrent);
  s.put(K_F, 1, -srcCurrent);
  s.put(K_F, 2, srcDrop);
  s.put(K_F, 3, -srcBC);          // fVec[li_Bra] -= srcBC
  s.put(K_DFDX, 0, 1.0);
  s.put(K_DFDX, 1, -1.0);
  s.put(K_DFDX, 2, 1.0);
  s.put(K_DFDX, 3, -1.0);
}

template <bool FUSED>
__global__ void __launch_bounds__(256) k_update(KArgs a) {
  eval<FUSED ? 1 : 0>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (FUSED) fused_writeout(a);
}

static int launch_update(Group& g, const xa_solver_state& ss, const xa_vectors& v, bool fused) {
  if (!v.nextSol) return fail(XA_ERR_ARG, "vsource update: nextSol required");
  KArgs a;
  fill_kargs(g, ss, v, fused, a);
  return launch_model_kernel(g, a, fused, (const void*)k_update<false>, (const void*)k_update<true>);
}

// Algorithmic bytes per instance: the waveform columns (17 doubles, two
// ints) and three node LIDs + voltages read; 4 F rows and 4 dFdx slots
// written (value + target each).
static double traffic(const Group&, int phase) {
  if (phase == 0) return 8 * (ND + 3) + 4 * (3 + NIF);
  if (phase == 1) return 16 * 4;
  return 16 * 4;
}

}  // namespace vsrc

const ModelImpl* model_vsrc() {
  static ModelImpl m = [] {
    ModelImpl x{};
    x.type = XA_VSOURCE;
    x.name = "vsource";
    x.mparams = vsrc::kModelParams;
    x.n_mparams = 0;
    x.n_mslots = 1;
    x.iparams = vsrc::kInstParams;
    x.n_iparams = sizeof(vsrc::kInstParams) / sizeof(ParamDef);
    x.n_islots = vsrc::NI;
    x.max_nodes = vsrc::NNODE;
    x.max_sto = 0;
    x.max_sta = 0;
    const int E[NK] = {4, 0, 0, 0, 4, 0, 0, 0};
    for (int k = 0; k < NK; ++k) x.max_entries[k] = E[k];
    x.setup = vsrc::setup;
    x.resolve_nodes = vsrc::resolve_nodes;
    x.stamps = vsrc::stamps;
    x.launch_update = vsrc::launch_update;
    x.traffic = vsrc::traffic;
    x.fused_block = 256;
    x.fused_at_scale = true;
    x.fused_small = true;
    return x;
  }();
  return &m;
}

}  // namespace xa
