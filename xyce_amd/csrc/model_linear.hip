// model_linear.hip -- linear two-terminal devices: resistor ("r", level 1)
// and capacitor ("c", level 1), so that the wires and loads of a netlist
// take the same device-side path as its transistors.
//
// Reference: src/DeviceModelPKG/OpenModels/src/
//   N_DEV_Resistor.C  parameter tables :170-262; Instance ctor :279-360
//     (R from RSH and L/W, default 1000); updateTemperature :848-865;
//     Master::updateState :1075-1090; loadDAEVectors :1100-1135;
//     loadDAEMatrices :1145-1181.
//   N_DEV_Capacitor.C parameter tables :85-165; processParams :196-244
//     (AGE, semiconductor C from CJ/CJSW); updateTemperature :271-280;
//     Instance ctor :440-606 (IC= adds one branch unknown);
//     Master::updateState :1821-1850; loadDAEVectors :1860-1960;
//     loadDAEMatrices :1970-2135.
// A capacitance given as an expression of solution variables is an input
// this interface cannot carry (values only), so it is outside the group.
#include <cmath>
#include <cstring>
#include <vector>

#include "xa_internal.h"
#include "xa_kernel.h"

namespace xa {

// ===========================================================================
// resistor
// ===========================================================================
namespace resistor {

enum MSlot { M_TC1, M_TC2, M_RSH, M_RMULT, M_DEFW, M_NARROW, M_TNOM, NM };
static const ParamDef kModelParams[] = {
  {"TC1", M_TC1, 0.0, 0}, {"TC2", M_TC2, 0.0, 0}, {"RSH", M_RSH, 0.0, 0},
  {"R", M_RMULT, 1.0, 0}, {"DEFW", M_DEFW, 1.e-5, 0}, {"NARROW", M_NARROW, 0.0, 0},
  {"TNOM", M_TNOM, 0.0, 1},
};
enum ISlot { I_R, I_L, I_W, I_TEMP, I_TC1, I_TC2, I_DTEMP, NI };
static const ParamDef kInstParams[] = {
  {"R", I_R, 1000.0, 0}, {"L", I_L, 0.0, 0}, {"W", I_W, 0.0, 0}, {"TEMP", I_TEMP, 0.0, 1},
  {"TC1", I_TC1, 0.0, 0}, {"TC2", I_TC2, 0.0, 0}, {"DTEMP", I_DTEMP, 0.0, 0},
};
struct DevCard { double unused; };
enum DField { D_G, ND };
enum IField { F_MODEL, NIF };
enum Node { N_POS, N_NEG, NNODE };

static int setup(Group& g) {
  const int64_t n = g.n;
  for (const Card& c : g.models)
    if (c.level != 1) return fail(XA_ERR_UNSUPPORTED, "resistor: level must be 1");
  g.n_dfield = ND;
  g.n_ifield = NIF;
  g.dfield.assign((size_t)ND * n, 0.0);
  g.ifield.assign((size_t)NIF * n, 0);
  g.n_state = 0;
  g.state_init.clear();
  for (int64_t i = 0; i < n; ++i) {
    const Card& c = g.models[g.inst_model[i]];
    g.n_ext[i] = 2;
    g.n_int[i] = 0;
    g.n_sta[i] = 0;
    g.n_sto[i] = g.lead[i] ? 1 : 0;
    // Instance ctor: temperature and coefficient defaults from the options
    // and the model; R from the sheet resistance when R is not given
    const double temp = g.igiven[I_TEMP] ? g.ival[I_TEMP][i] : g.opt.temp;
    const double width = g.igiven[I_W] ? g.ival[I_W][i] : c.v[M_DEFW];
    const double length = g.ival[I_L][i];
    const double tc1 = g.igiven[I_TC1] ? g.ival[I_TC1][i] : c.v[M_TC1];
    const double tc2 = g.igiven[I_TC2] ? g.ival[I_TC2][i] : c.v[M_TC2];
    const double tnom = c.has(M_TNOM) ? c.v[M_TNOM] : g.opt.tnom;
    double R = g.ival[I_R][i];
    if (!g.igiven[I_R]) {
      if (c.has(M_RSH) && g.igiven[I_L] && c.v[M_RSH] != 0 && length != 0)
        R = c.v[M_RSH] * (length - c.v[M_NARROW]) / (width - c.v[M_NARROW]);
      else
        R = 1000.0;
    }
    // updateTemperature
    const double difference = temp - tnom;
    const double factor = c.v[M_RMULT] * (1.0 + tc1 * difference + tc2 * difference * difference);
    g.dfield[D_G * n + i] = (R * factor != 0.0) ? 1.0 / (R * factor) : 0.0;
    g.ifield[F_MODEL * n + i] = g.inst_model[i];
  }
  g.mcard_stride = sizeof(DevCard) / sizeof(double);
  g.mcard.assign(g.models.size() * g.mcard_stride, 0.0);
  return XA_OK;
}

static void resolve_nodes(const Group&, int64_t, const int32_t* reg, int32_t* full) {
  full[N_POS] = reg[0];
  full[N_NEG] = reg[1];
}

static const Stamp kF[] = {{N_POS, 0}, {N_NEG, 0}};
static const Stamp kDF[] = {{N_POS, N_POS}, {N_POS, N_NEG}, {N_NEG, N_POS}, {N_NEG, N_NEG}};

static void stamps(const Group& g, int64_t i, int kind, std::vector<Stamp>& out) {
  out.clear();
  switch (kind) {
    case K_F: out.assign(kF, kF + 2); break;
    case K_DFDX: out.assign(kDF, kDF + 4); break;
    case K_LEADF: if (g.lead[i]) out.push_back({0, 0}); break;
    default: break;
  }
}

template <int MODE>
__device__ __forceinline__ void eval(const KArgs& a, int64_t i) {
  if (i >= a.n) return;
  const double G = dfld(a, D_G, i);
  // Master::updateState
  const double i0 = (solv(a.v.nextSol, node(a, N_POS, i)) - solv(a.v.nextSol, node(a, N_NEG, i))) * G;
  // Master::loadDAEVectors / loadDAEMatrices
  Sink<MODE> s(a, i);
  s.put(K_F, 0, i0);
  s.put(K_F, 1, -i0);
  s.put(K_LEADF, 0, i0);
  s.put(K_DFDX, 0, G);
  s.put(K_DFDX, 1, -G);
  s.put(K_DFDX, 2, -G);
  s.put(K_DFDX, 3, G);
}

template <bool FUSED>
__global__ void __launch_bounds__(256) k_update(KArgs a) {
  eval<FUSED ? 1 : 0>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (FUSED) fused_writeout(a);
}

static int launch_update(Group& g, const xa_solver_state& ss, const xa_vectors& v, bool fused) {
  if (!v.nextSol) return fail(XA_ERR_ARG, "resistor update: nextSol required");
  KArgs a;
  fill_kargs(g, ss, v, fused, a);
  return launch_model_kernel(g, a, fused, (const void*)k_update<false>, (const void*)k_update<true>);
}

// Algorithmic bytes per instance: G and two node LIDs + voltages read; the
// loads write 2 F rows and 4 dFdx slots (value + target each).
static double traffic(const Group&, int phase) {
  if (phase == 0) return 8 * (1 + 2) + 4 * 2;
  if (phase == 1) return 16 * 2;
  return 16 * 4;
}

}  // namespace resistor

// ===========================================================================
// capacitor
// ===========================================================================
namespace capacitor {

enum MSlot { M_CMULT, M_CJ, M_CJSW, M_DEFW, M_NARROW, M_TC1, M_TC2, M_TNOM, NM };
static const ParamDef kModelParams[] = {
  {"C", M_CMULT, 1.0, 0}, {"CJ", M_CJ, 0.0, 0}, {"CJSW", M_CJSW, 0.0, 0},
  {"DEFW", M_DEFW, 1.e-6, 0}, {"NARROW", M_NARROW, 0.0, 0}, {"TC1", M_TC1, 0.0, 0},
  {"TC2", M_TC2, 0.0, 0}, {"TNOM", M_TNOM, 0.0, 1},
};
enum ISlot { I_C, I_IC, I_L, I_W, I_AGE, I_D, I_TEMP, I_TC1, I_TC2, NI };
static const ParamDef kInstParams[] = {
  {"C", I_C, 1.e-6, 0}, {"IC", I_IC, 0.0, 0}, {"L", I_L, 1.0, 0}, {"W", I_W, 1.e-6, 0},
  {"AGE", I_AGE, 0.0, 0}, {"D", I_D, 0.0233, 0}, {"TEMP", I_TEMP, 0.0, 1},
  {"TC1", I_TC1, 0.0, 0}, {"TC2", I_TC2, 0.0, 0},
};
struct DevCard { double unused; };
enum DField { D_C, D_IC, ND };
enum IField { F_MODEL, F_ICGIVEN, NIF };
enum Node { N_POS, N_NEG, N_BRA, NNODE };

static int setup(Group& g) {
  const int64_t n = g.n;
  for (const Card& c : g.models)
    if (c.level != 1) return fail(XA_ERR_UNSUPPORTED, "capacitor: level must be 1");
  if (!g.igiven[I_C] && g.igiven[I_AGE])
    return fail(XA_ERR_ARG, "capacitor: AGE given without C");
  if (!g.igiven[I_C] && !g.igiven[I_L])
    return fail(XA_ERR_ARG, "capacitor: neither C nor L given");
  g.n_dfield = ND;
  g.n_ifield = NIF;
  g.dfield.assign((size_t)ND * n, 0.0);
  g.ifield.assign((size_t)NIF * n, 0);
  g.n_state = 0;
  g.state_init.clear();
  const bool ic = g.igiven[I_IC];
  for (int64_t i = 0; i < n; ++i) {
    const Card& c = g.models[g.inst_model[i]];
    g.n_ext[i] = 2;
    g.n_int[i] = ic ? 1 : 0;   // IC= adds the branch unknown
    g.n_sta[i] = 1;            // the charge
    g.n_sto[i] = g.lead[i] ? 1 : 0;
    const double width = g.igiven[I_W] ? g.ival[I_W][i] : c.v[M_DEFW];
    const double length = g.ival[I_L][i];
    const double temp = g.igiven[I_TEMP] ? g.ival[I_TEMP][i] : g.opt.temp;
    const double tc1 = g.igiven[I_TC1] ? g.ival[I_TC1][i] : c.v[M_TC1];
    const double tc2 = g.igiven[I_TC2] ? g.ival[I_TC2][i] : c.v[M_TC2];
    const double tnom = c.has(M_TNOM) ? c.v[M_TNOM] : g.opt.tnom;
    // processParams: aging of a given C, or the semiconductor capacitance
    double baseCap = g.ival[I_C][i];
    const double age = g.ival[I_AGE][i];
    if (g.igiven[I_AGE] && age >= 1) baseCap = baseCap * (1 - g.ival[I_D][i] * log10(age));
    if (!g.igiven[I_C]) {
      const double narrow = c.v[M_NARROW];
      baseCap = c.v[M_CJ] * (length - narrow) * (width - narrow) +
                2 * c.v[M_CJSW] * (length + width - 2 * narrow);
    }
    // updateTemperature
    const double difference = temp - tnom;
    const double factor = c.v[M_CMULT] * (1.0 + tc1 * difference + tc2 * difference * difference);
    g.dfield[D_C * n + i] = baseCap * factor;
    g.dfield[D_IC * n + i] = g.ival[I_IC][i];
    g.ifield[F_MODEL * n + i] = g.inst_model[i];
    g.ifield[F_ICGIVEN * n + i] = ic ? 1 : 0;
  }
  g.mcard_stride = sizeof(DevCard) / sizeof(double);
  g.mcard.assign(g.models.size() * g.mcard_stride, 0.0);
  return XA_OK;
}

static void resolve_nodes(const Group& g, int64_t i, const int32_t* reg, int32_t* full) {
  full[N_POS] = reg[0];
  full[N_NEG] = reg[1];
  full[N_BRA] = g.n_int[i] ? reg[2] : -1;
}

// With IC= the F rows and the branch entries of dFdx are in every plan; the
// load puts 0 where the reference's load skips an entry for the current
// solver state (x + 0 = x for every entry a load can hold: sums that start
// from +0 are never -0 under round-to-nearest).
static const Stamp kF[] = {{N_POS, 0}, {N_NEG, 0}, {N_BRA, 0}};
static const Stamp kQ[] = {{N_POS, 0}, {N_NEG, 0}};
static const Stamp kDF[] = {{N_POS, N_BRA}, {N_NEG, N_BRA}, {N_BRA, N_POS}, {N_BRA, N_NEG},
                            {N_BRA, N_BRA}};
static const Stamp kDQ[] = {{N_POS, N_POS}, {N_POS, N_NEG}, {N_NEG, N_POS}, {N_NEG, N_NEG}};

static void stamps(const Group& g, int64_t i, int kind, std::vector<Stamp>& out) {
  out.clear();
  const bool ic = g.n_int[i] != 0;
  switch (kind) {
    case K_F: if (ic) out.assign(kF, kF + 3); break;
    case K_Q: out.assign(kQ, kQ + 2); break;
    case K_DFDX: if (ic) out.assign(kDF, kDF + 5); break;
    case K_DQDX: out.assign(kDQ, kDQ + 4); break;
    case K_LEADF: case K_LEADQ: if (g.lead[i]) out.push_back({0, 0}); break;
    default: break;
  }
}

template <int MODE>
__device__ __forceinline__ void eval(const KArgs& a, int64_t i) {
  if (i >= a.n) return;
  const double C = dfld(a, D_C, i);
  const bool ic = ifld(a, F_ICGIVEN, i) != 0;
  const bool dcop = a.ss.dcopFlag != 0;
  const double Vpos = solv(a.v.nextSol, node(a, N_POS, i));
  const double Vneg = solv(a.v.nextSol, node(a, N_NEG, i));
  // Master::updateState
  double vcap = Vpos - Vneg;
  if (dcop && ic) vcap = dfld(a, D_IC, i);
  const double q0 = C * vcap;
  const int32_t lq = stalid(a, 0, i);
  if (lq >= 0) a.v.nextSta[lq] = q0;
  // Master::loadDAEVectors
  Sink<MODE> s(a, i);
  if (ic) {
    const double ibra = dcop ? solv(a.v.nextSol, node(a, N_BRA, i)) : 0.0;
    s.put(K_F, 0, ibra);
    s.put(K_F, 1, -ibra);
    s.put(K_F, 2, dcop ? (Vpos - Vneg - dfld(a, D_IC, i)) : 0.0);
  }
  s.put(K_Q, 0, q0);
  s.put(K_Q, 1, -q0);
  s.put(K_LEADF, 0, 0.0);   // the lead's last F assignment is 0 in every state
  s.put(K_LEADQ, 0, q0);
  // Master::loadDAEMatrices
  if (ic) {
    const double one = dcop ? 1.0 : 0.0;
    s.put(K_DFDX, 0, one);
    s.put(K_DFDX, 1, -one);
    s.put(K_DFDX, 2, one);
    s.put(K_DFDX, 3, -one);
    s.put(K_DFDX, 4, dcop ? 0.0 : 1.0);
  }
  const double Cq = (ic && dcop) ? 0.0 : C;
  s.put(K_DQDX, 0, Cq);
  s.put(K_DQDX, 1, -Cq);
  s.put(K_DQDX, 2, -Cq);
  s.put(K_DQDX, 3, Cq);
}

template <bool FUSED>
__global__ void __launch_bounds__(256) k_update(KArgs a) {
  eval<FUSED ? 1 : 0>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (FUSED) fused_writeout(a);
}

static int launch_update(Group& g, const xa_solver_state& ss, const xa_vectors& v, bool fused) {
  if (!v.nextSol || !v.nextSta) return fail(XA_ERR_ARG, "capacitor update: nextSol/nextSta required");
  KArgs a;
  fill_kargs(g, ss, v, fused, a);
  return launch_model_kernel(g, a, fused, (const void*)k_update<false>, (const void*)k_update<true>);
}

// Algorithmic bytes per instance (no IC): C, two node LIDs + voltages, the
// charge's LID and write; the loads write 2 Q rows and 4 dQdx slots.
static double traffic(const Group&, int phase) {
  if (phase == 0) return 8 * (1 + 2 + 1) + 4 * 3;
  if (phase == 1) return 16 * 2;
  return 16 * 4;
}

}  // namespace capacitor

const ModelImpl* model_resistor() {
  static ModelImpl m = [] {
    ModelImpl x{};
    x.type = XA_RESISTOR;
    x.name = "resistor";
    x.mparams = resistor::kModelParams;
    x.n_mparams = sizeof(resistor::kModelParams) / sizeof(ParamDef);
    x.n_mslots = resistor::NM;
    x.iparams = resistor::kInstParams;
    x.n_iparams = sizeof(resistor::kInstParams) / sizeof(ParamDef);
    x.n_islots = resistor::NI;
    x.max_nodes = resistor::NNODE;
    x.max_sto = 1;
    x.max_sta = 0;
    const int E[NK] = {2, 0, 0, 0, 4, 0, 1, 0};
    for (int k = 0; k < NK; ++k) x.max_entries[k] = E[k];
    x.setup = resistor::setup;
    x.resolve_nodes = resistor::resolve_nodes;
    x.stamps = resistor::stamps;
    x.launch_update = resistor::launch_update;
    x.traffic = resistor::traffic;
    x.fused_block = 256;
    x.fused_at_scale = true;
    x.fused_small = true;
    return x;
  }();
  return &m;
}

const ModelImpl* model_capacitor() {
  static ModelImpl m = [] {
    ModelImpl x{};
    x.type = XA_CAPACITOR;
    x.name = "capacitor";
    x.mparams = capacitor::kModelParams;
    x.n_mparams = sizeof(capacitor::kModelParams) / sizeof(ParamDef);
    x.n_mslots = capacitor::NM;
    x.iparams = capacitor::kInstParams;
    x.n_iparams = sizeof(capacitor::kInstParams) / sizeof(ParamDef);
    x.n_islots = capacitor::NI;
    x.max_nodes = capacitor::NNODE;
    x.max_sto = 1;
    x.max_sta = 1;
    const int E[NK] = {3, 2, 0, 0, 5, 4, 1, 1};
    for (int k = 0; k < NK; ++k) x.max_entries[k] = E[k];
    x.setup = capacitor::setup;
    x.resolve_nodes = capacitor::resolve_nodes;
    x.stamps = capacitor::stamps;
    x.launch_update = capacitor::launch_update;
    x.traffic = capacitor::traffic;
    x.fused_block = 256;
    x.fused_at_scale = true;
    x.fused_small = true;
    return x;
  }();
  return &m;
}

}  // namespace xa
