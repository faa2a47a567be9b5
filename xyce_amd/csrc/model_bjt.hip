// model_bjt.hip -- Gummel-Poon bipolar transistor (level 1): host-side
// parameter processing and the per-instance evaluation + stamp kernel.
//
// Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_BJT.C
//   parameters :85-500, Instance ctor :555-1038 (node counts :1003-1036),
//   registerLIDs :1061-1176, registerStateLIDs :1255-1270,
//   registerStoreLIDs :1281-1300, jacStamp_RB_RC_RE_ :867-906,
//   updateTemperature :1603-1773, lambertWCurrent :1784-1800,
//   oldDAEExcessPhaseCalculation1/2 :2370-2466,
//   updateIntermediateVars :2476-3207, Model::processParams :3290-3361,
//   Model ctor :3394-3798, Master::updateState :3891-3937,
//   Master::loadDAEVectors :3978-4142, Master::loadDAEMatrices :4152-4260.
// newExcessPhase = 1 takes the new-DAE excess phase: two internal unknowns
// per instance (i_fx and its derivative, :1032-1036), the currents from
// auxDAECalculations (:1928-1980) and the Master's excess-phase rows
// (:3997-4260).  updateSecondaryState only fills instance currents
// that no load reads, so it launches nothing.
#include <cmath>
#include <cstring>
#include "xa_internal.h"
#include "xa_kernel.h"

namespace xa {
namespace bjt {

// ---- parameter tables ---------------------------------------------------
// Every alias has a slot of its own: the reference reconciles the given
// flags of an alias group before it reads the card (Model ctor :3521-3773),
// so the flag of the name actually used is the one that counts.
enum MSlot {
  P_TNOM, P_IS, P_BF, P_BFM, P_NF, P_VA, P_VBF, P_VAF, P_IKF, P_IK, P_JBF, P_ISE, P_JLE, P_NE,
  P_NLE, P_BR, P_BRM, P_NR, P_VAR, P_VB, P_VRB, P_BV, P_IKR, P_JBR, P_ISC, P_JLC, P_NC, P_RB,
  P_IRB, P_JRB, P_IOB, P_RBM, P_RE, P_RC, P_CJE, P_VJE, P_PE, P_MJE, P_ME, P_TF, P_XTF, P_VTF,
  P_ITF, P_JTF, P_PTF, P_CJC, P_VJC, P_PC, P_MJC, P_MC, P_XCJC, P_CDIS, P_TR, P_CJS, P_CCS,
  P_CSUB, P_VJS, P_PS, P_PSUB, P_MJS, P_MS, P_ESUB, P_XTB, P_TB, P_TCB, P_EG, P_XTI, P_PT,
  P_KF, P_AF, P_FC, P_C2, P_C4, P_NK, P_NKF, P_TYPE, NM
};
static const ParamDef kModelParams[] = {
  {"TNOM", P_TNOM, 0.0, 1}, {"IS", P_IS, 1e-16, 0}, {"BF", P_BF, 100.0, 0},
  {"BFM", P_BFM, 100.0, 0}, {"NF", P_NF, 1.0, 0}, {"VA", P_VA, 0.0, 0}, {"VBF", P_VBF, 0.0, 0},
  {"VAF", P_VAF, 0.0, 0}, {"IKF", P_IKF, 0.0, 0}, {"IK", P_IK, 0.0, 0}, {"JBF", P_JBF, 0.0, 0},
  {"ISE", P_ISE, 0.0, 0}, {"JLE", P_JLE, 0.0, 0}, {"NE", P_NE, 1.5, 0}, {"NLE", P_NLE, 1.5, 0},
  {"BR", P_BR, 1.0, 0}, {"BRM", P_BRM, 1.0, 0}, {"NR", P_NR, 1.0, 0}, {"VAR", P_VAR, 0.0, 0},
  {"VB", P_VB, 0.0, 0}, {"VRB", P_VRB, 0.0, 0}, {"BV", P_BV, 0.0, 0}, {"IKR", P_IKR, 0.0, 0},
  {"JBR", P_JBR, 0.0, 0}, {"ISC", P_ISC, 0.0, 0}, {"JLC", P_JLC, 0.0, 0}, {"NC", P_NC, 2.0, 0},
  {"RB", P_RB, 0.0, 0}, {"IRB", P_IRB, 0.0, 0}, {"JRB", P_JRB, 0.0, 0}, {"IOB", P_IOB, 0.0, 0},
  {"RBM", P_RBM, 0.0, 0}, {"RE", P_RE, 0.0, 0}, {"RC", P_RC, 0.0, 0}, {"CJE", P_CJE, 0.0, 0},
  {"VJE", P_VJE, 0.75, 0}, {"PE", P_PE, 0.75, 0}, {"MJE", P_MJE, 0.33, 0}, {"ME", P_ME, 0.33, 0},
  {"TF", P_TF, 0.0, 0}, {"XTF", P_XTF, 0.0, 0}, {"VTF", P_VTF, 0.0, 0}, {"ITF", P_ITF, 0.0, 0},
  {"JTF", P_JTF, 0.0, 0}, {"PTF", P_PTF, 0.0, 0}, {"CJC", P_CJC, 0.0, 0},
  {"VJC", P_VJC, 0.75, 0}, {"PC", P_PC, 0.75, 0}, {"MJC", P_MJC, 0.33, 0}, {"MC", P_MC, 0.33, 0},
  {"XCJC", P_XCJC, 1.0, 0}, {"CDIS", P_CDIS, 1.0, 0}, {"TR", P_TR, 0.0, 0},
  {"CJS", P_CJS, 0.0, 0}, {"CCS", P_CCS, 0.0, 0}, {"CSUB", P_CSUB, 0.0, 0},
  {"VJS", P_VJS, 0.75, 0}, {"PS", P_PS, 0.75, 0}, {"PSUB", P_PSUB, 0.75, 0},
  {"MJS", P_MJS, 0.0, 0}, {"MS", P_MS, 0.0, 0}, {"ESUB", P_ESUB, 0.0, 0}, {"XTB", P_XTB, 0.0, 0},
  {"TB", P_TB, 0.0, 0}, {"TCB", P_TCB, 0.0, 0}, {"EG", P_EG, 1.11, 0}, {"XTI", P_XTI, 3.0, 0},
  {"PT", P_PT, 3.0, 0}, {"KF", P_KF, 0.0, 0}, {"AF", P_AF, 1.0, 0}, {"FC", P_FC, 0.5, 0},
  {"C2", P_C2, 0.0, 0}, {"C4", P_C4, 0.0, 0}, {"NK", P_NK, 0.5, 0}, {"NKF", P_NKF, 0.5, 0},
  {"TYPE", P_TYPE, 1.0, 0},
};
enum ISlot { I_AREA, I_IC1, I_IC2, I_TEMP, I_OFF, I_LAMBERTW, NI };
static const ParamDef kInstParams[] = {
  {"AREA", I_AREA, 1.0, 0}, {"IC1", I_IC1, 0.0, 0}, {"IC2", I_IC2, 0.0, 0},
  {"TEMP", I_TEMP, 0.0, 1}, {"OFF", I_OFF, 0.0, 0}, {"LAMBERTW", I_LAMBERTW, 0.0, 0},
};

// ---- device-side layout ------------------------------------------------
struct DevCard {   // model values the evaluation reads
  double dtype, NF, NR, NE, NC, rollOffExp;
  double invRollOffF, invRollOffR, invEarlyF, invEarlyR;   // from the raw values
  double TF, XTF, vbcFac, ITF, XCJC, FC, VJC, MJE, MJC, f2, f3, f6, f7;
  double fxBE, fxBC;   // 1 - (1 - FC)^(1 - MJ) of each junction
  double TR, CJS, VJS, MJS, tdExcess, gEmit, gColl, RBM, RB, IRB;
};
enum DField {
  D_AREA, D_VT, D_TVCRIT, D_TSATCUR, D_TBETAF, D_TBETAR, D_TBELEAK, D_TBCLEAK, D_TBECAP,
  D_TBEPOT, D_TBCCAP, D_TBCPOT, D_ICVBE, D_ICVCE, ND
};
enum IField { F_MODEL, F_FLAGS, F_INTNODES, NIF };
enum Flag { FL_OFF = 1, FL_ICGIVEN = 2, FL_LAMBERTW = 4 };
enum IntNode { IN_C = 1, IN_B = 2, IN_E = 4 };
// N_IFX, N_DIFX: the new-DAE excess phase's unknowns (i_fx and its time
// derivative), present only with DeviceOptions newExcessPhase
enum Node { N_C, N_B, N_E, N_S, N_CP, N_BP, N_EP, N_IFX, N_DIFX, NNODE };
constexpr int16_t N_NONE = -1;
enum Sto { T_vBE, T_vBC, T_capeqCB, T_ib, T_ie, T_ic, T_is, NSTO };
enum Sta { A_qBEdiff, A_qBEdep, A_qCS, A_qBCdiff, A_qBCdep, A_qBX, A_CEXBC, NSTA };

// ---- host setup ----------------------------------------------------------
struct HostModel {
  int TYPE;
  double TNOM, satCur, betaF, emissionCoeffF, earlyVoltF, rollOffF, leakBECurrent;
  double leakBEEmissionCoeff, betaR, emissionCoeffR, earlyVoltR, rollOffR, leakBCCurrent;
  double leakBCEmissionCoeff, baseResist, baseCurrHalfResist, minBaseResist, emitterResist;
  double collectorResist, depCapBE, potBE, juncExpBE, transTimeF, transTimeBiasCoeffF;
  double transTimeHighCurrF, excessPhase, depCapBC, potBC, juncExpBC, baseFracBCCap;
  double transTimeR, CJS, potSubst, expSubst, betaExp, energyGap, tempExpIS, depCapCoeff;
  double c2, c4, rollOffExp, collectorConduct, emitterConduct, transTimeVBCFac;
  double excessPhaseFac, f2, f3, f6, f7;
};

// the value of an alias group: the one name given, else the first's default
static int alias(const Card& c, std::initializer_list<int> slots, double& out) {
  int given = 0;
  out = c.v[*slots.begin()];
  for (int s : slots)
    if (c.has(s)) { out = c.v[s]; ++given; }
  if (given > 1) return fail(XA_ERR_ARG, "bjt: a model value is given under two alias names");
  return XA_OK;
}

// Model ctor (:3394-3798) and Model::processParams (:3290-3361)
static int model_init(HostModel& m, const Card& c, const xa_device_options& opt) {
  int rc = XA_OK;
  auto A = [&](double& dst, std::initializer_list<int> s) { if (rc == XA_OK) rc = alias(c, s, dst); };
  m.TYPE = c.v[P_TYPE] < 0 ? -1 : 1;
  m.TNOM = c.has(P_TNOM) ? c.v[P_TNOM] : opt.tnom;
  m.satCur = c.v[P_IS];
  A(m.betaF, {P_BF, P_BFM});
  m.emissionCoeffF = c.v[P_NF];
  A(m.earlyVoltF, {P_VAF, P_VA, P_VBF});
  A(m.rollOffF, {P_IKF, P_IK, P_JBF});
  A(m.leakBECurrent, {P_ISE, P_JLE});
  A(m.leakBEEmissionCoeff, {P_NE, P_NLE});
  A(m.betaR, {P_BR, P_BRM});
  m.emissionCoeffR = c.v[P_NR];
  A(m.earlyVoltR, {P_VAR, P_VB, P_VRB, P_BV});
  A(m.rollOffR, {P_IKR, P_JBR});
  A(m.leakBCCurrent, {P_ISC, P_JLC});
  m.leakBCEmissionCoeff = c.v[P_NC];
  m.baseResist = c.v[P_RB];
  A(m.baseCurrHalfResist, {P_IRB, P_JRB, P_IOB});
  m.minBaseResist = c.v[P_RBM];
  m.emitterResist = c.v[P_RE];
  m.collectorResist = c.v[P_RC];
  m.depCapBE = c.v[P_CJE];
  A(m.potBE, {P_VJE, P_PE});
  A(m.juncExpBE, {P_MJE, P_ME});
  m.transTimeF = c.v[P_TF];
  m.transTimeBiasCoeffF = c.v[P_XTF];
  const double transTimeFVBC = c.v[P_VTF];
  A(m.transTimeHighCurrF, {P_ITF, P_JTF});
  m.excessPhase = c.v[P_PTF];
  m.depCapBC = c.v[P_CJC];
  A(m.potBC, {P_VJC, P_PC});
  A(m.juncExpBC, {P_MJC, P_MC});
  A(m.baseFracBCCap, {P_XCJC, P_CDIS});
  m.transTimeR = c.v[P_TR];
  A(m.CJS, {P_CJS, P_CCS, P_CSUB});
  A(m.potSubst, {P_VJS, P_PS, P_PSUB});
  A(m.expSubst, {P_MJS, P_MS, P_ESUB});
  A(m.betaExp, {P_XTB, P_TB, P_TCB});
  m.energyGap = c.v[P_EG];
  A(m.tempExpIS, {P_XTI, P_PT});
  m.depCapCoeff = c.v[P_FC];
  m.c2 = c.v[P_C2];
  m.c4 = c.v[P_C4];
  A(m.rollOffExp, {P_NK, P_NKF});
  if (rc != XA_OK) return rc;
  // NK is the only name the roll-off exponent is read from (:3300)
  if (!c.has(P_NK)) m.rollOffExp = 0.5;
  // processParams
  if (!c.has(P_ISE) && c.has(P_C2)) m.leakBECurrent = m.c2 * m.satCur;
  if (!c.has(P_ISC) && c.has(P_C4)) m.leakBCCurrent = m.c4 * m.satCur;
  if (!c.has(P_RBM)) m.minBaseResist = m.baseResist;
  m.collectorConduct = (m.collectorResist != 0.0) ? 1.0 / m.collectorResist : 0;
  m.emitterConduct = (m.emitterResist != 0.0) ? 1.0 / m.emitterResist : 0;
  if (c.has(P_VTF) && transTimeFVBC != 0.0) m.transTimeVBCFac = 1.0 / (transTimeFVBC * 1.44);
  else m.transTimeVBCFac = 0.0;
  m.excessPhaseFac = (m.excessPhase / (180.0 / M_PI)) * m.transTimeF;
  if (c.has(P_FC)) {
    if (m.depCapCoeff > 0.9999) m.depCapCoeff = 0.9999;
  } else {
    m.depCapCoeff = 0.5;
  }
  const double xfc = log(1.0 - m.depCapCoeff);
  m.f2 = exp((1.0 + m.juncExpBE) * xfc);
  m.f3 = 1.0 - m.depCapCoeff * (1.0 + m.juncExpBE);
  m.f6 = exp((1.0 + m.juncExpBC) * xfc);
  m.f7 = 1.0 - m.depCapCoeff * (1.0 + m.juncExpBC);
  return XA_OK;
}

static int setup(Group& g) {
  const int64_t n = g.n;
  if (g.opt.newExcessPhase) {
    // the excess-phase rows and columns follow each kind's base entries
    const int E[NK] = {8, 7, 5, 3, 32, 16, 4, 4};
    for (int k = 0; k < NK; ++k) g.ent[k] = E[k];
  }
  std::vector<HostModel> hm(g.models.size());
  for (size_t k = 0; k < g.models.size(); ++k) {
    if (g.models[k].level != 1) return fail(XA_ERR_UNSUPPORTED, "bjt: level must be 1");
    const int rc = model_init(hm[k], g.models[k], g.opt);
    if (rc != XA_OK) return rc;
  }
  g.n_dfield = ND;
  g.n_ifield = NIF;
  g.dfield.assign((size_t)ND * n, 0.0);
  g.ifield.assign((size_t)NIF * n, 0);
  g.n_state = 0;
  g.state_init.clear();
  auto D = [&](int f, int64_t i) -> double& { return g.dfield[f * n + i]; };
  auto I = [&](int f, int64_t i) -> int32_t& { return g.ifield[f * n + i]; };
  auto iv = [&](int s, int64_t i) { return g.ival[s][i]; };
  const bool icGiven = g.igiven[I_IC1] || g.igiven[I_IC2];
  for (int64_t i = 0; i < n; ++i) {
    const HostModel& m = hm[g.inst_model[i]];
    const double TEMP = g.igiven[I_TEMP] ? iv(I_TEMP, i) : g.opt.temp;
    g.n_ext[i] = 4;
    g.n_int[i] = (m.collectorResist == 0.0 ? 0 : 1) + (m.emitterResist == 0.0 ? 0 : 1) +
                 (m.baseResist == 0.0 ? 0 : 1) + (g.opt.newExcessPhase ? 2 : 0);   // :1032-1036
    g.n_sta[i] = NSTA;
    g.n_sto[i] = 3 + (g.lead[i] ? 4 : 0);

    // Instance::updateTemperature (:1603-1773)
    const double vt = TEMP * kKoverQ;
    const double fact2 = TEMP / kRefTemp;
    const double egfet = kEg0 - (kAlphaEg * TEMP * TEMP) / (TEMP + kBetaEg);
    const double arg = -egfet / (2.0 * kBoltz * TEMP) + kEg300 / (kBoltz * (2.0 * kRefTemp));
    const double pbfact = -2.0 * vt * (1.5 * log(fact2) + kQ * arg);
    const double ratlog = log(TEMP / m.TNOM);
    const double ratio1 = TEMP / m.TNOM - 1.0;
    const double factlog = ratio1 * m.energyGap / vt + m.tempExpIS * ratlog;
    const double factor = exp(factlog);
    const double bfactor = exp(ratlog * m.betaExp);
    const double fact1 = m.TNOM / kRefTemp;
    double pbo = (m.potBE - pbfact) / fact1;
    double gmaold = (m.potBE - pbo) / pbo;
    double tBECap = m.depCapBE / (1.0 + m.juncExpBE * (4.e-4 * (m.TNOM - kRefTemp) - gmaold));
    const double tBEPot = fact2 * pbo + pbfact;
    double gmanew = (tBEPot - pbo) / pbo;
    tBECap *= 1.0 + m.juncExpBE * (4.e-4 * (TEMP - kRefTemp) - gmanew);
    pbo = (m.potBC - pbfact) / fact1;
    gmaold = (m.potBC - pbo) / pbo;
    double tBCCap = m.depCapBC / (1.0 + m.juncExpBC * (4.e-4 * (m.TNOM - kRefTemp) - gmaold));
    const double tBCPot = fact2 * pbo + pbfact;
    gmanew = (tBCPot - pbo) / pbo;
    tBCCap *= 1.0 + m.juncExpBC * (4.e-4 * (TEMP - kRefTemp) - gmanew);
    const double tVCrit = vt * log(vt / (kRoot2 * m.satCur));
    D(D_AREA, i) = iv(I_AREA, i);
    D(D_VT, i) = vt;
    D(D_TVCRIT, i) = tVCrit;
    D(D_TSATCUR, i) = m.satCur * factor;
    D(D_TBETAF, i) = m.betaF * bfactor;
    D(D_TBETAR, i) = m.betaR * bfactor;
    D(D_TBELEAK, i) = m.leakBECurrent * exp(factlog / m.leakBEEmissionCoeff) / bfactor;
    D(D_TBCLEAK, i) = m.leakBCCurrent * exp(factlog / m.leakBCEmissionCoeff) / bfactor;
    D(D_TBECAP, i) = tBECap;
    D(D_TBEPOT, i) = tBEPot;
    D(D_TBCCAP, i) = tBCCap;
    D(D_TBCPOT, i) = tBCPot;
    D(D_ICVBE, i) = iv(I_IC1, i);
    D(D_ICVCE, i) = iv(I_IC2, i);
    I(F_MODEL, i) = g.inst_model[i];
    I(F_FLAGS, i) = (iv(I_OFF, i) != 0.0 ? FL_OFF : 0) | (icGiven ? FL_ICGIVEN : 0) |
                    (iv(I_LAMBERTW, i) != 0.0 ? FL_LAMBERTW : 0);
    I(F_INTNODES, i) = (m.collectorResist != 0.0 ? IN_C : 0) | (m.baseResist != 0.0 ? IN_B : 0) |
                       (m.emitterResist != 0.0 ? IN_E : 0);
  }
  g.mcard_stride = sizeof(DevCard) / sizeof(double);
  g.mcard.assign(g.models.size() * g.mcard_stride, 0.0);
  for (size_t k = 0; k < hm.size(); ++k) {
    const HostModel& m = hm[k];
    const double xfc = log(1.0 - m.depCapCoeff);
    DevCard c{};
    c.dtype = m.TYPE;
    c.NF = m.emissionCoeffF;
    c.NR = m.emissionCoeffR;
    c.NE = m.leakBEEmissionCoeff;
    c.NC = m.leakBCEmissionCoeff;
    c.rollOffExp = m.rollOffExp;
    c.invRollOffF = (m.rollOffF != 0) ? 1 / m.rollOffF : 0;
    c.invRollOffR = (m.rollOffR != 0) ? 1 / m.rollOffR : 0;
    c.invEarlyF = (m.earlyVoltF != 0) ? 1 / m.earlyVoltF : 0;
    c.invEarlyR = (m.earlyVoltR != 0) ? 1 / m.earlyVoltR : 0;
    c.TF = m.transTimeF;
    c.XTF = m.transTimeBiasCoeffF;
    c.vbcFac = m.transTimeVBCFac;
    c.ITF = m.transTimeHighCurrF;
    c.XCJC = m.baseFracBCCap;
    c.FC = m.depCapCoeff;
    c.VJC = m.potBC;
    c.MJE = m.juncExpBE;
    c.MJC = m.juncExpBC;
    c.f2 = m.f2; c.f3 = m.f3; c.f6 = m.f6; c.f7 = m.f7;
    c.fxBE = 1.0 - exp((1.0 - m.juncExpBE) * xfc);
    c.fxBC = 1.0 - exp((1.0 - m.juncExpBC) * xfc);
    c.TR = m.transTimeR;
    c.CJS = m.CJS;
    c.VJS = m.potSubst;
    c.MJS = m.expSubst;
    c.tdExcess = m.excessPhaseFac;
    c.gEmit = m.emitterConduct;
    c.gColl = m.collectorConduct;
    c.RBM = m.minBaseResist;
    c.RB = m.baseResist;
    c.IRB = m.baseCurrHalfResist;
    std::memcpy(&g.mcard[k * g.mcard_stride], &c, sizeof(c));
  }
  return XA_OK;
}

static void resolve_nodes(const Group& g, int64_t i, const int32_t* reg, int32_t* full) {
  const int in = g.ifield[F_INTNODES * g.n + i];
  full[N_C] = reg[0];
  full[N_B] = reg[1];
  full[N_E] = reg[2];
  full[N_S] = reg[3];
  int loc = 4;
  full[N_CP] = (in & IN_C) ? reg[loc++] : reg[0];
  full[N_BP] = (in & IN_B) ? reg[loc++] : reg[1];
  full[N_EP] = (in & IN_E) ? reg[loc++] : reg[2];
  // :1149-1153
  full[N_IFX] = g.opt.newExcessPhase ? reg[loc++] : -1;
  full[N_DIFX] = g.opt.newExcessPhase ? reg[loc++] : -1;
}

// Stamp tables in the order of the reference's loads.
static const Stamp kStF[] = {{N_C, 0}, {N_B, 0}, {N_E, 0}, {N_CP, 0}, {N_BP, 0}, {N_EP, 0}};
static const Stamp kStQ[] = {{N_B, 0}, {N_S, 0}, {N_CP, 0}, {N_BP, 0}, {N_EP, 0}};
static const Stamp kStLim[] = {{N_CP, 0}, {N_BP, 0}, {N_EP, 0}};
static const Stamp kStDF[] = {
  {N_C, N_C}, {N_C, N_CP}, {N_B, N_B}, {N_B, N_CP}, {N_B, N_BP}, {N_B, N_EP}, {N_E, N_E},
  {N_E, N_EP}, {N_CP, N_C}, {N_CP, N_CP}, {N_CP, N_BP}, {N_CP, N_EP}, {N_BP, N_B},
  {N_BP, N_CP}, {N_BP, N_BP}, {N_BP, N_EP}, {N_EP, N_E}, {N_EP, N_CP}, {N_EP, N_BP},
  {N_EP, N_EP}};
static const Stamp kStDQ[] = {
  {N_B, N_B}, {N_B, N_CP}, {N_S, N_S}, {N_S, N_CP}, {N_CP, N_B}, {N_CP, N_S}, {N_CP, N_CP},
  {N_CP, N_BP}, {N_BP, N_CP}, {N_BP, N_BP}, {N_BP, N_EP}, {N_EP, N_CP}, {N_EP, N_BP},
  {N_EP, N_EP}};
// lead currents: store slots ib, ie, ic, is
static const Stamp kStLead[] = {{T_ib, 0}, {T_ie, 0}, {T_ic, 0}, {T_is, 0}};
// newExcessPhase (:908-931): the base lists, then the excess-phase entries --
// F and Q rows Ifx, dIfx; the limiter's Ifx (operating point) and dIfx
// (transient) rows; dF/dx (C',Ifx) (E',Ifx), the transient rows
// (Ifx,dIfx) (dIfx,C') (dIfx,B') (dIfx,E') (dIfx,dIfx) (dIfx,Ifx) and the
// operating point's (Ifx,C') (Ifx,B') (Ifx,E') (Ifx,Ifx); dQ/dx (Ifx,Ifx)
// (dIfx,dIfx).  An entry the current mode does not stamp is loaded as 0.
static const Stamp kStFX[] = {{N_C, 0}, {N_B, 0}, {N_E, 0}, {N_CP, 0}, {N_BP, 0}, {N_EP, 0},
                              {N_IFX, 0}, {N_DIFX, 0}};
static const Stamp kStQX[] = {{N_B, 0}, {N_S, 0}, {N_CP, 0}, {N_BP, 0}, {N_EP, 0},
                              {N_IFX, 0}, {N_DIFX, 0}};
static const Stamp kStFLimX[] = {{N_CP, 0}, {N_BP, 0}, {N_EP, 0}, {N_IFX, 0}, {N_DIFX, 0}};
static const Stamp kStDFX[] = {
  {N_C, N_C}, {N_C, N_CP}, {N_B, N_B}, {N_B, N_CP}, {N_B, N_BP}, {N_B, N_EP}, {N_E, N_E},
  {N_E, N_EP}, {N_CP, N_C}, {N_CP, N_CP}, {N_CP, N_BP}, {N_CP, N_EP}, {N_BP, N_B},
  {N_BP, N_CP}, {N_BP, N_BP}, {N_BP, N_EP}, {N_EP, N_E}, {N_EP, N_CP}, {N_EP, N_BP},
  {N_EP, N_EP},
  {N_CP, N_IFX}, {N_EP, N_IFX}, {N_IFX, N_DIFX}, {N_DIFX, N_CP}, {N_DIFX, N_BP},
  {N_DIFX, N_EP}, {N_DIFX, N_DIFX}, {N_DIFX, N_IFX}, {N_IFX, N_CP}, {N_IFX, N_BP},
  {N_IFX, N_EP}, {N_IFX, N_IFX}};
static const Stamp kStDQX[] = {
  {N_B, N_B}, {N_B, N_CP}, {N_S, N_S}, {N_S, N_CP}, {N_CP, N_B}, {N_CP, N_S}, {N_CP, N_CP},
  {N_CP, N_BP}, {N_BP, N_CP}, {N_BP, N_BP}, {N_BP, N_EP}, {N_EP, N_CP}, {N_EP, N_BP},
  {N_EP, N_EP}, {N_IFX, N_IFX}, {N_DIFX, N_DIFX}};

static double card_td(const Group& g, int64_t i) {
  return reinterpret_cast<const DevCard*>(&g.mcard[(size_t)g.inst_model[i] * g.mcard_stride])
      ->tdExcess;
}

static void stamps(const Group& g, int64_t i, int kind, std::vector<Stamp>& out) {
  out.clear();
  auto add = [&](const Stamp* s, int m) { out.assign(s, s + m); };
  if (g.opt.newExcessPhase) {
    // an instance without excess phase (PTF 0) loads its F rows and the
    // unit dF/dx diagonals of Ifx and dIfx (:4226-4231), zeros in its other
    // excess-phase dF/dx entries (they are in the jacStamp all the same)
    const bool td = card_td(g, i) != 0;
    switch (kind) {
      case K_F: add(kStFX, 8); break;
      case K_Q: add(kStQX, 7); if (!td) out[5].row = out[6].row = N_NONE; break;
      case K_FLIM:
        if (g.opt.voltageLimiterFlag) {
          add(kStFLimX, 5);
          if (!td) out[3].row = out[4].row = N_NONE;
        }
        break;
      case K_QLIM: if (g.opt.voltageLimiterFlag) add(kStLim, 3); break;
      case K_DFDX: add(kStDFX, 32); break;   // the pattern is the jacStamp's, td or not
      case K_DQDX: add(kStDQX, 16); if (!td) out[14].row = out[15].row = N_NONE; break;
      case K_LEADF:
      case K_LEADQ: if (g.lead[i]) add(kStLead, 4); break;
    }
    return;
  }
  switch (kind) {
    case K_F: add(kStF, 6); break;
    case K_Q: add(kStQ, 5); break;
    case K_FLIM:
    case K_QLIM: if (g.opt.voltageLimiterFlag) add(kStLim, 3); break;
    case K_DFDX: add(kStDF, 20); break;
    case K_DQDX: add(kStDQ, 14); break;
    case K_LEADF:
    case K_LEADQ: if (g.lead[i]) add(kStLead, 4); break;
  }
}

// lambertWCurrent (:1784-1800)
XA_HD void lambertw_current(double& Id, double& Gd, double Vd, double Vte, double Isat) {
  const double RS = 1.0e-12;
  double arg1 = (Vd + Isat * RS) / Vte;
  arg1 = dmin(kMaxExpArg, arg1);
  const double evd = exp(arg1);
  double w, xi;
  int ierr;
  lambertw(Isat * RS * evd / Vte, w, ierr, xi);
  Id = -Isat + Vte * (w) / RS;
  Gd = w / ((1 + w) * RS);
}

// Junction current and conductance of one junction: exponential (or
// Lambert-W) branch above -5 vt, chord conductance below (:2697-2800).
XA_HD void junction(bool lw, double v, double vtJ, double vtL, double isat, double ileak,
                    double gmin, double& i, double& gj, double& il, double& gl) {
  if (v > -5.0 * vtJ) {
    if (lw) {
      lambertw_current(i, gj, v, vtJ, isat);
      if (ileak == 0.0) { il = 0.0; gl = 0.0; }
      else lambertw_current(il, gl, v, vtL, ileak);
    } else {
      const double ev = exp(dmin(kMaxExpArg, v / vtJ));
      i = isat * (ev - 1.0) + gmin * v;
      gj = isat * ev / vtJ + gmin;
      if (ileak == 0.0) {
        il = 0.0; gl = 0.0;
      } else {
        const double evl = exp(dmin(kMaxExpArg, v / vtL));
        il = ileak * (evl - 1.0);
        gl = ileak * evl / vtL;
      }
    }
  } else {
    gj = -isat / v + gmin;
    i = gj * v;
    gl = -ileak / v;
    il = gl * v;
  }
}

// Depletion charge and capacitance of a junction with the linear extension
// above fc (:2900-2990).
XA_HD void depletion(double v, double cz, double pot, double mj, double fcp, double tF,
                     double fA, double fB, double& q, double& c) {
  if (cz == 0.0) {
    q = 0.0; c = 0.0;
  } else if (v < fcp) {
    const double arg = 1.0 - v / pot;
    const double sarg = exp(-mj * log(arg));
    q = pot * cz * (1.0 - arg * sarg) / (1.0 - mj);
    c = cz * sarg;
  } else {
    const double czf = cz / fA;
    q = cz * tF + czf * (fB * (v - fcp) + (mj / (2.0 * pot)) * (v * v - fcp * fcp));
    c = czf * (fB + mj * v / pot);
  }
}

// ---- device evaluation ---------------------------------------------------
template <bool FUSED>
__device__ __forceinline__ void eval(const KArgs& a, int64_t i) {
  if (i >= a.n) return;
  const DevCard& mc = *reinterpret_cast<const DevCard*>(a.mcard + (int64_t)ifld(a, F_MODEL, i) * a.mstride);
  const xa_solver_state& ss = a.ss;
  const double gmin = a.opt.gmin;
  const double T = mc.dtype;
  const int flags = ifld(a, F_FLAGS, i);
  const bool OFF = (flags & FL_OFF) != 0;
  const int32_t lC = node(a, N_C, i), lB = node(a, N_B, i), lE = node(a, N_E, i);
  const int32_t lS = node(a, N_S, i), lCP = node(a, N_CP, i), lBP = node(a, N_BP, i);
  const int32_t lEP = node(a, N_EP, i);
  const int32_t tvbe = stolid(a, T_vBE, i), tvbc = stolid(a, T_vBC, i);
  const int32_t tcap = stolid(a, T_capeqCB, i);
  const double AREA = dfld(a, D_AREA, i), vt = dfld(a, D_VT, i), tVCrit = dfld(a, D_TVCRIT, i);

  // ---- Instance::updateIntermediateVars ----
  const double v_emit = solv(a.v.nextSol, lE), v_emitP = solv(a.v.nextSol, lEP);
  const double v_base = solv(a.v.nextSol, lB), v_baseP = solv(a.v.nextSol, lBP);
  const double v_coll = solv(a.v.nextSol, lC), v_collP = solv(a.v.nextSol, lCP);
  const double v_subst = solv(a.v.nextSol, lS);
  const double vEEp = v_emit - v_emitP, vBBp = v_base - v_baseP, vCCp = v_coll - v_collP;
  double vBE = T * (v_baseP - v_emitP);
  double vBC = T * (v_baseP - v_collP);
  double vBX = T * (v_base - v_collP);
  const double vCS = T * (v_subst - v_collP);
  const double vBE_orig = vBE, vBC_orig = vBC;
  int origFlag = 1, offFlag = 0;
  if (ss.initJctFlag && !OFF && a.opt.voltageLimiterFlag) {
    if (flags & FL_ICGIVEN) {
      vBE = T * dfld(a, D_ICVBE, i);
      const double vCE = T * dfld(a, D_ICVCE, i);
      vBC = vBE - vCE;
      vBX = vBC;
      origFlag = 0;
    } else if (ss.inputOPFlag) {
      // a ground node reads its flag as 0
      const double* fl = a.v.flagSol;
      if (solv(fl, lE) == 0 || solv(fl, lEP) == 0 || solv(fl, lB) == 0 || solv(fl, lBP) == 0 ||
          solv(fl, lC) == 0 || solv(fl, lCP) == 0 || solv(fl, lS) == 0) {
        vBC = 0;
        vBE = tVCrit;
        vBX = vBC;
        origFlag = 0;
      }
    } else {
      vBE = tVCrit;
      vBX = vBC;
      origFlag = 0;
    }
  } else if ((ss.initFixFlag || ss.initJctFlag) && OFF) {
    vBX = vBC = vBE = 0.0;
    offFlag = 1;
  }
  double vBE_old, vBC_old, capeqCB;
  if (ss.newtonIter == 0) {
    if (!ss.dcopFlag || (ss.locaEnabledFlag && ss.dcopFlag)) {
      vBE_old = a.v.currSto[tvbe];
      vBC_old = a.v.currSto[tvbc];
      capeqCB = a.v.currSto[tcap];
    } else {
      vBE_old = vBE;
      vBC_old = vBC;
      capeqCB = 0.0;
    }
  } else {
    vBE_old = a.v.nextSto[tvbe];
    vBC_old = a.v.nextSto[tvbc];
    capeqCB = a.v.nextSto[tcap];
  }
  if (a.opt.voltageLimiterFlag && !(ss.initFixFlag && OFF) && ss.newtonIter >= 0) {
    int icheck = 0, ichk1 = 1;
    vBE = pnjlim(vBE, vBE_old, vt, tVCrit, &icheck);
    vBC = pnjlim(vBC, vBC_old, vt, tVCrit, &ichk1);
    if (ichk1 == 1) icheck = 1;
    if (icheck == 1) origFlag = 0;
  }
  const double csat = dfld(a, D_TSATCUR, i) * AREA;
  const double vtF = vt * mc.NF, vtR = vt * mc.NR, vtE = vt * mc.NE, vtC = vt * mc.NC;
  const double iLeakBE = dfld(a, D_TBELEAK, i) * AREA, iLeakBC = dfld(a, D_TBCLEAK, i) * AREA;
  const bool lw = (flags & FL_LAMBERTW) != 0;
  double iBE, gBE, iBEleak, gBEleak, iBC, gBC, iBCleak, gBCleak;
  junction(lw, vBE, vtF, vtE, csat, iLeakBE, gmin, iBE, gBE, iBEleak, gBEleak);
  junction(lw, vBC, vtR, vtC, csat, iLeakBC, gmin, iBC, gBC, iBCleak, gBCleak);

  // base charge
  const double rofF = mc.invRollOffF / AREA, rofR = mc.invRollOffR / AREA;
  const double q1 = 1.0 / (1.0 - mc.invEarlyF * vBC - mc.invEarlyR * vBE);
  double qB, dqBdvEp, dqBdvCp;
  if (rofF == 0.0 && rofR == 0.0) {
    qB = q1;
    const double q1_qB = q1 * qB;
    dqBdvEp = q1_qB * mc.invEarlyR;
    dqBdvCp = q1_qB * mc.invEarlyF;
  } else {
    const double q2 = rofF * iBE + rofR * iBC;
    if (q2 >= 0) {
      const double ar = 1.0 + 4.0 * q2;
      double sqarg = 1.0, fg = 0.0, rg = 0.0;
      if (fabs(ar) > 0.0) sqarg = pow(ar, mc.rollOffExp);
      if (ar != 0) fg = rofF * gBE * 2 * mc.rollOffExp * sqarg / ar;
      if (ar != 0) rg = rofR * gBC * 2 * mc.rollOffExp * sqarg / ar;
      qB = 0.5 * q1 * (1.0 + sqarg);
      dqBdvEp = q1 * (qB * mc.invEarlyR + fg);
      dqBdvCp = q1 * (qB * mc.invEarlyF + rg);
    } else {
      qB = q1;
      dqBdvEp = q1 * qB * mc.invEarlyR;
      dqBdvCp = q1 * qB * mc.invEarlyF;
    }
  }
  const double dqBdvBp = -(dqBdvEp + dqBdvCp);
  const double invqB = 1.0 / qB;

  // charges and capacitances
  const double tBEPot = dfld(a, D_TBEPOT, i), tBCPot = dfld(a, D_TBCPOT, i);
  const double ctot = dfld(a, D_TBCCAP, i) * AREA;
  const double czBC = ctot * mc.XCJC;
  const double czBX = ctot - czBC;
  const double czBE = dfld(a, D_TBECAP, i) * AREA;
  const double czCS = mc.CJS * AREA;
  const double fcpc = mc.FC * mc.VJC;
  const double fcpe = mc.FC * tBEPot;
  double iBEhighCurr = iBE, gBEhighCurr = gBE;
  if (mc.TF != 0.0 && vBE > 0.0 && ((!ss.dcopFlag) || ss.tranopFlag || ss.acopFlag)) {
    double argtf = 0.0, arg2 = 0.0, arg3 = 0.0;
    if (mc.XTF != 0.0) {
      argtf = mc.XTF;
      if (mc.vbcFac != 0.0) argtf *= exp(vBC * mc.vbcFac);
      arg2 = argtf;
      if (mc.ITF != 0.0) {
        const double tmp = iBEhighCurr / (iBEhighCurr + mc.ITF * AREA);
        argtf *= tmp * tmp;
        arg2 = argtf * (3.0 - 2.0 * tmp);
      }
      arg3 = iBEhighCurr * argtf * mc.vbcFac;
    }
    iBEhighCurr *= (1.0 + argtf) / qB;
    gBEhighCurr = (gBEhighCurr * (1.0 + arg2) - iBEhighCurr * dqBdvEp) / qB;
    capeqCB = mc.TF * (arg3 - iBEhighCurr * dqBdvCp) / qB;
  }
  double qBEdep, capBEdep, qBCdep, capBCdep, qBX, capBX, qCS, capCS;
  const double tF1 = tBEPot * mc.fxBE / (1.0 - mc.MJE);
  const double tF5 = tBCPot * mc.fxBC / (1.0 - mc.MJC);
  depletion(vBE, czBE, tBEPot, mc.MJE, fcpe, tF1, mc.f2, mc.f3, qBEdep, capBEdep);
  const double qBEdiff = (mc.TF == 0.0) ? 0.0 : mc.TF * iBEhighCurr;
  const double capBEdiff = (mc.TF == 0.0) ? 0.0 : mc.TF * gBEhighCurr;
  depletion(vBC, czBC, tBCPot, mc.MJC, fcpc, tF5, mc.f6, mc.f7, qBCdep, capBCdep);
  const double qBCdiff = (mc.TR == 0.0) ? 0.0 : mc.TR * iBC;
  const double capBCdiff = (mc.TR == 0.0) ? 0.0 : mc.TR * gBC;
  depletion(vBX, czBX, tBCPot, mc.MJC, fcpc, tF5, mc.f6, mc.f7, qBX, capBX);
  if (czCS == 0.0) {
    qCS = 0.0;
    capCS = 0.0;
  } else if (vCS < 0.0) {
    const double arg = 1.0 - vCS / mc.VJS;
    const double sarg = exp(-mc.MJS * log(arg));
    qCS = mc.VJS * czCS * (1.0 - arg * sarg) / (1.0 - mc.MJS);
    capCS = czCS * sarg;
  } else {
    qCS = vCS * czCS * (1.0 + mc.MJS * vCS / (2.0 * mc.VJS));
    capCS = czCS * (1.0 + mc.MJS * vCS / mc.VJS);
  }

  // old-style excess phase (:2370-2466)
  double iEX = iBE, gEX = gBE, iC = 0.0;
  const double td = mc.tdExcess;
  if (!ss.dcopFlag && td != 0) {
    const int32_t acx = stalid(a, A_CEXBC, i);
    double currCexbc, lastCexbc;
    if (ss.beginIntegrationFlag) {
      currCexbc = lastCexbc = iBE / qB;
      a.v.currSta[acx] = currCexbc;
      a.v.lastSta[acx] = lastCexbc;
    } else {
      currCexbc = a.v.currSta[acx];
      lastCexbc = a.v.lastSta[acx];
    }
    const double dt0 = ss.currTimeStep, dt1 = ss.lastTimeStep;
    double arg1 = dt0 / td;
    const double arg2 = 3.0 * arg1;
    arg1 = arg2 * arg1;
    const double denom = 1.0 + arg1 + arg2;
    const double phaseScalar = arg1 / denom;
    iC = ((currCexbc) * (1 + dt0 / dt1 + arg2) - (lastCexbc) * dt0 / dt1) / denom;
    iEX = iBE * phaseScalar;
    gEX = gBE * phaseScalar;
    a.v.nextSta[acx] = iC + iEX / qB;
  }
  double iCE = (iEX - iBC) / qB;
  const double tBetaF = dfld(a, D_TBETAF, i), tBetaR = dfld(a, D_TBETAR, i);
  iC += iCE - iBC / tBetaR - iBCleak;
  const double iB = iBE / tBetaF + iBEleak + iBC / tBetaR + iBCleak;
  double iE = -iC - iB;
  double diCEdvEp = invqB * (iCE * dqBdvEp - gEX);
  double diCEdvCp = invqB * (iCE * dqBdvCp + gBC);
  double diCEdvBp = invqB * (iCE * dqBdvBp + gEX - gBC);
  const bool nep = a.opt.newExcessPhase != 0;
  double i_fx = 0.0, di_fx = 0.0, diBEdvBp = 0.0, diBEdvCp = 0.0, diBEdvEp = 0.0;
  if (nep) {
    // auxDAECalculations (:1928-1980): the new-DAE currents take i_fx from
    // the solution; the old calculation above still updates its state
    const int32_t lX = node(a, N_IFX, i), lD = node(a, N_DIFX, i);
    i_fx = solv(a.v.nextSol, lX);
    di_fx = solv(a.v.nextSol, lD);
    if (td != 0 && !ss.dcopFlag) iCE = i_fx - iBC / qB;
    else iCE = (iBE - iBC) / qB;
    iC = iCE - iBC / tBetaR - iBCleak;
    iE = -iC - iB;
    if (td != 0) {
      if (!ss.dcopFlag) {
        diCEdvBp = invqB * (-invqB * iBC * dqBdvBp - gBC);
        diCEdvCp = invqB * (-invqB * iBC * dqBdvCp + gBC);
        diCEdvEp = -invqB * invqB * iBC * dqBdvEp;
      } else {
        diCEdvBp = invqB * (iCE * dqBdvBp + gBE - gBC);
        diCEdvCp = invqB * (iCE * dqBdvCp + gBC);
        diCEdvEp = invqB * (iCE * dqBdvEp - gBE);
      }
      diBEdvBp = invqB * (invqB * iBE * dqBdvBp + gBE);
      diBEdvCp = invqB * invqB * iBE * dqBdvCp;
      diBEdvEp = invqB * (invqB * iBE * dqBdvEp - gBE);
    } else {
      diCEdvBp = invqB * (iCE * dqBdvBp + gBE - gBC);
      diCEdvCp = invqB * (iCE * dqBdvCp + gBC);
      diCEdvEp = invqB * (iCE * dqBdvEp - gBE);
    }
  }
  const double gEpr = mc.gEmit * AREA, gCpr = mc.gColl * AREA;
  double gX;
  {
    const double rBpr = mc.RBM / AREA;
    const double rBpi = mc.RB / AREA - rBpr;
    const double xjrB = mc.IRB * AREA;
    gX = rBpr + rBpi / qB;
    if (fabs(xjrB) > 0.0) {
      double arg1 = dmax(iB / xjrB, 1.0e-09);
      const double arg2 = (-1.0 + sqrt(1.0 + 14.59025 * arg1)) / 2.4317 / sqrt(arg1);
      arg1 = tan(arg2);
      gX = rBpr + 3.0 * rBpi * (arg1 - arg2) / arg2 / arg1 / arg1;
    }
    if (fabs(gX) > 0.0) gX = 1.0 / gX;
  }
  const double diBrdvB = gX, diBrdvCp = 0.0, diBrdvEp = 0.0, diBrdvBp = -gX;
  const double gBEtot = gBE / tBetaF + gBEleak;
  const double gBCtot = gBC / tBetaR + gBCleak;

  // ---- Master::updateState ----
  {
    double* sto = a.v.nextSto;
    sto[tvbe] = vBE;
    sto[tvbc] = vBC;
    sto[tcap] = capeqCB;
    double* sta = a.v.nextSta;
    const int32_t s0 = stalid(a, A_qBEdiff, i), s1 = stalid(a, A_qBEdep, i);
    const int32_t s2 = stalid(a, A_qCS, i), s3 = stalid(a, A_qBCdiff, i);
    const int32_t s4 = stalid(a, A_qBCdep, i), s5 = stalid(a, A_qBX, i);
    sta[s0] = qBEdiff; sta[s1] = qBEdep; sta[s3] = qBCdiff; sta[s4] = qBCdep;
    sta[s5] = qBX; sta[s2] = qCS;
    if (!ss.dcopFlag && ss.initTranFlag && ss.newtonIter == 0) {
      double* cs = a.v.currSta;
      cs[s0] = qBEdiff; cs[s1] = qBEdep; cs[s3] = qBCdiff; cs[s4] = qBCdep;
      cs[s5] = qBX; cs[s2] = qCS;
    }
  }

  // ---- Master::loadDAEVectors ----  (a -= b is stamped as += -b)
  Sink<FUSED> s(a, i);
  s.put(K_F, 0, -(-vCCp * gCpr));
  s.put(K_F, 1, -(-vBBp * gX));
  s.put(K_F, 2, -(-vEEp * gEpr));
  s.put(K_F, 3, -(vCCp * gCpr + T * (-iC)));
  s.put(K_F, 4, -(vBBp * gX - T * (iB)));
  s.put(K_F, 5, -(vEEp * gEpr + T * (-iE)));
  if (nep) {
    // :3997-4024 (the operating point's assignment of the dIfx entry adds
    // zero: no other device writes that row)
    if (td != 0) {
      if (!ss.dcopFlag) {
        s.put(K_F, 6, -di_fx);
        s.put(K_F, 7, 3 * di_fx * td + 3 * i_fx - 3 * iBE / qB);
      } else {
        s.put(K_F, 6, i_fx - iBE / qB);
        s.put(K_F, 7, 0.0);
      }
    } else {
      s.put(K_F, 6, i_fx);
      s.put(K_F, 7, di_fx);
    }
    if (td != 0) {   // :4033-4046
      s.put(K_Q, 5, i_fx);
      s.put(K_Q, 6, !ss.dcopFlag ? di_fx * td * td : 0.0);
    }
  }
  s.put(K_Q, 0, -(-T * qBX));
  s.put(K_Q, 1, -(-T * qCS));
  s.put(K_Q, 2, -(T * (qCS + qBX + qBCdep + qBCdiff)));
  s.put(K_Q, 3, -(-T * (qBEdep + qBEdiff + qBCdep + qBCdiff)));
  s.put(K_Q, 4, -(T * (qBEdep + qBEdiff)));
  if (a.opt.voltageLimiterFlag) {
    // corrections are loaded only when a voltage moved; otherwise the
    // reference adds 0.0
    double Cp_f = 0.0, Bp_f = 0.0, Ep_f = 0.0, Cp_q = 0.0, Bp_q = 0.0, Ep_q = 0.0;
    if (!origFlag || offFlag) {
      const double vbe_diff = vBE - vBE_orig, vbc_diff = vBC - vBC_orig;
      const double vce_diff = vbe_diff - vbc_diff;
      Cp_f = +diCEdvBp * vbe_diff + diCEdvCp * vce_diff - gBCtot * vbc_diff;
      Cp_f *= T;
      Bp_f = gBEtot * vbe_diff + gBCtot * vbc_diff;
      Bp_f *= T;
      Ep_f = -diCEdvCp * vce_diff - (diCEdvBp + gBEtot) * vbe_diff;
      Ep_f *= T;
      Cp_q = -(capBCdep + capBCdiff) * vbc_diff;
      Cp_q *= T;
      Bp_q = (capBEdep + capBEdiff) * vbe_diff + (capBCdiff + capBCdep + capeqCB) * vbc_diff;
      Bp_q *= T;
      Ep_q = -capeqCB * vbc_diff - (capBEdiff + capBEdep) * vbe_diff;
      Ep_q *= T;
    }
    s.put(K_FLIM, 0, Cp_f);
    s.put(K_FLIM, 1, Bp_f);
    s.put(K_FLIM, 2, Ep_f);
    if (nep && td != 0) {
      // :4069-4100: the dIfx row in transient, the Ifx row at the operating
      // point
      double ex = 0.0;
      if (!origFlag || offFlag) {
        const double vbe_diff = vBE - vBE_orig, vbc_diff = vBC - vBC_orig;
        const double vce_diff = vbe_diff - vbc_diff;
        ex = !ss.dcopFlag ? -3 * (diBEdvBp * vbe_diff + diBEdvCp * vce_diff)
                          : (diBEdvBp * vbe_diff + diBEdvCp * vce_diff);
        ex *= T;
      }
      s.put(K_FLIM, 3, ss.dcopFlag ? ex : 0.0);
      s.put(K_FLIM, 4, ss.dcopFlag ? 0.0 : ex);
    }
    s.put(K_QLIM, 0, Cp_q);
    s.put(K_QLIM, 1, Bp_q);
    s.put(K_QLIM, 2, Ep_q);
  }
  // lead currents (stored, not added): ib, ie, ic, is
  s.put(K_LEADF, 0, T * (iB));
  s.put(K_LEADF, 1, T * (iE));
  s.put(K_LEADF, 2, T * (iC));
  s.put(K_LEADF, 3, 0.0);
  s.put(K_LEADQ, 0, T * (qBX + qBEdep + qBEdiff + qBCdep + qBCdiff));
  s.put(K_LEADQ, 1, -T * (qBEdep + qBEdiff));
  s.put(K_LEADQ, 2, -T * (qCS + qBX + qBCdep + qBCdiff));
  s.put(K_LEADQ, 3, T * qCS);

  // ---- Master::loadDAEMatrices ----
  s.put(K_DFDX, 0, gCpr);
  s.put(K_DFDX, 1, -gCpr);
  s.put(K_DFDX, 2, diBrdvB);
  s.put(K_DFDX, 3, diBrdvCp);
  s.put(K_DFDX, 4, diBrdvBp);
  s.put(K_DFDX, 5, diBrdvEp);
  s.put(K_DFDX, 6, gEpr);
  s.put(K_DFDX, 7, -gEpr);
  s.put(K_DFDX, 8, -gCpr);
  s.put(K_DFDX, 9, diCEdvCp + gBCtot + gCpr);
  s.put(K_DFDX, 10, diCEdvBp - gBCtot);
  s.put(K_DFDX, 11, diCEdvEp);
  s.put(K_DFDX, 12, -diBrdvB);
  s.put(K_DFDX, 13, -diBrdvCp - gBCtot);
  s.put(K_DFDX, 14, -diBrdvBp + gBEtot + gBCtot);
  s.put(K_DFDX, 15, -diBrdvEp - gBEtot);
  s.put(K_DFDX, 16, -gEpr);
  s.put(K_DFDX, 17, -diCEdvCp);
  s.put(K_DFDX, 18, -diCEdvBp - gBEtot);
  s.put(K_DFDX, 19, gBEtot + gEpr + diCEdvBp + diCEdvCp);
  if (nep) {
    // :4177-4233 (an entry the mode does not stamp takes 0)
    const bool tr = td != 0 && !ss.dcopFlag, dc = td != 0 && ss.dcopFlag;
    s.put(K_DFDX, 20, tr ? 1.0 * T : 0.0);
    s.put(K_DFDX, 21, tr ? -1.0 * T : 0.0);
    s.put(K_DFDX, 22, tr ? -1.0 : 0.0);
    s.put(K_DFDX, 23, tr ? -3 * diBEdvCp * T : 0.0);
    s.put(K_DFDX, 24, tr ? -3 * diBEdvBp * T : 0.0);
    s.put(K_DFDX, 25, tr ? -3 * diBEdvEp * T : 0.0);
    s.put(K_DFDX, 26, tr ? 3 * td : 1.0);
    s.put(K_DFDX, 27, tr ? 3.0 : 0.0);
    s.put(K_DFDX, 28, dc ? -diBEdvCp * T : 0.0);
    s.put(K_DFDX, 29, dc ? -diBEdvBp * T : 0.0);
    s.put(K_DFDX, 30, dc ? -diBEdvEp * T : 0.0);
    s.put(K_DFDX, 31, tr ? 0.0 : 1.0);
  }
  s.put(K_DQDX, 0, capBX);
  s.put(K_DQDX, 1, -capBX);
  s.put(K_DQDX, 2, capCS);
  s.put(K_DQDX, 3, -capCS);
  s.put(K_DQDX, 4, -capBX);
  s.put(K_DQDX, 5, -capCS);
  s.put(K_DQDX, 6, capCS + capBX + capBCdep + capBCdiff);
  s.put(K_DQDX, 7, -capBCdep - capBCdiff);
  s.put(K_DQDX, 8, -capBCdiff - capBCdep - capeqCB);
  s.put(K_DQDX, 9, capBEdiff + capBEdep + capBCdiff + capBCdep + capeqCB);
  s.put(K_DQDX, 10, -capBEdiff - capBEdep);
  s.put(K_DQDX, 11, capeqCB);
  s.put(K_DQDX, 12, -capBEdiff - capBEdep - capeqCB);
  s.put(K_DQDX, 13, capBEdiff + capBEdep);
  if (nep && td != 0) {   // :4253-4260
    s.put(K_DQDX, 14, !ss.dcopFlag ? 1.0 : 0.0);
    s.put(K_DQDX, 15, !ss.dcopFlag ? 1 * td * td : 0.0);
  }
}

template <bool FUSED>
__global__ void __launch_bounds__(256) k_update(KArgs a) {
  eval<FUSED>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (FUSED) fused_writeout(a);
}

static bool has_excess_phase(const Group& g) {
  for (size_t k = 0; k < g.models.size(); ++k)
    if (reinterpret_cast<const DevCard*>(&g.mcard[k * g.mcard_stride])->tdExcess != 0) return true;
  return false;
}

static int launch_update(Group& g, const xa_solver_state& ss, const xa_vectors& v, bool fused) {
  if (!v.nextSol || !v.nextSto || !v.nextSta)
    return fail(XA_ERR_ARG, "bjt update: nextSol/nextSto/nextSta required");
  if (ss.newtonIter == 0 && (!ss.dcopFlag || ss.locaEnabledFlag) && !v.currSto)
    return fail(XA_ERR_ARG, "bjt update: currSto required");
  if (!ss.dcopFlag && ss.initTranFlag && ss.newtonIter == 0 && !v.currSta)
    return fail(XA_ERR_ARG, "bjt update: currSta required");
  if (!ss.dcopFlag && has_excess_phase(g) && (!v.currSta || !v.lastSta))
    return fail(XA_ERR_ARG, "bjt update: currSta/lastSta required with excess phase (PTF)");
  if (ss.inputOPFlag && ss.initJctFlag && !v.flagSol)
    return fail(XA_ERR_ARG, "bjt update: flagSol required with inputOPFlag");
  KArgs a;
  fill_kargs(g, ss, v, fused, a);
  return launch_model_kernel(g, a, fused, (const void*)k_update<false>, (const void*)k_update<true>);
}

// Algorithmic bytes per instance (no internal nodes, transient, no excess
// phase): 14 SoA columns, 4 node values + 7 LIDs, 3 store LIDs + 3 old
// stores read + 3 written, 7 state LIDs + 6 states written, model index and
// flags.
static double traffic(const Group& g, int phase) {
  (void)g;
  if (phase == 0) return 8.0 * (ND + 4 + 3 + 3 + 6) + 4.0 * (7 + 3 + 7 + 2);
  if (phase == 1) return 16.0 * (6 /*F*/ + 5 /*Q*/ + 3 + 3 /*limiter*/);
  return 16.0 * (20 + 14);
}

}  // namespace bjt

const ModelImpl* model_bjt() {
  static ModelImpl m = [] {
    ModelImpl x{};
    x.type = XA_BJT;
    x.name = "bjt";
    x.mparams = bjt::kModelParams;
    x.n_mparams = sizeof(bjt::kModelParams) / sizeof(ParamDef);
    x.n_mslots = bjt::NM;
    x.iparams = bjt::kInstParams;
    x.n_iparams = sizeof(bjt::kInstParams) / sizeof(ParamDef);
    x.n_islots = bjt::NI;
    x.max_nodes = bjt::NNODE;
    x.max_sto = bjt::NSTO;
    x.max_sta = bjt::NSTA;
    const int E[NK] = {6, 5, 3, 3, 20, 14, 4, 4};
    for (int k = 0; k < NK; ++k) x.max_entries[k] = E[k];
    x.setup = bjt::setup;
    x.resolve_nodes = bjt::resolve_nodes;
    x.stamps = bjt::stamps;
    x.launch_update = bjt::launch_update;
    x.launch_secondary = nullptr;
    x.skip_kind = nullptr;
    x.state_names = nullptr;
    x.n_state = 0;
    x.traffic = bjt::traffic;
    x.fused_block = 64;
    x.fused_at_scale = true;
    x.fused_small = true;
    return x;
  }();
  return &m;
}

}  // namespace xa
