// model_diode.hip -- junction diode (levels 1 and 2): host-side parameter
// processing and the per-instance evaluation kernel.
//
// Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_Diode.C
//   parameters :70-318, Model::processParams :1816-1839, Model ctor :1870-1940,
//   Instance ctor :338-456, registerLIDs :479-500, jacStamp :440-452,
//   updateTemperature :1376-1593, updateIntermediateVars :917-1366,
//   lambertW* :1604-1802, Master::updateState :2051-2068,
//   Master::loadDAEVectors :2078-2122, Master::loadDAEMatrices :2132-2171.
#include <cmath>
#include <cstring>
#include "xa_internal.h"
#include "xa_kernel.h"

namespace xa {
namespace diode {

// ---- parameter tables ---------------------------------------------------
enum MSlot {
  M_IS, M_RS, M_N, M_ISR, M_NR, M_IKF, M_TT, M_CJO, M_VJ, M_M, M_EG, M_XTI,
  M_TIKF, M_TBV1, M_TBV2, M_TRS1, M_TRS2, M_FC, M_BV, M_IBV, M_IRF, M_NBV,
  M_IBVL, M_NBVL, M_TNOM, M_KF, M_AF, NM
};
static const ParamDef kModelParams[] = {
  {"IS", M_IS, 1.0e-14, 0}, {"JS", M_IS, 1.0e-14, 0}, {"RS", M_RS, 0.0, 0},
  {"N", M_N, 1.0, 0}, {"ISR", M_ISR, 0.0, 0}, {"NR", M_NR, 2.0, 0},
  {"IKF", M_IKF, 0.0, 0}, {"TT", M_TT, 0.0, 0}, {"CJO", M_CJO, 0.0, 0},
  {"CJ", M_CJO, 0.0, 0}, {"CJ0", M_CJO, 0.0, 0}, {"VJ", M_VJ, 1.0, 0},
  {"M", M_M, 0.5, 0}, {"EG", M_EG, 1.11, 0}, {"XTI", M_XTI, 3.0, 0},
  {"TIKF", M_TIKF, 0.0, 0}, {"TBV1", M_TBV1, 0.0, 0}, {"TBV2", M_TBV2, 0.0, 0},
  {"TRS1", M_TRS1, 0.0, 0}, {"TRS2", M_TRS2, 0.0, 0}, {"FC", M_FC, 0.5, 0},
  {"BV", M_BV, 1e99, 0}, {"VB", M_BV, 1e99, 0}, {"IBV", M_IBV, 1.0e-3, 0},
  {"IRF", M_IRF, 1.0, 0}, {"NBV", M_NBV, 1.0, 0}, {"IBVL", M_IBVL, 0.0, 0},
  {"NBVL", M_NBVL, 1.0, 0}, {"TNOM", M_TNOM, 0.0, 1}, {"KF", M_KF, 0.0, 0},
  {"AF", M_AF, 1.0, 0},
};
enum ISlot { I_AREA, I_IC, I_TEMP, I_OFF, I_LAMBERTW, NI };
static const ParamDef kInstParams[] = {
  {"AREA", I_AREA, 1.0, 0}, {"IC", I_IC, 0.0, 0}, {"TEMP", I_TEMP, 0.0, 1},
  {"OFF", I_OFF, 0.0, 0}, {"LAMBERTW", I_LAMBERTW, 0.0, 0},
};

// ---- device-side layout ------------------------------------------------
struct DevCard {  // model constants the evaluation reads
  double M, BV, IBV, NBV, IBVL, NBVL, N, NR, TT, F2, F3, RS, CJO;
  double BVGiven, level;
};
enum DField {
  D_AREA, D_INITCOND, D_TEMP, D_TJCTPOT, D_TJCTCAP, D_TDEPCAP, D_TSATCUR,
  D_TVCRIT, D_TF1, D_TBRKDWNV, D_TSATCURR, D_TIKF, D_TCOND, D_TIRF, ND
};
enum IField { F_MODEL, F_OFF, F_ICGIVEN, F_LAMBERTW, NIF };
enum State { S_VD, NS };
static const char* const kStateNames[NS] = {"Vd"};
enum Node { N_POS, N_NEG, N_PRI, NNODE };

// ---- host setup ----------------------------------------------------------
struct HostModel {
  double v[NM];
  int level;
  bool BVGiven;
  double COND, F2, F3;
};

static int setup(Group& g) {
  const int64_t n = g.n;
  std::vector<HostModel> hm(g.models.size());
  for (size_t k = 0; k < g.models.size(); ++k) {
    const Card& c = g.models[k];
    HostModel& m = hm[k];
    if (c.level != 1 && c.level != 2)
      return fail(XA_ERR_UNSUPPORTED, "diode: level must be 1 or 2");
    for (int s = 0; s < NM; ++s) m.v[s] = c.v[s];
    m.level = c.level;
    m.BVGiven = c.has(M_BV);
    if (!c.has(M_TNOM)) m.v[M_TNOM] = g.opt.tnom;
    // Model::processParams (N_DEV_Diode.C:1816-1839)
    if (m.v[M_M] > 0.9) m.v[M_M] = 0.9;
    if (m.v[M_EG] < 0.1) m.v[M_EG] = 0.1;
    if (m.v[M_FC] > 0.95) m.v[M_FC] = 0.95;
    m.COND = (m.v[M_RS] == 0.0) ? 0.0 : 1.0 / m.v[M_RS];
    const double xfc = log(1.0 - m.v[M_FC]);
    m.F2 = exp((1.0 + m.v[M_M]) * xfc);
    m.F3 = 1.0 - m.v[M_FC] * (1.0 + m.v[M_M]);
  }
  g.n_dfield = ND;
  g.n_ifield = NIF;
  g.dfield.assign((size_t)ND * n, 0.0);
  g.ifield.assign((size_t)NIF * n, 0);
  g.n_state = NS;
  g.state_init.assign((size_t)NS * n, 0.0);
  auto D = [&](int f, int64_t i) -> double& { return g.dfield[f * n + i]; };
  auto I = [&](int f, int64_t i) -> int32_t& { return g.ifield[f * n + i]; };
  for (int64_t i = 0; i < n; ++i) {
    HostModel& m = hm[g.inst_model[i]];
    const double area = g.ival[I_AREA][i];
    const int lambertW = g.igiven[I_LAMBERTW] ? (int)g.ival[I_LAMBERTW][i] : g.opt.lambertWFlag;
    const double Temp = g.igiven[I_TEMP] ? g.ival[I_TEMP][i] : g.opt.temp;
    // Instance ctor: node/state counts, and the RS tweak applied to the
    // owning model for Lambert-W instances (N_DEV_Diode.C:436-445)
    g.n_ext[i] = 2;
    g.n_int[i] = (m.v[M_RS] == 0.0 || lambertW == 1) ? 0 : 1;
    g.n_sta[i] = (m.v[M_CJO] == 0.0) ? 1 : 0;
    g.n_sto[i] = 1 + (g.lead[i] ? 1 : 0);
    if (m.v[M_RS] == 0.0 && lambertW == 1) m.v[M_RS] = 1.0e-12;

    // Instance::updateTemperature (N_DEV_Diode.C:1376-1593)
    const double TNOM = m.v[M_TNOM];
    const double vtnom = kKoverQ * TNOM;
    const double xfc = log(1.0 - m.v[M_FC]);
    const double vt = kKoverQ * Temp;
    const double fact2 = Temp / kRefTemp;
    const double egfet = kEg0 - (kAlphaEg * Temp * Temp) / (Temp + kBetaEg);
    const double arg = -egfet / (2.0 * kBoltz * Temp) + kEg300 / (kBoltz * (kRefTemp + kRefTemp));
    const double pbfact = -2.0 * vt * (1.5 * log(fact2) + kQ * arg);
    const double egfet1 = kEg0 - (kAlphaEg * TNOM * TNOM) / (TNOM + kBetaEg);
    const double arg1 = -egfet1 / (2.0 * kBoltz * TNOM) + kEg300 / (2.0 * kBoltz * kRefTemp);
    const double fact1 = TNOM / kRefTemp;
    const double pbfact1 = -2.0 * vtnom * (1.5 * log(fact1) + kQ * arg1);
    const double pbo = (m.v[M_VJ] - pbfact1) / fact1;
    const double gmaold = (m.v[M_VJ] - pbo) / pbo;
    double tJctCap = m.v[M_CJO] / (1.0 + m.v[M_M] * (4.0e-4 * (TNOM - kRefTemp) - gmaold));
    double tJctPot = pbfact + fact2 * pbo;
    const double gmanew = (tJctPot - pbo) / pbo;
    tJctCap *= 1.0 + m.v[M_M] * (4.0e-4 * (Temp - kRefTemp) - gmanew);
    const double tSatCur = m.v[M_IS] * exp(((Temp / TNOM) - 1.0) * m.v[M_EG] / (m.v[M_N] * vt) +
                                           m.v[M_XTI] / m.v[M_N] * log(Temp / TNOM));
    const double tF1 = tJctPot * (1.0 - exp((1.0 - m.v[M_M]) * xfc)) / (1.0 - m.v[M_M]);
    const double tDepCap = m.v[M_FC] * tJctPot;
    const double vte = m.v[M_N] * vt;
    const double tVcrit = vte * log(vte / (kRoot2 * tSatCur));
    double tCOND = m.COND;
    const double tIRF = m.v[M_IRF] * pow(fact2, 1.6);
    double tSatCurR = 0.0, tIKF = 0.0, tempBV;
    if (m.level == 2) {
      tSatCurR = m.v[M_ISR] * exp((Temp / TNOM - 1.0) * m.v[M_EG] / (m.v[M_NR] * vt) +
                                  m.v[M_XTI] / m.v[M_NR] * log(Temp / TNOM));
      tIKF = m.v[M_IKF] * (1 + m.v[M_TIKF] * (Temp - TNOM));
      tempBV = m.v[M_BV] * (1 + (Temp - TNOM) * (m.v[M_TBV1] + m.v[M_TBV2] * (Temp - TNOM)));
      const double tRS = m.v[M_RS] * (1 + (Temp - TNOM) * (m.v[M_TRS1] + m.v[M_TRS2] * (Temp - TNOM)));
      tCOND = 0.0;
      if (tRS != 0.0) tCOND = 1.0 / tRS;
      tJctPot = (m.v[M_VJ] - egfet1) * fact2 - 3 * vt * log(fact2) + egfet;
      tJctCap = m.v[M_CJO] / (1.0 + m.v[M_M] * (4.0e-4 * (Temp - TNOM) + (1 - tJctPot / m.v[M_VJ])));
    } else {
      tempBV = m.v[M_BV];
    }
    double tBrkdwnV = 0.0;
    if (m.BVGiven) {
      // breakdown knee by fixed-count Picard iteration (N_DEV_Diode.C:1534-1581)
      const double cbv = m.v[M_IBV];
      double a1 = 3.0 * vte / (XA_CONSTE * tempBV);
      a1 = a1 * a1 * a1;
      const double lo = tSatCur * tIRF * (1 - a1);
      const double hi = tSatCur * tIRF * (1 - 1.0 / (XA_CONSTE * XA_CONSTE * XA_CONSTE)) * exp(-1.0 * (3.0 * vte - tempBV) / vte);
      if (cbv >= hi) {
        tBrkdwnV = 3.0 * vte;
      } else if (cbv <= lo) {
        tBrkdwnV = tempBV;
      } else {
        double xbv = tempBV;
        for (int it = 0; it < 8; ++it) {
          double a2 = 3.0 * vte / (XA_CONSTE * xbv);
          a2 = a2 * a2 * a2;
          xbv = tempBV - vte * log(cbv / (tSatCur * tIRF)) + vte * log(1 - a2);
        }
        tBrkdwnV = xbv;
      }
    }
    D(D_AREA, i) = area;
    D(D_INITCOND, i) = g.ival[I_IC][i];
    D(D_TEMP, i) = Temp;
    D(D_TJCTPOT, i) = tJctPot;
    D(D_TJCTCAP, i) = tJctCap;
    D(D_TDEPCAP, i) = tDepCap;
    D(D_TSATCUR, i) = tSatCur;
    D(D_TVCRIT, i) = tVcrit;
    D(D_TF1, i) = tF1;
    D(D_TBRKDWNV, i) = tBrkdwnV;
    D(D_TSATCURR, i) = tSatCurR;
    D(D_TIKF, i) = tIKF;
    D(D_TCOND, i) = tCOND;
    D(D_TIRF, i) = tIRF;
    I(F_MODEL, i) = g.inst_model[i];
    I(F_OFF, i) = (int32_t)g.ival[I_OFF][i];
    I(F_ICGIVEN, i) = g.igiven[I_IC] ? 1 : 0;
    I(F_LAMBERTW, i) = lambertW;
  }
  // device cards (after the Lambert-W RS tweak, as the reference's model is)
  g.mcard_stride = sizeof(DevCard) / sizeof(double);
  g.mcard.assign(g.models.size() * g.mcard_stride, 0.0);
  for (size_t k = 0; k < hm.size(); ++k) {
    const HostModel& m = hm[k];
    DevCard c{m.v[M_M], m.v[M_BV], m.v[M_IBV], m.v[M_NBV], m.v[M_IBVL], m.v[M_NBVL],
              m.v[M_N], m.v[M_NR], m.v[M_TT], m.F2, m.F3, m.v[M_RS], m.v[M_CJO],
              m.BVGiven ? 1.0 : 0.0, (double)m.level};
    std::memcpy(&g.mcard[k * g.mcard_stride], &c, sizeof(c));
  }
  return XA_OK;
}

static void resolve_nodes(const Group& g, int64_t i, const int32_t* reg, int32_t* full) {
  full[N_POS] = reg[0];
  full[N_NEG] = reg[1];
  full[N_PRI] = g.n_int[i] ? reg[2] : reg[0];   // RS = 0: Pri is Pos
}

// Stamp tables in the order of the reference's loads.
static const Stamp kF[] = {{N_POS, 0}, {N_NEG, 0}, {N_PRI, 0}};
static const Stamp kQ[] = {{N_NEG, 0}, {N_PRI, 0}};
static const Stamp kLim[] = {{N_NEG, 0}, {N_PRI, 0}};
static const Stamp kDF[] = {{N_POS, N_POS}, {N_POS, N_PRI}, {N_NEG, N_NEG}, {N_NEG, N_PRI},
                            {N_PRI, N_POS}, {N_PRI, N_NEG}, {N_PRI, N_PRI}};
static const Stamp kDQ[] = {{N_NEG, N_NEG}, {N_NEG, N_PRI}, {N_PRI, N_NEG}, {N_PRI, N_PRI}};

static void stamps(const Group& g, int64_t i, int kind, std::vector<Stamp>& out) {
  out.clear();
  auto add = [&](const Stamp* s, int m) { out.assign(s, s + m); };
  switch (kind) {
    case K_F: add(kF, 3); break;
    case K_Q: add(kQ, 2); break;
    case K_FLIM: case K_QLIM: if (g.opt.voltageLimiterFlag) add(kLim, 2); break;
    case K_DFDX: add(kDF, 7); break;
    case K_DQDX: add(kDQ, 4); break;
    case K_LEADF: if (g.lead[i]) out.push_back({1, 0}); break;
    case K_LEADQ:
      if (g.lead[i] && g.models[g.inst_model[i]].v[M_CJO] != 0.0) out.push_back({1, 0});
      break;
  }
}

// ---- device evaluation ---------------------------------------------------
template <bool FUSED>
__device__ __forceinline__ void eval(const KArgs& a, int64_t i) {
  if (i >= a.n) return;
  const DevCard mc = *reinterpret_cast<const DevCard*>(a.mcard + (int64_t)ifld(a, F_MODEL, i) * a.mstride);
  const int32_t lPos = node(a, N_POS, i), lNeg = node(a, N_NEG, i), lPri = node(a, N_PRI, i);
  const int32_t liVd = stolid(a, 0, i);
  const xa_solver_state& ss = a.ss;
  const double gmin = a.opt.gmin;

  // Master::updateState stores the previous Vd before evaluating.
  const double VdPrev = st(a, S_VD, i);
  if (liVd >= 0) a.v.nextSto[liVd] = VdPrev;

  const double Area = dfld(a, D_AREA, i);
  const double tSatCur = dfld(a, D_TSATCUR, i);
  const double tVcrit = dfld(a, D_TVCRIT, i);
  const double tBrkdwnV = dfld(a, D_TBRKDWNV, i);
  const double Vpp = solv(a.v.nextSol, lPri);
  const double Vn = solv(a.v.nextSol, lNeg);
  const double Vp = solv(a.v.nextSol, lPos);
  double Vd = Vpp - Vn;
  double Isat = tSatCur * Area;
  const double IsatR = dfld(a, D_TSATCURR, i) * Area;
  const double Vt = kKoverQ * dfld(a, D_TEMP, i);
  const double Vte = mc.N * Vt;
  const double VteR = mc.NR * Vt;
  const double Gspr = dfld(a, D_TCOND, i) * Area;
  const double Vd_orig = Vd;
  const int lambertW = ifld(a, F_LAMBERTW, i);

  double Vd_old;
  if (ss.newtonIter == 0) {
    if (ss.initJctFlag && a.opt.voltageLimiterFlag) {
      if (ifld(a, F_ICGIVEN, i)) {
        Vd = dfld(a, D_INITCOND, i);
      } else if (ifld(a, F_OFF, i)) {
        Vd = 0.0;
      } else if (ss.inputOPFlag) {
        const double* fl = a.v.flagSol;
        const bool unset = !fl || lPos < 0 || lNeg < 0 || lPri < 0 ||
                           fl[lPos] == 0 || fl[lNeg] == 0 || fl[lPri] == 0;
        if (unset) Vd = tVcrit;
      } else {
        Vd = tVcrit;
      }
    }
    Vd_old = Vd;
    if (!ss.dcopFlag || (ss.locaEnabledFlag && ss.dcopFlag))
      Vd_old = (liVd >= 0) ? a.v.currSto[liVd] : VdPrev;
  } else {
    Vd_old = VdPrev;   // nextSto[li_storevd], written above
  }

  if (a.opt.voltageLimiterFlag && ss.newtonIter >= 0) {
    int ichk = 0;
    if (mc.BVGiven != 0.0 && Vd < dmin(0.0, -mc.BV + 10.0 * Vte)) {
      double Vdtmp = -(mc.BV + Vd);
      Vdtmp = pnjlim(Vdtmp, -(Vd_old + mc.BV), Vte, tVcrit, &ichk);
      Vd = -(Vdtmp + mc.BV);
    } else {
      Vd = pnjlim(Vd, Vd_old, Vte, tVcrit, &ichk);
    }
  }

  double Id = 0.0, Gd = 0.0, Vc = Vd;
  const int level = (int)mc.level;
  if (level == 1 && (lambertW == 1 || lambertW == 2)) {
    // Lambert-W current forms (N_DEV_Diode.C:1604-1802); Icd is zero there.
    const double RS = (lambertW == 1) ? mc.RS : 1.0e-15;
    double w, xi;
    int ierr;
    if (Vd >= -3.0 * Vte) {
      double e1 = (Vd + Isat * RS) / Vte;
      e1 = dmin(kMaxExpArg, e1);
      const double evd = exp(e1);
      lambertw(Isat * RS * evd / Vte, w, ierr, xi);
      Id = -Isat + Vte * w / RS;
      Gd = w / ((1 + w) * RS);
    } else if (!tBrkdwnV || Vd >= -tBrkdwnV) {
      const double FF1 = (Vd + tBrkdwnV) / (-3.0 * Vte + tBrkdwnV);
      const double FF2 = (0 - FF1) * (-2);   // (1/2 - FF1)*(-2) with integer 1/2 == 0
      const double ar = Vte / RS;
      double a1 = (Isat / ar - 3);
      const double a2 = FF1 * a1;
      a1 = dmin(kMaxExpArg, a2);
      const double evd = exp(a2);
      lambertw(Isat * evd / ar, w, ierr, xi);
      Id = -Isat * FF1 + FF2 * w * ar;
      const double GdFF1 = 1 / (-3.0 * Vte + tBrkdwnV);
      const double GdFF2 = 2 * GdFF1;
      const double GdW = ar * w * GdFF1 * a1 / (1 + w);
      Gd = -Isat * GdFF1 + GdFF2 * ar * w + FF2 * GdW;
    } else {
      double e1 = (-Vd - tBrkdwnV) / Vte;
      e1 = dmin(kMaxExpArg, e1);
      const double evd = exp(e1);
      lambertw(Isat * RS * evd / Vte, w, ierr, xi);
      Id = -Vte * w / RS;
      Gd = w / ((1 + w) * RS);
    }
    Vc = (lambertW == 1) ? Vd - (Id + 0.0) * RS : Vd;
  } else {
    const double IRfactor = (Vd >= -3.0 * Vte) ? 1.0 : dfld(a, D_TIRF, i);
    Isat *= IRfactor;
    if (Vd >= -3.0 * Vte) {
      double e1 = dmin(kMaxExpArg, Vd / Vte);
      double evd = exp(e1);
      if (level == 1) {
        Id = Isat * (evd - 1.0) + gmin * Vd;
        Gd = Isat * evd / Vte + gmin;
      } else {
        const double Inorm = Isat * (evd - 1.0) + gmin * Vd;
        const double Gd1 = Isat * evd / Vte + gmin;
        e1 = dmin(kMaxExpArg, Vd / VteR);
        evd = exp(e1);
        const double Irec = IsatR * (evd - 1.0);
        const double Gd2 = IsatR * evd / VteR;
        double Khi = 1, DKhi = 0;
        const double tIKF = dfld(a, D_TIKF, i);
        if (tIKF > 0) {
          Khi = sqrt(tIKF / (tIKF + Inorm));
          DKhi = 0.5 * Khi * Gd1 / (tIKF + Inorm);
        }
        double Kgen = 0, DKgen = 0;
        if (Irec != 0) {
          const double tJctPot = dfld(a, D_TJCTPOT, i);
          Kgen = sqrt(pow(((1 - Vd / tJctPot) * (1 - Vd / tJctPot) + 0.005), mc.M));
          DKgen = -mc.M * (1 - Vd / tJctPot) / (tJctPot * Kgen);
        }
        Id = Inorm * Khi + Irec * Kgen;
        Gd = Gd1 * Khi + Inorm * DKhi + Gd2 * Kgen + Irec * DKgen;
      }
    } else if (!tBrkdwnV || Vd >= -tBrkdwnV) {
      double ar = 3.0 * Vte / (Vd * XA_CONSTE);
      ar = ar * ar * ar;
      Id = -Isat * (1.0 + ar) + gmin * Vd;
      Gd = Isat * 3.0 * ar / Vd + gmin;
    } else {
      double e1 = dmin(kMaxExpArg, -(tBrkdwnV + Vd) / Vte);
      const double evrev = exp(e1);
      double a2 = 3.0 * Vte / (XA_CONSTE * tBrkdwnV);
      a2 = a2 * a2 * a2;
      const double IsatB = Isat * (1 - a2);
      Id = -IsatB * evrev + gmin * Vd;
      Gd = IsatB * evrev / Vte + gmin;
    }
    Vc = Vd;
  }

  double Qd = 0.0, Cd = 0.0;
  const double tJctCap = dfld(a, D_TJCTCAP, i);
  if (tJctCap != 0.0) {
    const double Czero = tJctCap * Area;
    const double tJctPot = dfld(a, D_TJCTPOT, i);
    const double tDepCap = dfld(a, D_TDEPCAP, i);
    if (Vc < tDepCap) {
      const double ar = 1.0 - Vc / tJctPot;
      double e1 = -mc.M * log(ar);
      e1 = dmin(kMaxExpArg, e1);
      const double sarg = exp(e1);
      Qd = mc.TT * Id + tJctPot * Czero * (1.0 - ar * sarg) / (1.0 - mc.M);
      Cd = mc.TT * Gd + Czero * sarg;
    } else {
      const double Czof2 = Czero / mc.F2;
      const double MotJctPot = mc.M / tJctPot;
      Qd = mc.TT * Id + Czero * dfld(a, D_TF1, i) +
           Czof2 * (mc.F3 * (Vc - tDepCap) + (0.5 * MotJctPot) * (Vc * Vc - tDepCap * tDepCap));
      Cd = mc.TT * Gd + Czof2 * (mc.F3 + MotJctPot * Vc);
    }
  }
  st(a, S_VD, i) = Vd;

  // ---- stamps (Master::loadDAEVectors / loadDAEMatrices) ----
  Sink<FUSED> s(a, i);
  const double Ir = Gspr * (Vp - Vpp);
  s.put(K_F, 0, Ir);
  s.put(K_F, 1, -Id);
  s.put(K_F, 2, -(-Id + Ir));
  s.put(K_Q, 0, -Qd);
  s.put(K_Q, 1, Qd);
  if (a.opt.voltageLimiterFlag) {
    const double Vd_diff = Vd - Vd_orig;
    const double Cd_Jdxp = -(Cd) * Vd_diff;
    const double Gd_Jdxp = -(Gd) * Vd_diff;
    s.put(K_FLIM, 0, Gd_Jdxp);
    s.put(K_FLIM, 1, -Gd_Jdxp);
    s.put(K_QLIM, 0, Cd_Jdxp);
    s.put(K_QLIM, 1, -Cd_Jdxp);
  }
  s.put(K_LEADF, 0, Id);
  s.put(K_LEADQ, 0, Qd);
  s.put(K_DFDX, 0, Gspr);
  s.put(K_DFDX, 1, -Gspr);
  s.put(K_DFDX, 2, Gd);
  s.put(K_DFDX, 3, -Gd);
  s.put(K_DFDX, 4, -Gspr);
  s.put(K_DFDX, 5, -Gd);
  s.put(K_DFDX, 6, Gspr + Gd);
  s.put(K_DQDX, 0, Cd);
  s.put(K_DQDX, 1, -Cd);
  s.put(K_DQDX, 2, -Cd);
  s.put(K_DQDX, 3, Cd);
}

template <bool FUSED>
__global__ void __launch_bounds__(256) k_update(KArgs a) {
  eval<FUSED>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (FUSED) fused_writeout(a);
}

static int launch_update(Group& g, const xa_solver_state& ss, const xa_vectors& v, bool fused) {
  if (!v.nextSol || !v.nextSto) return fail(XA_ERR_ARG, "diode update: nextSol/nextSto required");
  if (ss.newtonIter == 0 && !v.currSto) return fail(XA_ERR_ARG, "diode update: currSto required");
  KArgs a;
  fill_kargs(g, ss, v, fused, a);
  return launch_model_kernel(g, a, fused, (const void*)k_update<false>, (const void*)k_update<true>);
}

// Algorithmic bytes per instance (default card: RS = 0, limiter on).
static double traffic(const Group& g, int phase) {
  const double nodes = 2, params = 9, lids = 4;
  if (phase == 0) return 8 * (nodes + params + 2 /*state r/w*/ + 1 /*sto w*/) + 4 * lids;
  if (phase == 1) return 16 * (2 /*F*/ + 2 /*Q*/ + (g.opt.voltageLimiterFlag ? 4 : 0));
  return 16 * (4 /*dFdx*/ + 4 /*dQdx*/);
}

}  // namespace diode

const ModelImpl* model_diode() {
  static ModelImpl m = [] {
    ModelImpl x{};
    x.type = XA_DIODE;
    x.name = "diode";
    x.mparams = diode::kModelParams;
    x.n_mparams = sizeof(diode::kModelParams) / sizeof(ParamDef);
    x.n_mslots = diode::NM;
    x.iparams = diode::kInstParams;
    x.n_iparams = sizeof(diode::kInstParams) / sizeof(ParamDef);
    x.n_islots = diode::NI;
    x.max_nodes = diode::NNODE;
    x.max_sto = 2;
    x.max_sta = 1;
    const int E[NK] = {3, 2, 2, 2, 7, 4, 1, 1};
    for (int k = 0; k < NK; ++k) x.max_entries[k] = E[k];
    x.setup = diode::setup;
    x.resolve_nodes = diode::resolve_nodes;
    x.stamps = diode::stamps;
    x.launch_update = diode::launch_update;
    x.launch_secondary = nullptr;
    x.skip_kind = nullptr;
    x.state_names = diode::kStateNames;
    x.n_state = diode::NS;
    x.traffic = diode::traffic;
    x.fused_block = 256;
    x.fused_at_scale = true;
    x.fused_small = true;
    return x;
  }();
  return &m;
}

}  // namespace xa
