// model_mosfet1.hip -- Level-1 (Shichman-Hodges) MOSFET: host-side
// parameter processing and the per-instance evaluation + stamp kernel.
//
// Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET1.C
//   parameters :71-262, Instance ctor :309-1256, registerLIDs :1278-1372,
//   registerStateLIDs :1452-1503, registerStoreLIDs :1513-1535,
//   updateIntermediateVars :2584-3393, updateTemperature :3403-3646,
//   Model::processParams :3752-3824, Model ctor :3856-3946,
//   Master::updateState :4069-4158, Master::loadDAEVectors :4168-4396,
//   Master::loadDAEMatrices :4407-4688.
// newMeyerFlag = 1 is rejected at setup (XA_ERR_UNSUPPORTED).
#include <cmath>
#include <cstring>
#include "xa_internal.h"
#include "xa_kernel.h"

namespace xa {
namespace mosfet1 {

// ---- parameter tables ---------------------------------------------------
enum MSlot {
  M_L, M_W, M_VTO, M_KP, M_GAMMA, M_PHI, M_LAMBDA, M_RD, M_RS, M_CBD, M_CBS, M_IS, M_PB,
  M_CGSO, M_CGDO, M_CGBO, M_RSH, M_CJ, M_MJ, M_CJSW, M_MJSW, M_JS, M_TOX, M_LD, M_UO, M_U0,
  M_FC, M_NSUB, M_NSS, M_TNOM, M_KF, M_AF, M_TPG, M_TYPE, NM
};
static const ParamDef kModelParams[] = {
  {"L", M_L, 1e-4, 0}, {"W", M_W, 1e-4, 0}, {"VTO", M_VTO, 0.0, 0}, {"KP", M_KP, 2e-5, 0},
  {"GAMMA", M_GAMMA, 0.0, 0}, {"PHI", M_PHI, 0.6, 0}, {"LAMBDA", M_LAMBDA, 0.0, 0},
  {"RD", M_RD, 0.0, 0}, {"RS", M_RS, 0.0, 0}, {"CBD", M_CBD, 0.0, 0}, {"CBS", M_CBS, 0.0, 0},
  {"IS", M_IS, 1e-14, 0}, {"PB", M_PB, 0.8, 0}, {"CGSO", M_CGSO, 0.0, 0},
  {"CGDO", M_CGDO, 0.0, 0}, {"CGBO", M_CGBO, 0.0, 0}, {"RSH", M_RSH, 0.0, 0},
  {"CJ", M_CJ, 0.0, 0}, {"MJ", M_MJ, 0.5, 0}, {"CJSW", M_CJSW, 0.0, 0}, {"MJSW", M_MJSW, 0.5, 0},
  {"JS", M_JS, 0.0, 0}, {"TOX", M_TOX, 1e-7, 0}, {"LD", M_LD, 0.0, 0}, {"UO", M_UO, 600.0, 0},
  {"U0", M_U0, 600.0, 0}, {"FC", M_FC, 0.5, 0}, {"NSUB", M_NSUB, 0.0, 0},
  {"NSS", M_NSS, 0.0, 0}, {"TNOM", M_TNOM, 27.0, 1}, {"KF", M_KF, 0.0, 0},
  {"AF", M_AF, 1.0, 0}, {"TPG", M_TPG, 0.0, 0}, {"TYPE", M_TYPE, 1.0, 0},
};
enum ISlot { I_TEMP, I_L, I_W, I_AD, I_AS, I_NRD, I_NRS, I_PD, I_PS, I_M, I_IC1, I_IC2, I_IC3, I_OFF, NI };
static const ParamDef kInstParams[] = {
  {"TEMP", I_TEMP, 0.0, 1}, {"L", I_L, 0.0, 0}, {"W", I_W, 0.0, 0}, {"AD", I_AD, 0.0, 0},
  {"AS", I_AS, 0.0, 0}, {"NRD", I_NRD, 1.0, 0}, {"NRS", I_NRS, 1.0, 0}, {"PD", I_PD, 0.0, 0},
  {"PS", I_PS, 0.0, 0}, {"M", I_M, 1.0, 0}, {"IC1", I_IC1, 0.0, 0}, {"IC2", I_IC2, 0.0, 0},
  {"IC3", I_IC3, 0.0, 0}, {"OFF", I_OFF, 0.0, 0},
};

// ---- device-side layout ------------------------------------------------
struct DevCard {   // model values the evaluation reads
  double dtype, gamma, lambda, MJ, MJSW;
};
enum DField {
  D_SRCSAT, D_DRNSAT, D_BETA, D_TPHI, D_TVBI, D_TVTO, D_VT, D_SVCRIT, D_DVCRIT, D_CBS, D_CBSSW,
  D_CBD, D_CBDSW, D_TDEPCAP, D_TBULKPOT, D_F2S, D_F3S, D_F4S, D_F2D, D_F3D, D_F4D, D_OXCAP,
  D_CGSOV, D_CGDOV, D_CGBOV, D_GD, D_GS, D_M, D_ICVDS, D_ICVGS, D_ICVBS, ND
};
enum IField { F_MODEL, F_OFF, F_ICGIVEN, NIF };
enum Node { N_D, N_G, N_S, N_B, N_DP, N_SP, NNODE };
enum Sto { T_vbd, T_vbs, T_vgs, T_vds, T_von, T_id, T_ig, T_is, T_ib, NSTO };
enum Sta { A_qgs, A_qgd, A_qgb, A_capgs, A_capgd, A_capgb, A_qbd, A_qbs, NSTA };

// ---- host setup ----------------------------------------------------------
struct HostModel {
  int dtype;
  double tnom, model_l, model_w, latDiff, jctSatCurDensity, jctSatCur, drainResistance;
  double sourceResistance, sheetResistance, transconductance, cgsoF, cgdoF, cgboF;
  double oxideCapFactor, vt0, capBD, capBS, bulkCapFactor, sideWallCapFactor, bulkJctPotential;
  double MJ, MJSW, fwdCapDepCoeff, phi, gamma, lambda, substrateDoping, gateType;
  double surfaceStateDensity, oxideThickness, surfaceMobility;
  bool capBDGiven, capBSGiven, bulkCapFactorGiven, sideWallCapFactorGiven, rshGiven;
  double fact1, vtnom, egfet1, pbfact1;
};

// Model ctor + Model::processParams (:3856-3946, :3752-3824)
static int model_init(HostModel& m, const Card& c, const xa_device_options& opt) {
  m.dtype = c.v[M_TYPE] < 0 ? -1 : 1;
  m.model_l = c.has(M_L) ? c.v[M_L] : opt.defl;
  m.model_w = c.has(M_W) ? c.v[M_W] : opt.defw;
  m.vt0 = c.v[M_VTO]; m.transconductance = c.v[M_KP]; m.gamma = c.v[M_GAMMA];
  m.phi = c.v[M_PHI]; m.lambda = c.v[M_LAMBDA]; m.drainResistance = c.v[M_RD];
  m.sourceResistance = c.v[M_RS]; m.capBD = c.v[M_CBD]; m.capBS = c.v[M_CBS];
  m.jctSatCur = c.v[M_IS]; m.bulkJctPotential = c.v[M_PB]; m.cgsoF = c.v[M_CGSO];
  m.cgdoF = c.v[M_CGDO]; m.cgboF = c.v[M_CGBO]; m.sheetResistance = c.v[M_RSH];
  m.bulkCapFactor = c.v[M_CJ]; m.MJ = c.v[M_MJ]; m.sideWallCapFactor = c.v[M_CJSW];
  m.MJSW = c.v[M_MJSW]; m.jctSatCurDensity = c.v[M_JS]; m.oxideThickness = c.v[M_TOX];
  m.latDiff = c.v[M_LD]; m.surfaceMobility = c.v[M_UO]; m.fwdCapDepCoeff = c.v[M_FC];
  m.substrateDoping = c.v[M_NSUB]; m.surfaceStateDensity = c.v[M_NSS];
  m.tnom = c.has(M_TNOM) ? c.v[M_TNOM] : opt.tnom;
  m.gateType = c.v[M_TPG];
  m.capBDGiven = c.has(M_CBD) || m.capBD != 0;
  m.capBSGiven = c.has(M_CBS) || m.capBS != 0;
  m.bulkCapFactorGiven = c.has(M_CJ);
  m.sideWallCapFactorGiven = c.has(M_CJSW);
  m.rshGiven = c.has(M_RSH);
  if (c.has(M_U0)) {
    if (c.has(M_UO)) return fail(XA_ERR_ARG, "mosfet1: both UO and U0 given");
    m.surfaceMobility = c.v[M_U0];
  }
  m.fact1 = m.tnom / kRefTemp;
  m.vtnom = m.tnom * kKoverQ;
  const double kt1 = kBoltz * m.tnom;
  m.egfet1 = 1.16 - (7.02e-4 * m.tnom * m.tnom) / (m.tnom + 1108);
  const double arg1 = -m.egfet1 / (kt1 + kt1) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
  m.pbfact1 = -2 * m.vtnom * (1.5 * log(m.fact1) + kQ * arg1);
  if (!c.has(M_TOX) || m.oxideThickness == 0) {
    m.oxideCapFactor = 0;
  } else {
    m.oxideCapFactor = 3.9 * 8.854214871e-12 / m.oxideThickness;
    if (!c.has(M_KP)) {
      if (!c.has(M_UO) && !c.has(M_U0)) m.surfaceMobility = 600;
      m.transconductance = m.surfaceMobility * m.oxideCapFactor * 1e-4;
    }
    if (c.has(M_NSUB)) {
      if (m.substrateDoping * 1e6 > 1.45e16) {
        if (!c.has(M_PHI)) {
          m.phi = 2 * m.vtnom * log(m.substrateDoping * 1e6 / 1.45e16);
          m.phi = dmax(0.1, m.phi);
        }
        const double fermis = m.dtype * .5 * m.phi;
        double wkfng = 3.2;
        if (!c.has(M_TPG)) m.gateType = 1;
        if (m.gateType != 0) {
          const double fermig = m.dtype * m.gateType * .5 * m.egfet1;
          wkfng = 3.25 + .5 * m.egfet1 - fermig;
        }
        const double wkfngs = wkfng - (3.25 + .5 * m.egfet1 + fermis);
        if (!c.has(M_GAMMA))
          m.gamma = sqrt(2 * 11.70 * 8.854214871e-12 * kQ * m.substrateDoping * 1e6) / m.oxideCapFactor;
        if (!c.has(M_VTO)) {
          if (!c.has(M_NSS)) m.surfaceStateDensity = 0;
          const double vfb = wkfngs - m.surfaceStateDensity * 1e4 * kQ / m.oxideCapFactor;
          m.vt0 = vfb + m.dtype * (m.gamma * sqrt(m.phi) + m.phi);
        }
      } else {
        return fail(XA_ERR_ARG, "mosfet1: NSUB < Ni");
      }
    }
  }
  return XA_OK;
}

static int setup(Group& g) {
  if (g.opt.newMeyerFlag) return fail(XA_ERR_UNSUPPORTED, "mosfet1: newMeyerFlag");
  const int64_t n = g.n;
  std::vector<HostModel> hm(g.models.size());
  for (size_t k = 0; k < g.models.size(); ++k) {
    if (g.models[k].level != 1) return fail(XA_ERR_UNSUPPORTED, "mosfet1: level must be 1");
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
  for (int64_t i = 0; i < n; ++i) {
    const HostModel& m = hm[g.inst_model[i]];
    // Instance ctor (:1168-1255)
    const double temp = g.igiven[I_TEMP] ? iv(I_TEMP, i) : g.opt.temp;
    const double l = g.igiven[I_L] ? iv(I_L, i) : m.model_l;
    const double w = g.igiven[I_W] ? iv(I_W, i) : m.model_w;
    const double drainSquares = iv(I_NRD, i), sourceSquares = iv(I_NRS, i);
    double gd = 0, gs = 0;
    if (m.drainResistance != 0) gd = 1 / m.drainResistance;
    else if (m.rshGiven && m.sheetResistance != 0) gd = 1 / (m.sheetResistance * drainSquares);
    if (m.sourceResistance != 0) gs = 1 / m.sourceResistance;
    else if (m.rshGiven && m.sheetResistance != 0) gs = 1 / (m.sheetResistance * sourceSquares);
    const double drainArea = g.igiven[I_AD] ? iv(I_AD, i) : g.opt.defad;
    const double sourceArea = g.igiven[I_AS] ? iv(I_AS, i) : g.opt.defas;
    const double drainPerimeter = iv(I_PD, i), sourcePerimeter = iv(I_PS, i);
    if (l - 2 * m.latDiff <= 0)
      return fail(XA_ERR_ARG, "mosfet1: effective channel length less than zero");
    const double Leff = l - 2 * m.latDiff;
    g.n_ext[i] = 4;
    g.n_int[i] = (gs == 0.0 ? 0 : 1) + (gd == 0.0 ? 0 : 1);
    g.n_sta[i] = NSTA;
    g.n_sto[i] = 5 + (g.lead[i] ? 4 : 0);

    // Instance::updateTemperature (:3403-3646)
    const double ratio = temp / m.tnom;
    const double vt = temp * kKoverQ;
    const double fact2 = temp / kRefTemp;
    const double kt = temp * kBoltz;
    const double egfet = 1.16 - (7.02e-4 * temp * temp) / (temp + 1108);
    const double arg0 = -egfet / (kt + kt) + 1.1150877 / (kBoltz * (kRefTemp + kRefTemp));
    const double pbfact = -2 * vt * (1.5 * log(fact2) + kQ * arg0);
    const double ratio4 = ratio * sqrt(ratio);
    const double tTransconductance = m.transconductance / ratio4;
    const double phio = (m.phi - m.pbfact1) / m.fact1;
    const double tPhi = fact2 * phio + pbfact;
    const double tVbi = m.vt0 - m.dtype * (m.gamma * sqrt(m.phi)) + .5 * (m.egfet1 - egfet) +
                        m.dtype * .5 * (tPhi - m.phi);
    const double tVto = tVbi + m.dtype * m.gamma * sqrt(tPhi);
    const double tSatCur = m.jctSatCur * exp(-egfet / vt + m.egfet1 / m.vtnom);
    const double tSatCurDens = m.jctSatCurDensity * exp(-egfet / vt + m.egfet1 / m.vtnom);
    const double pbo = (m.bulkJctPotential - m.pbfact1) / m.fact1;
    const double gmaold = (m.bulkJctPotential - pbo) / pbo;
    double capfact = 1 / (1 + m.MJ * (4e-4 * (m.tnom - kRefTemp) - gmaold));
    double tCbd = m.capBD * capfact, tCbs = m.capBS * capfact, tCj = m.bulkCapFactor * capfact;
    capfact = 1 / (1 + m.MJSW * (4e-4 * (m.tnom - kRefTemp) - gmaold));
    double tCjsw = m.sideWallCapFactor * capfact;
    const double tBulkPot = fact2 * pbo + pbfact;
    const double gmanew = (tBulkPot - pbo) / pbo;
    capfact = (1 + m.MJ * (4e-4 * (temp - kRefTemp) - gmanew));
    tCbd *= capfact;
    tCbs *= capfact;
    tCj *= capfact;
    capfact = (1 + m.MJSW * (4e-4 * (temp - kRefTemp) - gmanew));
    tCjsw *= capfact;
    const double tDepCap = m.fwdCapDepCoeff * tBulkPot;
    double drainVcrit, sourceVcrit;
    if ((m.jctSatCurDensity == 0) || (drainArea == 0) || (sourceArea == 0)) {
      sourceVcrit = drainVcrit = vt * log(vt / (kRoot2 * m.jctSatCur));
    } else {
      drainVcrit = vt * log(vt / (kRoot2 * m.jctSatCurDensity * drainArea));
      sourceVcrit = vt * log(vt / (kRoot2 * m.jctSatCurDensity * sourceArea));
    }
    const double czbd = m.capBDGiven ? tCbd : (m.bulkCapFactorGiven ? tCj * drainArea : 0);
    const double czbdsw = m.sideWallCapFactorGiven ? tCjsw * drainPerimeter : 0;
    const double arg = 1 - m.fwdCapDepCoeff;
    const double sarg = exp((-m.MJ) * log(arg));
    const double sargsw = exp((-m.MJSW) * log(arg));
    const double f2d = czbd * (1 - m.fwdCapDepCoeff * (1 + m.MJ)) * sarg / arg +
                       czbdsw * (1 - m.fwdCapDepCoeff * (1 + m.MJSW)) * sargsw / arg;
    const double f3d = czbd * m.MJ * sarg / arg / tBulkPot + czbdsw * m.MJSW * sargsw / arg / tBulkPot;
    const double f4d = czbd * tBulkPot * (1 - arg * sarg) / (1 - m.MJ) +
                       czbdsw * tBulkPot * (1 - arg * sargsw) / (1 - m.MJSW) -
                       f3d / 2 * (tDepCap * tDepCap) - tDepCap * f2d;
    const double czbs = m.capBSGiven ? tCbs : (m.bulkCapFactorGiven ? tCj * sourceArea : 0);
    const double czbssw = m.sideWallCapFactorGiven ? tCjsw * sourcePerimeter : 0;
    const double f2s = czbs * (1 - m.fwdCapDepCoeff * (1 + m.MJ)) * sarg / arg +
                       czbssw * (1 - m.fwdCapDepCoeff * (1 + m.MJSW)) * sargsw / arg;
    const double f3s = czbs * m.MJ * sarg / arg / tBulkPot + czbssw * m.MJSW * sargsw / arg / tBulkPot;
    const double f4s = czbs * tBulkPot * (1 - arg * sarg) / (1 - m.MJ) +
                       czbssw * tBulkPot * (1 - arg * sargsw) / (1 - m.MJSW) -
                       f3s / 2 * (tDepCap * tDepCap) - tDepCap * f2s;
    // saturation currents and beta as updateIntermediateVars forms them (:2621-2631)
    double DrainSatCur, SourceSatCur;
    if ((tSatCurDens == 0) || (drainArea == 0) || (sourceArea == 0)) {
      DrainSatCur = tSatCur;
      SourceSatCur = tSatCur;
    } else {
      DrainSatCur = tSatCurDens * drainArea;
      SourceSatCur = tSatCurDens * sourceArea;
    }
    D(D_SRCSAT, i) = SourceSatCur;
    D(D_DRNSAT, i) = DrainSatCur;
    D(D_BETA, i) = tTransconductance * w / Leff;
    D(D_TPHI, i) = tPhi;
    D(D_TVBI, i) = tVbi;
    D(D_TVTO, i) = tVto;
    D(D_VT, i) = vt;
    D(D_SVCRIT, i) = sourceVcrit;
    D(D_DVCRIT, i) = drainVcrit;
    D(D_CBS, i) = czbs;
    D(D_CBSSW, i) = czbssw;
    D(D_CBD, i) = czbd;
    D(D_CBDSW, i) = czbdsw;
    D(D_TDEPCAP, i) = tDepCap;
    D(D_TBULKPOT, i) = tBulkPot;
    D(D_F2S, i) = f2s; D(D_F3S, i) = f3s; D(D_F4S, i) = f4s;
    D(D_F2D, i) = f2d; D(D_F3D, i) = f3d; D(D_F4D, i) = f4d;
    D(D_OXCAP, i) = m.oxideCapFactor * Leff * w;
    D(D_CGSOV, i) = m.cgsoF * w;
    D(D_CGDOV, i) = m.cgdoF * w;
    D(D_CGBOV, i) = m.cgboF * Leff;
    D(D_GD, i) = gd;
    D(D_GS, i) = gs;
    D(D_M, i) = iv(I_M, i);
    D(D_ICVDS, i) = iv(I_IC1, i);
    D(D_ICVGS, i) = iv(I_IC2, i);
    D(D_ICVBS, i) = iv(I_IC3, i);
    I(F_MODEL, i) = g.inst_model[i];
    I(F_OFF, i) = iv(I_OFF, i) != 0.0;
    I(F_ICGIVEN, i) = (g.igiven[I_IC1] || g.igiven[I_IC2] || g.igiven[I_IC3]) ? 1 : 0;
  }
  g.mcard_stride = sizeof(DevCard) / sizeof(double);
  g.mcard.assign(g.models.size() * g.mcard_stride, 0.0);
  for (size_t k = 0; k < hm.size(); ++k) {
    const DevCard c{(double)hm[k].dtype, hm[k].gamma, hm[k].lambda, hm[k].MJ, hm[k].MJSW};
    std::memcpy(&g.mcard[k * g.mcard_stride], &c, sizeof(c));
  }
  return XA_OK;
}

static void resolve_nodes(const Group& g, int64_t i, const int32_t* reg, int32_t* full) {
  const int64_t n = g.n;
  full[N_D] = reg[0];
  full[N_G] = reg[1];
  full[N_S] = reg[2];
  full[N_B] = reg[3];
  int loc = 4;
  full[N_DP] = g.dfield[D_GD * n + i] != 0.0 ? reg[loc++] : reg[0];
  full[N_SP] = g.dfield[D_GS * n + i] != 0.0 ? reg[loc++] : reg[2];
}

// Stamp tables in the order of the reference's loads; a row of -1 marks an
// entry the instance does not load.
constexpr int16_t N_NONE = -1;
static const Stamp kStF[] = {{N_D, 0}, {N_G, 0}, {N_S, 0}, {N_B, 0}, {N_DP, 0}, {N_SP, 0}};
static const Stamp kStQ[] = {{N_G, 0}, {N_B, 0}, {N_DP, 0}, {N_SP, 0}};
static const Stamp kStFLim[] = {{N_B, 0}, {N_DP, 0}, {N_SP, 0}};
static const Stamp kStQLim[] = {{N_G, 0}, {N_B, 0}, {N_DP, 0}, {N_SP, 0}};
static const Stamp kStDF[] = {
  {N_D, N_D}, {N_D, N_DP}, {N_S, N_S}, {N_S, N_SP}, {N_B, N_B}, {N_B, N_DP}, {N_B, N_SP},
  {N_DP, N_D}, {N_DP, N_G}, {N_DP, N_B}, {N_DP, N_DP}, {N_DP, N_SP},
  {N_SP, N_G}, {N_SP, N_S}, {N_SP, N_B}, {N_SP, N_DP}, {N_SP, N_SP}};
static const Stamp kStDQ[] = {
  {N_G, N_G}, {N_G, N_B}, {N_G, N_DP}, {N_G, N_SP}, {N_B, N_G}, {N_B, N_B}, {N_B, N_DP},
  {N_B, N_SP}, {N_DP, N_G}, {N_DP, N_B}, {N_DP, N_DP}, {N_SP, N_G}, {N_SP, N_B}, {N_SP, N_SP}};
// lead currents: store slots id, is, ig, ib (F) and id, is, ig, ib (Q)
static const Stamp kStLead[] = {{T_id, 0}, {T_is, 0}, {T_ig, 0}, {T_ib, 0}};

static void stamps(const Group& g, int64_t i, int kind, std::vector<Stamp>& out) {
  out.clear();
  const int64_t n = g.n;
  const bool dex = g.dfield[D_GD * n + i] != 0.0, sex = g.dfield[D_GS * n + i] != 0.0;
  auto add = [&](const Stamp* s, int m) { out.assign(s, s + m); };
  switch (kind) {
    case K_F:
      add(kStF, 6);
      if (!dex) out[0].row = N_NONE;
      if (!sex) out[2].row = N_NONE;
      break;
    case K_Q: add(kStQ, 4); break;
    case K_FLIM: add(kStFLim, 3); break;
    case K_QLIM: add(kStQLim, 4); break;
    case K_DFDX: add(kStDF, 17); break;
    case K_DQDX: add(kStDQ, 14); break;
    case K_LEADF: if (g.lead[i]) add(kStLead, 4); break;
    case K_LEADQ:
      if (g.lead[i]) {
        add(kStLead, 4);
        if (dex) out[0].row = N_NONE;   // Q of a drain lead with RD is not stored
        if (sex) out[1].row = N_NONE;
      }
      break;
  }
}

// ---- device evaluation ---------------------------------------------------
template <bool FUSED>
__device__ __forceinline__ void eval(const KArgs& a, int64_t i) {
  if (i >= a.n) return;
  const DevCard mc = *reinterpret_cast<const DevCard*>(a.mcard + (int64_t)ifld(a, F_MODEL, i) * a.mstride);
  const xa_solver_state& ss = a.ss;
  const double gmin = a.opt.gmin;
  const double dtype = mc.dtype;
  const int32_t lD = node(a, N_D, i), lG = node(a, N_G, i), lS = node(a, N_S, i);
  const int32_t lB = node(a, N_B, i), lDP = node(a, N_DP, i), lSP = node(a, N_SP, i);
  const int32_t tvbd = stolid(a, T_vbd, i), tvbs = stolid(a, T_vbs, i), tvgs = stolid(a, T_vgs, i);
  const int32_t tvds = stolid(a, T_vds, i), tvon = stolid(a, T_von, i);
  const bool OFF = ifld(a, F_OFF, i) != 0;
  const double vt = dfld(a, D_VT, i);
  const double gd = dfld(a, D_GD, i), gs = dfld(a, D_GS, i), Mp = dfld(a, D_M, i);

  // ---- Instance::updateIntermediateVars ----
  const double Vd = solv(a.v.nextSol, lD), Vg = solv(a.v.nextSol, lG);
  const double Vs = solv(a.v.nextSol, lS), Vb = solv(a.v.nextSol, lB);
  const double Vsp = solv(a.v.nextSol, lSP), Vdp = solv(a.v.nextSol, lDP);
  const double Vddp = Vd - Vdp, Vssp = Vs - Vsp, Vbsp = Vb - Vsp, Vgsp = Vg - Vsp;
  double Vgb = Vg - Vb;
  const double Vdpsp = Vdp - Vsp;
  double vbs = dtype * Vbsp, vgs = dtype * Vgsp, vds = dtype * Vdpsp;
  double vbd = vbs - vds, vgd = vgs - vds;
  int origFlag = 1;
  const double vgs_orig = vgs, vds_orig = vds, vbs_orig = vbs, vbd_orig = vbd, vgd_orig = vgd;
  if (ss.initJctFlag && !OFF && a.opt.voltageLimiterFlag) {
    if (ifld(a, F_ICGIVEN, i)) {
      vds = dtype * dfld(a, D_ICVDS, i);
      vgs = dtype * dfld(a, D_ICVGS, i);
      vbs = dtype * dfld(a, D_ICVBS, i);
      vbd = vbs - vds;
      vgd = vgs - vds;
      origFlag = 0;
    } else {
      bool set = true;
      if (ss.inputOPFlag) {
        // drain, gate, source tested for == 0; the internal nodes and the
        // bulk for != 0 (:2696-2698); a ground node reads its flag as 0
        const double* fl = a.v.flagSol;
        set = solv(fl, lD) == 0 || solv(fl, lG) == 0 || solv(fl, lS) == 0 || solv(fl, lSP) != 0 ||
              solv(fl, lDP) != 0 || solv(fl, lB) != 0;
      }
      if (set) {
        vbs = -1;
        vgs = dtype * dfld(a, D_TVTO, i);
        vds = 0;
        vbd = vbs - vds;
        vgd = vgs - vds;
      }
    }
  } else if ((ss.initFixFlag || ss.initJctFlag) && OFF) {
    vbs = vgs = vds = 0;
    vbd = vgd = 0;
  }
  double vbs_old, vbd_old, vgs_old, vds_old, vgd_old, Von;
  if (ss.newtonIter == 0) {
    if (!ss.dcopFlag || (ss.locaEnabledFlag && ss.dcopFlag)) {
      const double* s = a.v.currSto;
      vbs_old = s[tvbs]; vbd_old = s[tvbd]; vgs_old = s[tvgs]; vds_old = s[tvds];
      Von = dtype * s[tvon];
    } else {
      vbs_old = vbs; vbd_old = vbd; vgs_old = vgs; vds_old = vds;
      Von = 0.0;
    }
    vgd_old = vgs_old - vds_old;
  } else {
    const double* s = a.v.nextSto;
    vbs_old = s[tvbs]; vbd_old = s[tvbd]; vgs_old = s[tvgs]; vds_old = s[tvds];
    Von = dtype * s[tvon];
    vgd_old = vgs_old - vds_old;
  }
  if (a.opt.voltageLimiterFlag && !(ss.initFixFlag && OFF)) {
    int Check = 1;
    if (vds_old >= 0) {
      vgs = fetlim(vgs, vgs_old, Von);
      vds = vgs - vgd;
      vds = limvds(vds, vds_old);
      vgd = vgs - vds;
    } else {
      vgd = fetlim(vgd, vgd_old, Von);
      vds = vgs - vgd;
      vds = -limvds(-vds, -vds_old);
      vgs = vgd + vds;
    }
    if (vds >= 0.0) {
      vbs = pnjlim(vbs, vbs_old, vt, dfld(a, D_SVCRIT, i), &Check);
      vbd = vbs - vds;
    } else {
      vbd = pnjlim(vbd, vbd_old, vt, dfld(a, D_DVCRIT, i), &Check);
      vbs = vbd + vds;
    }
  }
  vbd = vbs - vds;
  vgd = vgs - vds;
  Vgb = vgs - vbs;
  if (vgs_orig != vgs || vds_orig != vds || vbs_orig != vbs || vbd_orig != vbd || vgd_orig != vgd)
    origFlag = 0;
  double gbs, cbs, gbd, cbd;
  {
    const double SourceSatCur = dfld(a, D_SRCSAT, i), DrainSatCur = dfld(a, D_DRNSAT, i);
    if (vbs <= 0) {
      gbs = SourceSatCur / vt;
      gbs += gmin;
      cbs = gbs * vbs;
    } else {
      const double evbs = exp(dmin(kMaxExpArg, vbs / vt));
      gbs = (SourceSatCur * evbs / vt + gmin);
      cbs = (SourceSatCur * (evbs - 1) + gmin * vbs);
    }
    if (vbd <= 0) {
      gbd = DrainSatCur / vt;
      gbd += gmin;
      cbd = gbd * vbd;
    } else {
      const double evbd = exp(dmin(kMaxExpArg, vbd / vt));
      gbd = (DrainSatCur * evbd / vt + gmin);
      cbd = (DrainSatCur * (evbd - 1) + gmin * vbd);
    }
  }
  const int mode = (vds >= 0) ? 1 : -1;
  double cdrain, gm, gds, gmbs, Vdsat;
  const double tPhi = dfld(a, D_TPHI, i);
  // the continuation below changes vds/vgs/vgd for the drain current and the
  // Meyer capacitances; the stored values are restored after them (:3326-3328)
  const double vds_s = vds, vgs_s = vgs, vgd_s = vgd;
  {
    if (ss.artParameterFlag) {
      // contVds / contVgst with vdsScaleMin = 0.3, vgstConst = 4.5
      const double alpha = ss.gainScale;
      vds = vds * ((1.0 - 0.3) * ss.nltermScale + 0.3);
      if (mode == 1) vgs = vgs * alpha + (1.0 - alpha) * 4.5;
      else vgd = vgd * alpha + (1.0 - alpha) * 4.5;
    }
    const double vbx = (mode == 1) ? vbs : vbd;
    double sarg;
    if (vbx <= 0) {
      sarg = sqrt(tPhi - vbx);
    } else {
      sarg = sqrt(tPhi);
      sarg = sarg - vbx / (sarg + sarg);
      sarg = dmax(0.0, sarg);
    }
    Von = (dfld(a, D_TVBI, i) * dtype) + mc.gamma * sarg;
    const double vgst = ((mode == 1) ? vgs : vgd) - Von;
    Vdsat = dmax(vgst, 0.0);
    const double arg = (sarg <= 0) ? 0 : mc.gamma / (sarg + sarg);
    if (vgst <= 0) {
      cdrain = 0; gm = 0; gds = 0; gmbs = 0;
    } else {
      const double Beta = dfld(a, D_BETA, i);
      const double betap = Beta * (1 + mc.lambda * (vds * mode));
      if (vgst <= (vds * mode)) {
        cdrain = betap * vgst * vgst * .5;
        gm = betap * vgst;
        gds = mc.lambda * Beta * vgst * vgst * .5;
        gmbs = gm * arg;
      } else {
        cdrain = betap * (vds * mode) * (vgst - .5 * (vds * mode));
        gm = betap * (vds * mode);
        gds = betap * (vgst - (vds * mode)) +
              mc.lambda * Beta * (vds * mode) * (vgst - .5 * (vds * mode));
        gmbs = gm * arg;
      }
    }
  }
  const double von = dtype * Von;
  // bulk junction charges (:3015-3132)
  double qbs, capbs, qbd, capbd;
  {
    const double tDepCap = dfld(a, D_TDEPCAP, i), tBulkPot = dfld(a, D_TBULKPOT, i);
    const double Cbs = dfld(a, D_CBS, i), Cbssw = dfld(a, D_CBSSW, i);
    const double Cbd = dfld(a, D_CBD, i), Cbdsw = dfld(a, D_CBDSW, i);
    double sarg, sargsw;
    if (Cbs != 0 || Cbssw != 0) {
      if (vbs < tDepCap) {
        const double arg = 1 - vbs / tBulkPot;
        if (mc.MJ == mc.MJSW) {
          if (mc.MJ == .5) sarg = sargsw = 1 / sqrt(arg);
          else sarg = sargsw = exp(-mc.MJ * log(arg));
        } else {
          sarg = (mc.MJ == .5) ? 1 / sqrt(arg) : exp(-mc.MJ * log(arg));
          sargsw = (mc.MJSW == .5) ? 1 / sqrt(arg) : exp(-mc.MJSW * log(arg));
        }
        qbs = tBulkPot * (Cbs * (1 - arg * sarg) / (1 - mc.MJ) + Cbssw * (1 - arg * sargsw) / (1 - mc.MJSW));
        capbs = Cbs * sarg + Cbssw * sargsw;
      } else {
        qbs = dfld(a, D_F4S, i) + vbs * (dfld(a, D_F2S, i) + vbs * (dfld(a, D_F3S, i) / 2));
        capbs = dfld(a, D_F2S, i) + dfld(a, D_F3S, i) * vbs;
      }
    } else {
      qbs = 0;
      capbs = 0;
    }
    if (Cbd != 0 || Cbdsw != 0) {
      if (vbd < tDepCap) {
        const double arg = 1 - vbd / tBulkPot;
        if (mc.MJ == .5 && mc.MJSW == .5) {
          sarg = sargsw = 1 / sqrt(arg);
        } else {
          sarg = (mc.MJ == .5) ? 1 / sqrt(arg) : exp(-mc.MJ * log(arg));
          sargsw = (mc.MJSW == .5) ? 1 / sqrt(arg) : exp(-mc.MJSW * log(arg));
        }
        qbd = tBulkPot * (Cbd * (1 - arg * sarg) / (1 - mc.MJ) + Cbdsw * (1 - arg * sargsw) / (1 - mc.MJSW));
        capbd = Cbd * sarg + Cbdsw * sargsw;
      } else {
        qbd = dfld(a, D_F4D, i) + vbd * (dfld(a, D_F2D, i) + vbd * dfld(a, D_F3D, i) / 2);
        capbd = dfld(a, D_F2D, i) + vbd * dfld(a, D_F3D, i);
      }
    } else {
      qbd = 0;
      capbd = 0;
    }
  }
  // Meyer gate capacitances (:3153-3268)
  double capgs, capgd, capgb;
  const double OxideCap = dfld(a, D_OXCAP, i);
  {
    // one call with the operands swapped by mode keeps the outputs in registers
    double c_src, c_drn;
    qmeyer(mode > 0 ? vgs : vgd, mode > 0 ? vgd : vgs, Von, Vdsat, c_src, c_drn, capgb, tPhi,
           OxideCap);
    capgs = mode > 0 ? c_src : c_drn;
    capgd = mode > 0 ? c_drn : c_src;
  }
  const int32_t acgs = stalid(a, A_capgs, i), acgd = stalid(a, A_capgd, i), acgb = stalid(a, A_capgb, i);
  double Capgs, Capgd, Capgb;
  if (ss.dcopFlag) {
    Capgs = 2.0 * capgs + dfld(a, D_CGSOV, i);
    Capgd = 2.0 * capgd + dfld(a, D_CGDOV, i);
    Capgb = 2.0 * capgb + dfld(a, D_CGBOV, i);
  } else {
    Capgs = (capgs + a.v.currSta[acgs] + dfld(a, D_CGSOV, i));
    Capgd = (capgd + a.v.currSta[acgd] + dfld(a, D_CGDOV, i));
    Capgb = (capgb + a.v.currSta[acgb] + dfld(a, D_CGBOV, i));
  }
  Capgs *= (Capgs < 0.0) ? -1.0 : 1.0;
  Capgd *= (Capgd < 0.0) ? -1.0 : 1.0;
  Capgb *= (Capgb < 0.0) ? -1.0 : 1.0;
  vds = vds_s;
  vgs = vgs_s;
  vgd = vgd_s;
  const double Idrain = gd * Vddp, Isource = gs * Vssp;
  double Gm, Gmbs, nrmsum, revsum, cdreq;
  if (mode >= 0) {
    Gm = gm; Gmbs = gmbs; nrmsum = Gm + Gmbs; revsum = 0; cdreq = dtype * cdrain;
  } else {
    Gm = -gm; Gmbs = -gmbs; nrmsum = 0; revsum = -(Gm + Gmbs); cdreq = -(dtype)*cdrain;
  }

  // ---- Master::updateState ----
  {
    double* sto = a.v.nextSto;
    sto[tvbd] = vbd; sto[tvbs] = vbs; sto[tvgs] = vgs; sto[tvds] = vds; sto[tvon] = von;
  }
  double qgs, qgd, qgb;
  const int32_t aqgs = stalid(a, A_qgs, i), aqgd = stalid(a, A_qgd, i), aqgb = stalid(a, A_qgb, i);
  if (ss.dcopFlag) {
    qgs = Capgs * vgs;
    qgd = Capgd * vgd;
    qgb = Capgb * Vgb;
  } else {
    const double* cs = a.v.currSta;
    const double* co = a.v.currSto;
    qgs = cs[aqgs];
    qgd = cs[aqgd];
    qgb = cs[aqgb];
    const double vgs1 = co[tvgs], vbs1 = co[tvbs], vds1 = co[tvds];
    const double vgb1 = vgs1 - vbs1, vgd1 = vgs1 - vds1;
    qgs += Capgs * (vgs - vgs1);
    qgd += Capgd * (vgd - vgd1);
    qgb += Capgb * ((vgs - vbs) - vgb1);
  }
  {
    double* sta = a.v.nextSta;
    sta[acgs] = capgs; sta[acgd] = capgd; sta[acgb] = capgb;
    sta[aqgs] = qgs; sta[aqgd] = qgd; sta[aqgb] = qgb;
    sta[stalid(a, A_qbd, i)] = qbd;
    sta[stalid(a, A_qbs, i)] = qbs;
  }

  // ---- Master::loadDAEVectors ----
  Sink<FUSED> s(a, i);
  const double ceqbs = dtype * (cbs), ceqbd = dtype * (cbd);
  const double ceqgb = 0.0, ceqgs = 0.0, ceqgd = 0.0;
  if (gd != 0.0) s.put(K_F, 0, Idrain * Mp);
  s.put(K_F, 1, (ceqgs + ceqgd + ceqgb) * Mp);
  if (gs != 0.0) s.put(K_F, 2, Isource * Mp);
  s.put(K_F, 3, (ceqbs + ceqbd - ceqgb) * Mp);
  s.put(K_F, 4, (-Idrain - (ceqbd - cdreq + ceqgd)) * Mp);
  s.put(K_F, 5, (-Isource - (ceqbs + cdreq + ceqgs)) * Mp);
  const double Qeqbs = dtype * (qbs), Qeqbd = dtype * (qbd), Qeqgb = dtype * (qgb);
  const double Qeqgs = dtype * (qgs), Qeqgd = dtype * (qgd);
  s.put(K_Q, 0, (Qeqgs + Qeqgd + Qeqgb) * Mp);
  s.put(K_Q, 1, (Qeqbs + Qeqbd - Qeqgb) * Mp);
  s.put(K_Q, 2, (-(Qeqbd + Qeqgd)) * Mp);
  s.put(K_Q, 3, (-(Qeqbs + Qeqgs)) * Mp);
  const bool charge = ss.tranopFlag || ss.acopFlag || ss.transientFlag;
  const double gcgd = charge ? Capgd : 0.0, gcgs = charge ? Capgs : 0.0, gcgb = charge ? Capgb : 0.0;
  const double gcbs = charge ? capbs : 0.0, gcbd = charge ? capbd : 0.0;
  // Limiter corrections are loaded only when a voltage moved; an instance
  // that did not adds -0.0, which leaves every sum unchanged.
  if (!origFlag) {
    const double gmin1 = gmin;
    const double dvbd = vbd - vbd_orig, dvbs = vbs - vbs_orig, dvds = vds - vds_orig;
    const double dvgx = (mode > 0) ? (vgs - vgs_orig) : (vgd - vgd_orig);
    const double dvbx = (mode > 0) ? (vbs - vbs_orig) : (vbd - vbd_orig);
    const double J4 = dtype * (+((gbd - gmin1)) * dvbd + ((gbs - gmin1)) * dvbs);
    const double J5 = dtype * (-((gbd - gmin1)) * dvbd + gds * dvds + Gm * dvgx + Gmbs * dvbx);
    const double J6 = dtype * (-((gbs - gmin1)) * dvbs - gds * dvds - Gm * dvgx - Gmbs * dvbx);
    s.put(K_FLIM, 0, J4 * Mp);
    s.put(K_FLIM, 1, J5 * Mp);
    s.put(K_FLIM, 2, J6 * Mp);
    const double Q2 = dtype * (gcgd * (vgd - vgd_orig) + gcgs * (vgs - vgs_orig) +
                               gcgb * (vgs - vgs_orig - vbs + vbs_orig));
    const double Q4 = dtype * (-(gcgb) * (vgs - vgs_orig - vbs + vbs_orig) + (gcgb) * (vbd - vbd_orig) +
                               (gcbs) * (vbs - vbs_orig));
    const double Q5 = dtype * (-(gcgd) * (vgd - vgd_orig) - (gcbd) * (vbd - vbd_orig));
    const double Q6 = dtype * (-gcgs * (vgs - vgs_orig) - (gcbs) * (vbs - vbs_orig));
    s.put(K_QLIM, 0, Q2 * Mp);
    s.put(K_QLIM, 1, Q4 * Mp);
    s.put(K_QLIM, 2, Q5 * Mp);
    s.put(K_QLIM, 3, Q6 * Mp);
  } else {
    for (int e = 0; e < 3; ++e) s.put(K_FLIM, e, -0.0);
    for (int e = 0; e < 4; ++e) s.put(K_QLIM, e, -0.0);
  }
  // lead currents (stored, not added)
  s.put(K_LEADF, 0, gd != 0.0 ? Idrain * Mp : (-Idrain - (ceqbd - cdreq + ceqgd)) * Mp);
  s.put(K_LEADF, 1, gs != 0.0 ? Isource * Mp : (-Isource - (ceqbs + cdreq + ceqgs)) * Mp);
  s.put(K_LEADF, 2, (ceqgs + ceqgd + ceqgb) * Mp);
  s.put(K_LEADF, 3, (ceqbs + ceqbd - ceqgb) * Mp);
  s.put(K_LEADQ, 0, (-(Qeqbd + Qeqgd)) * Mp);
  s.put(K_LEADQ, 1, (-(Qeqbs + Qeqgs)) * Mp);
  s.put(K_LEADQ, 2, (Qeqgs + Qeqgd + Qeqgb) * Mp);
  s.put(K_LEADQ, 3, (Qeqbs + Qeqbd - Qeqgb) * Mp);

  // ---- Master::loadDAEMatrices ----
  s.put(K_DFDX, 0, gd * Mp);
  s.put(K_DFDX, 1, -(gd * Mp));
  s.put(K_DFDX, 2, gs * Mp);
  s.put(K_DFDX, 3, -(gs * Mp));
  s.put(K_DFDX, 4, (gbs + gbd) * Mp);
  s.put(K_DFDX, 5, -(gbd * Mp));
  s.put(K_DFDX, 6, -(gbs * Mp));
  s.put(K_DFDX, 7, -(gd * Mp));
  s.put(K_DFDX, 8, (Gm)*Mp);
  s.put(K_DFDX, 9, (-gbd + Gmbs) * Mp);
  s.put(K_DFDX, 10, (gd + gds + gbd + revsum) * Mp);
  s.put(K_DFDX, 11, (-gds - nrmsum) * Mp);
  s.put(K_DFDX, 12, -((Gm)*Mp));
  s.put(K_DFDX, 13, -(gs * Mp));
  s.put(K_DFDX, 14, -((gbs + Gmbs) * Mp));
  s.put(K_DFDX, 15, -((gds + revsum) * Mp));
  s.put(K_DFDX, 16, (gs + gds + gbs + nrmsum) * Mp);
  s.put(K_DQDX, 0, (gcgd + gcgs + gcgb) * Mp);
  s.put(K_DQDX, 1, -(gcgb * Mp));
  s.put(K_DQDX, 2, -(gcgd * Mp));
  s.put(K_DQDX, 3, -(gcgs * Mp));
  s.put(K_DQDX, 4, -(gcgb * Mp));
  s.put(K_DQDX, 5, (+gcbs + gcbd + gcgb) * Mp);
  s.put(K_DQDX, 6, -(+gcbd * Mp));
  s.put(K_DQDX, 7, -(+gcbs * Mp));
  s.put(K_DQDX, 8, -gcgd * Mp);
  s.put(K_DQDX, 9, -gcbd * Mp);
  s.put(K_DQDX, 10, (+gcbd + gcgd) * Mp);
  s.put(K_DQDX, 11, -(gcgs * Mp));
  s.put(K_DQDX, 12, -(+gcbs * Mp));
  s.put(K_DQDX, 13, (+gcbs + gcgs) * Mp);
}

template <bool FUSED>
__global__ void __launch_bounds__(256) k_update(KArgs a) {
  eval<FUSED>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (FUSED) fused_writeout(a);
}

static int launch_update(Group& g, const xa_solver_state& ss, const xa_vectors& v, bool fused) {
  if (!v.nextSol || !v.nextSto || !v.nextSta)
    return fail(XA_ERR_ARG, "mosfet1 update: nextSol/nextSto/nextSta required");
  if ((!ss.dcopFlag || (ss.newtonIter == 0 && ss.locaEnabledFlag)) && (!v.currSto || !v.currSta))
    return fail(XA_ERR_ARG, "mosfet1 update: currSto/currSta required");
  if (ss.newtonIter == 0 && !ss.dcopFlag && !v.currSto)
    return fail(XA_ERR_ARG, "mosfet1 update: currSto required");
  if (ss.inputOPFlag && ss.initJctFlag && !v.flagSol)
    return fail(XA_ERR_ARG, "mosfet1 update: flagSol required with inputOPFlag");
  KArgs a;
  fill_kargs(g, ss, v, fused, a);
  return launch_model_kernel(g, a, fused, (const void*)k_update<false>, (const void*)k_update<true>);
}

// Algorithmic bytes per instance (no internal nodes, transient): 31 SoA
// columns, 4 node values + 6 LIDs, 5 store LIDs + 4 old stores read + 5
// written, 8 state LIDs + 6 old states read + 8 written, model index.
static double traffic(const Group& g, int phase) {
  (void)g;
  if (phase == 0) return 8.0 * (ND + 4 + 4 + 5 + 6 + 8) + 4.0 * (6 + 5 + 8 + 1);
  if (phase == 1) return 16.0 * (4 /*F*/ + 4 /*Q*/);
  return 16.0 * (17 + 14);
}

}  // namespace mosfet1

const ModelImpl* model_mosfet1() {
  static ModelImpl m = [] {
    ModelImpl x{};
    x.type = XA_MOSFET1;
    x.name = "mosfet1";
    x.mparams = mosfet1::kModelParams;
    x.n_mparams = sizeof(mosfet1::kModelParams) / sizeof(ParamDef);
    x.n_mslots = mosfet1::NM;
    x.iparams = mosfet1::kInstParams;
    x.n_iparams = sizeof(mosfet1::kInstParams) / sizeof(ParamDef);
    x.n_islots = mosfet1::NI;
    x.max_nodes = mosfet1::NNODE;
    x.max_sto = mosfet1::NSTO;
    x.max_sta = mosfet1::NSTA;
    const int E[NK] = {6, 4, 3, 4, 17, 14, 4, 4};
    for (int k = 0; k < NK; ++k) x.max_entries[k] = E[k];
    x.setup = mosfet1::setup;
    x.resolve_nodes = mosfet1::resolve_nodes;
    x.stamps = mosfet1::stamps;
    x.launch_update = mosfet1::launch_update;
    x.launch_secondary = nullptr;
    x.skip_kind = nullptr;
    x.state_names = nullptr;
    x.n_state = 0;
    x.traffic = mosfet1::traffic;
    x.fused_block = 64;
    x.fused_at_scale = false;
    x.fused_small = true;
    return x;
  }();
  return &m;
}

}  // namespace xa
