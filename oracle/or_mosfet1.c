/* or_mosfet1.c -- oracle restatement of the Level-1 (Shichman-Hodges)
 * MOSFET, test infrastructure only.
 * Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_MOSFET1.C
 *   parameters :71-262, Instance ctor :309-1256 (node counts :1246-1254,
 *   drain/source conductances :1182-1221, overlap caps :1239-1243),
 *   registerLIDs :1278-1372, registerStateLIDs :1452-1503,
 *   registerStoreLIDs :1513-1535, jacStamp (non-newMeyer) :1119-1165,
 *   updateIntermediateVars :2584-3393, updateTemperature :3403-3646,
 *   Model::processParams :3752-3824, Model ctor :3856-3946,
 *   Master::updateState :4069-4158, Master::loadDAEVectors :4168-4396,
 *   Master::loadDAEMatrices :4407-4688.
 * newMeyerFlag = 1 (the extra charge-derivative nodes) is not restated and is
 * rejected at setup. */
#include <stdio.h>
#include "or_internal.h"
#include "or_eval_libm.h"

/* ---- parameter tables --------------------------------------------------- */
enum {
  M_L, M_W, M_VTO, M_KP, M_GAMMA, M_PHI, M_LAMBDA, M_RD, M_RS, M_CBD, M_CBS, M_IS, M_PB,
  M_CGSO, M_CGDO, M_CGBO, M_RSH, M_CJ, M_MJ, M_CJSW, M_MJSW, M_JS, M_TOX, M_LD, M_UO, M_U0,
  M_FC, M_NSUB, M_NSS, M_TNOM, M_KF, M_AF, M_TPG, M_TYPE, NM1_
};
static const or_param mpar[] = {
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
enum { I_TEMP, I_L, I_W, I_AD, I_AS, I_NRD, I_NRS, I_PD, I_PS, I_M, I_IC1, I_IC2, I_IC3, I_OFF, NI1_ };
static const or_param ipar[] = {
  {"TEMP", I_TEMP, 0.0, 1}, {"L", I_L, 0.0, 0}, {"W", I_W, 0.0, 0}, {"AD", I_AD, 0.0, 0},
  {"AS", I_AS, 0.0, 0}, {"NRD", I_NRD, 1.0, 0}, {"NRS", I_NRS, 1.0, 0}, {"PD", I_PD, 0.0, 0},
  {"PS", I_PS, 0.0, 0}, {"M", I_M, 1.0, 0}, {"IC1", I_IC1, 0.0, 0}, {"IC2", I_IC2, 0.0, 0},
  {"IC3", I_IC3, 0.0, 0}, {"OFF", I_OFF, 0.0, 0},
};

typedef struct {
  int dtype;
  double tnom, model_l, model_w, latDiff, jctSatCurDensity, jctSatCur, drainResistance;
  double sourceResistance, sheetResistance, transconductance, gateSourceOverlapCapFactor;
  double gateDrainOverlapCapFactor, gateBulkOverlapCapFactor, oxideCapFactor, vt0, capBD, capBS;
  double bulkCapFactor, sideWallCapFactor, bulkJctPotential, bulkJctBotGradingCoeff;
  double bulkJctSideGradingCoeff, fwdCapDepCoeff, phi, gamma, lambda, substrateDoping;
  double gateType, surfaceStateDensity, oxideThickness, surfaceMobility, surfaceMobility0;
  int capBDGiven, capBSGiven, bulkCapFactorGiven, sideWallCapFactorGiven, sheetResistanceGiven;
  double fact1, vtnom, egfet1, pbfact1;
} m1model;

typedef struct {
  m1model* model;
  double l, w, drainArea, sourceArea, drainSquares, sourceSquares, drainPerimeter;
  double sourcePerimeter, numberParallel, icVDS, icVGS, icVBS, temp;
  int IC_GIVEN, OFF, loadLeadCurrent;
  double drainConductance, sourceConductance, EffectiveLength;
  double GateSourceOverlapCap, GateDrainOverlapCap, GateBulkOverlapCap, OxideCap;
  double vt, tTransconductance, tSurfMob, tPhi, tVbi, tVto, tSatCur, tSatCurDens;
  double tCbd, tCbs, tCj, tCjsw, tBulkPot, tDepCap, drainVcrit, sourceVcrit;
  double Cbd, Cbdsw, Cbs, Cbssw, f2d, f3d, f4d, f2s, f3s, f4s;
  /* evaluation results */
  double Vd, Vg, Vs, Vb, Vsp, Vdp, Vddp, Vssp, Vbsp, Vbdp, Vgsp, Vgdp, Vgb, Vdpsp;
  double vbs, vgs, vds, vbd, vgd, vbs_orig, vgs_orig, vds_orig, vbd_orig, vgd_orig;
  double vbs_old, vbd_old, vgs_old, vds_old, vgd_old;
  int origFlag, limitedFlag, mode;
  double DrainSatCur, SourceSatCur, gbs, cbs, gbd, cbd, cdrain, gm, gds, gmbs, von, vdsat, cd;
  double qbs, capbs, qbd, capbd, capgs, capgd, capgb, Capgs, Capgd, Capgb;
  double dcap[9];
  double Idrain, Isource, Gm, Gmbs, nrmsum, revsum, cdreq, qgs, qgd, qgb;
  /* LIDs and slots */
  int li_Drain, li_Gate, li_Source, li_Bulk, li_DrainPrime, li_SourcePrime;
  int li_store_vbd, li_store_vbs, li_store_vgs, li_store_vds, li_store_von;
  int li_store_dev_id, li_store_dev_ig, li_store_dev_is, li_store_dev_ib;
  int li_state_qgs, li_state_qgd, li_state_qgb, li_state_capgs, li_state_capgd, li_state_capgb;
  int li_state_qbd, li_state_qbs;
  int64_t fDD, fDDp, fSS, fSSp, fBB, fBDp, fBSp, fDpD, fDpG, fDpB, fDpDp, fDpSp;
  int64_t fSpG, fSpS, fSpB, fSpDp, fSpSp;
  int64_t qGG, qGB, qGDp, qGSp, qBG, qBB, qBDp, qBSp, qDpG, qDpB, qDpDp, qSpG, qSpB, qSpSp;
} m1inst;

/* Model::processParams (:3752-3824) */
static int model_process(m1model* m, const or_card* c) {
  double wkfngs, wkfng, fermig, fermis, vfb, kt1, arg1;
  m->fact1 = m->tnom / OR_REFTEMP;
  m->vtnom = m->tnom * OR_KoverQ;
  kt1 = OR_BOLTZ * m->tnom;
  m->egfet1 = 1.16 - (7.02e-4 * m->tnom * m->tnom) / (m->tnom + 1108);
  arg1 = -m->egfet1 / (kt1 + kt1) + 1.1150877 / (OR_BOLTZ * (OR_REFTEMP + OR_REFTEMP));
  m->pbfact1 = -2 * m->vtnom * (1.5 * log(m->fact1) + OR_Q * arg1);
  if (!c->given[M_TOX] || m->oxideThickness == 0) {
    m->oxideCapFactor = 0;
  } else {
    m->oxideCapFactor = 3.9 * 8.854214871e-12 / m->oxideThickness;
    if (!c->given[M_KP]) {
      if (!c->given[M_UO] && !c->given[M_U0]) m->surfaceMobility = 600;
      m->transconductance = m->surfaceMobility * m->oxideCapFactor * 1e-4;
    }
    if (c->given[M_NSUB]) {
      if (m->substrateDoping * 1e6 > 1.45e16) {
        if (!c->given[M_PHI]) {
          m->phi = 2 * m->vtnom * log(m->substrateDoping * 1e6 / 1.45e16);
          m->phi = or_max(0.1, m->phi);
        }
        fermis = m->dtype * .5 * m->phi;
        wkfng = 3.2;
        if (!c->given[M_TPG]) m->gateType = 1;
        if (m->gateType != 0) {
          fermig = m->dtype * m->gateType * .5 * m->egfet1;
          wkfng = 3.25 + .5 * m->egfet1 - fermig;
        }
        wkfngs = wkfng - (3.25 + .5 * m->egfet1 + fermis);
        if (!c->given[M_GAMMA])
          m->gamma = sqrt(2 * 11.70 * 8.854214871e-12 * OR_Q * m->substrateDoping * 1e6) /
                     m->oxideCapFactor;
        if (!c->given[M_VTO]) {
          if (!c->given[M_NSS]) m->surfaceStateDensity = 0;
          vfb = wkfngs - m->surfaceStateDensity * 1e4 * OR_Q / m->oxideCapFactor;
          m->vt0 = vfb + m->dtype * (m->gamma * sqrt(m->phi) + m->phi);
        }
      } else {
        return or_fail(XA_ERR_ARG, "mosfet1: NSUB < Ni");
      }
    }
  }
  return XA_OK;
}

/* Instance::updateTemperature (:3403-3646) */
static void update_temperature(m1inst* I, double temp_tmp) {
  const m1model* m = I->model;
  double czbd, czbdsw, czbs, czbssw, arg, sarg, sargsw, ratio, ratio4, fact2, kt, egfet;
  double pbfact, capfact, phio, pbo, gmanew, gmaold, tnom;
  if (temp_tmp != -999.0) I->temp = temp_tmp;
  tnom = m->tnom;
  ratio = I->temp / tnom;
  I->vt = I->temp * OR_KoverQ;
  ratio = I->temp / tnom;
  fact2 = I->temp / OR_REFTEMP;
  kt = I->temp * OR_BOLTZ;
  egfet = 1.16 - (7.02e-4 * I->temp * I->temp) / (I->temp + 1108);
  arg = -egfet / (kt + kt) + 1.1150877 / (OR_BOLTZ * (OR_REFTEMP + OR_REFTEMP));
  pbfact = -2 * I->vt * (1.5 * log(fact2) + OR_Q * arg);
  ratio4 = ratio * sqrt(ratio);
  I->tTransconductance = m->transconductance / ratio4;
  I->tSurfMob = m->surfaceMobility / ratio4;
  phio = (m->phi - m->pbfact1) / m->fact1;
  I->tPhi = fact2 * phio + pbfact;
  I->tVbi = m->vt0 - m->dtype * (m->gamma * sqrt(m->phi)) + .5 * (m->egfet1 - egfet) +
            m->dtype * .5 * (I->tPhi - m->phi);
  I->tVto = I->tVbi + m->dtype * m->gamma * sqrt(I->tPhi);
  I->tSatCur = m->jctSatCur * exp(-egfet / I->vt + m->egfet1 / m->vtnom);
  I->tSatCurDens = m->jctSatCurDensity * exp(-egfet / I->vt + m->egfet1 / m->vtnom);
  pbo = (m->bulkJctPotential - m->pbfact1) / m->fact1;
  gmaold = (m->bulkJctPotential - pbo) / pbo;
  capfact = 1 / (1 + m->bulkJctBotGradingCoeff * (4e-4 * (m->tnom - OR_REFTEMP) - gmaold));
  I->tCbd = m->capBD * capfact;
  I->tCbs = m->capBS * capfact;
  I->tCj = m->bulkCapFactor * capfact;
  capfact = 1 / (1 + m->bulkJctSideGradingCoeff * (4e-4 * (m->tnom - OR_REFTEMP) - gmaold));
  I->tCjsw = m->sideWallCapFactor * capfact;
  I->tBulkPot = fact2 * pbo + pbfact;
  gmanew = (I->tBulkPot - pbo) / pbo;
  capfact = (1 + m->bulkJctBotGradingCoeff * (4e-4 * (I->temp - OR_REFTEMP) - gmanew));
  I->tCbd *= capfact;
  I->tCbs *= capfact;
  I->tCj *= capfact;
  capfact = (1 + m->bulkJctSideGradingCoeff * (4e-4 * (I->temp - OR_REFTEMP) - gmanew));
  I->tCjsw *= capfact;
  I->tDepCap = m->fwdCapDepCoeff * I->tBulkPot;
  if ((m->jctSatCurDensity == 0) || (I->drainArea == 0) || (I->sourceArea == 0)) {
    I->sourceVcrit = I->drainVcrit = I->vt * log(I->vt / (sqrt(2.0) * m->jctSatCur));
  } else {
    I->drainVcrit = I->vt * log(I->vt / (sqrt(2.0) * m->jctSatCurDensity * I->drainArea));
    I->sourceVcrit = I->vt * log(I->vt / (sqrt(2.0) * m->jctSatCurDensity * I->sourceArea));
  }
  if (m->capBDGiven) czbd = I->tCbd;
  else if (m->bulkCapFactorGiven) czbd = I->tCj * I->drainArea;
  else czbd = 0;
  if (m->sideWallCapFactorGiven) czbdsw = I->tCjsw * I->drainPerimeter;
  else czbdsw = 0;
  arg = 1 - m->fwdCapDepCoeff;
  sarg = exp((-m->bulkJctBotGradingCoeff) * log(arg));
  sargsw = exp((-m->bulkJctSideGradingCoeff) * log(arg));
  I->Cbd = czbd;
  I->Cbdsw = czbdsw;
  I->f2d = czbd * (1 - m->fwdCapDepCoeff * (1 + m->bulkJctBotGradingCoeff)) * sarg / arg +
           czbdsw * (1 - m->fwdCapDepCoeff * (1 + m->bulkJctSideGradingCoeff)) * sargsw / arg;
  I->f3d = czbd * m->bulkJctBotGradingCoeff * sarg / arg / I->tBulkPot +
           czbdsw * m->bulkJctSideGradingCoeff * sargsw / arg / I->tBulkPot;
  I->f4d = czbd * I->tBulkPot * (1 - arg * sarg) / (1 - m->bulkJctBotGradingCoeff) +
           czbdsw * I->tBulkPot * (1 - arg * sargsw) / (1 - m->bulkJctSideGradingCoeff) -
           I->f3d / 2 * (I->tDepCap * I->tDepCap) - I->tDepCap * I->f2d;
  if (m->capBSGiven) czbs = I->tCbs;
  else if (m->bulkCapFactorGiven) czbs = I->tCj * I->sourceArea;
  else czbs = 0;
  if (m->sideWallCapFactorGiven) czbssw = I->tCjsw * I->sourcePerimeter;
  else czbssw = 0;
  arg = 1 - m->fwdCapDepCoeff;
  sarg = exp((-m->bulkJctBotGradingCoeff) * log(arg));
  sargsw = exp((-m->bulkJctSideGradingCoeff) * log(arg));
  I->Cbs = czbs;
  I->Cbssw = czbssw;
  I->f2s = czbs * (1 - m->fwdCapDepCoeff * (1 + m->bulkJctBotGradingCoeff)) * sarg / arg +
           czbssw * (1 - m->fwdCapDepCoeff * (1 + m->bulkJctSideGradingCoeff)) * sargsw / arg;
  I->f3s = czbs * m->bulkJctBotGradingCoeff * sarg / arg / I->tBulkPot +
           czbssw * m->bulkJctSideGradingCoeff * sargsw / arg / I->tBulkPot;
  I->f4s = czbs * I->tBulkPot * (1 - arg * sarg) / (1 - m->bulkJctBotGradingCoeff) +
           czbssw * I->tBulkPot * (1 - arg * sargsw) / (1 - m->bulkJctSideGradingCoeff) -
           I->f3s / 2 * (I->tDepCap * I->tDepCap) - I->tDepCap * I->f2s;
}

static int setup(struct or_group_s* g) {
  int k;
  int64_t i;
  m1model* mods = (m1model*)calloc(g->nmodels, sizeof(m1model));
  m1inst* in = (m1inst*)calloc(g->n > 0 ? g->n : 1, sizeof(m1inst));
  int rc;
  g->mdata = mods;
  g->inst = in;
  if (g->opt.newMeyerFlag) return or_fail(XA_ERR_UNSUPPORTED, "mosfet1: newMeyerFlag");
  for (k = 0; k < g->nmodels; ++k) {
    const or_card* c = &g->models[k];
    m1model* m = &mods[k];
    if (c->level != 1) return or_fail(XA_ERR_UNSUPPORTED, "mosfet1: level must be 1");
    /* Model ctor (:3856-3946): table defaults, then the given values */
    m->dtype = (c->v[M_TYPE] < 0) ? -1 : 1;
    m->model_l = c->v[M_L]; m->model_w = c->v[M_W]; m->vt0 = c->v[M_VTO];
    m->transconductance = c->v[M_KP]; m->gamma = c->v[M_GAMMA]; m->phi = c->v[M_PHI];
    m->lambda = c->v[M_LAMBDA]; m->drainResistance = c->v[M_RD];
    m->sourceResistance = c->v[M_RS]; m->capBD = c->v[M_CBD]; m->capBS = c->v[M_CBS];
    m->jctSatCur = c->v[M_IS]; m->bulkJctPotential = c->v[M_PB];
    m->gateSourceOverlapCapFactor = c->v[M_CGSO]; m->gateDrainOverlapCapFactor = c->v[M_CGDO];
    m->gateBulkOverlapCapFactor = c->v[M_CGBO]; m->sheetResistance = c->v[M_RSH];
    m->bulkCapFactor = c->v[M_CJ]; m->bulkJctBotGradingCoeff = c->v[M_MJ];
    m->sideWallCapFactor = c->v[M_CJSW]; m->bulkJctSideGradingCoeff = c->v[M_MJSW];
    m->jctSatCurDensity = c->v[M_JS]; m->oxideThickness = c->v[M_TOX];
    m->latDiff = c->v[M_LD]; m->surfaceMobility = c->v[M_UO]; m->surfaceMobility0 = c->v[M_U0];
    m->fwdCapDepCoeff = c->v[M_FC]; m->substrateDoping = c->v[M_NSUB];
    m->surfaceStateDensity = c->v[M_NSS]; m->tnom = c->v[M_TNOM]; m->gateType = c->v[M_TPG];
    m->capBDGiven = c->given[M_CBD]; m->capBSGiven = c->given[M_CBS];
    m->bulkCapFactorGiven = c->given[M_CJ]; m->sideWallCapFactorGiven = c->given[M_CJSW];
    m->sheetResistanceGiven = c->given[M_RSH];
    if (!c->given[M_L]) m->model_l = g->opt.defl;
    if (!c->given[M_W]) m->model_w = g->opt.defw;
    if (!c->given[M_TNOM]) m->tnom = g->opt.tnom;
    if (m->capBD != 0) m->capBDGiven = 1;
    if (m->capBS != 0) m->capBSGiven = 1;
    if (c->given[M_U0]) {
      if (c->given[M_UO]) return or_fail(XA_ERR_ARG, "mosfet1: both UO and U0 given");
      m->surfaceMobility = m->surfaceMobility0;
    }
    if ((rc = model_process(m, c)) != XA_OK) return rc;
  }
  for (i = 0; i < g->n; ++i) {
    m1inst* I = &in[i];
    m1model* m = &mods[g->inst_model[i]];
    I->model = m;
    I->temp = g->ival[I_TEMP][i]; I->l = g->ival[I_L][i]; I->w = g->ival[I_W][i];
    I->drainArea = g->ival[I_AD][i]; I->sourceArea = g->ival[I_AS][i];
    I->drainSquares = g->ival[I_NRD][i]; I->sourceSquares = g->ival[I_NRS][i];
    I->drainPerimeter = g->ival[I_PD][i]; I->sourcePerimeter = g->ival[I_PS][i];
    I->numberParallel = g->ival[I_M][i];
    I->icVDS = g->ival[I_IC1][i]; I->icVGS = g->ival[I_IC2][i]; I->icVBS = g->ival[I_IC3][i];
    I->IC_GIVEN = g->igiven[I_IC1] || g->igiven[I_IC2] || g->igiven[I_IC3];
    I->OFF = g->ival[I_OFF][i] != 0.0;
    I->loadLeadCurrent = g->lead[i];
    if (!g->igiven[I_TEMP]) I->temp = g->opt.temp;
    if (!g->igiven[I_L]) I->l = m->model_l;
    if (!g->igiven[I_W]) I->w = m->model_w;
    if (m->drainResistance != 0) I->drainConductance = 1 / m->drainResistance;
    else if (m->sheetResistanceGiven) {
      if (m->sheetResistance != 0) I->drainConductance = 1 / (m->sheetResistance * I->drainSquares);
      else I->drainConductance = 0;
    } else I->drainConductance = 0;
    if (m->sourceResistance != 0) I->sourceConductance = 1 / m->sourceResistance;
    else if (m->sheetResistanceGiven) {
      if (m->sheetResistance != 0)
        I->sourceConductance = 1 / (m->sheetResistance * I->sourceSquares);
      else I->sourceConductance = 0;
    } else I->sourceConductance = 0;
    if (!g->igiven[I_AD]) I->drainArea = g->opt.defad;
    if (!g->igiven[I_AS]) I->sourceArea = g->opt.defas;
    if (I->l - 2 * m->latDiff <= 0)
      return or_fail(XA_ERR_ARG, "mosfet1: effective channel length less than zero");
    I->EffectiveLength = I->l - 2 * m->latDiff;
    I->GateSourceOverlapCap = m->gateSourceOverlapCapFactor * I->w;
    I->GateDrainOverlapCap = m->gateDrainOverlapCapFactor * I->w;
    I->GateBulkOverlapCap = m->gateBulkOverlapCapFactor * I->EffectiveLength;
    I->OxideCap = m->oxideCapFactor * I->EffectiveLength * I->w;
    g->n_ext[i] = 4;
    g->n_int[i] = ((I->sourceConductance == 0.0) ? 0 : 1) + ((I->drainConductance == 0.0) ? 0 : 1);
    g->n_sta[i] = 8;
    g->n_sto[i] = 5 + (I->loadLeadCurrent ? 4 : 0);
    update_temperature(I, I->temp);   /* processParams -> updateTemperature(temp) */
  }
  return XA_OK;
}

static void register_lids(struct or_group_s* g, int64_t i, const int32_t* nodes,
                          const int32_t* sta, const int32_t* sto) {
  m1inst* I = &((m1inst*)g->inst)[i];
  int loc = 4;
  I->li_Drain = nodes[0];
  I->li_Gate = nodes[1];
  I->li_Source = nodes[2];
  I->li_Bulk = nodes[3];
  I->li_DrainPrime = I->drainConductance ? nodes[loc++] : I->li_Drain;
  I->li_SourcePrime = I->sourceConductance ? nodes[loc++] : I->li_Source;
  I->li_state_qgs = sta ? sta[0] : -1;
  I->li_state_qgd = sta ? sta[1] : -1;
  I->li_state_qgb = sta ? sta[2] : -1;
  I->li_state_capgs = sta ? sta[3] : -1;
  I->li_state_capgd = sta ? sta[4] : -1;
  I->li_state_capgb = sta ? sta[5] : -1;
  I->li_state_qbd = sta ? sta[6] : -1;
  I->li_state_qbs = sta ? sta[7] : -1;
  I->li_store_vbd = sto[0];
  I->li_store_vbs = sto[1];
  I->li_store_vgs = sto[2];
  I->li_store_vds = sto[3];
  I->li_store_von = sto[4];
  I->li_store_dev_id = I->loadLeadCurrent ? sto[5] : -1;
  I->li_store_dev_ig = I->loadLeadCurrent ? sto[6] : -1;
  I->li_store_dev_is = I->loadLeadCurrent ? sto[7] : -1;
  I->li_store_dev_ib = I->loadLeadCurrent ? sto[8] : -1;
}

/* jacStamp_DC_SC (:1124-1152) with drain'/source' folded into drain/source
 * when their resistance is zero (jacStampMap) */
static int64_t jac_pairs(struct or_group_s* g, int64_t i, int32_t* r, int32_t* c) {
  const m1inst* I = &((m1inst*)g->inst)[i];
  const int D = I->li_Drain, G = I->li_Gate, S = I->li_Source, B = I->li_Bulk;
  const int DP = I->li_DrainPrime, SP = I->li_SourcePrime;
  const int32_t pr[20] = {D, D, G, G, G, G, S, S, B, B, B, B, DP, DP, DP, DP, DP, SP, SP, SP};
  const int32_t pc[20] = {D, DP, G, B, DP, SP, S, SP, G, B, DP, SP, D, G, B, DP, SP, G, S, B};
  int64_t k, n = 0;
  for (k = 0; k < 20; ++k) {
    if (pr[k] < 0 || pc[k] < 0) continue;
    if (r) { r[n] = pr[k]; c[n] = pc[k]; }
    ++n;
  }
  /* the source' row also holds (SP, DP) and (SP, SP) */
  if (SP >= 0 && DP >= 0) { if (r) { r[n] = SP; c[n] = DP; } ++n; }
  if (SP >= 0) { if (r) { r[n] = SP; c[n] = SP; } ++n; }
  return n;
}

static int register_jac(struct or_group_s* g, int64_t i) {
  m1inst* I = &((m1inst*)g->inst)[i];
  const int D = I->li_Drain, G = I->li_Gate, S = I->li_Source, B = I->li_Bulk;
  const int DP = I->li_DrainPrime, SP = I->li_SourcePrime;
  I->fDD = or_slot(g, D, D); I->fDDp = or_slot(g, D, DP);
  I->fSS = or_slot(g, S, S); I->fSSp = or_slot(g, S, SP);
  I->fBB = or_slot(g, B, B); I->fBDp = or_slot(g, B, DP); I->fBSp = or_slot(g, B, SP);
  I->fDpD = or_slot(g, DP, D); I->fDpG = or_slot(g, DP, G); I->fDpB = or_slot(g, DP, B);
  I->fDpDp = or_slot(g, DP, DP); I->fDpSp = or_slot(g, DP, SP);
  I->fSpG = or_slot(g, SP, G); I->fSpS = or_slot(g, SP, S); I->fSpB = or_slot(g, SP, B);
  I->fSpDp = or_slot(g, SP, DP); I->fSpSp = or_slot(g, SP, SP);
  I->qGG = or_slot(g, G, G); I->qGB = or_slot(g, G, B); I->qGDp = or_slot(g, G, DP);
  I->qGSp = or_slot(g, G, SP); I->qBG = or_slot(g, B, G); I->qBB = or_slot(g, B, B);
  I->qBDp = or_slot(g, B, DP); I->qBSp = or_slot(g, B, SP); I->qDpG = or_slot(g, DP, G);
  I->qDpB = or_slot(g, DP, B); I->qDpDp = or_slot(g, DP, DP); I->qSpG = or_slot(g, SP, G);
  I->qSpB = or_slot(g, SP, B); I->qSpSp = or_slot(g, SP, SP);
  {
    const int64_t* s = &I->fDD;
    int k;
    for (k = 0; k < 31; ++k) if (s[k] < -1) return XA_ERR_ARG;
  }
  return XA_OK;
}

static double sol(const double* x, int lid) { return lid >= 0 ? x[lid] : 0.0; }

/* Instance::updateIntermediateVars (:2584-3393), newMeyerFlag = 0 */
static int update_intermediate(struct or_group_s* g, m1inst* I, const xa_solver_state* ss,
                               const xa_vectors* v) {
  const m1model* m = I->model;
  const double gmin = g->opt.gmin;
  double Von, Vdsat, Beta, evbs, evbd, sarg, sargsw, arg, capgs_old, capgd_old, capgb_old;
  double vgs_save, vgd_save, vds_save;
  int Check = 1;
  if ((I->tSatCurDens == 0) || (I->drainArea == 0) || (I->sourceArea == 0)) {
    I->DrainSatCur = I->tSatCur;
    I->SourceSatCur = I->tSatCur;
  } else {
    I->DrainSatCur = I->tSatCurDens * I->drainArea;
    I->SourceSatCur = I->tSatCurDens * I->sourceArea;
  }
  Beta = I->tTransconductance * I->w / I->EffectiveLength;
  I->Vd = sol(v->nextSol, I->li_Drain);
  I->Vg = sol(v->nextSol, I->li_Gate);
  I->Vs = sol(v->nextSol, I->li_Source);
  I->Vb = sol(v->nextSol, I->li_Bulk);
  I->Vsp = sol(v->nextSol, I->li_SourcePrime);
  I->Vdp = sol(v->nextSol, I->li_DrainPrime);
  I->Vddp = I->Vd - I->Vdp;
  I->Vssp = I->Vs - I->Vsp;
  I->Vbsp = I->Vb - I->Vsp;
  I->Vbdp = I->Vb - I->Vdp;
  I->Vgsp = I->Vg - I->Vsp;
  I->Vgdp = I->Vg - I->Vdp;
  I->Vgb = I->Vg - I->Vb;
  I->Vdpsp = I->Vdp - I->Vsp;
  I->vbs = m->dtype * I->Vbsp;
  I->vgs = m->dtype * I->Vgsp;
  I->vds = m->dtype * I->Vdpsp;
  I->vbd = I->vbs - I->vds;
  I->vgd = I->vgs - I->vds;
  I->origFlag = 1;
  I->limitedFlag = 0;
  I->vgs_orig = I->vgs;
  I->vds_orig = I->vds;
  I->vbs_orig = I->vbs;
  I->vbd_orig = I->vbd;
  I->vgd_orig = I->vgd;
  if (ss->initJctFlag && !I->OFF && g->opt.voltageLimiterFlag) {
    if (I->IC_GIVEN) {
      I->vds = m->dtype * I->icVDS;
      I->vgs = m->dtype * I->icVGS;
      I->vbs = m->dtype * I->icVBS;
      I->vbd = I->vbs - I->vds;
      I->vgd = I->vgs - I->vds;
      I->origFlag = 0;
    } else {
      int set = 1;
      if (ss->inputOPFlag) {
        /* the reference tests drain, gate and source for == 0 but the
         * internal nodes and the bulk for != 0 (:2696-2698) */
        const double* fl = v->flagSol;
        /* a ground node reads its flag as 0 */
        set = sol(fl, I->li_Drain) == 0 || sol(fl, I->li_Gate) == 0 || sol(fl, I->li_Source) == 0 ||
              sol(fl, I->li_SourcePrime) != 0 || sol(fl, I->li_DrainPrime) != 0 ||
              sol(fl, I->li_Bulk) != 0;
      }
      if (set) {
        I->vbs = -1;
        I->vgs = m->dtype * I->tVto;
        I->vds = 0;
        I->vbd = I->vbs - I->vds;
        I->vgd = I->vgs - I->vds;
      }
    }
  } else if ((ss->initFixFlag || ss->initJctFlag) && I->OFF) {
    I->vbs = I->vgs = I->vds = 0;
    I->vbd = I->vgd = 0;
  }
  if (ss->newtonIter == 0) {
    if (!(ss->dcopFlag) || (ss->locaEnabledFlag && ss->dcopFlag)) {
      I->vbs_old = v->currSto[I->li_store_vbs];
      I->vbd_old = v->currSto[I->li_store_vbd];
      I->vgs_old = v->currSto[I->li_store_vgs];
      I->vds_old = v->currSto[I->li_store_vds];
      Von = m->dtype * v->currSto[I->li_store_von];
    } else {
      I->vbs_old = I->vbs;
      I->vbd_old = I->vbd;
      I->vgs_old = I->vgs;
      I->vds_old = I->vds;
      Von = 0.0;
    }
    I->vgd_old = I->vgs_old - I->vds_old;
  } else {
    I->vbs_old = v->nextSto[I->li_store_vbs];
    I->vbd_old = v->nextSto[I->li_store_vbd];
    I->vgs_old = v->nextSto[I->li_store_vgs];
    I->vds_old = v->nextSto[I->li_store_vds];
    Von = m->dtype * v->nextSto[I->li_store_von];
    I->vgd_old = I->vgs_old - I->vds_old;
  }
  if (g->opt.voltageLimiterFlag) {
    if (!(ss->initFixFlag && I->OFF)) {
      if (I->vds_old >= 0) {
        I->vgs = or_fetlim(I->vgs, I->vgs_old, Von);
        I->vds = I->vgs - I->vgd;
        I->vds = or_limvds(I->vds, I->vds_old);
        I->vgd = I->vgs - I->vds;
      } else {
        I->vgd = or_fetlim(I->vgd, I->vgd_old, Von);
        I->vds = I->vgs - I->vgd;
        I->vds = -or_limvds(-I->vds, -I->vds_old);
        I->vgs = I->vgd + I->vds;
      }
      if (I->vds >= 0.0) {
        I->vbs = or_pnjlim(I->vbs, I->vbs_old, I->vt, I->sourceVcrit, &Check);
        I->vbd = I->vbs - I->vds;
      } else {
        I->vbd = or_pnjlim(I->vbd, I->vbd_old, I->vt, I->drainVcrit, &Check);
        I->vbs = I->vbd + I->vds;
      }
      if (Check == 1) I->limitedFlag = 1;
    }
  }
  I->vbd = I->vbs - I->vds;
  I->vgd = I->vgs - I->vds;
  I->Vgb = I->vgs - I->vbs;
  if (I->vgs_orig != I->vgs || I->vds_orig != I->vds || I->vbs_orig != I->vbs ||
      I->vbd_orig != I->vbd || I->vgd_orig != I->vgd)
    I->origFlag = 0;
  if (I->vbs <= 0) {
    I->gbs = I->SourceSatCur / I->vt;
    I->gbs += gmin;
    I->cbs = I->gbs * I->vbs;
  } else {
    evbs = exp(or_min(OR_MAX_EXP_ARG, I->vbs / I->vt));
    I->gbs = (I->SourceSatCur * evbs / I->vt + gmin);
    I->cbs = (I->SourceSatCur * (evbs - 1) + gmin * I->vbs);
  }
  if (I->vbd <= 0) {
    I->gbd = I->DrainSatCur / I->vt;
    I->gbd += gmin;
    I->cbd = I->gbd * I->vbd;
  } else {
    evbd = exp(or_min(OR_MAX_EXP_ARG, I->vbd / I->vt));
    I->gbd = (I->DrainSatCur * evbd / I->vt + gmin);
    I->cbd = (I->DrainSatCur * (evbd - 1) + gmin * I->vbd);
  }
  I->mode = (I->vds >= 0) ? 1 : -1;
  {
    double betap, vgst;
    const double vbx = (I->mode == 1) ? I->vbs : I->vbd;
    vds_save = I->vds;
    vgs_save = I->vgs;
    vgd_save = I->vgd;
    if (ss->artParameterFlag) {
      /* continuation on the gain: contVds / contVgst (N_DEV_DeviceSupport.C)
       * with the reference's defaults vdsScaleMin = 0.3, vgstConst = 4.5 */
      const double alpha = ss->gainScale;
      const double vdsScaleMin = 0.3, vgstConst = 4.5;
      const double nlscale = ss->nltermScale;
      I->vds = I->vds * ((1.0 - vdsScaleMin) * nlscale + vdsScaleMin);
      if (I->mode == 1) I->vgs = I->vgs * alpha + (1.0 - alpha) * vgstConst;
      else I->vgd = I->vgd * alpha + (1.0 - alpha) * vgstConst;
    }
    if (vbx <= 0) {
      sarg = sqrt(I->tPhi - vbx);
    } else {
      sarg = sqrt(I->tPhi);
      sarg = sarg - vbx / (sarg + sarg);
      sarg = or_max(0.0, sarg);
    }
    Von = (I->tVbi * m->dtype) + m->gamma * sarg;
    vgst = ((I->mode == 1) ? I->vgs : I->vgd) - Von;
    Vdsat = or_max(vgst, 0.0);
    if (sarg <= 0) arg = 0;
    else arg = m->gamma / (sarg + sarg);
    if (vgst <= 0) {
      I->cdrain = 0;
      I->gm = 0;
      I->gds = 0;
      I->gmbs = 0;
    } else {
      betap = Beta * (1 + m->lambda * (I->vds * I->mode));
      if (vgst <= (I->vds * I->mode)) {
        I->cdrain = betap * vgst * vgst * .5;
        I->gm = betap * vgst;
        I->gds = m->lambda * Beta * vgst * vgst * .5;
        I->gmbs = I->gm * arg;
      } else {
        I->cdrain = betap * (I->vds * I->mode) * (vgst - .5 * (I->vds * I->mode));
        I->gm = betap * (I->vds * I->mode);
        I->gds = betap * (vgst - (I->vds * I->mode)) +
                 m->lambda * Beta * (I->vds * I->mode) * (vgst - .5 * (I->vds * I->mode));
        I->gmbs = I->gm * arg;
      }
    }
  }
  I->von = m->dtype * Von;
  I->vdsat = m->dtype * Vdsat;
  I->cd = I->mode * I->cdrain - I->cbd;
  if (I->Cbs != 0 || I->Cbssw != 0) {
    if (I->vbs < I->tDepCap) {
      arg = 1 - I->vbs / I->tBulkPot;
      if (m->bulkJctBotGradingCoeff == m->bulkJctSideGradingCoeff) {
        if (m->bulkJctBotGradingCoeff == .5) sarg = sargsw = 1 / sqrt(arg);
        else sarg = sargsw = exp(-m->bulkJctBotGradingCoeff * log(arg));
      } else {
        if (m->bulkJctBotGradingCoeff == .5) sarg = 1 / sqrt(arg);
        else sarg = exp(-m->bulkJctBotGradingCoeff * log(arg));
        if (m->bulkJctSideGradingCoeff == .5) sargsw = 1 / sqrt(arg);
        else sargsw = exp(-m->bulkJctSideGradingCoeff * log(arg));
      }
      I->qbs = I->tBulkPot * (I->Cbs * (1 - arg * sarg) / (1 - m->bulkJctBotGradingCoeff) +
                              I->Cbssw * (1 - arg * sargsw) / (1 - m->bulkJctSideGradingCoeff));
      I->capbs = I->Cbs * sarg + I->Cbssw * sargsw;
    } else {
      I->qbs = I->f4s + I->vbs * (I->f2s + I->vbs * (I->f3s / 2));
      I->capbs = I->f2s + I->f3s * I->vbs;
    }
  } else {
    I->qbs = 0;
    I->capbs = 0;
  }
  if (I->Cbd != 0 || I->Cbdsw != 0) {
    if (I->vbd < I->tDepCap) {
      arg = 1 - I->vbd / I->tBulkPot;
      if (m->bulkJctBotGradingCoeff == .5 && m->bulkJctSideGradingCoeff == .5) {
        sarg = sargsw = 1 / sqrt(arg);
      } else {
        if (m->bulkJctBotGradingCoeff == .5) sarg = 1 / sqrt(arg);
        else sarg = exp(-m->bulkJctBotGradingCoeff * log(arg));
        if (m->bulkJctSideGradingCoeff == .5) sargsw = 1 / sqrt(arg);
        else sargsw = exp(-m->bulkJctSideGradingCoeff * log(arg));
      }
      I->qbd = I->tBulkPot * (I->Cbd * (1 - arg * sarg) / (1 - m->bulkJctBotGradingCoeff) +
                              I->Cbdsw * (1 - arg * sargsw) / (1 - m->bulkJctSideGradingCoeff));
      I->capbd = I->Cbd * sarg + I->Cbdsw * sargsw;
    } else {
      I->qbd = I->f4d + I->vbd * (I->f2d + I->vbd * I->f3d / 2);
      I->capbd = I->f2d + I->vbd * I->f3d;
    }
  } else {
    I->qbd = 0;
    I->capbd = 0;
  }
  if (I->mode > 0) {
    or_qmeyer(I->vgs, I->vgd, I->Vgb, Von, Vdsat, &I->capgs, &I->capgd, &I->capgb, I->tPhi,
              I->OxideCap);
    or_qmeyerderivs(I->vgs, I->vgd, I->Vgb, Von, Vdsat, I->dcap, I->tPhi, I->OxideCap, m->dtype);
  } else {
    or_qmeyer(I->vgd, I->vgs, I->Vgb, Von, Vdsat, &I->capgd, &I->capgs, &I->capgb, I->tPhi,
              I->OxideCap);
    or_qmeyerderivs(I->vgd, I->vgs, I->Vgb, Von, Vdsat, I->dcap, I->tPhi, I->OxideCap, m->dtype);
  }
  if (ss->dcopFlag) {
    I->Capgs = 2.0 * I->capgs + I->GateSourceOverlapCap;
    I->Capgd = 2.0 * I->capgd + I->GateDrainOverlapCap;
    I->Capgb = 2.0 * I->capgb + I->GateBulkOverlapCap;
  } else {
    capgs_old = v->currSta[I->li_state_capgs];
    capgd_old = v->currSta[I->li_state_capgd];
    capgb_old = v->currSta[I->li_state_capgb];
    I->Capgs = (I->capgs + capgs_old + I->GateSourceOverlapCap);
    I->Capgd = (I->capgd + capgd_old + I->GateDrainOverlapCap);
    I->Capgb = (I->capgb + capgb_old + I->GateBulkOverlapCap);
  }
  I->Capgs *= (I->Capgs < 0.0) ? -1.0 : 1.0;
  I->Capgd *= (I->Capgd < 0.0) ? -1.0 : 1.0;
  I->Capgb *= (I->Capgb < 0.0) ? -1.0 : 1.0;
  I->Idrain = I->drainConductance * I->Vddp;
  I->Isource = I->sourceConductance * I->Vssp;
  if (I->mode >= 0) {
    I->Gm = I->gm;
    I->Gmbs = I->gmbs;
    I->nrmsum = I->Gm + I->Gmbs;
    I->revsum = 0;
    I->cdreq = m->dtype * I->cdrain;
  } else {
    I->Gm = -I->gm;
    I->Gmbs = -I->gmbs;
    I->nrmsum = 0;
    I->revsum = -(I->Gm + I->Gmbs);
    I->cdreq = -(m->dtype) * I->cdrain;
  }
  I->vds = vds_save;
  I->vgs = vgs_save;
  I->vgd = vgd_save;
  return 1;
}

/* Master::updateState (:4069-4158) */
static int update_state(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  m1inst* in = (m1inst*)g->inst;
  int ok = 1;
  /* instances are independent here; one thread is the serial order */
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1) reduction(&& : ok)
  for (i = 0; i < g->n; ++i) {
    m1inst* I = &in[i];
    double* sto = v->nextSto;
    double* sta = v->nextSta;
    ok = update_intermediate(g, I, ss, v) && ok;
    sto[I->li_store_vbd] = I->vbd;
    sto[I->li_store_vbs] = I->vbs;
    sto[I->li_store_vgs] = I->vgs;
    sto[I->li_store_vds] = I->vds;
    sto[I->li_store_von] = I->von;
    sta[I->li_state_capgs] = I->capgs;
    sta[I->li_state_capgd] = I->capgd;
    sta[I->li_state_capgb] = I->capgb;
    if (ss->dcopFlag) {
      I->qgs = I->Capgs * I->vgs;
      I->qgd = I->Capgd * I->vgd;
      I->qgb = I->Capgb * I->Vgb;
    } else {
      double vgs1, vbs1, vds1, vgb1, vgd1;
      I->qgs = v->currSta[I->li_state_qgs];
      I->qgd = v->currSta[I->li_state_qgd];
      I->qgb = v->currSta[I->li_state_qgb];
      vgs1 = v->currSto[I->li_store_vgs];
      vbs1 = v->currSto[I->li_store_vbs];
      vds1 = v->currSto[I->li_store_vds];
      vgb1 = vgs1 - vbs1;
      vgd1 = vgs1 - vds1;
      I->qgs += I->Capgs * (I->vgs - vgs1);
      I->qgd += I->Capgd * (I->vgd - vgd1);
      I->qgb += I->Capgb * ((I->vgs - I->vbs) - vgb1);
    }
    sta[I->li_state_qgs] = I->qgs;
    sta[I->li_state_qgd] = I->qgd;
    sta[I->li_state_qgb] = I->qgb;
    sta[I->li_state_qbd] = I->qbd;
    sta[I->li_state_qbs] = I->qbs;
  }
  return ok ? XA_OK : XA_ERR_DEVICE;
}

/* Master::loadDAEVectors (:4168-4396), newMeyerFlag = 0 */
static int load_vectors(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  m1inst* in = (m1inst*)g->inst;
  const double gmin1 = g->opt.gmin;
  for (i = 0; i < g->n; ++i) {
    m1inst* mi = &in[i];
    const int Dtype = mi->model->dtype;
    const double Mp = mi->numberParallel;
    double ceqbs, ceqbd, ceqgb, ceqgs, ceqgd, Qeqbs, Qeqbd, Qeqgb, Qeqgs, Qeqgd, coef;
    ceqbs = Dtype * (mi->cbs);
    ceqbd = Dtype * (mi->cbd);
    ceqgb = 0.0;
    ceqgs = 0.0;
    ceqgd = 0.0;
    if (v->f) {
      if (mi->drainConductance != 0.0) OR_ADD(v->f, mi->li_Drain, mi->Idrain * Mp);
      coef = (ceqgs + ceqgd + ceqgb);
      OR_ADD(v->f, mi->li_Gate, coef * Mp);
      if (mi->sourceConductance != 0.0) OR_ADD(v->f, mi->li_Source, mi->Isource * Mp);
      coef = ceqbs + ceqbd - ceqgb;
      OR_ADD(v->f, mi->li_Bulk, coef * Mp);
      coef = -mi->Idrain - (ceqbd - mi->cdreq + ceqgd);
      OR_ADD(v->f, mi->li_DrainPrime, coef * Mp);
      coef = -mi->Isource - (ceqbs + mi->cdreq + ceqgs);
      OR_ADD(v->f, mi->li_SourcePrime, coef * Mp);
    }
    Qeqbs = Dtype * (mi->qbs);
    Qeqbd = Dtype * (mi->qbd);
    Qeqgb = Dtype * (mi->qgb);
    Qeqgs = Dtype * (mi->qgs);
    Qeqgd = Dtype * (mi->qgd);
    if (v->q) {
      coef = (Qeqgs + Qeqgd + Qeqgb);
      OR_ADD(v->q, mi->li_Gate, coef * Mp);
      coef = Qeqbs + Qeqbd - Qeqgb;
      OR_ADD(v->q, mi->li_Bulk, coef * Mp);
      coef = -(Qeqbd + Qeqgd);
      OR_ADD(v->q, mi->li_DrainPrime, coef * Mp);
      coef = -(Qeqbs + Qeqgs);
      OR_ADD(v->q, mi->li_SourcePrime, coef * Mp);
    }
    if (!mi->origFlag) {
      const double dvbd = mi->vbd - mi->vbd_orig, dvbs = mi->vbs - mi->vbs_orig;
      const double dvds = mi->vds - mi->vds_orig;
      const double dvgx = (mi->mode > 0) ? (mi->vgs - mi->vgs_orig) : (mi->vgd - mi->vgd_orig);
      const double dvbx = (mi->mode > 0) ? (mi->vbs - mi->vbs_orig) : (mi->vbd - mi->vbd_orig);
      double gcgd, gcgs, gcgb, gcbs, gcbd;
      const double J4 = Dtype * (+((mi->gbd - gmin1)) * dvbd + ((mi->gbs - gmin1)) * dvbs);
      const double J5 = Dtype * (-((mi->gbd - gmin1)) * dvbd + mi->gds * dvds + mi->Gm * dvgx +
                                 mi->Gmbs * dvbx);
      const double J6 = Dtype * (-((mi->gbs - gmin1)) * dvbs - mi->gds * dvds - mi->Gm * dvgx -
                                 mi->Gmbs * dvbx);
      if (v->dFdxdVp) {
        OR_ADD(v->dFdxdVp, mi->li_Bulk, J4 * Mp);
        OR_ADD(v->dFdxdVp, mi->li_DrainPrime, J5 * Mp);
        OR_ADD(v->dFdxdVp, mi->li_SourcePrime, J6 * Mp);
      }
      if (ss->tranopFlag || ss->acopFlag || ss->transientFlag) {
        gcgd = mi->Capgd; gcgs = mi->Capgs; gcgb = mi->Capgb;
        gcbs = mi->capbs; gcbd = mi->capbd;
      } else {
        gcgd = 0.0; gcgs = 0.0; gcgb = 0.0; gcbs = 0.0; gcbd = 0.0;
      }
      {
        const double Q2 = Dtype * (gcgd * (mi->vgd - mi->vgd_orig) + gcgs * (mi->vgs - mi->vgs_orig) +
                                   gcgb * (mi->vgs - mi->vgs_orig - mi->vbs + mi->vbs_orig));
        const double Q4 = Dtype * (-(gcgb) * (mi->vgs - mi->vgs_orig - mi->vbs + mi->vbs_orig) +
                                   (gcgb) * (mi->vbd - mi->vbd_orig) + (gcbs) * (mi->vbs - mi->vbs_orig));
        const double Q5 = Dtype * (-(gcgd) * (mi->vgd - mi->vgd_orig) - (gcbd) * (mi->vbd - mi->vbd_orig));
        const double Q6 = Dtype * (-gcgs * (mi->vgs - mi->vgs_orig) - (gcbs) * (mi->vbs - mi->vbs_orig));
        if (v->dQdxdVp) {
          OR_ADD(v->dQdxdVp, mi->li_Gate, Q2 * Mp);
          OR_ADD(v->dQdxdVp, mi->li_Bulk, Q4 * Mp);
          OR_ADD(v->dQdxdVp, mi->li_DrainPrime, Q5 * Mp);
          OR_ADD(v->dQdxdVp, mi->li_SourcePrime, Q6 * Mp);
        }
      }
    }
    if (mi->loadLeadCurrent) {
      double* LF = v->storeLeadF;
      double* LQ = v->storeLeadQ;
      if (LF && LQ) {
        if (mi->drainConductance != 0.0) {
          OR_SET(LF, mi->li_store_dev_id, mi->Idrain * Mp);
        } else {
          OR_SET(LF, mi->li_store_dev_id, (-mi->Idrain - (ceqbd - mi->cdreq + ceqgd)) * Mp);
          OR_SET(LQ, mi->li_store_dev_id, (-(Qeqbd + Qeqgd)) * Mp);
        }
        if (mi->sourceConductance != 0.0) {
          OR_SET(LF, mi->li_store_dev_is, mi->Isource * Mp);
        } else {
          OR_SET(LF, mi->li_store_dev_is, (-mi->Isource - (ceqbs + mi->cdreq + ceqgs)) * Mp);
          OR_SET(LQ, mi->li_store_dev_is, (-(Qeqbs + Qeqgs)) * Mp);
        }
        OR_SET(LF, mi->li_store_dev_ig, (ceqgs + ceqgd + ceqgb) * Mp);
        OR_SET(LQ, mi->li_store_dev_ig, (Qeqgs + Qeqgd + Qeqgb) * Mp);
        OR_SET(LF, mi->li_store_dev_ib, (ceqbs + ceqbd - ceqgb) * Mp);
        OR_SET(LQ, mi->li_store_dev_ib, (Qeqbs + Qeqbd - Qeqgb) * Mp);
      }
    }
  }
  return XA_OK;
}

/* Master::loadDAEMatrices (:4407-4688), newMeyerFlag = 0 */
static int load_matrices(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  m1inst* in = (m1inst*)g->inst;
  double* F = v->dFdx;
  double* Q = v->dQdx;
  (void)ss;
  for (i = 0; i < g->n; ++i) {
    m1inst* mi = &in[i];
    const double Mp = mi->numberParallel;
    if (F) {
      OR_ADD(F, mi->fDD, mi->drainConductance * Mp);
      OR_SUB(F, mi->fDDp, mi->drainConductance * Mp);
      OR_ADD(F, mi->fSS, mi->sourceConductance * Mp);
      OR_SUB(F, mi->fSSp, mi->sourceConductance * Mp);
      OR_ADD(F, mi->fBB, (mi->gbs + mi->gbd) * Mp);
      OR_SUB(F, mi->fBDp, mi->gbd * Mp);
      OR_SUB(F, mi->fBSp, mi->gbs * Mp);
      OR_SUB(F, mi->fDpD, mi->drainConductance * Mp);
      OR_ADD(F, mi->fDpG, (mi->Gm) * Mp);
      OR_ADD(F, mi->fDpB, (-mi->gbd + mi->Gmbs) * Mp);
      OR_ADD(F, mi->fDpDp, (mi->drainConductance + mi->gds + mi->gbd + mi->revsum) * Mp);
      OR_ADD(F, mi->fDpSp, (-mi->gds - mi->nrmsum) * Mp);
      OR_SUB(F, mi->fSpG, (mi->Gm) * Mp);
      OR_SUB(F, mi->fSpS, mi->sourceConductance * Mp);
      OR_SUB(F, mi->fSpB, (mi->gbs + mi->Gmbs) * Mp);
      OR_SUB(F, mi->fSpDp, (mi->gds + mi->revsum) * Mp);
      OR_ADD(F, mi->fSpSp, (mi->sourceConductance + mi->gds + mi->gbs + mi->nrmsum) * Mp);
    }
    if (Q) {
      double gcgd = 0.0, gcgs = 0.0, gcgb = 0.0, gcbs = 0.0, gcbd = 0.0;
      if (ss->tranopFlag || ss->acopFlag || ss->transientFlag) {
        gcgd = mi->Capgd; gcgs = mi->Capgs; gcgb = mi->Capgb;
        gcbs = mi->capbs; gcbd = mi->capbd;
      }
      OR_ADD(Q, mi->qGG, (gcgd + gcgs + gcgb) * Mp);
      OR_SUB(Q, mi->qGB, gcgb * Mp);
      OR_SUB(Q, mi->qGDp, gcgd * Mp);
      OR_SUB(Q, mi->qGSp, gcgs * Mp);
      OR_SUB(Q, mi->qBG, gcgb * Mp);
      OR_ADD(Q, mi->qBB, (+gcbs + gcbd + gcgb) * Mp);
      OR_SUB(Q, mi->qBDp, +gcbd * Mp);
      OR_SUB(Q, mi->qBSp, +gcbs * Mp);
      OR_ADD(Q, mi->qDpG, -gcgd * Mp);
      OR_ADD(Q, mi->qDpB, -gcbd * Mp);
      OR_ADD(Q, mi->qDpDp, (+gcbd + gcgd) * Mp);
      OR_SUB(Q, mi->qSpG, gcgs * Mp);
      OR_SUB(Q, mi->qSpB, +gcbs * Mp);
      OR_ADD(Q, mi->qSpSp, (+gcbs + gcgs) * Mp);
    }
  }
  return XA_OK;
}

static int read_var(struct or_group_s* g, const char* name, double* out) {
  int64_t i;
  m1inst* in = (m1inst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    const m1inst* I = &in[i];
    if (!strcmp(name, "cdrain")) out[i] = I->cdrain;
    else if (!strcmp(name, "gm")) out[i] = I->gm;
    else if (!strcmp(name, "gds")) out[i] = I->gds;
    else if (!strcmp(name, "gmbs")) out[i] = I->gmbs;
    else if (!strcmp(name, "von")) out[i] = I->von;
    else if (!strcmp(name, "vdsat")) out[i] = I->vdsat;
    else if (!strcmp(name, "mode")) out[i] = I->mode;
    else if (!strcmp(name, "origFlag")) out[i] = I->origFlag;
    else if (!strcmp(name, "Capgs")) out[i] = I->Capgs;
    else if (!strcmp(name, "Capgd")) out[i] = I->Capgd;
    else if (!strcmp(name, "Capgb")) out[i] = I->Capgb;
    else if (!strcmp(name, "tVto")) out[i] = I->tVto;
    else if (!strcmp(name, "tPhi")) out[i] = I->tPhi;
    else if (!strcmp(name, "OxideCap")) out[i] = I->OxideCap;
    else if (!strcmp(name, "tCj")) out[i] = I->tCj;
    else if (!strcmp(name, "tCjsw")) out[i] = I->tCjsw;
    else if (!strcmp(name, "tBulkPot")) out[i] = I->tBulkPot;
    else return or_fail(XA_ERR_ARG, "mosfet1: unknown intermediate");
  }
  return XA_OK;
}

const or_model_ops or_mosfet1_ops = {
  "mosfet1", mpar, sizeof(mpar) / sizeof(mpar[0]), NM1_, ipar, sizeof(ipar) / sizeof(ipar[0]),
  NI1_, setup, register_lids, jac_pairs, register_jac, update_state, NULL, load_vectors,
  load_matrices, read_var, NULL, sizeof(m1inst),
};
