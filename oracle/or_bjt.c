/* or_bjt.c -- oracle restatement of the Gummel-Poon BJT (level 1), test
 * infrastructure only.
 * Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_BJT.C
 *   parameters :85-500, Instance ctor :555-1038 (node counts :1003-1036),
 *   registerLIDs :1061-1176, registerStateLIDs :1255-1270,
 *   registerStoreLIDs :1281-1300, updateTemperature :1603-1773,
 *   lambertWCurrent :1784-1800, oldDAEExcessPhaseCalculation1/2 :2370-2466,
 *   updateIntermediateVars :2476-3207, Model::processParams :3290-3361,
 *   Model ctor :3394-3798, Master::updateState :3891-3937,
 *   Master::updateSecondaryState :3947-3968, Master::loadDAEVectors :3978-4142,
 *   Master::loadDAEMatrices :4152-4260.
 * newExcessPhase = 1 (the new-DAE excess phase: two internal unknowns, the
 * phase-shifted current i_fx and its derivative, :908-931, :1032-1036,
 * :1149-1153, auxDAECalculations :1928-1980, the Master's excess-phase rows
 * :3997-4024, :4033-4046, :4069-4100, :4177-4260) is restated; external
 * internal nodes are not.
 *
 * Aliases (BF/BFM, VAF/VA/VBF, ...) write the same model value but keep
 * their own given flag: the reference reconciles the alias flags before it
 * reads the card (Model ctor :3521-3773 precede setModParams :3784), so only
 * the flag of the name actually used is set.  Giving two aliases of one value
 * is rejected here (the reference keeps the netlist-last one). */
#include <stdio.h>
#include "or_internal.h"
#include "or_eval_libm.h"

enum {
  P_TNOM, P_IS, P_BF, P_BFM, P_NF, P_VA, P_VBF, P_VAF, P_IKF, P_IK, P_JBF, P_ISE, P_JLE, P_NE,
  P_NLE, P_BR, P_BRM, P_NR, P_VAR, P_VB, P_VRB, P_BV, P_IKR, P_JBR, P_ISC, P_JLC, P_NC, P_RB,
  P_IRB, P_JRB, P_IOB, P_RBM, P_RE, P_RC, P_CJE, P_VJE, P_PE, P_MJE, P_ME, P_TF, P_XTF, P_VTF,
  P_ITF, P_JTF, P_PTF, P_CJC, P_VJC, P_PC, P_MJC, P_MC, P_XCJC, P_CDIS, P_TR, P_CJS, P_CCS,
  P_CSUB, P_VJS, P_PS, P_PSUB, P_MJS, P_MS, P_ESUB, P_XTB, P_TB, P_TCB, P_EG, P_XTI, P_PT,
  P_KF, P_AF, P_FC, P_C2, P_C4, P_NK, P_NKF, P_TYPE, NMB_
};
static const or_param mpar[] = {
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
enum { Q_AREA, Q_IC1, Q_IC2, Q_TEMP, Q_OFF, Q_LAMBERTW, NIB_ };
static const or_param ipar[] = {
  {"AREA", Q_AREA, 1.0, 0}, {"IC1", Q_IC1, 0.0, 0}, {"IC2", Q_IC2, 0.0, 0},
  {"TEMP", Q_TEMP, 0.0, 1}, {"OFF", Q_OFF, 0.0, 0}, {"LAMBERTW", Q_LAMBERTW, 0.0, 0},
};

typedef struct {
  int TYPE;
  double TNOM, satCur, betaF, emissionCoeffF, earlyVoltF, rollOffF, leakBECurrent;
  double leakBEEmissionCoeff, betaR, emissionCoeffR, earlyVoltR, rollOffR, leakBCCurrent;
  double leakBCEmissionCoeff, baseResist, baseCurrHalfResist, minBaseResist, emitterResist;
  double collectorResist, depCapBE, potBE, juncExpBE, transTimeF, transTimeBiasCoeffF;
  double transTimeFVBC, transTimeHighCurrF, excessPhase, depCapBC, potBC, juncExpBC;
  double baseFracBCCap, transTimeR, CJS, potSubst, expSubst, betaExp, energyGap, tempExpIS;
  double depCapCoeff, c2, c4, rollOffExp;
  int leakBECurrentGiven, leakBCCurrentGiven, c2Given, c4Given, minBaseResistGiven;
  int VAFgiven, IKFgiven, VARgiven, IKRgiven;
  double invEarlyVoltF, invRollOffF, invEarlyVoltR, invRollOffR, collectorConduct;
  double emitterConduct, transTimeVBCFac, excessPhaseFac, f2, f3, f6, f7;
} bmodel;

typedef struct {
  bmodel* model;
  double AREA, icVBE, icVCE, TEMP;
  int IC_GIVEN, OFF, lambertWFlag, loadLeadCurrent;
  /* temperature */
  double vt, tBECap, tBEPot, tBCCap, tBCPot, tDepCap, tF1, tF4, tF5, tVCrit, tSatCur;
  double tBetaF, tBetaR, tBELeakCur, tBCLeakCur, tleakBEEmissionCoeff, tleakBCEmissionCoeff;
  double tRollOffExp, tInvRollOffF, tInvRollOffR, tInvEarlyVoltF, tInvEarlyVoltR;
  /* evaluation */
  double vEEp, vBBp, vCCp, vBE, vBC, vBX, vCS, vBE_orig, vBC_orig, vBE_old, vBC_old, capeqCB;
  int origFlag, offFlag;
  double iBE, gBE, iBEleak, gBEleak, iBC, gBC, iBCleak, gBCleak, qB, invqB;
  double dqBdvEp, dqBdvCp, dqBdvBp, iBEhighCurr, gBEhighCurr, geqCB;
  double qBEdep, capBEdep, qBEdiff, capBEdiff, qBCdep, capBCdep, qBCdiff, capBCdiff;
  double qBX, capBX, qCS, capCS, currCexbc, lastCexbc, nextCexbc, phaseScalar;
  double gCapBEdiff, gCapBEdep, gCapBCdiff, gCapBCdep, gCapBX, gCapCS;
  double iC, iCE, iB, iE, diCEdvEp, diCEdvCp, diCEdvBp, gEpr, gCpr, gX;
  double diBrdvB, diBrdvCp, diBrdvEp, diBrdvBp, gBEtot, gBCtot;
  double iBEdiff, iBEdep, iCS, iBCdiff, iBCdep, iBX;
  double diBEdvEp, diBEdvCp, diBEdvBp;   /* new excess phase (auxDAECalculations) */
  /* LIDs and slots */
  int li_Coll, li_Base, li_Emit, li_Subst, li_CollP, li_BaseP, li_EmitP, li_Ifx, li_dIfx;
  int li_qstateBEdiff, li_qstateBEdep, li_qstateCS, li_qstateBCdiff, li_qstateBCdep;
  int li_qstateBX, li_istateCEXBC;
  int li_storevBE, li_storevBC, li_store_capeqCB;
  int li_store_dev_ib, li_store_dev_ie, li_store_dev_ic, li_store_dev_is;
  int64_t f_CC, f_CCp, f_BB, f_BCp, f_BBp, f_BEp, f_EE, f_EEp, f_CpC, f_CpCp, f_CpBp, f_CpEp;
  int64_t f_BpB, f_BpCp, f_BpBp, f_BpEp, f_EpE, f_EpCp, f_EpBp, f_EpEp;
  int64_t q_BB, q_BCp, q_SS, q_SCp, q_CpB, q_CpS, q_CpCp, q_CpBp, q_BpCp, q_BpBp, q_BpEp;
  int64_t q_EpCp, q_EpBp, q_EpEp;
  /* new excess phase slots (-1 when the option is off) */
  int64_t f_CpIfx, f_EpIfx, f_IfxdIfx, f_dIfxCp, f_dIfxBp, f_dIfxEp, f_dIfxdIfx, f_dIfxIfx;
  int64_t f_IfxCp, f_IfxBp, f_IfxEp, f_IfxIfx, q_IfxIfx, q_dIfxdIfx;
} binst;

/* one value from a group of alias names; more than one given is an error */
static int alias(const or_card* c, const int* slots, int n, double* out) {
  int k, given = 0;
  *out = c->v[slots[0]];
  for (k = 0; k < n; ++k)
    if (c->given[slots[k]]) { *out = c->v[slots[k]]; ++given; }
  if (given > 1) return or_fail(XA_ERR_ARG, "bjt: a model value is given under two alias names");
  return XA_OK;
}

/* Model ctor (:3394-3798) and Model::processParams (:3290-3361) */
static int model_init(bmodel* m, const or_card* c, const xa_device_options* opt) {
#define ALIAS(dst, ...) do { const int s_[] = {__VA_ARGS__}; \
    if (alias(c, s_, (int)(sizeof(s_) / sizeof(s_[0])), &(dst)) != XA_OK) return XA_ERR_ARG; } while (0)
  double xfc;
  memset(m, 0, sizeof(*m));
  m->TYPE = (c->v[P_TYPE] < 0) ? -1 : 1;
  m->TNOM = c->given[P_TNOM] ? c->v[P_TNOM] : opt->tnom;
  m->satCur = c->v[P_IS];
  ALIAS(m->betaF, P_BF, P_BFM);
  m->emissionCoeffF = c->v[P_NF];
  ALIAS(m->earlyVoltF, P_VAF, P_VA, P_VBF);
  ALIAS(m->rollOffF, P_IKF, P_IK, P_JBF);
  ALIAS(m->leakBECurrent, P_ISE, P_JLE);
  ALIAS(m->leakBEEmissionCoeff, P_NE, P_NLE);
  ALIAS(m->betaR, P_BR, P_BRM);
  m->emissionCoeffR = c->v[P_NR];
  ALIAS(m->earlyVoltR, P_VAR, P_VB, P_VRB, P_BV);
  ALIAS(m->rollOffR, P_IKR, P_JBR);
  ALIAS(m->leakBCCurrent, P_ISC, P_JLC);
  m->leakBCEmissionCoeff = c->v[P_NC];
  m->baseResist = c->v[P_RB];
  ALIAS(m->baseCurrHalfResist, P_IRB, P_JRB, P_IOB);
  m->minBaseResist = c->v[P_RBM];
  m->emitterResist = c->v[P_RE];
  m->collectorResist = c->v[P_RC];
  m->depCapBE = c->v[P_CJE];
  ALIAS(m->potBE, P_VJE, P_PE);
  ALIAS(m->juncExpBE, P_MJE, P_ME);
  m->transTimeF = c->v[P_TF];
  m->transTimeBiasCoeffF = c->v[P_XTF];
  m->transTimeFVBC = c->v[P_VTF];
  ALIAS(m->transTimeHighCurrF, P_ITF, P_JTF);
  m->excessPhase = c->v[P_PTF];
  m->depCapBC = c->v[P_CJC];
  ALIAS(m->potBC, P_VJC, P_PC);
  ALIAS(m->juncExpBC, P_MJC, P_MC);
  ALIAS(m->baseFracBCCap, P_XCJC, P_CDIS);
  m->transTimeR = c->v[P_TR];
  ALIAS(m->CJS, P_CJS, P_CCS, P_CSUB);
  ALIAS(m->potSubst, P_VJS, P_PS, P_PSUB);
  ALIAS(m->expSubst, P_MJS, P_MS, P_ESUB);
  ALIAS(m->betaExp, P_XTB, P_TB, P_TCB);
  m->energyGap = c->v[P_EG];
  ALIAS(m->tempExpIS, P_XTI, P_PT);
  m->depCapCoeff = c->v[P_FC];
  m->c2 = c->v[P_C2];
  m->c4 = c->v[P_C4];
  ALIAS(m->rollOffExp, P_NK, P_NKF);
#undef ALIAS
  m->leakBECurrentGiven = c->given[P_ISE];
  m->leakBCCurrentGiven = c->given[P_ISC];
  m->c2Given = c->given[P_C2];
  m->c4Given = c->given[P_C4];
  m->minBaseResistGiven = c->given[P_RBM];
  m->VAFgiven = c->given[P_VAF];
  m->IKFgiven = c->given[P_IKF];
  m->VARgiven = c->given[P_VAR];
  m->IKRgiven = c->given[P_IKR];
  if (!c->given[P_NK]) m->rollOffExp = 0.5;
  /* processParams */
  if (!m->leakBECurrentGiven && m->c2Given) m->leakBECurrent = m->c2 * m->satCur;
  if (!m->leakBCCurrentGiven && m->c4Given) m->leakBCCurrent = m->c4 * m->satCur;
  if (!m->minBaseResistGiven) m->minBaseResist = m->baseResist;
  m->invEarlyVoltF = (m->VAFgiven && (m->earlyVoltF != 0.0)) ? 1.0 / m->earlyVoltF : 0.0;
  m->invRollOffF = (m->IKFgiven && (m->rollOffF != 0.0)) ? 1.0 / m->rollOffF : 0.0;
  m->invEarlyVoltR = (m->VARgiven && (m->earlyVoltR != 0.0)) ? 1.0 / m->earlyVoltR : 0.0;
  m->invRollOffR = (m->IKRgiven && (m->rollOffR != 0.0)) ? 1.0 / m->rollOffR : 0.0;
  m->collectorConduct = (m->collectorResist != 0.0) ? 1.0 / m->collectorResist : 0;
  m->emitterConduct = (m->emitterResist != 0.0) ? 1.0 / m->emitterResist : 0;
  if (c->given[P_VTF] && (m->transTimeFVBC != 0.0)) m->transTimeVBCFac = 1.0 / (m->transTimeFVBC * 1.44);
  else m->transTimeVBCFac = 0.0;
  m->excessPhaseFac = (m->excessPhase / (180.0 / M_PI)) * m->transTimeF;
  if (c->given[P_FC]) {
    if (m->depCapCoeff > 0.9999) m->depCapCoeff = 0.9999;
  } else {
    m->depCapCoeff = 0.5;
  }
  xfc = log(1.0 - m->depCapCoeff);
  m->f2 = exp((1.0 + m->juncExpBE) * xfc);
  m->f3 = 1.0 - m->depCapCoeff * (1.0 + m->juncExpBE);
  m->f6 = exp((1.0 + m->juncExpBC) * xfc);
  m->f7 = 1.0 - m->depCapCoeff * (1.0 + m->juncExpBC);
  return XA_OK;
}

/* Instance::updateTemperature (:1603-1773) */
static void update_temperature(binst* I, double temp) {
  bmodel* m = I->model;
  double fact2, egfet, arg, pbfact, ratlog, ratio1, factlog, factor, bfactor, fact1, pbo, gmaold;
  double gmanew, xfc;
  if (temp != -999.0) I->TEMP = temp;
  I->vt = I->TEMP * OR_KoverQ;
  fact2 = I->TEMP / OR_REFTEMP;
  egfet = OR_Eg0 - (OR_alphaEg * I->TEMP * I->TEMP) / (I->TEMP + OR_betaEg);
  arg = -egfet / (2.0 * OR_BOLTZ * I->TEMP) + OR_Eg300 / (OR_BOLTZ * (2.0 * OR_REFTEMP));
  pbfact = -2.0 * I->vt * (1.5 * log(fact2) + OR_Q * arg);
  ratlog = log(I->TEMP / m->TNOM);
  ratio1 = I->TEMP / m->TNOM - 1.0;
  factlog = ratio1 * m->energyGap / I->vt + m->tempExpIS * ratlog;
  factor = exp(factlog);
  bfactor = exp(ratlog * m->betaExp);
  fact1 = m->TNOM / OR_REFTEMP;
  pbo = (m->potBE - pbfact) / fact1;
  gmaold = (m->potBE - pbo) / pbo;
  I->tBECap = m->depCapBE / (1.0 + m->juncExpBE * (4.e-4 * (m->TNOM - OR_REFTEMP) - gmaold));
  I->tBEPot = fact2 * pbo + pbfact;
  gmanew = (I->tBEPot - pbo) / pbo;
  I->tBECap *= 1.0 + m->juncExpBE * (4.e-4 * (I->TEMP - OR_REFTEMP) - gmanew);
  pbo = (m->potBC - pbfact) / fact1;
  gmaold = (m->potBC - pbo) / pbo;
  I->tBCCap = m->depCapBC / (1.0 + m->juncExpBC * (4.e-4 * (m->TNOM - OR_REFTEMP) - gmaold));
  I->tBCPot = fact2 * pbo + pbfact;
  gmanew = (I->tBCPot - pbo) / pbo;
  I->tBCCap *= 1.0 + m->juncExpBC * (4.e-4 * (I->TEMP - OR_REFTEMP) - gmanew);
  I->tDepCap = m->depCapCoeff * I->tBEPot;
  xfc = log(1.0 - m->depCapCoeff);
  I->tF1 = I->tBEPot * (1.0 - exp((1.0 - m->juncExpBE) * xfc)) / (1.0 - m->juncExpBE);
  I->tF4 = m->depCapCoeff * I->tBCPot;
  I->tF5 = I->tBCPot * (1.0 - exp((1.0 - m->juncExpBC) * xfc)) / (1.0 - m->juncExpBC);
  I->tVCrit = I->vt * log(I->vt / (sqrt(2.0) * m->satCur));
  I->tSatCur = m->satCur * factor;
  I->tBetaF = m->betaF * bfactor;
  I->tBetaR = m->betaR * bfactor;
  if (!m->leakBECurrentGiven && m->c2Given) m->leakBECurrent = m->c2 * m->satCur;
  if (!m->leakBCCurrentGiven && m->c4Given) m->leakBCCurrent = m->c4 * m->satCur;
  I->tBELeakCur = m->leakBECurrent * exp(factlog / m->leakBEEmissionCoeff) / bfactor;
  I->tBCLeakCur = m->leakBCCurrent * exp(factlog / m->leakBCEmissionCoeff) / bfactor;
  I->tleakBEEmissionCoeff = m->leakBEEmissionCoeff;
  I->tleakBCEmissionCoeff = m->leakBCEmissionCoeff;
  I->tRollOffExp = m->rollOffExp;
  I->tInvRollOffF = (m->rollOffF != 0) ? 1 / m->rollOffF : 0;
  I->tInvRollOffR = (m->rollOffR != 0) ? 1 / m->rollOffR : 0;
  I->tInvEarlyVoltF = (m->earlyVoltF != 0) ? 1 / m->earlyVoltF : 0;
  I->tInvEarlyVoltR = (m->earlyVoltR != 0) ? 1 / m->earlyVoltR : 0;
}

static int setup(struct or_group_s* g) {
  int k, rc;
  int64_t i;
  bmodel* mods = (bmodel*)calloc(g->nmodels, sizeof(bmodel));
  binst* in = (binst*)calloc(g->n > 0 ? g->n : 1, sizeof(binst));
  g->mdata = mods;
  g->inst = in;
  for (k = 0; k < g->nmodels; ++k) {
    if (g->models[k].level != 1) return or_fail(XA_ERR_UNSUPPORTED, "bjt: level must be 1");
    if ((rc = model_init(&mods[k], &g->models[k], &g->opt)) != XA_OK) return rc;
  }
  for (i = 0; i < g->n; ++i) {
    binst* I = &in[i];
    bmodel* m = &mods[g->inst_model[i]];
    I->model = m;
    I->AREA = g->ival[Q_AREA][i];
    I->icVBE = g->ival[Q_IC1][i];
    I->icVCE = g->ival[Q_IC2][i];
    I->IC_GIVEN = g->igiven[Q_IC1] || g->igiven[Q_IC2];
    I->TEMP = g->igiven[Q_TEMP] ? g->ival[Q_TEMP][i] : g->opt.temp;
    I->OFF = g->ival[Q_OFF][i] != 0.0;
    I->lambertWFlag = g->ival[Q_LAMBERTW][i] != 0.0;
    I->loadLeadCurrent = g->lead[i];
    g->n_ext[i] = 4;
    g->n_int[i] = (m->collectorResist == 0.0 ? 0 : 1) + (m->emitterResist == 0.0 ? 0 : 1) +
                  (m->baseResist == 0.0 ? 0 : 1);
    /* the excess-phase current and its derivative (:1032-1036) */
    if (g->opt.newExcessPhase) g->n_int[i] += 2;
    g->n_sta[i] = 7;
    g->n_sto[i] = 3 + (I->loadLeadCurrent ? 4 : 0);
    update_temperature(I, I->TEMP);   /* processParams -> updateTemperature(TEMP) */
  }
  return XA_OK;
}

static void register_lids(struct or_group_s* g, int64_t i, const int32_t* nodes,
                          const int32_t* sta, const int32_t* sto) {
  binst* I = &((binst*)g->inst)[i];
  const bmodel* m = I->model;
  int loc = 4;
  I->li_Coll = nodes[0];
  I->li_Base = nodes[1];
  I->li_Emit = nodes[2];
  I->li_Subst = nodes[3];
  I->li_CollP = (m->collectorResist == 0.0) ? I->li_Coll : nodes[loc++];
  I->li_BaseP = (m->baseResist == 0.0) ? I->li_Base : nodes[loc++];
  I->li_EmitP = (m->emitterResist == 0.0) ? I->li_Emit : nodes[loc++];
  /* :1149-1153 */
  I->li_Ifx = g->opt.newExcessPhase ? nodes[loc++] : -1;
  I->li_dIfx = g->opt.newExcessPhase ? nodes[loc++] : -1;
  I->li_qstateBEdiff = sta ? sta[0] : -1;
  I->li_qstateBEdep = sta ? sta[1] : -1;
  I->li_qstateCS = sta ? sta[2] : -1;
  I->li_qstateBCdiff = sta ? sta[3] : -1;
  I->li_qstateBCdep = sta ? sta[4] : -1;
  I->li_qstateBX = sta ? sta[5] : -1;
  I->li_istateCEXBC = sta ? sta[6] : -1;
  I->li_storevBE = sto[0];
  I->li_storevBC = sto[1];
  I->li_store_capeqCB = sto[2];
  I->li_store_dev_ib = I->loadLeadCurrent ? sto[3] : -1;
  I->li_store_dev_ie = I->loadLeadCurrent ? sto[4] : -1;
  I->li_store_dev_ic = I->loadLeadCurrent ? sto[5] : -1;
  I->li_store_dev_is = I->loadLeadCurrent ? sto[6] : -1;
}

/* jacStamp_RB_RC_RE_ (:867-906) with the primed nodes folded into their
 * terminals when the resistance is zero (jacStampMap); with the new excess
 * phase (:908-931) the C' and E' rows gain an Ifx column and the Ifx and
 * dIfx rows follow */
static int64_t jac_pairs(struct or_group_s* g, int64_t i, int32_t* r, int32_t* c) {
  const binst* I = &((binst*)g->inst)[i];
  const int C = I->li_Coll, B = I->li_Base, E = I->li_Emit, S = I->li_Subst;
  const int Cp = I->li_CollP, Bp = I->li_BaseP, Ep = I->li_EmitP;
  const int X = I->li_Ifx, D = I->li_dIfx;
  const int nep = g->opt.newExcessPhase != 0;
  const int32_t pr0[24] = {C, C, B, B, B, B, E, E, S, S, Cp, Cp, Cp, Cp, Cp, Cp,
                           Bp, Bp, Bp, Bp, Ep, Ep, Ep, Ep};
  const int32_t pc0[24] = {C, Cp, B, Cp, Bp, Ep, E, Ep, S, Cp, C, B, S, Cp, Bp, Ep,
                           B, Cp, Bp, Ep, E, Cp, Bp, Ep};
  const int32_t pr1[36] = {C, C, B, B, B, B, E, E, S, S, Cp, Cp, Cp, Cp, Cp, Cp, Cp,
                           Bp, Bp, Bp, Bp, Ep, Ep, Ep, Ep, Ep,
                           X, X, X, X, X, D, D, D, D, D};
  const int32_t pc1[36] = {C, Cp, B, Cp, Bp, Ep, E, Ep, S, Cp, C, B, S, Cp, Bp, Ep, X,
                           B, Cp, Bp, Ep, E, Cp, Bp, Ep, X,
                           Cp, Bp, Ep, X, D, Cp, Bp, Ep, X, D};
  const int32_t* pr = nep ? pr1 : pr0;
  const int32_t* pc = nep ? pc1 : pc0;
  const int np = nep ? 36 : 24;
  int64_t k, n = 0;
  for (k = 0; k < np; ++k) {
    if (pr[k] < 0 || pc[k] < 0) continue;
    if (r) { r[n] = pr[k]; c[n] = pc[k]; }
    ++n;
  }
  return n;
}

static int register_jac(struct or_group_s* g, int64_t i) {
  binst* I = &((binst*)g->inst)[i];
  const int C = I->li_Coll, B = I->li_Base, E = I->li_Emit, S = I->li_Subst;
  const int Cp = I->li_CollP, Bp = I->li_BaseP, Ep = I->li_EmitP;
  I->f_CC = or_slot(g, C, C); I->f_CCp = or_slot(g, C, Cp); I->f_BB = or_slot(g, B, B);
  I->f_BCp = or_slot(g, B, Cp); I->f_BBp = or_slot(g, B, Bp); I->f_BEp = or_slot(g, B, Ep);
  I->f_EE = or_slot(g, E, E); I->f_EEp = or_slot(g, E, Ep); I->f_CpC = or_slot(g, Cp, C);
  I->f_CpCp = or_slot(g, Cp, Cp); I->f_CpBp = or_slot(g, Cp, Bp); I->f_CpEp = or_slot(g, Cp, Ep);
  I->f_BpB = or_slot(g, Bp, B); I->f_BpCp = or_slot(g, Bp, Cp); I->f_BpBp = or_slot(g, Bp, Bp);
  I->f_BpEp = or_slot(g, Bp, Ep); I->f_EpE = or_slot(g, Ep, E); I->f_EpCp = or_slot(g, Ep, Cp);
  I->f_EpBp = or_slot(g, Ep, Bp); I->f_EpEp = or_slot(g, Ep, Ep);
  I->q_BB = or_slot(g, B, B); I->q_BCp = or_slot(g, B, Cp); I->q_SS = or_slot(g, S, S);
  I->q_SCp = or_slot(g, S, Cp); I->q_CpB = or_slot(g, Cp, B); I->q_CpS = or_slot(g, Cp, S);
  I->q_CpCp = or_slot(g, Cp, Cp); I->q_CpBp = or_slot(g, Cp, Bp); I->q_BpCp = or_slot(g, Bp, Cp);
  I->q_BpBp = or_slot(g, Bp, Bp); I->q_BpEp = or_slot(g, Bp, Ep); I->q_EpCp = or_slot(g, Ep, Cp);
  I->q_EpBp = or_slot(g, Ep, Bp); I->q_EpEp = or_slot(g, Ep, Ep);
  if (g->opt.newExcessPhase) {
    const int X = I->li_Ifx, D = I->li_dIfx;
    I->f_CpIfx = or_slot(g, Cp, X); I->f_EpIfx = or_slot(g, Ep, X);
    I->f_IfxdIfx = or_slot(g, X, D); I->f_dIfxCp = or_slot(g, D, Cp);
    I->f_dIfxBp = or_slot(g, D, Bp); I->f_dIfxEp = or_slot(g, D, Ep);
    I->f_dIfxdIfx = or_slot(g, D, D); I->f_dIfxIfx = or_slot(g, D, X);
    I->f_IfxCp = or_slot(g, X, Cp); I->f_IfxBp = or_slot(g, X, Bp);
    I->f_IfxEp = or_slot(g, X, Ep); I->f_IfxIfx = or_slot(g, X, X);
    I->q_IfxIfx = or_slot(g, X, X); I->q_dIfxdIfx = or_slot(g, D, D);
  } else {
    I->f_CpIfx = I->f_EpIfx = I->f_IfxdIfx = I->f_dIfxCp = I->f_dIfxBp = I->f_dIfxEp = -1;
    I->f_dIfxdIfx = I->f_dIfxIfx = I->f_IfxCp = I->f_IfxBp = I->f_IfxEp = I->f_IfxIfx = -1;
    I->q_IfxIfx = I->q_dIfxdIfx = -1;
  }
  {
    const int64_t* s = &I->f_CC;
    int k;
    for (k = 0; k < 48; ++k) if (s[k] < -1) return XA_ERR_ARG;
  }
  return XA_OK;
}

static double sol(const double* x, int lid) { return lid >= 0 ? x[lid] : 0.0; }

/* lambertWCurrent (:1784-1800) */
static void lambertw_current(double* Id, double* Gd, double Vd, double Vte, double Isat) {
  const double RS = 1.0e-12;
  double arg1 = (Vd + Isat * RS) / Vte, evd, w, xi;
  int ierr;
  arg1 = or_min(OR_MAX_EXP_ARG, arg1);
  evd = exp(arg1);
  or_lambertw(Isat * RS * evd / Vte, &w, &ierr, &xi);
  *Id = -Isat + Vte * (w) / RS;
  *Gd = w / ((1 + w) * RS);
}

/* Instance::updateIntermediateVars (:2476-3207) */
static int update_intermediate(struct or_group_s* g, binst* I, const xa_solver_state* ss,
                               const xa_vectors* v) {
  const bmodel* m = I->model;
  const double gmin = g->opt.gmin;
  const double* solVec = v->nextSol;
  double v_emit, v_emitP, v_base, v_baseP, v_coll, v_collP, v_subst, q1, q2;
  double csat, vtF, vtR, vtE, vtC, iLeakBE, iLeakBC, rofF, rofR;
  double arg, sarg, ctot, czBC, czBX, czBE, czCS, fcpc, fcpe;
  double iEX_tmp = 0.0, gEX_tmp = 0.0, iC_tmp = 0.0;
  v_emit = sol(solVec, I->li_Emit);
  v_emitP = sol(solVec, I->li_EmitP);
  v_base = sol(solVec, I->li_Base);
  v_baseP = sol(solVec, I->li_BaseP);
  v_coll = sol(solVec, I->li_Coll);
  v_collP = sol(solVec, I->li_CollP);
  v_subst = sol(solVec, I->li_Subst);
  I->vEEp = (v_emit - v_emitP);
  I->vBBp = (v_base - v_baseP);
  I->vCCp = (v_coll - v_collP);
  I->vBE = m->TYPE * (v_baseP - v_emitP);
  I->vBC = m->TYPE * (v_baseP - v_collP);
  I->vBX = m->TYPE * (v_base - v_collP);
  I->vCS = m->TYPE * (v_subst - v_collP);
  I->vBE_orig = I->vBE;
  I->vBC_orig = I->vBC;
  I->origFlag = 1;
  I->offFlag = 0;
  if (ss->initJctFlag && !I->OFF && g->opt.voltageLimiterFlag) {
    if (I->IC_GIVEN) {
      double vCE;
      I->vBE = m->TYPE * I->icVBE;
      vCE = m->TYPE * I->icVCE;
      I->vBC = I->vBE - vCE;
      I->vBX = I->vBC;
      I->origFlag = 0;
    } else if (ss->inputOPFlag) {
      const double* fl = v->flagSol;
      /* a ground node reads its flag as 0 */
      if (sol(fl, I->li_Emit) == 0 || sol(fl, I->li_EmitP) == 0 || sol(fl, I->li_Base) == 0 ||
          sol(fl, I->li_BaseP) == 0 || sol(fl, I->li_Coll) == 0 || sol(fl, I->li_CollP) == 0 ||
          sol(fl, I->li_Subst) == 0) {
        I->vBC = 0;
        I->vBE = I->tVCrit;
        I->vBX = I->vBC;
        I->origFlag = 0;
      }
    } else {
      I->vBE = I->tVCrit;
      I->vBX = I->vBC;
      I->origFlag = 0;
    }
  } else if ((ss->initFixFlag || ss->initJctFlag) && I->OFF) {
    I->vBX = I->vBC = I->vBE = 0.0;
    I->offFlag = 1;
  }
  if (ss->newtonIter == 0) {
    if (!(ss->dcopFlag) || (ss->locaEnabledFlag && ss->dcopFlag)) {
      I->vBE_old = v->currSto[I->li_storevBE];
      I->vBC_old = v->currSto[I->li_storevBC];
      I->capeqCB = v->currSto[I->li_store_capeqCB];
    } else {
      I->vBE_old = I->vBE;
      I->vBC_old = I->vBC;
      I->capeqCB = 0.0;
    }
  } else {
    I->vBE_old = v->nextSto[I->li_storevBE];
    I->vBC_old = v->nextSto[I->li_storevBC];
    I->capeqCB = v->nextSto[I->li_store_capeqCB];
  }
  if (g->opt.voltageLimiterFlag && !(ss->initFixFlag && I->OFF)) {
    int ichk1, icheck;
    if (ss->newtonIter >= 0) {
      ichk1 = 1;
      I->vBE = or_pnjlim(I->vBE, I->vBE_old, I->vt, I->tVCrit, &icheck);
      I->vBC = or_pnjlim(I->vBC, I->vBC_old, I->vt, I->tVCrit, &ichk1);
      if (ichk1 == 1) icheck = 1;
      if (icheck == 1) I->origFlag = 0;
    }
  }
  csat = I->tSatCur * I->AREA;
  vtF = I->vt * m->emissionCoeffF;
  vtR = I->vt * m->emissionCoeffR;
  vtE = I->vt * I->tleakBEEmissionCoeff;
  vtC = I->vt * I->tleakBCEmissionCoeff;
  iLeakBE = I->tBELeakCur * I->AREA;
  iLeakBC = I->tBCLeakCur * I->AREA;
  if (I->lambertWFlag) {
    if (I->vBE > -5.0 * vtF) {
      lambertw_current(&I->iBE, &I->gBE, I->vBE, vtF, csat);
      if (iLeakBE == 0.0) I->iBEleak = I->gBEleak = 0.0;
      else lambertw_current(&I->iBEleak, &I->gBEleak, I->vBE, vtE, iLeakBE);
    } else {
      I->gBE = -csat / I->vBE + gmin;
      I->iBE = I->gBE * I->vBE;
      I->gBEleak = -iLeakBE / I->vBE;
      I->iBEleak = I->gBEleak * I->vBE;
    }
    if (I->vBC > -5.0 * vtR) {
      lambertw_current(&I->iBC, &I->gBC, I->vBC, vtR, csat);
      if (iLeakBC == 0.0) I->iBCleak = I->gBCleak = 0.0;
      else lambertw_current(&I->iBCleak, &I->gBCleak, I->vBC, vtC, iLeakBC);
    } else {
      I->gBC = -csat / I->vBC + gmin;
      I->iBC = I->gBC * I->vBC;
      I->gBCleak = -iLeakBC / I->vBC;
      I->iBCleak = I->gBCleak * I->vBC;
    }
  } else {
    /* bjtArtParameterFlag takes Xexp, which is exp itself at the default
     * exp_order of 100 (N_DEV_DeviceSupport.C Xexp) */
    if (I->vBE > -5.0 * vtF) {
      double arg1 = or_min(OR_MAX_EXP_ARG, I->vBE / vtF);
      double evBE = exp(arg1), devBE = evBE;
      I->iBE = csat * (evBE - 1.0) + gmin * I->vBE;
      I->gBE = csat * devBE / vtF + gmin;
      if (iLeakBE == 0.0) {
        I->iBEleak = I->gBEleak = 0.0;
      } else {
        double a2 = or_min(OR_MAX_EXP_ARG, I->vBE / vtE);
        double ev = exp(a2), dev = ev;
        I->iBEleak = iLeakBE * (ev - 1.0);
        I->gBEleak = iLeakBE * dev / vtE;
      }
    } else {
      I->gBE = -csat / I->vBE + gmin;
      I->iBE = I->gBE * I->vBE;
      I->gBEleak = -iLeakBE / I->vBE;
      I->iBEleak = I->gBEleak * I->vBE;
    }
    if (I->vBC > -5.0 * vtR) {
      double arg1 = or_min(OR_MAX_EXP_ARG, I->vBC / vtR);
      double evBC = exp(arg1), devBC = evBC;
      I->iBC = csat * (evBC - 1.0) + gmin * I->vBC;
      I->gBC = csat * devBC / vtR + gmin;
      if (iLeakBC == 0.0) {
        I->iBCleak = I->gBCleak = 0.0;
      } else {
        double a2 = or_min(OR_MAX_EXP_ARG, I->vBC / vtC);
        double ev = exp(a2), dev = ev;
        I->iBCleak = iLeakBC * (ev - 1.0);
        I->gBCleak = iLeakBC * dev / vtC;
      }
    } else {
      I->gBC = -csat / I->vBC + gmin;
      I->iBC = I->gBC * I->vBC;
      I->gBCleak = -iLeakBC / I->vBC;
      I->iBCleak = I->gBCleak * I->vBC;
    }
  }
  rofF = I->tInvRollOffF / I->AREA;
  rofR = I->tInvRollOffR / I->AREA;
  q1 = 1.0 / (1.0 - I->tInvEarlyVoltF * I->vBC - I->tInvEarlyVoltR * I->vBE);
  if (rofF == 0.0 && rofR == 0.0) {
    double q1_qB;
    I->qB = q1;
    q1_qB = q1 * I->qB;
    I->dqBdvEp = q1_qB * I->tInvEarlyVoltR;
    I->dqBdvCp = q1_qB * I->tInvEarlyVoltF;
  } else {
    q2 = rofF * I->iBE + rofR * I->iBC;
    if (q2 >= 0) {
      double a = 1.0 + 4.0 * q2, sqarg = 1.0, fg = 0.0, rg = 0.0;
      if (fabs(a) > 0.0) sqarg = pow(a, I->tRollOffExp);
      if (a != 0) fg = rofF * I->gBE * 2 * I->tRollOffExp * sqarg / a;
      if (a != 0) rg = rofR * I->gBC * 2 * I->tRollOffExp * sqarg / a;
      I->qB = 0.5 * q1 * (1.0 + sqarg);
      I->dqBdvEp = q1 * (I->qB * I->tInvEarlyVoltR + fg);
      I->dqBdvCp = q1 * (I->qB * I->tInvEarlyVoltF + rg);
    } else {
      I->qB = q1;
      I->dqBdvEp = q1 * I->qB * I->tInvEarlyVoltR;
      I->dqBdvCp = q1 * I->qB * I->tInvEarlyVoltF;
    }
  }
  I->dqBdvBp = -(I->dqBdvEp + I->dqBdvCp);
  I->invqB = 1.0 / I->qB;
  ctot = I->tBCCap * I->AREA;
  czBC = ctot * m->baseFracBCCap;
  czBX = ctot - czBC;
  czBE = I->tBECap * I->AREA;
  czCS = m->CJS * I->AREA;
  fcpc = m->depCapCoeff * m->potBC;
  fcpe = I->tDepCap;
  I->iBEhighCurr = I->iBE;
  I->gBEhighCurr = I->gBE;
  if (m->transTimeF != 0.0 && I->vBE > 0.0 &&
      ((!ss->dcopFlag) || ss->tranopFlag || ss->acopFlag)) {
    double argtf = 0.0, arg2 = 0.0, arg3 = 0.0;
    if (m->transTimeBiasCoeffF != 0.0) {
      argtf = m->transTimeBiasCoeffF;
      if (m->transTimeVBCFac != 0.0) argtf *= exp(I->vBC * m->transTimeVBCFac);
      arg2 = argtf;
      if (m->transTimeHighCurrF != 0.0) {
        double tmp = I->iBEhighCurr / (I->iBEhighCurr + m->transTimeHighCurrF * I->AREA);
        argtf *= tmp * tmp;
        arg2 = argtf * (3.0 - 2.0 * tmp);
      }
      arg3 = I->iBEhighCurr * argtf * m->transTimeVBCFac;
    }
    I->iBEhighCurr *= (1.0 + argtf) / I->qB;
    I->gBEhighCurr = (I->gBEhighCurr * (1.0 + arg2) - I->iBEhighCurr * I->dqBdvEp) / I->qB;
    I->capeqCB = m->transTimeF * (arg3 - I->iBEhighCurr * I->dqBdvCp) / I->qB;
    I->geqCB = I->capeqCB * ss->pdt;
  }
  if (czBE == 0.0) {
    I->qBEdep = 0.0;
    I->capBEdep = 0.0;
  } else if (I->vBE < fcpe) {
    arg = 1.0 - I->vBE / I->tBEPot;
    sarg = exp(-m->juncExpBE * log(arg));
    I->qBEdep = I->tBEPot * czBE * (1.0 - arg * sarg) / (1.0 - m->juncExpBE);
    I->capBEdep = czBE * sarg;
  } else {
    double czBEf2 = czBE / m->f2;
    I->qBEdep = czBE * I->tF1 + czBEf2 * (m->f3 * (I->vBE - fcpe) +
                (m->juncExpBE / (2.0 * I->tBEPot)) * (I->vBE * I->vBE - fcpe * fcpe));
    I->capBEdep = czBEf2 * (m->f3 + m->juncExpBE * I->vBE / I->tBEPot);
  }
  if (m->transTimeF == 0.0) {
    I->qBEdiff = I->capBEdiff = 0.0;
  } else {
    I->qBEdiff = m->transTimeF * I->iBEhighCurr;
    I->capBEdiff = m->transTimeF * I->gBEhighCurr;
  }
  if (czBC == 0.0) {
    I->qBCdep = 0.0;
    I->capBCdep = 0.0;
  } else if (I->vBC < fcpc) {
    arg = 1.0 - I->vBC / I->tBCPot;
    sarg = exp(-m->juncExpBC * log(arg));
    I->qBCdep = I->tBCPot * czBC * (1.0 - arg * sarg) / (1.0 - m->juncExpBC);
    I->capBCdep = czBC * sarg;
  } else {
    double czBCf2 = czBC / m->f6;
    I->qBCdep = czBC * I->tF5 + czBCf2 * (m->f7 * (I->vBC - fcpc) +
                (m->juncExpBC / (2.0 * I->tBCPot)) * (I->vBC * I->vBC - fcpc * fcpc));
    I->capBCdep = czBCf2 * (m->f7 + m->juncExpBC * I->vBC / I->tBCPot);
  }
  if (m->transTimeR == 0.0) {
    I->qBCdiff = I->capBCdiff = 0.0;
  } else {
    I->qBCdiff = m->transTimeR * I->iBC;
    I->capBCdiff = m->transTimeR * I->gBC;
  }
  if (czBX == 0.0) {
    I->qBX = 0.0;
    I->capBX = 0.0;
  } else if (I->vBX < fcpc) {
    arg = 1.0 - I->vBX / I->tBCPot;
    sarg = exp(-m->juncExpBC * log(arg));
    I->qBX = I->tBCPot * czBX * (1.0 - arg * sarg) / (1.0 - m->juncExpBC);
    I->capBX = czBX * sarg;
  } else {
    double czBXf2 = czBX / m->f6;
    I->qBX = czBX * I->tF5 + czBXf2 * (m->f7 * (I->vBX - fcpc) +
             (m->juncExpBC / (2.0 * I->tBCPot)) * (I->vBX * I->vBX - fcpc * fcpc));
    I->capBX = czBXf2 * (m->f7 + m->juncExpBC * I->vBX / I->tBCPot);
  }
  if (czCS == 0.0) {
    I->qCS = 0.0;
    I->capCS = 0.0;
  } else if (I->vCS < 0.0) {
    arg = 1.0 - I->vCS / m->potSubst;
    sarg = exp(-m->expSubst * log(arg));
    I->qCS = m->potSubst * czCS * (1.0 - arg * sarg) / (1.0 - m->expSubst);
    I->capCS = czCS * sarg;
  } else {
    I->qCS = I->vCS * czCS * (1.0 + m->expSubst * I->vCS / (2.0 * m->potSubst));
    I->capCS = czCS * (1.0 + m->expSubst * I->vCS / m->potSubst);
  }
  /* oldDAEExcessPhaseCalculation1 (:2370-2389) */
  {
    const double td = m->excessPhaseFac;
    I->currCexbc = I->lastCexbc = 0.0;
    if (!(ss->dcopFlag) && td != 0 && ss->beginIntegrationFlag) {
      I->currCexbc = I->lastCexbc = I->iBE / I->qB;
      v->currSta[I->li_istateCEXBC] = I->currCexbc;
      v->lastSta[I->li_istateCEXBC] = I->lastCexbc;
    }
  }
  /* oldDAEExcessPhaseCalculation2 (:2406-2466) */
  {
    const double td = m->excessPhaseFac;
    const double dt0 = ss->currTimeStep, dt1 = ss->lastTimeStep;
    iEX_tmp = I->iBE;
    gEX_tmp = I->gBE;
    iC_tmp = 0.0;
    if (!(ss->dcopFlag) && td != 0) {
      double arg1 = dt0 / td, arg2 = 3.0 * arg1, denom;
      arg1 = arg2 * arg1;
      denom = 1.0 + arg1 + arg2;
      I->phaseScalar = arg1 / denom;
      if (!ss->beginIntegrationFlag) {
        I->currCexbc = v->currSta[I->li_istateCEXBC];
        I->lastCexbc = v->lastSta[I->li_istateCEXBC];
      }
      iC_tmp = ((I->currCexbc) * (1 + dt0 / dt1 + arg2) - (I->lastCexbc) * dt0 / dt1) / denom;
      iEX_tmp = I->iBE * I->phaseScalar;
      gEX_tmp = I->gBE * I->phaseScalar;
      I->nextCexbc = iC_tmp + iEX_tmp / I->qB;
      v->nextSta[I->li_istateCEXBC] = I->nextCexbc;
    }
  }
  if (g->opt.newExcessPhase) {
    /* auxDAECalculations (:1928-1980): the currents from the excess-phase
     * unknown i_fx, which the new-DAE form carries in the solution */
    const double td = m->excessPhaseFac;
    if (td != 0 && !(ss->dcopFlag)) {
      const double i_fx = sol(solVec, I->li_Ifx);
      I->iCE = i_fx - I->iBC / I->qB;
    } else {
      I->iCE = (I->iBE - I->iBC) / I->qB;
    }
    I->iC = I->iCE - I->iBC / I->tBetaR - I->iBCleak;
    I->iB = I->iBE / I->tBetaF + I->iBEleak + I->iBC / I->tBetaR + I->iBCleak;
    I->iE = -I->iC - I->iB;
    if (td != 0) {
      if (!(ss->dcopFlag)) {
        I->diCEdvBp = I->invqB * (-I->invqB * I->iBC * I->dqBdvBp - I->gBC);
        I->diCEdvCp = I->invqB * (-I->invqB * I->iBC * I->dqBdvCp + I->gBC);
        I->diCEdvEp = -I->invqB * I->invqB * I->iBC * I->dqBdvEp;
      } else {
        I->diCEdvBp = I->invqB * (I->iCE * I->dqBdvBp + I->gBE - I->gBC);
        I->diCEdvCp = I->invqB * (I->iCE * I->dqBdvCp + I->gBC);
        I->diCEdvEp = I->invqB * (I->iCE * I->dqBdvEp - I->gBE);
      }
      I->diBEdvBp = I->invqB * (I->invqB * I->iBE * I->dqBdvBp + I->gBE);
      I->diBEdvCp = I->invqB * I->invqB * I->iBE * I->dqBdvCp;
      I->diBEdvEp = I->invqB * (I->invqB * I->iBE * I->dqBdvEp - I->gBE);
    } else {
      I->diCEdvBp = I->invqB * (I->iCE * I->dqBdvBp + I->gBE - I->gBC);
      I->diCEdvCp = I->invqB * (I->iCE * I->dqBdvCp + I->gBC);
      I->diCEdvEp = I->invqB * (I->iCE * I->dqBdvEp - I->gBE);
      I->diBEdvBp = I->diBEdvCp = I->diBEdvEp = 0.0;
    }
  } else {
  if (!(ss->dcopFlag)) {
    double pdt = ss->pdt;
    if (ss->initTranFlag && ss->newtonIter == 0) pdt = 0.0;
    I->gCapBEdiff = I->capBEdiff * pdt;
    I->gCapBEdep = I->capBEdep * pdt;
    I->gCapBCdiff = I->capBCdiff * pdt;
    I->gCapBCdep = I->capBCdep * pdt;
    I->gCapBX = I->capBX * pdt;
    I->gCapCS = I->capCS * pdt;
  } else {
    I->gCapBEdiff = I->gCapBEdep = I->gCapBCdiff = I->gCapBCdep = I->gCapBX = I->gCapCS = 0.0;
  }
  {
    double iEX, gEX;
    I->iC = iC_tmp;
    iEX = iEX_tmp;
    gEX = gEX_tmp;
    I->iCE = (iEX - I->iBC) / I->qB;
    I->iC += I->iCE - I->iBC / I->tBetaR - I->iBCleak;
    I->iB = I->iBE / I->tBetaF + I->iBEleak + I->iBC / I->tBetaR + I->iBCleak;
    I->iE = -I->iC - I->iB;
    I->diCEdvEp = I->invqB * (I->iCE * I->dqBdvEp - gEX);
    I->diCEdvCp = I->invqB * (I->iCE * I->dqBdvCp + I->gBC);
    I->diCEdvBp = I->invqB * (I->iCE * I->dqBdvBp + gEX - I->gBC);
  }
  }
  I->gEpr = m->emitterConduct * I->AREA;
  I->gCpr = m->collectorConduct * I->AREA;
  {
    double rBpr = m->minBaseResist / I->AREA;
    double rBpi = m->baseResist / I->AREA - rBpr;
    double xjrB = m->baseCurrHalfResist * I->AREA;
    I->gX = rBpr + rBpi / I->qB;
    if (fabs(xjrB) > 0.0) {
      double arg1 = or_max(I->iB / xjrB, 1.0e-09);
      double arg2 = (-1.0 + sqrt(1.0 + 14.59025 * arg1)) / 2.4317 / sqrt(arg1);
      arg1 = tan(arg2);
      I->gX = rBpr + 3.0 * rBpi * (arg1 - arg2) / arg2 / arg1 / arg1;
    }
    if (fabs(I->gX) > 0.0) I->gX = 1.0 / I->gX;
  }
  I->diBrdvB = I->gX;
  I->diBrdvCp = 0.0;
  I->diBrdvEp = 0.0;
  I->diBrdvBp = -I->gX;
  I->gBEtot = I->gBE / I->tBetaF + I->gBEleak;
  I->gBCtot = I->gBC / I->tBetaR + I->gBCleak;
  return 1;
}

/* Master::updateState (:3891-3937) */
static int update_state(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  binst* in = (binst*)g->inst;
  int ok = 1;
  /* instances are independent here; one thread is the serial order */
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1) reduction(&& : ok)
  for (i = 0; i < g->n; ++i) {
    binst* bi = &in[i];
    double* sto = v->nextSto;
    double* sta = v->nextSta;
    ok = update_intermediate(g, bi, ss, v) && ok;
    sto[bi->li_storevBE] = bi->vBE;
    sto[bi->li_storevBC] = bi->vBC;
    sto[bi->li_store_capeqCB] = bi->capeqCB;
    sta[bi->li_qstateBEdiff] = bi->qBEdiff;
    sta[bi->li_qstateBEdep] = bi->qBEdep;
    sta[bi->li_qstateBCdiff] = bi->qBCdiff;
    sta[bi->li_qstateBCdep] = bi->qBCdep;
    sta[bi->li_qstateBX] = bi->qBX;
    sta[bi->li_qstateCS] = bi->qCS;
    if (!(ss->dcopFlag) && (ss->initTranFlag) && ss->newtonIter == 0) {
      double* cs = v->currSta;
      cs[bi->li_qstateBEdiff] = bi->qBEdiff;
      cs[bi->li_qstateBEdep] = bi->qBEdep;
      cs[bi->li_qstateBCdiff] = bi->qBCdiff;
      cs[bi->li_qstateBCdep] = bi->qBCdep;
      cs[bi->li_qstateBX] = bi->qBX;
      cs[bi->li_qstateCS] = bi->qCS;
    }
  }
  return ok ? XA_OK : XA_ERR_DEVICE;
}

/* Master::updateSecondaryState (:3947-3968) */
static int update_secondary(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  binst* in = (binst*)g->inst;
  const double* d = v->nextStaDeriv;
  (void)ss;
  if (!d) return XA_OK;
  for (i = 0; i < g->n; ++i) {
    binst* bi = &in[i];
    bi->iBEdiff = d[bi->li_qstateBEdiff];
    bi->iBEdep = d[bi->li_qstateBEdep];
    bi->iCS = d[bi->li_qstateCS];
    bi->iBCdiff = d[bi->li_qstateBCdiff];
    bi->iBCdep = d[bi->li_qstateBCdep];
    bi->iBX = d[bi->li_qstateBX];
  }
  return XA_OK;
}

/* Master::loadDAEVectors (:3978-4142) */
static int load_vectors(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  binst* in = (binst*)g->inst;
  (void)ss;
  for (i = 0; i < g->n; ++i) {
    binst* bi = &in[i];
    const int T = bi->model->TYPE;
    const double vbe_diff = bi->vBE - bi->vBE_orig;
    const double vbc_diff = bi->vBC - bi->vBC_orig;
    const double vce_diff = vbe_diff - vbc_diff;
    if (v->f) {
      OR_SUB(v->f, bi->li_Coll, -bi->vCCp * bi->gCpr);
      OR_SUB(v->f, bi->li_Base, -bi->vBBp * bi->gX);
      OR_SUB(v->f, bi->li_Emit, -bi->vEEp * bi->gEpr);
      OR_SUB(v->f, bi->li_CollP, bi->vCCp * bi->gCpr + T * (-bi->iC));
      OR_SUB(v->f, bi->li_BaseP, bi->vBBp * bi->gX - T * (bi->iB));
      OR_SUB(v->f, bi->li_EmitP, bi->vEEp * bi->gEpr + T * (-bi->iE));
      if (g->opt.newExcessPhase) {
        /* excess phase rows (:3997-4024); the dc operating point's
         * assignment of the dIfx entry is an addition of zero onto the
         * zeroed vector (no other device writes that row) */
        const double td = bi->model->excessPhaseFac;
        const double i_fx = sol(v->nextSol, bi->li_Ifx), di_fx = sol(v->nextSol, bi->li_dIfx);
        if (td != 0) {
          if (!(ss->dcopFlag)) {
            OR_ADD(v->f, bi->li_Ifx, -di_fx);
            OR_ADD(v->f, bi->li_dIfx, 3 * di_fx * td + 3 * i_fx - 3 * bi->iBE / bi->qB);
          } else {
            OR_ADD(v->f, bi->li_Ifx, i_fx - bi->iBE / bi->qB);
            OR_ADD(v->f, bi->li_dIfx, 0.0);
          }
        } else {
          OR_ADD(v->f, bi->li_Ifx, i_fx);
          OR_ADD(v->f, bi->li_dIfx, di_fx);
        }
      }
    }
    if (v->q) {
      OR_SUB(v->q, bi->li_Base, -T * bi->qBX);
      OR_SUB(v->q, bi->li_Subst, -T * bi->qCS);
      OR_SUB(v->q, bi->li_CollP, T * (bi->qCS + bi->qBX + bi->qBCdep + bi->qBCdiff));
      OR_SUB(v->q, bi->li_BaseP, -T * (bi->qBEdep + bi->qBEdiff + bi->qBCdep + bi->qBCdiff));
      OR_SUB(v->q, bi->li_EmitP, T * (bi->qBEdep + bi->qBEdiff));
      if (g->opt.newExcessPhase && bi->model->excessPhaseFac != 0) {
        /* :4033-4046 */
        const double td = bi->model->excessPhaseFac;
        OR_ADD(v->q, bi->li_Ifx, sol(v->nextSol, bi->li_Ifx));
        OR_ADD(v->q, bi->li_dIfx, !(ss->dcopFlag) ? sol(v->nextSol, bi->li_dIfx) * td * td : 0.0);
      }
    }
    if (g->opt.voltageLimiterFlag) {
      double Cp_f = 0.0, Bp_f = 0.0, Ep_f = 0.0, Cp_q = 0.0, Ep_q = 0.0, Bp_q = 0.0;
      double dIfx_f = 0.0, Ifx_f = 0.0;
      const int nep = g->opt.newExcessPhase && bi->model->excessPhaseFac != 0;
      if (nep && (!bi->origFlag || bi->offFlag)) {
        /* :4069-4081 */
        if (or_mag) {
          const double t = fabs(bi->diBEdvBp * vbe_diff) + fabs(bi->diBEdvCp * vce_diff);
          if (!(ss->dcopFlag)) dIfx_f = 3 * t; else Ifx_f = t;
        } else if (!(ss->dcopFlag)) {
          dIfx_f = -3 * (bi->diBEdvBp * vbe_diff + bi->diBEdvCp * vce_diff);
          dIfx_f *= T;
        } else {
          Ifx_f = (bi->diBEdvBp * vbe_diff + bi->diBEdvCp * vce_diff);
          Ifx_f *= T;
        }
      }
      if (or_mag && (!bi->origFlag || bi->offFlag)) {
        /* magnitude mode: the rounding scale of each correction is the sum
         * of its terms' magnitudes (they cancel near convergence) */
        Cp_f = fabs(bi->diCEdvBp * vbe_diff) + fabs(bi->diCEdvCp * vce_diff) +
               fabs(bi->gBCtot * vbc_diff);
        Bp_f = fabs(bi->gBEtot * vbe_diff) + fabs(bi->gBCtot * vbc_diff);
        Ep_f = fabs(bi->diCEdvCp * vce_diff) + fabs((bi->diCEdvBp + bi->gBEtot) * vbe_diff);
      } else if (!bi->origFlag || bi->offFlag) {
        Cp_f = +bi->diCEdvBp * vbe_diff + bi->diCEdvCp * vce_diff - bi->gBCtot * vbc_diff;
        Cp_f *= T;
        Bp_f = bi->gBEtot * vbe_diff + bi->gBCtot * vbc_diff;
        Bp_f *= T;
        Ep_f = -bi->diCEdvCp * vce_diff - (bi->diCEdvBp + bi->gBEtot) * vbe_diff;
        Ep_f *= T;
      }
      if (v->dFdxdVp) {
        OR_ADD(v->dFdxdVp, bi->li_CollP, Cp_f);
        OR_ADD(v->dFdxdVp, bi->li_BaseP, Bp_f);
        OR_ADD(v->dFdxdVp, bi->li_EmitP, Ep_f);
        if (nep) {
          if (!(ss->dcopFlag)) OR_ADD(v->dFdxdVp, bi->li_dIfx, dIfx_f);
          else OR_ADD(v->dFdxdVp, bi->li_Ifx, Ifx_f);
        }
      }
      if (or_mag && (!bi->origFlag || bi->offFlag)) {
        Cp_q = fabs((bi->capBCdep + bi->capBCdiff) * vbc_diff);
        Bp_q = fabs((bi->capBEdep + bi->capBEdiff) * vbe_diff) +
               fabs((bi->capBCdiff + bi->capBCdep + bi->capeqCB) * vbc_diff);
        Ep_q = fabs(bi->capeqCB * vbc_diff) + fabs((bi->capBEdiff + bi->capBEdep) * vbe_diff);
      } else if (!bi->origFlag || bi->offFlag) {
        Cp_q = -(bi->capBCdep + bi->capBCdiff) * vbc_diff;
        Cp_q *= T;
        Bp_q = (bi->capBEdep + bi->capBEdiff) * vbe_diff +
               (bi->capBCdiff + bi->capBCdep + bi->capeqCB) * vbc_diff;
        Bp_q *= T;
        Ep_q = -bi->capeqCB * vbc_diff - (bi->capBEdiff + bi->capBEdep) * vbe_diff;
        Ep_q *= T;
      }
      if (v->dQdxdVp) {
        OR_ADD(v->dQdxdVp, bi->li_CollP, Cp_q);
        OR_ADD(v->dQdxdVp, bi->li_BaseP, Bp_q);
        OR_ADD(v->dQdxdVp, bi->li_EmitP, Ep_q);
      }
    }
    if (bi->loadLeadCurrent && v->storeLeadF && v->storeLeadQ) {
      double* LF = v->storeLeadF;
      double* LQ = v->storeLeadQ;
      OR_SET(LQ, bi->li_store_dev_ic, -T * (bi->qCS + bi->qBX + bi->qBCdep + bi->qBCdiff));
      OR_SET(LQ, bi->li_store_dev_ib, T * (bi->qBX + bi->qBEdep + bi->qBEdiff + bi->qBCdep + bi->qBCdiff));
      OR_SET(LQ, bi->li_store_dev_ie, -T * (bi->qBEdep + bi->qBEdiff));
      OR_SET(LQ, bi->li_store_dev_is, T * bi->qCS);
      OR_SET(LF, bi->li_store_dev_ic, T * (bi->iC));
      OR_SET(LF, bi->li_store_dev_is, 0);
      OR_SET(LF, bi->li_store_dev_ie, T * (bi->iE));
      OR_SET(LF, bi->li_store_dev_ib, T * (bi->iB));
    }
  }
  return XA_OK;
}

/* Master::loadDAEMatrices (:4152-4260) */
static int load_matrices(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  binst* in = (binst*)g->inst;
  double* F = v->dFdx;
  double* Q = v->dQdx;
  for (i = 0; i < g->n; ++i) {
    binst* bi = &in[i];
    const int T = bi->model->TYPE;
    const double td = bi->model->excessPhaseFac;
    const int nep = g->opt.newExcessPhase && td != 0;
    if (F) {
      OR_ADD(F, bi->f_CC, bi->gCpr);
      OR_SUB(F, bi->f_CCp, bi->gCpr);
      OR_ADD(F, bi->f_BB, bi->diBrdvB);
      OR_ADD(F, bi->f_BCp, bi->diBrdvCp);
      OR_ADD(F, bi->f_BBp, bi->diBrdvBp);
      OR_ADD(F, bi->f_BEp, bi->diBrdvEp);
      OR_ADD(F, bi->f_EE, bi->gEpr);
      OR_SUB(F, bi->f_EEp, bi->gEpr);
      OR_SUB(F, bi->f_CpC, bi->gCpr);
      OR_ADD(F, bi->f_CpCp, bi->diCEdvCp + bi->gBCtot + bi->gCpr);
      OR_ADD(F, bi->f_CpBp, bi->diCEdvBp - bi->gBCtot);
      OR_ADD(F, bi->f_CpEp, bi->diCEdvEp);
      if (nep && !(ss->dcopFlag)) OR_ADD(F, bi->f_CpIfx, 1.0 * T);
      OR_SUB(F, bi->f_BpB, bi->diBrdvB);
      OR_ADD(F, bi->f_BpCp, -bi->diBrdvCp - bi->gBCtot);
      OR_ADD(F, bi->f_BpBp, -bi->diBrdvBp + bi->gBEtot + bi->gBCtot);
      OR_ADD(F, bi->f_BpEp, -bi->diBrdvEp - bi->gBEtot);
      OR_SUB(F, bi->f_EpE, bi->gEpr);
      OR_ADD(F, bi->f_EpCp, -bi->diCEdvCp);
      OR_ADD(F, bi->f_EpBp, -bi->diCEdvBp - bi->gBEtot);
      OR_ADD(F, bi->f_EpEp, bi->gBEtot + bi->gEpr + bi->diCEdvBp + bi->diCEdvCp);
      if (nep && !(ss->dcopFlag)) OR_ADD(F, bi->f_EpIfx, -1.0 * T);
      if (g->opt.newExcessPhase) {
        /* :4196-4233 */
        if (td != 0) {
          if (!(ss->dcopFlag)) {
            OR_ADD(F, bi->f_IfxdIfx, -1);
            OR_ADD(F, bi->f_dIfxCp, -3 * bi->diBEdvCp * T);
            OR_ADD(F, bi->f_dIfxBp, -3 * bi->diBEdvBp * T);
            OR_ADD(F, bi->f_dIfxEp, -3 * bi->diBEdvEp * T);
            OR_ADD(F, bi->f_dIfxdIfx, 3 * td);
            OR_ADD(F, bi->f_dIfxIfx, 3);
          } else {
            OR_ADD(F, bi->f_IfxCp, -bi->diBEdvCp * T);
            OR_ADD(F, bi->f_IfxBp, -bi->diBEdvBp * T);
            OR_ADD(F, bi->f_IfxEp, -bi->diBEdvEp * T);
            OR_ADD(F, bi->f_IfxIfx, 1);
            OR_ADD(F, bi->f_dIfxdIfx, 1);
          }
        } else {
          OR_ADD(F, bi->f_IfxIfx, 1);
          OR_ADD(F, bi->f_dIfxdIfx, 1);
        }
      }
    }
    if (Q) {
      OR_ADD(Q, bi->q_BB, bi->capBX);
      OR_ADD(Q, bi->q_BCp, -bi->capBX);
      OR_ADD(Q, bi->q_SS, bi->capCS);
      OR_SUB(Q, bi->q_SCp, bi->capCS);
      OR_SUB(Q, bi->q_CpB, bi->capBX);
      OR_SUB(Q, bi->q_CpS, bi->capCS);
      OR_ADD(Q, bi->q_CpCp, bi->capCS + bi->capBX + bi->capBCdep + bi->capBCdiff);
      OR_ADD(Q, bi->q_CpBp, -bi->capBCdep - bi->capBCdiff);
      OR_ADD(Q, bi->q_BpCp, -bi->capBCdiff - bi->capBCdep - bi->capeqCB);
      OR_ADD(Q, bi->q_BpBp, bi->capBEdiff + bi->capBEdep + bi->capBCdiff + bi->capBCdep + bi->capeqCB);
      OR_ADD(Q, bi->q_BpEp, -bi->capBEdiff - bi->capBEdep);
      OR_ADD(Q, bi->q_EpCp, bi->capeqCB);
      OR_ADD(Q, bi->q_EpBp, -bi->capBEdiff - bi->capBEdep - bi->capeqCB);
      OR_ADD(Q, bi->q_EpEp, bi->capBEdiff + bi->capBEdep);
      if (nep && !(ss->dcopFlag)) {
        /* :4253-4260 */
        OR_ADD(Q, bi->q_IfxIfx, 1);
        OR_ADD(Q, bi->q_dIfxdIfx, 1 * td * td);
      }
    }
  }
  return XA_OK;
}

static int read_var(struct or_group_s* g, const char* name, double* out) {
  int64_t i;
  binst* in = (binst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    const binst* I = &in[i];
    if (!strcmp(name, "iC")) out[i] = I->iC;
    else if (!strcmp(name, "iB")) out[i] = I->iB;
    else if (!strcmp(name, "iE")) out[i] = I->iE;
    else if (!strcmp(name, "qB")) out[i] = I->qB;
    else if (!strcmp(name, "iBE")) out[i] = I->iBE;
    else if (!strcmp(name, "iBC")) out[i] = I->iBC;
    else if (!strcmp(name, "vBE")) out[i] = I->vBE;
    else if (!strcmp(name, "vBC")) out[i] = I->vBC;
    else if (!strcmp(name, "gX")) out[i] = I->gX;
    else if (!strcmp(name, "capeqCB")) out[i] = I->capeqCB;
    else if (!strcmp(name, "origFlag")) out[i] = I->origFlag;
    else if (!strcmp(name, "tSatCur")) out[i] = I->tSatCur;
    else if (!strcmp(name, "tBetaF")) out[i] = I->tBetaF;
    else if (!strcmp(name, "tBECap")) out[i] = I->tBECap;
    else if (!strcmp(name, "tBEPot")) out[i] = I->tBEPot;
    else if (!strcmp(name, "tBCCap")) out[i] = I->tBCCap;
    else if (!strcmp(name, "tBCPot")) out[i] = I->tBCPot;
    else return or_fail(XA_ERR_ARG, "bjt: unknown intermediate");
  }
  return XA_OK;
}

const or_model_ops or_bjt_ops = {
  "bjt", mpar, sizeof(mpar) / sizeof(mpar[0]), NMB_, ipar, sizeof(ipar) / sizeof(ipar[0]), NIB_,
  setup, register_lids, jac_pairs, register_jac, update_state, update_secondary, load_vectors,
  load_matrices, read_var, NULL, sizeof(binst),
};
