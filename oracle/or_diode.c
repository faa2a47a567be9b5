/* or_diode.c -- oracle restatement of the Xyce junction diode (test
 * infrastructure only).
 * Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_Diode.C
 *   parameter tables :70-318; Instance ctor :338-456; registerLIDs :479-500;
 *   registerStoreLIDs :571-590; registerJacLIDs :640-687 (via slot search);
 *   updateIntermediateVars :917-1366; updateTemperature :1376-1593;
 *   lambertWCurrent :1604-1657; lambertWLinearReverseBias :1668-1718;
 *   lambertWBreakdownCurrent :1729-1802; Model::processParams :1816-1839;
 *   Model ctor :1870-1940; Master::updateState :2051-2068;
 *   Master::loadDAEVectors :2078-2122; Master::loadDAEMatrices :2132-2171.
 */
#include "or_internal.h"
#include "or_eval_libm.h"

enum { IS_, RS_, N_, ISR_, NR_, IKF_, TT_, CJO_, VJ_, M_, EG_, XTI_, TIKF_, TBV1_, TBV2_,
       TRS1_, TRS2_, FC_, BV_, IBV_, IRF_, NBV_, IBVL_, NBVL_, TNOM_, KF_, AF_, NMS_ };
static const or_param mpar[] = {
  {"IS", IS_, 1.0e-14, 0}, {"JS", IS_, 1.0e-14, 0}, {"RS", RS_, 0.0, 0}, {"N", N_, 1.0, 0},
  {"ISR", ISR_, 0.0, 0}, {"NR", NR_, 2.0, 0}, {"IKF", IKF_, 0.0, 0}, {"TT", TT_, 0.0, 0},
  {"CJO", CJO_, 0.0, 0}, {"CJ", CJO_, 0.0, 0}, {"CJ0", CJO_, 0.0, 0}, {"VJ", VJ_, 1.0, 0},
  {"M", M_, 0.5, 0}, {"EG", EG_, 1.11, 0}, {"XTI", XTI_, 3.0, 0}, {"TIKF", TIKF_, 0.0, 0},
  {"TBV1", TBV1_, 0.0, 0}, {"TBV2", TBV2_, 0.0, 0}, {"TRS1", TRS1_, 0.0, 0},
  {"TRS2", TRS2_, 0.0, 0}, {"FC", FC_, 0.5, 0}, {"BV", BV_, 1e99, 0}, {"VB", BV_, 1e99, 0},
  {"IBV", IBV_, 1.0e-3, 0}, {"IRF", IRF_, 1.0, 0}, {"NBV", NBV_, 1.0, 0},
  {"IBVL", IBVL_, 0.0, 0}, {"NBVL", NBVL_, 1.0, 0}, {"TNOM", TNOM_, 0.0, 1},
  {"KF", KF_, 0.0, 0}, {"AF", AF_, 1.0, 0},
};
enum { AREA_, IC_, TEMP_, OFF_, LAMBERTW_, NIS_ };
static const or_param ipar[] = {
  {"AREA", AREA_, 1.0, 0}, {"IC", IC_, 0.0, 0}, {"TEMP", TEMP_, 0.0, 1},
  {"OFF", OFF_, 0.0, 0}, {"LAMBERTW", LAMBERTW_, 0.0, 0},
};

typedef struct {
  double IS, RS, COND, N, ISR, NR, IKF, TT, CJO, VJ, M, EG, XTI, TIKF, TBV1, TBV2, TRS1, TRS2;
  double FC, BV, IBV, IRF, NBV, IBVL, NBVL, F2, F3, TNOM, KF, AF;
  int BVGiven, level;
} dmodel;

typedef struct {
  dmodel* model;
  int off; double Area, InitCond, Temp; int lambertWFlag, InitCondGiven;
  double tJctPot, tJctCap, tDepCap, tSatCur, tVcrit, tF1, tBrkdwnV, tSatCurR, tIKF, tRS, tCOND, tIRF;
  double Id, Gd, Cd, Gcd, Qd, Icd, Gspr, Vpp, Vp, Vn, Vc, Vd, Vd_old, Vd_orig;
  int newtonIterOld, origFlag, loadLeadCurrent;
  int li_Pos, li_Neg, li_Pri, li_storevd, li_store_dev_i;
  int64_t fPosPos, fPosPri, fNegNeg, fNegPri, fPriPos, fPriNeg, fPriPri;
  int64_t qNegNeg, qNegPri, qPriNeg, qPriPri;
} dinst;

/* Model::processParams, N_DEV_Diode.C:1816-1839 */
static void model_process(dmodel* m) {
  double xfc;
  if (m->M > 0.9) m->M = 0.9;
  if (m->EG < 0.1) m->EG = 0.1;
  if (m->FC > 0.95) m->FC = 0.95;
  if (m->RS == 0.0) m->COND = 0.0;
  else m->COND = 1.0 / m->RS;
  xfc = log(1.0 - m->FC);
  m->F2 = exp((1.0 + m->M) * xfc);
  m->F3 = 1.0 - m->FC * (1.0 + m->M);
}

/* Instance::updateTemperature, N_DEV_Diode.C:1376-1593 (default branch
 * without Xyce_BREAKDOWN_ORIGINAL) */
static void update_temperature(dinst* d, double temp) {
  const dmodel* m = d->model;
  double vtnom = OR_KoverQ * m->TNOM;
  double xfc = log(1.0 - m->FC);
  double TNOM, vt, fact2, egfet, arg, pbfact, egfet1, arg1, fact1, pbfact1, pbo, gmaold, gmanew;
  double vte, tempBV;
  if (temp != -999.0) d->Temp = temp;
  TNOM = m->TNOM;
  vt = OR_KoverQ * d->Temp;
  fact2 = d->Temp / OR_REFTEMP;
  egfet = OR_Eg0 - (OR_alphaEg * d->Temp * d->Temp) / (d->Temp + OR_betaEg);
  arg = -egfet / (2.0 * OR_BOLTZ * d->Temp) + OR_Eg300 / (OR_BOLTZ * (OR_REFTEMP + OR_REFTEMP));
  pbfact = -2.0 * vt * (1.5 * log(fact2) + OR_Q * arg);
  egfet1 = OR_Eg0 - (OR_alphaEg * m->TNOM * m->TNOM) / (m->TNOM + OR_betaEg);
  arg1 = -egfet1 / (2.0 * OR_BOLTZ * m->TNOM) + OR_Eg300 / (2.0 * OR_BOLTZ * OR_REFTEMP);
  fact1 = m->TNOM / OR_REFTEMP;
  pbfact1 = -2.0 * vtnom * (1.5 * log(fact1) + OR_Q * arg1);
  pbo = (m->VJ - pbfact1) / fact1;
  gmaold = (m->VJ - pbo) / pbo;
  d->tJctCap = m->CJO / (1.0 + m->M * (4.0e-4 * (m->TNOM - OR_REFTEMP) - gmaold));
  d->tJctPot = pbfact + fact2 * pbo;
  gmanew = (d->tJctPot - pbo) / pbo;
  d->tJctCap *= 1.0 + m->M * (4.0e-4 * (d->Temp - OR_REFTEMP) - gmanew);
  d->tSatCur = m->IS * exp(((d->Temp / m->TNOM) - 1.0) * m->EG / (m->N * vt) +
                           m->XTI / m->N * log(d->Temp / m->TNOM));
  d->tF1 = d->tJctPot * (1.0 - exp((1.0 - m->M) * xfc)) / (1.0 - m->M);
  d->tDepCap = m->FC * d->tJctPot;
  vte = m->N * vt;
  d->tVcrit = vte * log(vte / (sqrt(2.0) * d->tSatCur));
  d->tRS = m->RS;
  d->tCOND = m->COND;
  d->tIRF = m->IRF * pow(fact2, 1.6);
  if (m->level == 2) {
    d->tSatCurR = m->ISR * exp((d->Temp / TNOM - 1.0) * m->EG / (m->NR * vt) +
                               m->XTI / m->NR * log(d->Temp / TNOM));
    d->tIKF = m->IKF * (1 + m->TIKF * (d->Temp - TNOM));
    tempBV = m->BV * (1 + (d->Temp - TNOM) * (m->TBV1 + m->TBV2 * (d->Temp - TNOM)));
    d->tRS = m->RS * (1 + (d->Temp - TNOM) * (m->TRS1 + m->TRS2 * (d->Temp - TNOM)));
    d->tCOND = 0.0;
    if (d->tRS != 0.0) d->tCOND = 1.0 / d->tRS;
    d->tJctPot = (m->VJ - egfet1) * fact2 - 3 * vt * log(fact2) + egfet;
    d->tJctCap = m->CJO / (1.0 + m->M * (4.0e-4 * (d->Temp - TNOM) + (1 - d->tJctPot / m->VJ)));
  } else {
    tempBV = m->BV;
  }
  if (m->BVGiven) {
    double IRFactor = d->tIRF, cbv = m->IBV, xbv, cthreshlow, cthreshhigh, arg2;
    int iter_count;
    double a1 = 3.0 * vte / (exp(1.0) * tempBV);
    a1 = a1 * a1 * a1;
    cthreshlow = d->tSatCur * IRFactor * (1 - a1);
    cthreshhigh = d->tSatCur * IRFactor * (1 - 1.0 / (exp(1.0) * exp(1.0) * exp(1.0))) *
                  exp(-1.0 * (3.0 * vte - tempBV) / vte);
    if (cbv >= cthreshhigh) {
      d->tBrkdwnV = 3.0 * vte;
    } else if (cbv <= cthreshlow) {
      d->tBrkdwnV = tempBV;
    } else {
      xbv = tempBV;
      for (iter_count = 0; iter_count < 8; iter_count++) {
        arg2 = 3.0 * vte / (exp(1.0) * xbv);
        arg2 = arg2 * arg2 * arg2;
        xbv = tempBV - vte * log(cbv / (d->tSatCur * IRFactor)) + vte * log(1 - arg2);
      }
      d->tBrkdwnV = xbv;
    }
  }
}

static int setup(struct or_group_s* g) {
  int k;
  int64_t i;
  dmodel* mods = (dmodel*)calloc(g->nmodels, sizeof(dmodel));
  dinst* in = (dinst*)calloc(g->n > 0 ? g->n : 1, sizeof(dinst));
  for (k = 0; k < g->nmodels; ++k) {
    const or_card* c = &g->models[k];
    dmodel* m = &mods[k];
    if (c->level != 1 && c->level != 2) return or_fail(XA_ERR_UNSUPPORTED, "diode level");
    m->IS = c->v[IS_]; m->RS = c->v[RS_]; m->N = c->v[N_]; m->ISR = c->v[ISR_]; m->NR = c->v[NR_];
    m->IKF = c->v[IKF_]; m->TT = c->v[TT_]; m->CJO = c->v[CJO_]; m->VJ = c->v[VJ_]; m->M = c->v[M_];
    m->EG = c->v[EG_]; m->XTI = c->v[XTI_]; m->TIKF = c->v[TIKF_]; m->TBV1 = c->v[TBV1_];
    m->TBV2 = c->v[TBV2_]; m->TRS1 = c->v[TRS1_]; m->TRS2 = c->v[TRS2_]; m->FC = c->v[FC_];
    m->BV = c->v[BV_]; m->IBV = c->v[IBV_]; m->IRF = c->v[IRF_]; m->NBV = c->v[NBV_];
    m->IBVL = c->v[IBVL_]; m->NBVL = c->v[NBVL_]; m->TNOM = c->v[TNOM_]; m->KF = c->v[KF_];
    m->AF = c->v[AF_];
    m->BVGiven = c->given[BV_];
    m->level = c->level;
    if (!c->given[TNOM_]) m->TNOM = g->opt.tnom;
    model_process(m);
  }
  for (i = 0; i < g->n; ++i) {
    dinst* d = &in[i];
    d->model = &mods[g->inst_model[i]];
    d->Area = g->ival[AREA_][i];
    d->InitCond = g->ival[IC_][i];
    d->InitCondGiven = g->igiven[IC_];
    d->off = (int)g->ival[OFF_][i];
    d->tIRF = 1.0;
    d->lambertWFlag = g->igiven[LAMBERTW_] ? (int)g->ival[LAMBERTW_][i] : g->opt.lambertWFlag;
    d->Temp = g->igiven[TEMP_] ? g->ival[TEMP_][i] : g->opt.temp;
    d->loadLeadCurrent = g->lead[i];
    g->n_ext[i] = 2;
    g->n_int[i] = 1;
    g->n_sta[i] = 0;
    if ((d->model->RS == 0.0) || (d->lambertWFlag == 1)) g->n_int[i] = 0;
    if (d->model->CJO == 0.0) g->n_sta[i] = 1;
    if ((d->model->RS == 0.0) && (d->lambertWFlag == 1)) d->model->RS = 1.0e-12;
    g->n_sto[i] = 1 + (d->loadLeadCurrent ? 1 : 0);
    update_temperature(d, d->Temp);   /* processParams -> updateTemperature(Temp) */
  }
  g->mdata = mods;
  g->inst = in;
  return XA_OK;
}

static void register_lids(struct or_group_s* g, int64_t i, const int32_t* nodes,
                          const int32_t* sta, const int32_t* sto) {
  dinst* d = &((dinst*)g->inst)[i];
  (void)sta;
  d->li_Pos = nodes[0];
  d->li_Neg = nodes[1];
  if (d->model->RS && (d->lambertWFlag != 1)) d->li_Pri = nodes[2];
  else d->li_Pri = d->li_Pos;
  d->li_storevd = sto ? sto[0] : -1;
  d->li_store_dev_i = (d->loadLeadCurrent && sto) ? sto[1] : -1;
}

/* jacStamp_RS / collapsed jacStamp (N_DEV_Diode.C:420-445): dFdx and dQdx use
 * the same pattern; pairs are listed once per matrix. */
static int64_t jac_pairs(struct or_group_s* g, int64_t i, int32_t* r, int32_t* c) {
  dinst* d = &((dinst*)g->inst)[i];
  int32_t pr[7] = {d->li_Pos, d->li_Pos, d->li_Neg, d->li_Neg, d->li_Pri, d->li_Pri, d->li_Pri};
  int32_t pc[7] = {d->li_Pos, d->li_Pri, d->li_Neg, d->li_Pri, d->li_Pos, d->li_Neg, d->li_Pri};
  int64_t k, m = 0;
  for (k = 0; k < 7; ++k) {
    if (pr[k] < 0 || pc[k] < 0) continue;
    if (r) { r[m] = pr[k]; c[m] = pc[k]; }
    ++m;
  }
  return m;
}

static int register_jac(struct or_group_s* g, int64_t i) {
  dinst* d = &((dinst*)g->inst)[i];
  d->fPosPos = or_slot(g, d->li_Pos, d->li_Pos);
  d->fPosPri = or_slot(g, d->li_Pos, d->li_Pri);
  d->fNegNeg = or_slot(g, d->li_Neg, d->li_Neg);
  d->fNegPri = or_slot(g, d->li_Neg, d->li_Pri);
  d->fPriPos = or_slot(g, d->li_Pri, d->li_Pos);
  d->fPriNeg = or_slot(g, d->li_Pri, d->li_Neg);
  d->fPriPri = or_slot(g, d->li_Pri, d->li_Pri);
  d->qNegNeg = d->fNegNeg; d->qNegPri = d->fNegPri;
  d->qPriNeg = d->fPriNeg; d->qPriPri = d->fPriPri;
  if (d->fPosPos < -1 || d->fPosPri < -1 || d->fNegNeg < -1 || d->fNegPri < -1 ||
      d->fPriPos < -1 || d->fPriNeg < -1 || d->fPriPri < -1)
    return XA_ERR_ARG;
  return XA_OK;
}

static double sol(const double* x, int lid) { return lid >= 0 ? x[lid] : 0.0; }

static void lambertWCurrent(dinst* d, double Isat, double Vte, double RS) {
  double arg1 = (d->Vd + Isat * RS) / Vte, evd, w, xi;
  int ierr;
  arg1 = or_min(OR_MAX_EXP_ARG, arg1);
  evd = exp(arg1);
  or_lambertw(Isat * RS * evd / Vte, &w, &ierr, &xi);
  d->Id = -Isat + Vte * (w) / RS;
  d->Gd = w / ((1 + w) * RS);
}

static void lambertWLinearReverseBias(dinst* d, double Isat, double Vte, double RS) {
  double FF1 = (d->Vd + d->tBrkdwnV) / (-3.0 * Vte + d->tBrkdwnV);
  double FF2 = (1 / 2 - FF1) * (-2);
  double arg = Vte / RS;
  double arg1 = (Isat / arg - 3);
  double arg2 = FF1 * arg1;
  double evd, w, xi, GdFF1, GdFF2, GdW;
  int ierr;
  arg1 = or_min(OR_MAX_EXP_ARG, arg2);
  evd = exp(arg2);
  or_lambertw(Isat * evd / arg, &w, &ierr, &xi);
  d->Id = -Isat * FF1 + FF2 * w * arg;
  GdFF1 = 1 / (-3.0 * Vte + d->tBrkdwnV);
  GdFF2 = 2 * GdFF1;
  GdW = arg * w * GdFF1 * arg1 / (1 + w);
  d->Gd = -Isat * GdFF1 + GdFF2 * arg * w + FF2 * GdW;
}

static void lambertWBreakdownCurrent(dinst* d, double Isat, double Vte, double RS) {
  double arg1 = (-d->Vd - d->tBrkdwnV) / Vte, evd, w, xi;
  int ierr;
  arg1 = or_min(OR_MAX_EXP_ARG, arg1);
  evd = exp(arg1);
  or_lambertw(Isat * RS * evd / Vte, &w, &ierr, &xi);
  d->Id = -Vte * w / RS;
  d->Gd = w / ((1 + w) * RS);
}

/* Instance::updateIntermediateVars, N_DEV_Diode.C:917-1366 */
static void update_intermediate(struct or_group_s* g, dinst* d, const xa_solver_state* ss,
                                const xa_vectors* v) {
  const dmodel* mm = d->model;
  const double* solVec = v->nextSol;
  const double gmin = g->opt.gmin;
  double M = mm->M, BV = mm->BV, N = mm->N, NR = mm->NR, TT = mm->TT, F2 = mm->F2, F3 = mm->F3;
  double Isat, IsatR, Vt, Vte, VteR;
  int level;
  d->Vpp = sol(solVec, d->li_Pri);
  d->Vn = sol(solVec, d->li_Neg);
  d->Vp = sol(solVec, d->li_Pos);
  d->Vd = d->Vpp - d->Vn;
  Isat = d->tSatCur * d->Area;
  IsatR = d->tSatCurR * d->Area;
  Vt = OR_KoverQ * d->Temp;
  Vte = N * Vt;
  VteR = NR * Vt;
  d->Gspr = d->tCOND * d->Area;
  d->Vd_orig = d->Vd;
  d->origFlag = 1;
  if (ss->newtonIter == 0) {
    d->newtonIterOld = 0;
    if (ss->initJctFlag && g->opt.voltageLimiterFlag) {
      if (d->InitCondGiven) { d->Vd = d->InitCond; d->origFlag = 0; }
      else if (d->off) { d->Vd = 0.0; d->origFlag = 0; }
      else {
        if (ss->inputOPFlag) {
          const double* fl = v->flagSol;
          if (!fl || d->li_Pos < 0 || d->li_Neg < 0 || d->li_Pri < 0 ||
              fl[d->li_Pos] == 0 || fl[d->li_Neg] == 0 || fl[d->li_Pri] == 0) {
            d->Vd = d->tVcrit; d->origFlag = 0;
          }
        } else {
          d->Vd = d->tVcrit; d->origFlag = 0;
        }
      }
    }
    d->Vd_old = d->Vd;
    if (!(ss->dcopFlag) || (ss->locaEnabledFlag && ss->dcopFlag))
      d->Vd_old = d->li_storevd >= 0 ? v->currSto[d->li_storevd] : 0.0;
  } else {
    d->Vd_old = d->li_storevd >= 0 ? v->nextSto[d->li_storevd] : 0.0;
  }
  if (g->opt.voltageLimiterFlag) {
    int ichk = 0;
    if (ss->newtonIter >= 0) {
      if (mm->BVGiven && (d->Vd < or_min(0.0, -BV + 10.0 * Vte))) {
        double Vdtmp = -(BV + d->Vd);
        Vdtmp = or_pnjlim(Vdtmp, -(d->Vd_old + BV), Vte, d->tVcrit, &ichk);
        d->Vd = -(Vdtmp + BV);
      } else {
        d->Vd = or_pnjlim(d->Vd, d->Vd_old, Vte, d->tVcrit, &ichk);
      }
      if (ichk) d->origFlag = 0;
    }
  }
  if (ss->newtonIter != 0 && ss->newtonIter != d->newtonIterOld) d->newtonIterOld = ss->newtonIter;
  level = mm->level;
  if (level == 1) {
    if (d->lambertWFlag == 1) {
      double RS = mm->RS;
      if (d->Vd >= -3.0 * Vte) lambertWCurrent(d, Isat, Vte, RS);
      else if (!d->tBrkdwnV || (d->Vd >= -d->tBrkdwnV)) lambertWLinearReverseBias(d, Isat, Vte, RS);
      else lambertWBreakdownCurrent(d, Isat, Vte, RS);
      d->Vc = d->Vd - (d->Id + d->Icd) * RS;
    } else if (d->lambertWFlag == 2) {
      if (d->Vd >= -3.0 * Vte) lambertWCurrent(d, Isat, Vte, 1.0e-15);
      else if (!d->tBrkdwnV || (d->Vd >= -d->tBrkdwnV)) lambertWLinearReverseBias(d, Isat, Vte, 1.0e-15);
      else lambertWBreakdownCurrent(d, Isat, Vte, 1.0e-15);
      d->Vc = d->Vd;
    } else {
      double IRfactor = (d->Vd >= -3.0 * Vte) ? 1.0 : d->tIRF;
      Isat *= IRfactor;
      if (d->Vd >= -3.0 * Vte) {
        double arg1 = d->Vd / Vte, evd;
        arg1 = or_min(OR_MAX_EXP_ARG, arg1);
        evd = exp(arg1);
        d->Id = Isat * (evd - 1.0) + gmin * d->Vd;
        d->Gd = Isat * evd / Vte + gmin;
      } else if (!d->tBrkdwnV || (d->Vd >= -d->tBrkdwnV)) {
        double arg = 3.0 * Vte / (d->Vd * exp(1.0));
        arg = arg * arg * arg;
        d->Id = -Isat * (1.0 + arg) + gmin * d->Vd;
        d->Gd = Isat * 3.0 * arg / d->Vd + gmin;
      } else {
        double arg1 = -(d->tBrkdwnV + d->Vd) / Vte, evrev, arg2, Isat_t;
        arg1 = or_min(OR_MAX_EXP_ARG, arg1);
        evrev = exp(arg1);
        arg2 = 3.0 * Vte / (exp(1.0) * d->tBrkdwnV);
        arg2 = arg2 * arg2 * arg2;
        Isat_t = Isat * (1 - arg2);
        d->Id = -Isat_t * evrev + gmin * d->Vd;
        d->Gd = Isat_t * evrev / Vte + gmin;
      }
      d->Vc = d->Vd;
    }
  } else if (level == 2) {
    double IRfactor = (d->Vd >= -3.0 * Vte) ? 1.0 : d->tIRF;
    Isat *= IRfactor;
    if (d->Vd >= -3.0 * Vte) {
      double arg1 = d->Vd / Vte, evd, Inorm, Gd1, Irec, Gd2, Khi, DKhi, Kgen, DKgen;
      arg1 = or_min(OR_MAX_EXP_ARG, arg1);
      evd = exp(arg1);
      Inorm = Isat * (evd - 1.0) + gmin * d->Vd;
      Gd1 = Isat * evd / Vte + gmin;
      arg1 = d->Vd / VteR;
      arg1 = or_min(OR_MAX_EXP_ARG, arg1);
      evd = exp(arg1);
      Irec = IsatR * (evd - 1.0);
      Gd2 = IsatR * evd / VteR;
      Khi = 1;
      DKhi = 0;
      if (d->tIKF > 0) {
        Khi = sqrt(d->tIKF / (d->tIKF + Inorm));
        DKhi = 0.5 * Khi * Gd1 / (d->tIKF + Inorm);
      }
      Kgen = 0;
      DKgen = 0;
      if (Irec != 0) {
        Kgen = sqrt(pow(((1 - d->Vd / d->tJctPot) * (1 - d->Vd / d->tJctPot) + 0.005), M));
        DKgen = -M * (1 - d->Vd / d->tJctPot) / (d->tJctPot * Kgen);
      }
      d->Id = Inorm * Khi + Irec * Kgen;
      d->Gd = Gd1 * Khi + Inorm * DKhi + Gd2 * Kgen + Irec * DKgen;
    } else if (!d->tBrkdwnV || (d->Vd >= -d->tBrkdwnV)) {
      double arg = 3.0 * Vte / (d->Vd * exp(1.0));
      arg = arg * arg * arg;
      d->Id = -Isat * (1.0 + arg) + gmin * d->Vd;
      d->Gd = Isat * 3.0 * arg / d->Vd + gmin;
    } else {
      double arg1 = -(d->tBrkdwnV + d->Vd) / Vte, evrev, arg2, Isat_t;
      arg1 = or_min(OR_MAX_EXP_ARG, arg1);
      evrev = exp(arg1);
      arg2 = 3.0 * Vte / (exp(1.0) * d->tBrkdwnV);
      arg2 = arg2 * arg2 * arg2;
      Isat_t = Isat * (1 - arg2);
      d->Id = -Isat_t * evrev + gmin * d->Vd;
      d->Gd = Isat_t * evrev / Vte + gmin;
    }
    d->Vc = d->Vd;
  }
  if (d->tJctCap != 0.0) {
    double Czero = d->tJctCap * d->Area;
    if (d->Vc < d->tDepCap) {
      double arg = 1.0 - d->Vc / d->tJctPot;
      double arg1 = -M * log(arg), sarg;
      arg1 = or_min(OR_MAX_EXP_ARG, arg1);
      sarg = exp(arg1);
      d->Qd = TT * d->Id + d->tJctPot * Czero * (1.0 - arg * sarg) / (1.0 - M);
      d->Cd = TT * d->Gd + Czero * sarg;
    } else {
      double Czof2 = Czero / F2;
      double MotJctPot = M / d->tJctPot;
      d->Qd = TT * d->Id + Czero * d->tF1 +
              Czof2 * (F3 * (d->Vc - d->tDepCap) + (0.5 * MotJctPot) *
                       (d->Vc * d->Vc - d->tDepCap * d->tDepCap));
      d->Cd = TT * d->Gd + Czof2 * (F3 + MotJctPot * d->Vc);
    }
  } else {
    d->Qd = 0.0;
    d->Cd = 0.0;
  }
}

/* Master::updateState, N_DEV_Diode.C:2051-2068 */
static int update_state(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  dinst* in = (dinst*)g->inst;
  /* instances are independent here; one thread is the serial order */
#pragma omp parallel for schedule(static) num_threads(or_threads) if (or_threads > 1)
  for (i = 0; i < g->n; ++i) {
    dinst* d = &in[i];
    if (d->li_storevd >= 0) v->nextSto[d->li_storevd] = d->Vd;
    update_intermediate(g, d, ss, v);
  }
  return XA_OK;
}

/* Master::loadDAEVectors, N_DEV_Diode.C:2078-2122 */
static int load_vectors(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  dinst* in = (dinst*)g->inst;
  (void)ss;
  for (i = 0; i < g->n; ++i) {
    dinst* di = &in[i];
    double Ir = di->Gspr * (di->Vp - di->Vpp);
    if (v->f) {
      OR_SUB(v->f, di->li_Pos, -Ir);
      OR_SUB(v->f, di->li_Neg, di->Id);
      OR_SUB(v->f, di->li_Pri, (-di->Id + Ir));
    }
    if (v->q) {
      OR_SUB(v->q, di->li_Neg, di->Qd);
      OR_SUB(v->q, di->li_Pri, -di->Qd);
    }
    if (g->opt.voltageLimiterFlag) {
      double Vd_diff = di->Vd - di->Vd_orig;
      double Cd_Jdxp = -(di->Cd) * Vd_diff;
      double Gd_Jdxp = -(di->Gd) * Vd_diff;
      if (v->dFdxdVp) {
        OR_ADD(v->dFdxdVp, di->li_Neg, Gd_Jdxp);
        OR_SUB(v->dFdxdVp, di->li_Pri, Gd_Jdxp);
      }
      if (v->dQdxdVp) {
        OR_ADD(v->dQdxdVp, di->li_Neg, Cd_Jdxp);
        OR_SUB(v->dQdxdVp, di->li_Pri, Cd_Jdxp);
      }
    }
    if (di->loadLeadCurrent && di->li_store_dev_i >= 0) {
      if (v->storeLeadF) OR_SET(v->storeLeadF, di->li_store_dev_i, di->Id);
      if (di->model->CJO != 0.0 && v->storeLeadQ) OR_SET(v->storeLeadQ, di->li_store_dev_i, di->Qd);
    }
  }
  return XA_OK;
}

#define ADD OR_ADD
#define SUB OR_SUB

/* Master::loadDAEMatrices, N_DEV_Diode.C:2132-2171 */
static int load_matrices(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  dinst* in = (dinst*)g->inst;
  (void)ss;
  for (i = 0; i < g->n; ++i) {
    dinst* di = &in[i];
    if (v->dFdx) {
      ADD(v->dFdx, di->fPosPos, di->Gspr);
      SUB(v->dFdx, di->fPosPri, di->Gspr);
      ADD(v->dFdx, di->fNegNeg, di->Gd);
      SUB(v->dFdx, di->fNegPri, di->Gd);
      SUB(v->dFdx, di->fPriPos, di->Gspr);
      SUB(v->dFdx, di->fPriNeg, di->Gd);
      ADD(v->dFdx, di->fPriPri, di->Gspr + di->Gd);
    }
    if (v->dQdx) {
      ADD(v->dQdx, di->qNegNeg, di->Cd);
      SUB(v->dQdx, di->qNegPri, di->Cd);
      SUB(v->dQdx, di->qPriNeg, di->Cd);
      ADD(v->dQdx, di->qPriPri, di->Cd);
    }
  }
  return XA_OK;
}

static int read_var(struct or_group_s* g, const char* name, double* out) {
  int64_t i;
  dinst* in = (dinst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    const dinst* d = &in[i];
    if (!strcmp(name, "Vd")) out[i] = d->Vd;
    else if (!strcmp(name, "Id")) out[i] = d->Id;
    else if (!strcmp(name, "Gd")) out[i] = d->Gd;
    else if (!strcmp(name, "Qd")) out[i] = d->Qd;
    else if (!strcmp(name, "Cd")) out[i] = d->Cd;
    else if (!strcmp(name, "Vd_orig")) out[i] = d->Vd_orig;
    else if (!strcmp(name, "tSatCur")) out[i] = d->tSatCur;
    else if (!strcmp(name, "tVcrit")) out[i] = d->tVcrit;
    else if (!strcmp(name, "tJctCap")) out[i] = d->tJctCap;
    else if (!strcmp(name, "tJctPot")) out[i] = d->tJctPot;
    else if (!strcmp(name, "tBrkdwnV")) out[i] = d->tBrkdwnV;
    else return or_fail(XA_ERR_ARG, "diode: unknown intermediate");
  }
  return XA_OK;
}

const or_model_ops or_diode_ops = {
  "diode", mpar, sizeof(mpar) / sizeof(mpar[0]), NMS_, ipar, sizeof(ipar) / sizeof(ipar[0]), NIS_,
  setup, register_lids, jac_pairs, register_jac, update_state, NULL, load_vectors,
  load_matrices, read_var, NULL, sizeof(dinst),
};
