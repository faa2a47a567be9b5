/* or_linear.c -- oracle restatement of the Xyce resistor and capacitor (test
 * infrastructure only).
 * Reference: src/DeviceModelPKG/OpenModels/src/
 *   N_DEV_Resistor.C  parameter tables :170-262; Instance ctor :279-360;
 *     registerStoreLIDs :500-510; updateTemperature :848-865;
 *     Master::updateState :1075-1090; Master::loadDAEVectors :1100-1135;
 *     Master::loadDAEMatrices :1145-1181.
 *   N_DEV_Capacitor.C parameter tables :85-165; processParams :196-244;
 *     updateTemperature :271-280; Instance ctor :440-606;
 *     Master::updateState :1821-1850; Master::loadDAEVectors :1860-1960;
 *     Master::loadDAEMatrices :1970-2135.
 * Loads go through per-instance CSR slots in the reference's statement
 * order, skipping what the reference skips (no zero placeholders).
 */
#include "or_internal.h"

static double sol(const double* x, int lid) { return lid >= 0 ? x[lid] : 0.0; }

/* ======================= resistor ======================= */
enum { R_TC1, R_TC2, R_RSH, R_RMULT, R_DEFW, R_NARROW, R_TNOM, R_NM };
static const or_param r_mpar[] = {
  {"TC1", R_TC1, 0.0, 0}, {"TC2", R_TC2, 0.0, 0}, {"RSH", R_RSH, 0.0, 0},
  {"R", R_RMULT, 1.0, 0}, {"DEFW", R_DEFW, 1.e-5, 0}, {"NARROW", R_NARROW, 0.0, 0},
  {"TNOM", R_TNOM, 0.0, 1},
};
enum { RI_R, RI_L, RI_W, RI_TEMP, RI_TC1, RI_TC2, RI_DTEMP, R_NI };
static const or_param r_ipar[] = {
  {"R", RI_R, 1000.0, 0}, {"L", RI_L, 0.0, 0}, {"W", RI_W, 0.0, 0}, {"TEMP", RI_TEMP, 0.0, 1},
  {"TC1", RI_TC1, 0.0, 0}, {"TC2", RI_TC2, 0.0, 0}, {"DTEMP", RI_DTEMP, 0.0, 0},
};

typedef struct {
  double R, G, i0, temp, tempCoeff1, tempCoeff2, width, length;
  int loadLeadCurrent;
  int32_t li_Pos, li_Neg, li_store_dev_i;
  int64_t pp, pn, np, nn;
} rinst;

static int r_setup(struct or_group_s* g) {
  int64_t i;
  rinst* in = (rinst*)calloc(g->n > 0 ? g->n : 1, sizeof(rinst));
  int k;
  for (k = 0; k < g->nmodels; ++k)
    if (g->models[k].level != 1) return or_fail(XA_ERR_UNSUPPORTED, "resistor level");
  for (i = 0; i < g->n; ++i) {
    rinst* r = &in[i];
    const or_card* c = &g->models[g->inst_model[i]];
    double tnom = c->given[R_TNOM] ? c->v[R_TNOM] : g->opt.tnom;
    double difference, factor;
    r->R = g->ival[RI_R][i];
    r->length = g->ival[RI_L][i];
    r->width = g->ival[RI_W][i];
    r->temp = g->igiven[RI_TEMP] ? g->ival[RI_TEMP][i] : g->opt.temp;
    if (!g->igiven[RI_W]) r->width = c->v[R_DEFW];
    r->tempCoeff1 = g->igiven[RI_TC1] ? g->ival[RI_TC1][i] : c->v[R_TC1];
    r->tempCoeff2 = g->igiven[RI_TC2] ? g->ival[RI_TC2][i] : c->v[R_TC2];
    if (!g->igiven[RI_R]) {
      if (c->given[R_RSH] && g->igiven[RI_L] && (c->v[R_RSH] != 0) && (r->length != 0))
        r->R = c->v[R_RSH] * (r->length - c->v[R_NARROW]) / (r->width - c->v[R_NARROW]);
      else
        r->R = 1000.0;
    }
    difference = r->temp - tnom;
    factor = c->v[R_RMULT] * (1.0 + r->tempCoeff1 * difference +
                              r->tempCoeff2 * difference * difference);
    if (r->R * factor != 0.0) r->G = 1.0 / (r->R * factor);
    else r->G = 0.0;
    r->loadLeadCurrent = g->lead[i];
    g->n_ext[i] = 2; g->n_int[i] = 0; g->n_sta[i] = 0;
    g->n_sto[i] = r->loadLeadCurrent ? 1 : 0;
  }
  g->inst = in;
  return XA_OK;
}

static void r_register_lids(struct or_group_s* g, int64_t i, const int32_t* nodes,
                            const int32_t* sta, const int32_t* sto) {
  rinst* r = &((rinst*)g->inst)[i];
  (void)sta;
  r->li_Pos = nodes[0];
  r->li_Neg = nodes[1];
  r->li_store_dev_i = (r->loadLeadCurrent && sto) ? sto[0] : -1;
}

static int64_t r_jac_pairs(struct or_group_s* g, int64_t i, int32_t* rr, int32_t* cc) {
  rinst* r = &((rinst*)g->inst)[i];
  int32_t pr[4] = {r->li_Pos, r->li_Pos, r->li_Neg, r->li_Neg};
  int32_t pc[4] = {r->li_Pos, r->li_Neg, r->li_Pos, r->li_Neg};
  int64_t k, m = 0;
  for (k = 0; k < 4; ++k) {
    if (pr[k] < 0 || pc[k] < 0) continue;
    if (rr) { rr[m] = pr[k]; cc[m] = pc[k]; }
    ++m;
  }
  return m;
}

static int r_register_jac(struct or_group_s* g, int64_t i) {
  rinst* r = &((rinst*)g->inst)[i];
  r->pp = or_slot(g, r->li_Pos, r->li_Pos);
  r->pn = or_slot(g, r->li_Pos, r->li_Neg);
  r->np = or_slot(g, r->li_Neg, r->li_Pos);
  r->nn = or_slot(g, r->li_Neg, r->li_Neg);
  if (r->pp < -1 || r->pn < -1 || r->np < -1 || r->nn < -1) return XA_ERR_ARG;
  return XA_OK;
}

/* Master::updateState, N_DEV_Resistor.C:1075-1090 */
static int r_update_state(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  rinst* in = (rinst*)g->inst;
  (void)ss;
  for (i = 0; i < g->n; ++i) {
    rinst* r = &in[i];
    double v_pos = sol(v->nextSol, r->li_Pos);
    double v_neg = sol(v->nextSol, r->li_Neg);
    r->i0 = (v_pos - v_neg) * r->G;
  }
  return XA_OK;
}

/* Master::loadDAEVectors, N_DEV_Resistor.C:1100-1135 */
static int r_load_vectors(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  rinst* in = (rinst*)g->inst;
  (void)ss;
  for (i = 0; i < g->n; ++i) {
    rinst* r = &in[i];
    if (v->f) {
      OR_ADD(v->f, r->li_Pos, r->i0);
      OR_ADD(v->f, r->li_Neg, -r->i0);
    }
    if (r->loadLeadCurrent && r->li_store_dev_i >= 0 && v->storeLeadF)
      OR_SET(v->storeLeadF, r->li_store_dev_i, r->i0);
  }
  return XA_OK;
}

/* Master::loadDAEMatrices, N_DEV_Resistor.C:1145-1181 */
static int r_load_matrices(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  rinst* in = (rinst*)g->inst;
  (void)ss;
  for (i = 0; i < g->n; ++i) {
    rinst* r = &in[i];
    if (!v->dFdx) continue;
    OR_ADD(v->dFdx, r->pp, r->G);
    OR_SUB(v->dFdx, r->pn, r->G);
    OR_SUB(v->dFdx, r->np, r->G);
    OR_ADD(v->dFdx, r->nn, r->G);
  }
  return XA_OK;
}

static int r_read(struct or_group_s* g, const char* name, double* out) {
  int64_t i;
  rinst* in = (rinst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    if (!strcmp(name, "G")) out[i] = in[i].G;
    else if (!strcmp(name, "R")) out[i] = in[i].R;
    else if (!strcmp(name, "i0")) out[i] = in[i].i0;
    else return or_fail(XA_ERR_ARG, "resistor: unknown intermediate");
  }
  return XA_OK;
}

const or_model_ops or_resistor_ops = {
  "resistor", r_mpar, sizeof(r_mpar) / sizeof(r_mpar[0]), R_NM, r_ipar,
  sizeof(r_ipar) / sizeof(r_ipar[0]), R_NI, r_setup, r_register_lids, r_jac_pairs,
  r_register_jac, r_update_state, NULL, r_load_vectors, r_load_matrices, r_read, NULL,
  sizeof(rinst),
};

/* ======================= capacitor ======================= */
enum { C_CMULT, C_CJ, C_CJSW, C_DEFW, C_NARROW, C_TC1, C_TC2, C_TNOM, C_NM };
static const or_param c_mpar[] = {
  {"C", C_CMULT, 1.0, 0}, {"CJ", C_CJ, 0.0, 0}, {"CJSW", C_CJSW, 0.0, 0},
  {"DEFW", C_DEFW, 1.e-6, 0}, {"NARROW", C_NARROW, 0.0, 0}, {"TC1", C_TC1, 0.0, 0},
  {"TC2", C_TC2, 0.0, 0}, {"TNOM", C_TNOM, 0.0, 1},
};
enum { CI_C, CI_IC, CI_L, CI_W, CI_AGE, CI_D, CI_TEMP, CI_TC1, CI_TC2, C_NI };
static const or_param c_ipar[] = {
  {"C", CI_C, 1.e-6, 0}, {"IC", CI_IC, 0.0, 0}, {"L", CI_L, 1.0, 0}, {"W", CI_W, 1.e-6, 0},
  {"AGE", CI_AGE, 0.0, 0}, {"D", CI_D, 0.0233, 0}, {"TEMP", CI_TEMP, 0.0, 1},
  {"TC1", CI_TC1, 0.0, 0}, {"TC2", CI_TC2, 0.0, 0},
};

typedef struct {
  double C, baseCap, IC, vcap, q0, temp, tempCoeff1, tempCoeff2, width, length;
  int ICGiven, loadLeadCurrent;
  int32_t li_Pos, li_Neg, li_Bra, li_QState, li_store_dev_i;
  int64_t qpp, qpn, qnp, qnn, fpb, fnb, fbp, fbn, fbb;
} cinst;

static int c_setup(struct or_group_s* g) {
  int64_t i;
  int k;
  cinst* in = (cinst*)calloc(g->n > 0 ? g->n : 1, sizeof(cinst));
  for (k = 0; k < g->nmodels; ++k)
    if (g->models[k].level != 1) return or_fail(XA_ERR_UNSUPPORTED, "capacitor level");
  if (!g->igiven[CI_C] && g->igiven[CI_AGE]) return or_fail(XA_ERR_ARG, "capacitor: AGE without C");
  if (!g->igiven[CI_C] && !g->igiven[CI_L]) return or_fail(XA_ERR_ARG, "capacitor: no C nor L");
  for (i = 0; i < g->n; ++i) {
    cinst* c = &in[i];
    const or_card* m = &g->models[g->inst_model[i]];
    double tnom = m->given[C_TNOM] ? m->v[C_TNOM] : g->opt.tnom;
    double difference, factor;
    c->C = g->ival[CI_C][i];
    c->IC = g->ival[CI_IC][i];
    c->ICGiven = g->igiven[CI_IC];
    c->length = g->ival[CI_L][i];
    c->width = g->igiven[CI_W] ? g->ival[CI_W][i] : m->v[C_DEFW];
    c->temp = g->igiven[CI_TEMP] ? g->ival[CI_TEMP][i] : g->opt.temp;
    c->tempCoeff1 = g->igiven[CI_TC1] ? g->ival[CI_TC1][i] : m->v[C_TC1];
    c->tempCoeff2 = g->igiven[CI_TC2] ? g->ival[CI_TC2][i] : m->v[C_TC2];
    /* processParams */
    c->baseCap = c->C;
    if (g->igiven[CI_AGE] && g->ival[CI_AGE][i] >= 1)
      c->baseCap = c->baseCap * (1 - g->ival[CI_D][i] * log10(g->ival[CI_AGE][i]));
    if (!g->igiven[CI_C])
      c->baseCap = c->C = m->v[C_CJ] * (c->length - m->v[C_NARROW]) * (c->width - m->v[C_NARROW]) +
                          2 * m->v[C_CJSW] * (c->length + c->width - 2 * m->v[C_NARROW]);
    /* updateTemperature */
    difference = c->temp - tnom;
    factor = m->v[C_CMULT] * (1.0 + c->tempCoeff1 * difference +
                              c->tempCoeff2 * difference * difference);
    c->C = c->baseCap * factor;
    c->loadLeadCurrent = g->lead[i];
    g->n_ext[i] = 2;
    g->n_int[i] = c->ICGiven ? 1 : 0;
    g->n_sta[i] = 1;
    g->n_sto[i] = c->loadLeadCurrent ? 1 : 0;
  }
  g->inst = in;
  return XA_OK;
}

static void c_register_lids(struct or_group_s* g, int64_t i, const int32_t* nodes,
                            const int32_t* sta, const int32_t* sto) {
  cinst* c = &((cinst*)g->inst)[i];
  c->li_Pos = nodes[0];
  c->li_Neg = nodes[1];
  c->li_Bra = c->ICGiven ? nodes[2] : -1;
  c->li_QState = sta ? sta[0] : -1;
  c->li_store_dev_i = (c->loadLeadCurrent && sto) ? sto[0] : -1;
}

/* jacStamp (2x2) or jacStamp_IC (3x3, full), N_DEV_Capacitor.C:440-470 */
static int64_t c_jac_pairs(struct or_group_s* g, int64_t i, int32_t* rr, int32_t* cc) {
  cinst* c = &((cinst*)g->inst)[i];
  int32_t lid[3] = {c->li_Pos, c->li_Neg, c->li_Bra};
  int nn = c->ICGiven ? 3 : 2, a, b;
  int64_t m = 0;
  for (a = 0; a < nn; ++a)
    for (b = 0; b < nn; ++b) {
      if (lid[a] < 0 || lid[b] < 0) continue;
      if (rr) { rr[m] = lid[a]; cc[m] = lid[b]; }
      ++m;
    }
  return m;
}

static int c_register_jac(struct or_group_s* g, int64_t i) {
  cinst* c = &((cinst*)g->inst)[i];
  c->qpp = or_slot(g, c->li_Pos, c->li_Pos);
  c->qpn = or_slot(g, c->li_Pos, c->li_Neg);
  c->qnp = or_slot(g, c->li_Neg, c->li_Pos);
  c->qnn = or_slot(g, c->li_Neg, c->li_Neg);
  c->fpb = or_slot(g, c->li_Pos, c->li_Bra);
  c->fnb = or_slot(g, c->li_Neg, c->li_Bra);
  c->fbp = or_slot(g, c->li_Bra, c->li_Pos);
  c->fbn = or_slot(g, c->li_Bra, c->li_Neg);
  c->fbb = or_slot(g, c->li_Bra, c->li_Bra);
  if (c->qpp < -1 || c->qpn < -1 || c->qnp < -1 || c->qnn < -1 || c->fpb < -1 ||
      c->fnb < -1 || c->fbp < -1 || c->fbn < -1 || c->fbb < -1)
    return XA_ERR_ARG;
  return XA_OK;
}

/* Master::updateState, N_DEV_Capacitor.C:1821-1850 */
static int c_update_state(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  cinst* in = (cinst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    cinst* c = &in[i];
    double v_pos = sol(v->nextSol, c->li_Pos);
    double v_neg = sol(v->nextSol, c->li_Neg);
    c->vcap = v_pos - v_neg;
    if (ss->dcopFlag && c->ICGiven) c->vcap = c->IC;
    c->q0 = c->C * c->vcap;
    if (c->li_QState >= 0 && v->nextSta) v->nextSta[c->li_QState] = c->q0;
  }
  return XA_OK;
}

/* Master::loadDAEVectors, N_DEV_Capacitor.C:1860-1960 */
static int c_load_vectors(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  cinst* in = (cinst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    cinst* c = &in[i];
    if (c->ICGiven) {
      double Vpos = 0.0, Vneg = 0.0, v_tmp = 0.0;
      if (ss->dcopFlag) {
        Vpos = sol(v->nextSol, c->li_Pos);
        Vneg = sol(v->nextSol, c->li_Neg);
        if (v->f) {
          OR_ADD(v->f, c->li_Pos, sol(v->nextSol, c->li_Bra));
          OR_ADD(v->f, c->li_Neg, -sol(v->nextSol, c->li_Bra));
        }
        if (c->loadLeadCurrent && c->li_store_dev_i >= 0 && v->storeLeadF)
          OR_SET(v->storeLeadF, c->li_store_dev_i, sol(v->nextSol, c->li_Bra));
        v_tmp = (Vpos - Vneg - c->IC);
      }
      if (v->f) OR_ADD(v->f, c->li_Bra, v_tmp);
    }
    if (v->q) {
      OR_ADD(v->q, c->li_Pos, c->q0);
      OR_ADD(v->q, c->li_Neg, -c->q0);
    }
    if (c->loadLeadCurrent && c->li_store_dev_i >= 0) {
      if (v->storeLeadQ) OR_SET(v->storeLeadQ, c->li_store_dev_i, c->q0);
      if (v->storeLeadF) OR_SET(v->storeLeadF, c->li_store_dev_i, 0);
    }
  }
  return XA_OK;
}

/* Master::loadDAEMatrices, N_DEV_Capacitor.C:1970-2135 */
static int c_load_matrices(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  cinst* in = (cinst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    cinst* c = &in[i];
    if (c->ICGiven && ss->dcopFlag) {
      if (v->dFdx) {
        OR_ADD(v->dFdx, c->fpb, 1.0);
        OR_ADD(v->dFdx, c->fnb, -1.0);
        OR_ADD(v->dFdx, c->fbp, 1.0);
        OR_ADD(v->dFdx, c->fbn, -1.0);
      }
    } else if (c->ICGiven) {
      if (v->dFdx) OR_ADD(v->dFdx, c->fbb, 1.0);
    }
    if (!(c->ICGiven && ss->dcopFlag) && v->dQdx) {
      OR_ADD(v->dQdx, c->qpp, c->C);
      OR_SUB(v->dQdx, c->qpn, c->C);
      OR_SUB(v->dQdx, c->qnp, c->C);
      OR_ADD(v->dQdx, c->qnn, c->C);
    }
  }
  return XA_OK;
}

static int c_read(struct or_group_s* g, const char* name, double* out) {
  int64_t i;
  cinst* in = (cinst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    if (!strcmp(name, "C")) out[i] = in[i].C;
    else if (!strcmp(name, "q0")) out[i] = in[i].q0;
    else if (!strcmp(name, "vcap")) out[i] = in[i].vcap;
    else return or_fail(XA_ERR_ARG, "capacitor: unknown intermediate");
  }
  return XA_OK;
}

const or_model_ops or_capacitor_ops = {
  "capacitor", c_mpar, sizeof(c_mpar) / sizeof(c_mpar[0]), C_NM, c_ipar,
  sizeof(c_ipar) / sizeof(c_ipar[0]), C_NI, c_setup, c_register_lids, c_jac_pairs,
  c_register_jac, c_update_state, NULL, c_load_vectors, c_load_matrices, c_read, NULL,
  sizeof(cinst),
};
