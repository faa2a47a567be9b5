/* or_vsrc.c -- oracle restatement of the Xyce independent voltage source
 * (test infrastructure only).
 *
 * Reference: src/DeviceModelPKG/OpenModels/src/N_DEV_Vsrc.C
 *   parameter table :74-210; Instance ctor :300-380 (DC source always, the
 *   transient waveform when TRANSIENTSOURCETYPE is given or HB is set up);
 *   registerLIDs :470-500 (Pos, Neg, one branch unknown); jacStamp and
 *   registerJacLIDs :560-600; Master::updateState :941-990;
 *   loadDAEVectors :1000-1020 (SEPARATE_F_AND_B not defined: the source
 *   value enters F); loadDAEMatrices :1030-1047.
 * and src/DeviceModelPKG/Core/src/N_DEV_SourceData.C (waveforms):
 *   SinData :380-470, ExpData :650-730, PulseData :900-1010, ConstData
 *   :2060-2075; the Src_index enumeration N_DEV_SourceData.h:61-72.
 *
 * Waveform parameters left "not given" take defaults from the solver state
 * (the first step and the final time); a NaN in an instance column marks
 * that instance's value as not given.
 */
#include "or_internal.h"
#include "or_eval_libm.h"

static double sol(const double* x, int lid) { return lid >= 0 ? x[lid] : 0.0; }

enum { V_NM = 1 };
static const or_param v_mpar[] = {{"_unused", 0, 0.0, 0}};
enum {
  VI_DCV0, VI_V0, VI_V1, VI_V2, VI_TD, VI_TR, VI_TF, VI_PW, VI_PER, VI_VA, VI_FREQ, VI_THETA,
  VI_PHASE, VI_TD1, VI_TAU1, VI_TD2, VI_TAU2, VI_TRAN, V_NI
};
static const or_param v_ipar[] = {
  {"DCV0", VI_DCV0, 0.0, 0}, {"V0", VI_V0, 0.0, 0}, {"V1", VI_V1, 0.0, 0},
  {"V2", VI_V2, 0.0, 0}, {"TD", VI_TD, 0.0, 0}, {"TR", VI_TR, 0.0, 0}, {"TF", VI_TF, 0.0, 0},
  {"PW", VI_PW, 0.0, 0}, {"PER", VI_PER, 0.0, 0}, {"VA", VI_VA, 0.0, 0},
  {"FREQ", VI_FREQ, 0.0, 0}, {"THETA", VI_THETA, 0.0, 0}, {"PHASE", VI_PHASE, 0.0, 0},
  {"TD1", VI_TD1, 0.0, 0}, {"TAU1", VI_TAU1, 0.0, 0}, {"TD2", VI_TD2, 0.0, 0},
  {"TAU2", VI_TAU2, 0.0, 0}, {"TRANSIENTSOURCETYPE", VI_TRAN, -1.0, 0},
};

/* Src_index (N_DEV_SourceData.h:61-72) */
enum { SRC_SIN = 0, SRC_EXP = 1, SRC_PULSE = 2, SRC_DC = 6 };

typedef struct {
  double p[V_NI];
  char given[V_NI];
  int tran;                       /* -1: no transient waveform */
  double source, v_pos, v_neg, i_bra, srcCurrent, srcDrop, srcBC, srcVoltage;
  int32_t li_Pos, li_Neg, li_Bra;
  int64_t s_pb, s_nb, s_bp, s_bn;
} vinst;

static int v_setup(struct or_group_s* g) {
  int64_t i;
  int k;
  vinst* in = (vinst*)calloc(g->n > 0 ? g->n : 1, sizeof(vinst));
  for (k = 0; k < g->nmodels; ++k)
    if (g->models[k].level != 1) return or_fail(XA_ERR_UNSUPPORTED, "vsource level");
  for (i = 0; i < g->n; ++i) {
    vinst* s = &in[i];
    for (k = 0; k < V_NI; ++k) {
      const double x = g->ival[k][i];
      s->given[k] = g->igiven[k] && x == x;
      s->p[k] = s->given[k] ? x : v_ipar[k].def;
    }
    s->tran = s->given[VI_TRAN] ? (int)s->p[VI_TRAN] : -1;
    if (s->tran != -1 && s->tran != SRC_SIN && s->tran != SRC_EXP && s->tran != SRC_PULSE &&
        s->tran != SRC_DC)
      return or_fail(XA_ERR_UNSUPPORTED, "vsource: waveform other than SIN/EXP/PULSE/DC");
    g->n_ext[i] = 2;
    g->n_int[i] = 1;
    g->n_sta[i] = 0;
    g->n_sto[i] = 0;
  }
  g->inst = in;
  return XA_OK;
}

static void v_register_lids(struct or_group_s* g, int64_t i, const int32_t* nodes,
                            const int32_t* sta, const int32_t* sto) {
  vinst* s = &((vinst*)g->inst)[i];
  (void)sta;
  (void)sto;
  s->li_Pos = nodes[0];
  s->li_Neg = nodes[1];
  s->li_Bra = nodes[2];
}

/* jacStamp: Pos row [Bra], Neg row [Bra], Bra row [Pos, Neg] */
static int64_t v_jac_pairs(struct or_group_s* g, int64_t i, int32_t* rr, int32_t* cc) {
  vinst* s = &((vinst*)g->inst)[i];
  int32_t pr[4] = {s->li_Pos, s->li_Neg, s->li_Bra, s->li_Bra};
  int32_t pc[4] = {s->li_Bra, s->li_Bra, s->li_Pos, s->li_Neg};
  int64_t k, m = 0;
  for (k = 0; k < 4; ++k) {
    if (pr[k] < 0 || pc[k] < 0) continue;
    if (rr) { rr[m] = pr[k]; cc[m] = pc[k]; }
    ++m;
  }
  return m;
}

static int v_register_jac(struct or_group_s* g, int64_t i) {
  vinst* s = &((vinst*)g->inst)[i];
  s->s_pb = or_slot(g, s->li_Pos, s->li_Bra);
  s->s_nb = or_slot(g, s->li_Neg, s->li_Bra);
  s->s_bp = or_slot(g, s->li_Bra, s->li_Pos);
  s->s_bn = or_slot(g, s->li_Bra, s->li_Neg);
  if (s->s_pb < -1 || s->s_nb < -1 || s->s_bp < -1 || s->s_bn < -1) return XA_ERR_ARG;
  return XA_OK;
}

/* SourceData::updateSource of the waveform at time t */
static double waveform(const vinst* s, const xa_solver_state* ss) {
  const double* p = s->p;
  const char* gv = s->given;
  const double tstep = ss->startingTimeStep, tstop = ss->finalTime;
  double time = ss->currTime;
  switch (s->tran) {
    case SRC_SIN: {             /* SinData::initializeSource / updateSource */
      const double FREQ = gv[VI_FREQ] ? p[VI_FREQ] : 1.0 / tstop;
      const double TD = gv[VI_TD] ? p[VI_TD] : 0.0;
      const double THETA = gv[VI_THETA] ? p[VI_THETA] : 0.0;
      const double mpi = M_PI;
      time -= TD;
      if (time <= 0) return p[VI_V0] + p[VI_VA] * sin(2.0 * mpi * (p[VI_PHASE] / 360));
      return p[VI_V0] + p[VI_VA] * sin(2.0 * mpi * (FREQ * time + p[VI_PHASE] / 360)) *
                            exp(-(time * THETA));
    }
    case SRC_EXP: {             /* ExpData */
      const double TD1 = gv[VI_TD1] ? p[VI_TD1] : 0.0;
      const double TAU1 = gv[VI_TAU1] ? p[VI_TAU1] : tstep;
      const double TD2 = gv[VI_TD2] ? p[VI_TD2] : TD1 + tstep;
      const double TAU2 = gv[VI_TAU2] ? p[VI_TAU2] : tstep;
      const double V1 = p[VI_V1], V2 = p[VI_V2];
      if (time <= TD1) return V1;
      if (time <= TD2) return V1 + (V2 - V1) * (1 - exp(-(time - TD1) / TAU1));
      return V1 + (V2 - V1) * (1 - exp(-(time - TD1) / TAU1)) +
             (V1 - V2) * (1 - exp(-(time - TD2) / TAU2));
    }
    case SRC_PULSE: {           /* PulseData */
      const double TD = gv[VI_TD] ? p[VI_TD] : 0.0;
      const double TR = gv[VI_TR] ? p[VI_TR] : tstep;
      const double TF = gv[VI_TF] ? p[VI_TF] : tstep;
      const double PW = gv[VI_PW] ? p[VI_PW] : tstop;
      const double PER = gv[VI_PER] ? p[VI_PER] : tstop;
      const double V1 = p[VI_V1], V2 = p[VI_V2], bpTol = ss->bpTol;
      double basetime;
      time -= TD;
      if (time > PER && PER != 0.0) {
        basetime = PER * floor(time / PER);
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
    default:                    /* ConstData */
      return p[VI_DCV0];
  }
}

/* Master::updateState, N_DEV_Vsrc.C:941-990 (scale_src = 0) */
static int v_update_state(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  vinst* in = (vinst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    vinst* s = &in[i];
    const int tran = (ss->HBspecified || ss->tranopFlag || ss->transientFlag) && s->tran != -1;
    s->source = tran ? waveform(s, ss) : s->p[VI_DCV0];
    s->v_pos = sol(v->nextSol, s->li_Pos);
    s->v_neg = sol(v->nextSol, s->li_Neg);
    s->i_bra = sol(v->nextSol, s->li_Bra);
    s->srcCurrent = s->i_bra;
    s->srcDrop = (s->v_pos - s->v_neg);
    s->srcBC = s->source;
    s->srcVoltage = s->srcDrop - s->srcBC;
  }
  return XA_OK;
}

/* Master::loadDAEVectors, N_DEV_Vsrc.C:1000-1020 */
static int v_load_vectors(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  vinst* in = (vinst*)g->inst;
  (void)ss;
  if (!v->f) return XA_OK;
  for (i = 0; i < g->n; ++i) {
    vinst* s = &in[i];
    OR_ADD(v->f, s->li_Pos, s->srcCurrent);
    OR_ADD(v->f, s->li_Neg, -s->srcCurrent);
    OR_ADD(v->f, s->li_Bra, s->srcDrop);
    OR_SUB(v->f, s->li_Bra, s->srcBC);
  }
  return XA_OK;
}

/* Master::loadDAEMatrices, N_DEV_Vsrc.C:1030-1047 */
static int v_load_matrices(struct or_group_s* g, const xa_solver_state* ss, const xa_vectors* v) {
  int64_t i;
  vinst* in = (vinst*)g->inst;
  (void)ss;
  if (!v->dFdx) return XA_OK;
  for (i = 0; i < g->n; ++i) {
    vinst* s = &in[i];
    OR_ADD(v->dFdx, s->s_pb, 1.0);
    OR_SUB(v->dFdx, s->s_nb, 1.0);
    OR_ADD(v->dFdx, s->s_bp, 1.0);
    OR_SUB(v->dFdx, s->s_bn, 1.0);
  }
  return XA_OK;
}

static int v_read(struct or_group_s* g, const char* name, double* out) {
  int64_t i;
  vinst* in = (vinst*)g->inst;
  for (i = 0; i < g->n; ++i) {
    if (!strcmp(name, "source")) out[i] = in[i].source;
    else if (!strcmp(name, "srcVoltage")) out[i] = in[i].srcVoltage;
    else return or_fail(XA_ERR_ARG, "vsource: unknown intermediate");
  }
  return XA_OK;
}

const or_model_ops or_vsrc_ops = {
  "vsource", v_mpar, 0, V_NM, v_ipar, sizeof(v_ipar) / sizeof(v_ipar[0]), V_NI, v_setup,
  v_register_lids, v_jac_pairs, v_register_jac, v_update_state, NULL, v_load_vectors,
  v_load_matrices, v_read, NULL, sizeof(vinst),
};
